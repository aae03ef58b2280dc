// Structural validation of a Laine-Karras word array before it is used on
// the device: every word the raymarch kernel can read must lie inside the
// array, so a corrupt or hostile .oct cannot make a kernel read out of bounds.
//
// Layout (written by src/VoxelOctree.cpp:148-204, read by :252-338): a
// descriptor holds the child offset in bits 31..18 (relative to its own index),
// bit 17 = far pointer (offset = bits31..18 << 32 | next word), bit 16 = two
// words per child slot, bits 15..8 valid mask, bits 7..0 non-leaf mask.
// The check addresses children exactly as the traversal does
// (src/VoxelOctree.cpp:282,290-293): a valid non-leaf child m at
// index + offset + popc(non-leaf bits below m) * (1+large), a valid leaf child
// m at index + offset + popc(valid bits below m). The reference builder writes
// non-leaf = valid or 0, where both reduce to the k-th valid child at
// index + offset + k*(1+large); other masks (non-leaf bits without a valid bit
// included) are checked at the addresses the traversal will read.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <memory>
#include <thread>
#include <vector>

#include "svo_internal.hpp"

namespace svo {

namespace {

struct Item {
    uint64_t idx, off;
    int level;
};

// Checks the subtree below `it` (descriptor at it.idx, child block at
// it.idx + it.off). Returns the depth it reaches (0 when it has no children
// of its own) or -1 if a pointer leaves the array, the tree is deeper than
// the traversal can follow, or a non-leaf slot is the child of two parents
// (bit set in `seen`, one bit per word, shared by all threads). Child
// pointers only point forward, so with that rule the check visits each word
// at most once. With `frontier`, nodes at level `split` are handed out
// instead of descended into.
int validate_from(const uint32_t *w, uint64_t n, Item root, int split, std::vector<Item> *frontier, uint64_t *seen) {
    std::vector<Item> stack;
    stack.push_back(root);
    int depth = 0;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.level >= 23) return -1;  // MaxScale bounds the traversal depth
        if (frontier && it.level == split) {
            frontier->push_back(it);
            continue;
        }
        uint32_t d = w[it.idx];
        uint32_t valid = (d >> 8) & 0xFFu, nonleaf = d & 0xFFu;
        uint64_t per = 1 + ((d >> 16) & 1u);
        uint64_t base = it.idx + it.off;
        if (base < it.idx) return -1;
        if (it.level + 2 > depth && valid) depth = it.level + 2;
        for (int m = 0; m < 8; ++m) {
            if (!((valid >> m) & 1u)) continue;
            const uint32_t below = (1u << m) - 1u;
            const bool inner = (nonleaf >> m) & 1u;
            // the traversal's own slot: non-leaf children count non-leaf bits
            // (times the slot width), leaf children count valid bits
            const uint64_t slot =
                inner ? base + uint64_t(__builtin_popcount(nonleaf & below)) * per
                      : base + uint64_t(__builtin_popcount(valid & below));
            if (slot < base || slot >= n) return -1;
            if (inner) {
                uint32_t cw = w[slot];
                uint64_t coff = cw >> 18;
                if (cw & 0x20000u) {
                    if (slot + 1 >= n) return -1;
                    coff = (coff << 32) | w[slot + 1];
                }
                if (coff == 0) return -1;  // a node cannot be its own child block
                // each non-leaf slot has one parent: a slot reached twice would
                // make the check (and nothing else) revisit shared subtrees
                const uint64_t bit = uint64_t(1) << (slot & 63);
                if (__atomic_fetch_or(&seen[slot >> 6], bit, __ATOMIC_RELAXED) & bit) return -1;
                stack.push_back({slot, coff, it.level + 1});
            }
        }
    }
    return depth;
}

}  // namespace

// Every child pointer reachable from the root stays inside the array, no
// non-leaf slot is shared, and the depth is bounded; returns the octree depth
// (levels of descriptors + 1), -1 for a malformed tree or -2 when the check's
// bitmap (one bit per word) cannot be allocated. Subtrees below level 3 are
// checked on all host threads.
int octree_validate(const uint32_t *w, uint64_t n) {
    if (n == 0) return -1;
    // the root's child offset, far form included: the traversal reads it like
    // any other descriptor's (src/VoxelOctree.cpp:276-280)
    uint64_t rootOff = w[0] >> 18;
    if (w[0] & 0x20000u) {
        if (n < 2) return -1;
        rootOff = (rootOff << 32) | w[1];
    }
    if ((w[0] & 0xFF00u) && rootOff == 0) return -1;
    std::unique_ptr<uint64_t, void (*)(void *)> seenMem(
        static_cast<uint64_t *>(std::calloc(size_t((n + 63) / 64), sizeof(uint64_t))), std::free);
    if (!seenMem) return -2;
    uint64_t *seen = seenMem.get();
    seen[0] = 1;  // the root
    std::vector<Item> frontier;
    const int split = 3;
    int depth = validate_from(w, n, {0, rootOff, 0}, split, &frontier, seen);
    if (depth < 0) return -1;
    depth = std::max(depth, 1);
    if (frontier.empty()) return depth;
    std::atomic<size_t> next{0};
    std::atomic<int> worst{depth};
    std::atomic<bool> bad{false};
    auto worker = [&]() {
        for (size_t i; !bad.load() && (i = next.fetch_add(1)) < frontier.size();) {
            int d = validate_from(w, n, frontier[i], split, nullptr, seen);
            if (d < 0) {
                bad.store(true);
                return;
            }
            int cur = worst.load();
            while (d > cur && !worst.compare_exchange_weak(cur, d)) {
            }
        }
    };
    const size_t nt = std::min<size_t>(frontier.size(), std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    try {
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(worker);
    } catch (...) {
    }
    worker();
    for (std::thread &t : pool) t.join();
    return bad.load() ? -1 : worst.load();
}

// ---- host memory for large word arrays ----

void *map_large(size_t bytes) {
    void *p = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (p == MAP_FAILED) return nullptr;
#ifdef MADV_HUGEPAGE
    (void)madvise(p, bytes, MADV_HUGEPAGE);
#endif
#ifdef MADV_POPULATE_WRITE
    // the kernel clears every page it populates: arrays of tens of MiB and
    // more are populated by a few threads, in 2 MiB-aligned parts
    constexpr size_t kPart = size_t(2) << 20;
    const size_t parts = (bytes + kPart - 1) / kPart;
    const size_t nt = std::min<size_t>({size_t(4), std::max(1u, std::thread::hardware_concurrency()), parts / 8 + 1});
    auto populate = [p, bytes, parts, nt](size_t k) {
        const size_t a = parts * k / nt * kPart, b = std::min(bytes, parts * (k + 1) / nt * kPart);
        if (a < b) (void)madvise(static_cast<char *>(p) + a, b - a, MADV_POPULATE_WRITE);
    };
    std::vector<std::thread> pool;
    try {
        for (size_t k = 1; k < nt; ++k) pool.emplace_back(populate, k);
    } catch (...) {
        pool.clear();  // no threads: populate the rest on this one
        for (size_t k = 1; k < nt; ++k) populate(k);
    }
    populate(0);
    for (std::thread &t : pool) t.join();
#endif
    return p;
}

void unmap_large(void *p, size_t bytes) {
    if (p) (void)munmap(p, bytes);
}

// ---- read-only file mappings ----

bool FileView::open(const char *path) {
    close();
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) return false;
    struct stat sb;
    bool ok = fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode);
    if (ok && sb.st_size > 0) {
        void *p = mmap(nullptr, size_t(sb.st_size), PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
        if (p == MAP_FAILED) ok = false;
        else {
            data_ = static_cast<const uint8_t *>(p);
            size_ = size_t(sb.st_size);
        }
    }
    ::close(fd);
    return ok;
}

void FileView::close() {
    if (data_) (void)munmap(const_cast<uint8_t *>(data_), size_);
    data_ = nullptr;
    size_ = 0;
}

}  // namespace svo
