// Randomised check of the first-fit range allocator (csrc/range_alloc.hpp)
// against a brute-force model: a byte map of the block. Each allocation must
// land on the lowest offset where the request fits in the model, never overlap
// a live range, and stay inside the block; freeing everything must leave one
// range that a whole-block request fits. Prints "ok <operations>" or exits 1.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "range_alloc.hpp"

static bool fits_at(const std::vector<char> &used, uint64_t off, uint64_t n) {
    if (off + n > used.size()) return false;
    for (uint64_t i = off; i < off + n; ++i)
        if (used[i]) return false;
    return true;
}

int main(int argc, char **argv) {
    const uint64_t size = 4096;
    const int ops = argc > 1 ? std::atoi(argv[1]) : 20000;
    std::mt19937_64 rng(12345);
    svo::RangeAlloc ra;
    ra.reset(size);
    std::vector<char> used(size, 0);
    std::map<uint64_t, uint64_t> live;  // offset -> length
    for (int k = 0; k < ops; ++k) {
        const bool doAlloc = live.empty() || (rng() % 100) < 55;
        if (doAlloc) {
            const uint64_t n = 1 + rng() % 300;
            uint64_t first = size;  // the model's lowest fitting offset
            for (uint64_t o = 0; o + n <= size; ++o)
                if (fits_at(used, o, n)) {
                    first = o;
                    break;
                }
            uint64_t off = 0;
            const bool got = ra.alloc(n, off);
            if (got != (first < size)) {
                std::printf("op %d: alloc(%llu) returned %d, model fits: %d\n", k, (unsigned long long)n, int(got),
                            int(first < size));
                return 1;
            }
            if (!got) continue;
            if (off != first) {
                std::printf("op %d: alloc(%llu) at %llu, first fit is %llu\n", k, (unsigned long long)n,
                            (unsigned long long)off, (unsigned long long)first);
                return 1;
            }
            for (uint64_t i = off; i < off + n; ++i) used[i] = 1;
            live[off] = n;
        } else {
            auto it = live.begin();
            std::advance(it, long(rng() % live.size()));
            ra.release(it->first, it->second);
            for (uint64_t i = it->first; i < it->first + it->second; ++i) used[i] = 0;
            live.erase(it);
        }
    }
    for (auto &r : live) ra.release(r.first, r.second);
    uint64_t off = 1;
    if (!ra.alloc(size, off) || off != 0) {
        std::printf("freed block does not merge back into one range\n");
        return 1;
    }
    std::printf("ok %d\n", ops);
    return 0;
}
