// Structural validation of a Laine-Karras word array before it is used on
// the device: every word the raymarch kernel can read must lie inside the
// array, so a corrupt or hostile .oct cannot make a kernel read out of bounds.
//
// Layout (written by src/VoxelOctree.cpp:199-265, read by :312-399): a
// descriptor holds the child offset in bits 31..18 (relative to its own index),
// bit 17 = far pointer (offset = bits31..18 << 32 | next word), bit 16 = two
// words per child slot, bits 15..8 valid mask, bits 7..0 non-leaf mask. The
// k-th valid child (ascending bit) sits at index + offset + k*(1+large).
#include <vector>

#include "svo_internal.hpp"

namespace svo {

int octree_validate(const uint32_t *w, uint64_t n) {
    if (n == 0) return -1;
    struct Item {
        uint64_t idx, off;
        int level;
    };
    std::vector<Item> stack;
    stack.push_back({0, w[0] >> 18, 0});
    int depth = 1;
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.level >= 23) return -1;  // MaxScale bounds the traversal depth
        uint32_t d = w[it.idx];
        uint32_t valid = (d >> 8) & 0xFFu, nonleaf = d & 0xFFu;
        uint64_t per = 1 + ((d >> 16) & 1u);
        uint64_t base = it.idx + it.off;
        if (base < it.idx) return -1;
        uint64_t k = 0;
        if (it.level + 2 > depth && valid) depth = it.level + 2;
        for (int m = 0; m < 8; ++m) {
            if (!((valid >> m) & 1u)) continue;
            uint64_t slot = base + k * per;
            k++;
            if (slot >= n) return -1;
            if ((nonleaf >> m) & 1u) {
                uint32_t cw = w[slot];
                uint64_t coff = cw >> 18;
                if (cw & 0x20000u) {
                    if (slot + 1 >= n) return -1;
                    coff = (coff << 32) | w[slot + 1];
                }
                if (coff == 0) return -1;  // a node cannot be its own child block
                stack.push_back({slot, coff, it.level + 1});
            }
        }
    }
    return depth;
}

}  // namespace svo
