// Octree traversal state machine shared by the raymarch kernels.
//
// Restates VoxelOctree::raymarch (src/VoxelOctree.cpp:207-346) in FP32 with
// the reference's operation order; ray_init is the set-up before the loop
// (:216-250), ray_step one pass of the loop body (:252-338). Splitting the
// loop into resumable steps lets a persistent wave hand a finished lane a new
// ray while the other lanes keep walking.
//
// The same source also compiles for the host (single-ray calls, where a
// device round trip costs far more than the walk itself): every operation is
// a correctly rounded FP32 add, multiply or divide, a comparison or integer
// bit work, so host and device give the same bits.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svo {
namespace march {

#define SVO_HD __host__ __device__ __forceinline__

constexpr int kMaxScale = 23;  // VoxelOctree::MaxScale, src/VoxelOctree.hpp:38

SVO_HD float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
SVO_HD float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
SVO_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
SVO_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
SVO_HD int popc32(uint32_t x) { return __builtin_popcount(x); }
// bit 0 of x (a v_bfe_u32 on the device)
SVO_HD uint32_t bit0(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ubfe(x, 0, 1);
#else
    return x & 1u;
#endif
}

// invSqrt, src/Util.hpp:47-58
SVO_HD float quake_rsqrt(float x) {
    float half = x * 0.5f;
    float y = u2f(0x5f3759dfu - (f2u(x) >> 1));
    return y * (1.5f - half * y * y);
}

// The restart stack has one entry per scale a ray can push from: depth - 1
// for a tree of `depth` (descriptor levels + 1). The top `reg` entries live in
// registers instead of LDS: entry 0, the root's, is always (word 0, the ray's
// initial maxT: nothing changes maxT while the ray is at the root), entries 1
// and 2 are register pairs. 160 KiB of LDS holds 32 waves' stacks of 10
// entries (8 bytes per lane), so trees up to depth 12 keep one entry in
// registers and deeper ones three (the 8192^3 terrain, depth 14: 26 -> 32
// waves per CU).
__host__ __device__ inline int reg_entries(int depth) { return depth > 12 ? 3 : 1; }
__host__ __device__ inline int lds_slots(int depth, int reg) { return depth - 1 - reg > 0 ? depth - 1 - reg : 1; }

// Per-lane restart stack in LDS, one entry per scale below the root's children.
// 64-bit word indices: [slot][stride] indices then [slot][stride] maxT.
template <typename Idx>
struct Stack {
    Idx *off;
    float *tmax;
    int stride;
    int slots;
    int lane;
    SVO_HD Stack(void *base, int stride_, int slots_, int lane_)
        : off(reinterpret_cast<Idx *>(base)),
          tmax(reinterpret_cast<float *>(reinterpret_cast<Idx *>(base) + size_t(slots_) * stride_)),
          stride(stride_), slots(slots_), lane(lane_) {}
    SVO_HD void push(int slot, Idx parent, float t) const {
        off[slot * stride + lane] = parent;
        tmax[slot * stride + lane] = t;
    }
    SVO_HD void pop(int slot, Idx &parent, float &t) const {
        parent = off[slot * stride + lane];
        t = tmax[slot * stride + lane];
    }
};

// 32-bit word indices: one 8-byte {index, maxT} entry per [slot][stride], so a
// push or pop is a single 64-bit LDS access (same 8 bytes per entry).
template <>
struct Stack<uint32_t> {
    uint2 *e;
    int stride;
    int slots;
    SVO_HD Stack(void *base, int stride_, int slots_, int lane_)
        : e(reinterpret_cast<uint2 *>(base) + lane_), stride(stride_), slots(slots_) {}
    SVO_HD void push(int slot, uint32_t parent, float t) const { e[slot * stride] = make_uint2(parent, f2u(t)); }
    SVO_HD void pop(int slot, uint32_t &parent, float &t) const {
        const uint2 v = e[slot * stride];
        parent = v.x;
        t = u2f(v.y);
    }
};

template <typename Idx>
struct Ray {
    float dTx, dTy, dTz, bTx, bTy, bTz;
    float minT, maxT, maxT0, posX, posY, posZ, scaleExp2, rayScale;  // maxT0: the root's maxT
    Idx parent1, parent2;  // stack entries 1 and 2 when they are held in registers
    float maxT1, maxT2;
    uint32_t current;
    Idx parent;
    int octantMask, idx, scale;
};

// Outcome of a finished ray.
enum : int { kRunning = -1, kMiss = 0, kLeaf = 1, kLod = 2 };

// Set-up before the loop (:216-250). The root descriptor is read here, the
// read the first loop pass makes (:254).
template <typename Idx>
SVO_HD void ray_init(Ray<Idx> &r, const uint32_t *__restrict__ tree, float ox, float oy, float oz, float dx, float dy,
                     float dz, float rayScale) {
    if (fabsf(dx) < 1e-4f) dx = 1e-4f;
    if (fabsf(dy) < 1e-4f) dy = 1e-4f;
    if (fabsf(dz) < 1e-4f) dz = 1e-4f;
    r.dTx = 1.0f / -fabsf(dx);
    r.dTy = 1.0f / -fabsf(dy);
    r.dTz = 1.0f / -fabsf(dz);
    r.bTx = r.dTx * ox;
    r.bTy = r.dTy * oy;
    r.bTz = r.dTz * oz;
    r.octantMask = 7;
    if (dx > 0.0f) { r.octantMask ^= 1; r.bTx = 3.0f * r.dTx - r.bTx; }
    if (dy > 0.0f) { r.octantMask ^= 2; r.bTy = 3.0f * r.dTy - r.bTy; }
    if (dz > 0.0f) { r.octantMask ^= 4; r.bTz = 3.0f * r.dTz - r.bTz; }
    float minT = smax(2.0f * r.dTx - r.bTx, smax(2.0f * r.dTy - r.bTy, 2.0f * r.dTz - r.bTz));
    r.maxT = smin(r.dTx - r.bTx, smin(r.dTy - r.bTy, r.dTz - r.bTz));
    minT = smax(minT, 0.0f);
    r.minT = minT;
    r.maxT0 = r.maxT;
    r.current = tree[0];
    r.parent = 0;
    r.idx = 0;
    r.posX = 1.0f;
    r.posY = 1.0f;
    r.posZ = 1.0f;
    r.scale = kMaxScale - 1;
    r.scaleExp2 = 0.5f;
    r.rayScale = rayScale;
    if (1.5f * r.dTx - r.bTx > minT) { r.idx ^= 1; r.posX = 1.5f; }
    if (1.5f * r.dTy - r.bTy > minT) { r.idx ^= 2; r.posY = 1.5f; }
    if (1.5f * r.dTz - r.bTz > minT) { r.idx ^= 4; r.posZ = 1.5f; }
}

// False when a ray coefficient is NaN (a NaN origin or direction component, or
// an infinite origin on an axis whose direction is infinite: -0 * inf). Every
// corner time is then NaN, no axis ever steps and the reference's loop does not
// end. For all other rays some axis steps on every advance (the minimum corner
// time equals at least one corner time), each node is entered at most once and
// the walk ends.
template <typename Idx>
SVO_HD bool ray_defined(const Ray<Idx> &r) {
    return r.dTx == r.dTx && r.dTy == r.dTy && r.dTz == r.dTz && r.bTx == r.bTx && r.bTy == r.bTy &&
           r.bTz == r.bTz;
}

// One pass of the loop body. Returns kRunning, or the outcome with tOut /
// normal / leaf set as the reference sets them. The descriptor of the node
// the ray is in (`current`) is read as soon as its index is known (init,
// push, pop) rather than at the top of the next pass; the value is the same.
//
//  Lod = false: the ray's rayScale is 0 (+0 or -0). maxTC is finite, so
//    maxTC * rayScale >= scaleExp2 (> 0) never holds and the test is dropped.
//  Normal = false: the caller needs only the hit code and t; a leaf hit
//    leaves `leaf` and `normal` unset and reads no leaf word.
//  Reg: how many top stack entries are held in registers (0, 1 or 3); the
//    LDS stack holds the others, entry e at LDS slot e - Reg. 0 keeps every
//    entry in LDS, for kernels where that measured faster.
//  Finite = true: the ray's origin and direction are finite (|origin| below
//    2^100). Then every corner t is finite and never -0 (pos*dT is a nonzero
//    product; a difference of equal nonzero values is +0), so the minimum of
//    the three corner times does not depend on the order of ties or on NaN
//    handling and one hardware min3 returns the reference's std::min result.
template <bool Lod, bool Finite, typename Idx, bool Normal = true, int Reg = 1>
SVO_HD int ray_step(Ray<Idx> &r, const uint32_t *__restrict__ tree, const Stack<Idx> &stk,
                                        uint32_t &normal, float &tOut, Idx &leaf) {
    const float cornerTX = r.posX * r.dTx - r.bTx;
    const float cornerTY = r.posY * r.dTy - r.bTy;
    const float cornerTZ = r.posZ * r.dTz - r.bTz;
    const float maxTC = Finite ? fminf(cornerTX, fminf(cornerTY, cornerTZ)) : smin(cornerTX, smin(cornerTY, cornerTZ));

    const int childShift = r.idx ^ r.octantMask;
    const uint32_t childMasks = r.current << childShift;

    if ((childMasks & 0x8000u) && r.minT <= r.maxT) {
        if (Lod && maxTC * r.rayScale >= r.scaleExp2) {
            tOut = maxTC;
            return kLod;
        }
        const float maxTV = smin(r.maxT, maxTC);
        const float half = r.scaleExp2 * 0.5f;
        const float centerTX = half * r.dTx + cornerTX;
        const float centerTY = half * r.dTy + cornerTY;
        const float centerTZ = half * r.dTz + cornerTZ;

        if (r.minT <= maxTV) {
            Idx childOffset = Idx(r.current >> 18);
            if (r.current & 0x20000u)
                childOffset = Idx((uint64_t(r.current >> 18) << 32) | uint64_t(tree[r.parent + 1]));
            if (!(childMasks & 0x80u)) {
                if (Normal) {
                    leaf = childOffset + r.parent +
                           Idx(popc32((((childMasks >> (8 + childShift)) << childShift) & 127u)));
                    normal = tree[leaf];
                }
                tOut = r.minT;
                return kLeaf;
            }
            // slot - Reg < stk.slots: a node at level L (slot L) pushes only when
            // it has a non-leaf child, and octree_validate then counted depth >=
            // L + 2 (every loaded or built tree is validated); the root's entry
            // (slot 0) is (0, maxT0) and is not stored
            const int slot = (kMaxScale - 1) - r.scale;
            static_assert(Reg == 0 || Reg == 1 || Reg == 3, "register-held stack entries: 0, 1 or 3");
            if (slot >= Reg) {
                stk.push(slot - Reg, r.parent, r.maxT);
            } else if (Reg == 3 && slot == 1) {
                r.parent1 = r.parent;
                r.maxT1 = r.maxT;
            } else if (Reg == 3 && slot == 2) {
                r.parent2 = r.parent;
                r.maxT2 = r.maxT;
            }

            const Idx siblingCount = Idx(popc32(childMasks & 127u));
            r.parent += childOffset + siblingCount;
            if (r.current & 0x10000u) r.parent += siblingCount;
            r.current = tree[r.parent];

            r.scale--;
            r.scaleExp2 = half;
            const bool cx = centerTX > r.minT, cy = centerTY > r.minT, cz = centerTZ > r.minT;
            if (cx) r.posX += r.scaleExp2;
            if (cy) r.posY += r.scaleExp2;
            if (cz) r.posZ += r.scaleExp2;
            r.idx = (cx ? 1 : 0) | (cy ? 2 : 0) | (cz ? 4 : 0);
            r.maxT = maxTV;
            return kRunning;
        }
    }

    const bool sx = cornerTX <= maxTC, sy = cornerTY <= maxTC, sz = cornerTZ <= maxTC;
    if (sx) r.posX -= r.scaleExp2;
    if (sy) r.posY -= r.scaleExp2;
    if (sz) r.posZ -= r.scaleExp2;
    const int stepMask = (sx ? 1 : 0) | (sy ? 2 : 0) | (sz ? 4 : 0);

    r.minT = maxTC;
    r.idx ^= stepMask;

    if ((r.idx & stepMask) != 0) {
        uint32_t diff = 0;
        if (stepMask & 1) diff |= f2u(r.posX) ^ f2u(r.posX + r.scaleExp2);
        if (stepMask & 2) diff |= f2u(r.posY) ^ f2u(r.posY + r.scaleExp2);
        if (stepMask & 4) diff |= f2u(r.posZ) ^ f2u(r.posZ + r.scaleExp2);
        // The reference takes the highest differing bit from the exponent of
        // float(int(diff)) (src/VoxelOctree.cpp:318-325). diff is non-zero (a
        // stepped coordinate in [1,2) changes by at least its ulp); below 2^23
        // the conversion is exact, so the exponent is 31 - clz(diff); from 2^23
        // up both forms give a scale >= 23 and the ray leaves the octree. The
        // reference then reads rayStack[scale] and leaves the loop; that read
        // cannot reach the result.
        if (diff >= (1u << kMaxScale)) return kMiss;
        r.scale = 31 - __builtin_clz(diff);
        r.scaleExp2 = u2f(uint32_t(r.scale - kMaxScale + 127) << 23);
        const int slot = (kMaxScale - 1) - r.scale;  // a slot pushed earlier: the new scale is above the current one
        if (slot >= Reg) {
            stk.pop(slot - Reg, r.parent, r.maxT);
        } else if (slot == 0) {
            r.parent = 0;
            r.maxT = r.maxT0;
        } else if (Reg == 3 && slot == 1) {
            r.parent = r.parent1;
            r.maxT = r.maxT1;
        } else if (Reg == 3) {
            r.parent = r.parent2;
            r.maxT = r.maxT2;
        }
        r.current = tree[r.parent];

        const uint32_t shX = f2u(r.posX) >> r.scale;
        const uint32_t shY = f2u(r.posY) >> r.scale;
        const uint32_t shZ = f2u(r.posZ) >> r.scale;
        r.posX = u2f(shX << r.scale);
        r.posY = u2f(shY << r.scale);
        r.posZ = u2f(shZ << r.scale);
        r.idx = int(bit0(shX) | (bit0(shY) << 1) | (bit0(shZ) << 2));
    }
    return kRunning;
}

}  // namespace march
}  // namespace svo
