// Octree traversal state machine shared by the raymarch kernels.
//
// Restates VoxelOctree::raymarch (src/VoxelOctree.cpp:207-346) in FP32 with
// the reference's operation order; ray_init is the set-up before the loop
// (:216-250), ray_step one pass of the loop body (:252-338). Splitting the
// loop into resumable steps lets a persistent wave hand a finished lane a new
// ray while the other lanes keep walking.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svo {
namespace march {

constexpr int kMaxScale = 23;  // VoxelOctree::MaxScale, src/VoxelOctree.hpp:38

__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// invSqrt, src/Util.hpp:47-58
__device__ __forceinline__ float quake_rsqrt(float x) {
    float half = x * 0.5f;
    float y = u2f(0x5f3759dfu - (f2u(x) >> 1));
    return y * (1.5f - half * y * y);
}

// Per-lane restart stack in LDS: [slot][stride] offsets then [slot][stride] maxT.
template <typename Idx>
struct Stack {
    Idx *off;
    float *tmax;
    int stride;
    int slots;
    int lane;
    __device__ __forceinline__ Stack(void *base, int stride_, int slots_, int lane_)
        : off(reinterpret_cast<Idx *>(base)),
          tmax(reinterpret_cast<float *>(reinterpret_cast<Idx *>(base) + size_t(slots_) * stride_)),
          stride(stride_), slots(slots_), lane(lane_) {}
};

template <typename Idx>
struct Ray {
    float dTx, dTy, dTz, bTx, bTy, bTz;
    float minT, maxT, posX, posY, posZ, scaleExp2, rayScale;
    uint32_t current;
    Idx parent;
    int octantMask, idx, scale;
};

// Outcome of a finished ray.
enum : int { kRunning = -1, kMiss = 0, kLeaf = 1, kLod = 2, kBad = 3 };

template <typename Idx>
__device__ __forceinline__ void ray_init(Ray<Idx> &r, float ox, float oy, float oz, float dx, float dy, float dz,
                                         float rayScale) {
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
    r.current = 0;
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

// One pass of the loop body. Returns kRunning, or the outcome with tOut /
// normal / leaf set as the reference sets them.
template <typename Idx>
__device__ __forceinline__ int ray_step(Ray<Idx> &r, const uint32_t *__restrict__ tree, const Stack<Idx> &stk,
                                        uint32_t &normal, float &tOut, Idx &leaf) {
    if (r.current == 0) r.current = tree[r.parent];

    const float cornerTX = r.posX * r.dTx - r.bTx;
    const float cornerTY = r.posY * r.dTy - r.bTy;
    const float cornerTZ = r.posZ * r.dTz - r.bTz;
    const float maxTC = smin(cornerTX, smin(cornerTY, cornerTZ));

    const int childShift = r.idx ^ r.octantMask;
    const uint32_t childMasks = r.current << childShift;

    if ((childMasks & 0x8000u) && r.minT <= r.maxT) {
        if (maxTC * r.rayScale >= r.scaleExp2) {
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
                leaf = childOffset + r.parent +
                       Idx(__popc((((childMasks >> (8 + childShift)) << childShift) & 127u)));
                normal = tree[leaf];
                tOut = r.minT;
                return kLeaf;
            }
            const int slot = (kMaxScale - 1) - r.scale;
            if (slot >= stk.slots) return kBad;
            stk.off[slot * stk.stride + stk.lane] = r.parent;
            stk.tmax[slot * stk.stride + stk.lane] = r.maxT;

            const Idx siblingCount = Idx(__popc(childMasks & 127u));
            r.parent += childOffset + siblingCount;
            if (r.current & 0x10000u) r.parent += siblingCount;

            r.idx = 0;
            r.scale--;
            r.scaleExp2 = half;
            if (centerTX > r.minT) { r.idx ^= 1; r.posX += r.scaleExp2; }
            if (centerTY > r.minT) { r.idx ^= 2; r.posY += r.scaleExp2; }
            if (centerTZ > r.minT) { r.idx ^= 4; r.posZ += r.scaleExp2; }
            r.maxT = maxTV;
            r.current = 0;
            return kRunning;
        }
    }

    int stepMask = 0;
    if (cornerTX <= maxTC) { stepMask ^= 1; r.posX -= r.scaleExp2; }
    if (cornerTY <= maxTC) { stepMask ^= 2; r.posY -= r.scaleExp2; }
    if (cornerTZ <= maxTC) { stepMask ^= 4; r.posZ -= r.scaleExp2; }

    r.minT = maxTC;
    r.idx ^= stepMask;

    if ((r.idx & stepMask) != 0) {
        uint32_t diff = 0;
        if (stepMask & 1) diff |= f2u(r.posX) ^ f2u(r.posX + r.scaleExp2);
        if (stepMask & 2) diff |= f2u(r.posY) ^ f2u(r.posY + r.scaleExp2);
        if (stepMask & 4) diff |= f2u(r.posZ) ^ f2u(r.posZ + r.scaleExp2);
        r.scale = int(f2u(float(int(diff))) >> 23) - 127;
        r.scaleExp2 = u2f(uint32_t(r.scale - kMaxScale + 127) << 23);
        // the reference reads rayStack[scale] even when scale >= 23 and then
        // leaves the loop; that read cannot reach the result
        if (r.scale >= kMaxScale) return kMiss;
        const int slot = (kMaxScale - 1) - r.scale;
        if (slot >= stk.slots) return kBad;
        r.parent = stk.off[slot * stk.stride + stk.lane];
        r.maxT = stk.tmax[slot * stk.stride + stk.lane];

        const uint32_t shX = f2u(r.posX) >> r.scale;
        const uint32_t shY = f2u(r.posY) >> r.scale;
        const uint32_t shZ = f2u(r.posZ) >> r.scale;
        r.posX = u2f(shX << r.scale);
        r.posY = u2f(shY << r.scale);
        r.posZ = u2f(shZ << r.scale);
        r.idx = int((shX & 1u) | ((shY & 1u) << 1) | ((shZ & 1u) << 2));
        r.current = 0;
    }
    return kRunning;
}

}  // namespace march
}  // namespace svo
