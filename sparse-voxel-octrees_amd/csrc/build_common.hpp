// Device arithmetic of the octree builder, shared by the builder kernels.
// Each function restates a reference routine with its single-precision
// operation order (compiled with -ffp-contract=off, correctly rounded
// division and square root), so voxel sets and leaf words are bit-identical.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace svo {
namespace dev {

__host__ __device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }  // std::min
__host__ __device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }  // std::max

// float -> int32 truncation as the reference's casts compile on x86-64
// (cvttss2si): NaN and out-of-range inputs give INT32_MIN.
__host__ __device__ __forceinline__ int32_t cvt_i32(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return int32_t(f);
}

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// invSqrt, src/Util.hpp:47-58
__device__ __forceinline__ float quake_rsqrt(float x) {
    float half = x * 0.5f;
    float y = u2f(0x5f3759dfu - (f2u(x) >> 1));
    return y * (1.5f - half * y * y);
}

// compressMaterial, src/Util.hpp:64-84
__device__ __forceinline__ uint32_t compress_material(const float n[3], float shade) {
    uint32_t face = 0;
    float dom = fabsf(n[0]);
    if (fabsf(n[1]) > dom) { dom = fabsf(n[1]); face = 1; }
    if (fabsf(n[2]) > dom) { dom = fabsf(n[2]); face = 2; }
    const uint32_t sign = n[face] < 0.0f;
    const float n1 = n[face == 2 ? 0 : face + 1] / dom;
    const float n2 = n[face == 0 ? 2 : face - 1] / dom;
    const int32_t ui = cvt_i32((n1 * 0.5f + 0.5f) * 2047.0f);
    const int32_t vi = cvt_i32((n2 * 0.5f + 0.5f) * 2047.0f);
    const int32_t ci = cvt_i32(shade * 127.0f);
    const uint32_t u = uint32_t(ui < 0x7FF ? ui : 0x7FF);
    const uint32_t v = uint32_t(vi < 0x7FF ? vi : 0x7FF);
    const uint32_t c = uint32_t(ci < 0x7F ? ci : 0x7F);
    return (sign << 31) | (face << 29) | (u << 18) | (v << 7) | c;
}

// decompressMaterial, src/Util.hpp:86-100 (words written by compress_material
// only, so the face field is 0..2)
__device__ __forceinline__ void decompress_material(uint32_t w, float n[3], float &shade) {
    const uint32_t sign = (w & 0x80000000u) >> 31;
    const uint32_t face = ((w & 0x60000000u) >> 29) % 3u;
    const uint32_t u = (w & 0x1FFC0000u) >> 18;
    const uint32_t v = (w & 0x0003FF80u) >> 7;
    const uint32_t c = (w & 0x0000007Fu);
    float a[3];
    a[face] = sign ? -1.0f : 1.0f;
    a[face == 2 ? 0 : face + 1] = float(u) * 4.8852e-4f * 2.0f - 1.0f;
    a[face == 0 ? 2 : face - 1] = float(v) * 4.8852e-4f * 2.0f - 1.0f;
    const float s = quake_rsqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    n[0] = a[0] * s;
    n[1] = a[1] * s;
    n[2] = a[2] * s;
    shade = float(c) * 1.0f / 127.0f;
}

// One edge-cross-axis test of triBoxOverlap: the projections pa, pb of two
// vertices span [lo, hi]; separating if outside [-rad, rad]. b_first keeps the
// reference's comparison order for the Z12 tests (tribox3.c:118-123).
__device__ __forceinline__ bool axis_separates(float pa, float pb, float rad, bool b_first) {
    float lo, hi;
    if (b_first) {
        if (pb < pa) { lo = pb; hi = pa; } else { lo = pa; hi = pb; }
    } else {
        if (pa < pb) { lo = pa; hi = pb; } else { lo = pb; hi = pa; }
    }
    return lo > rad || hi < -rad;
}

// triBoxOverlap, src/third-party/tribox3.c:132-206 (Akenine-Moller SAT: nine
// edge-axis tests, the AABB test, then the triangle plane)
__device__ __forceinline__ bool tri_box(const float c[3], const float h[3], const float tv[9]) {
    float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        v0[k] = tv[k] - c[k];
        v1[k] = tv[3 + k] - c[k];
        v2[k] = tv[6 + k] - c[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        e0[k] = v1[k] - v0[k];
        e1[k] = v2[k] - v1[k];
        e2[k] = v0[k] - v2[k];
    }
    float fx, fy, fz, rad;
    fx = fabsf(e0[0]); fy = fabsf(e0[1]); fz = fabsf(e0[2]);
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(e0[2] * v0[1] - e0[1] * v0[2], e0[2] * v2[1] - e0[1] * v2[2], rad, false)) return false;
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(-e0[2] * v0[0] + e0[0] * v0[2], -e0[2] * v2[0] + e0[0] * v2[2], rad, false)) return false;
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(e0[1] * v1[0] - e0[0] * v1[1], e0[1] * v2[0] - e0[0] * v2[1], rad, true)) return false;

    fx = fabsf(e1[0]); fy = fabsf(e1[1]); fz = fabsf(e1[2]);
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(e1[2] * v0[1] - e1[1] * v0[2], e1[2] * v2[1] - e1[1] * v2[2], rad, false)) return false;
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(-e1[2] * v0[0] + e1[0] * v0[2], -e1[2] * v2[0] + e1[0] * v2[2], rad, false)) return false;
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(e1[1] * v0[0] - e1[0] * v0[1], e1[1] * v1[0] - e1[0] * v1[1], rad, false)) return false;

    fx = fabsf(e2[0]); fy = fabsf(e2[1]); fz = fabsf(e2[2]);
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(e2[2] * v0[1] - e2[1] * v0[2], e2[2] * v1[1] - e2[1] * v1[2], rad, false)) return false;
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(-e2[2] * v0[0] + e2[0] * v0[2], -e2[2] * v1[0] + e2[0] * v1[2], rad, false)) return false;
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(e2[1] * v1[0] - e2[0] * v1[1], e2[1] * v2[0] - e2[0] * v2[1], rad, true)) return false;

#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float lo = v0[k], hi = v0[k];
        if (v1[k] < lo) lo = v1[k];
        if (v1[k] > hi) hi = v1[k];
        if (v2[k] < lo) lo = v2[k];
        if (v2[k] > hi) hi = v2[k];
        if (lo > h[k] || hi < -h[k]) return false;
    }
    float nrm[3];
    nrm[0] = e0[1] * e1[2] - e0[2] * e1[1];
    nrm[1] = e0[2] * e1[0] - e0[0] * e1[2];
    nrm[2] = e0[0] * e1[1] - e0[1] * e1[0];
    float vmin[3], vmax[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (nrm[k] > 0.0f) { vmin[k] = -h[k] - v0[k]; vmax[k] = h[k] - v0[k]; }
        else { vmin[k] = h[k] - v0[k]; vmax[k] = -h[k] - v0[k]; }
    }
    if (nrm[0] * vmin[0] + nrm[1] * vmin[1] + nrm[2] * vmin[2] > 0.0f) return false;
    return nrm[0] * vmax[0] + nrm[1] * vmax[1] + nrm[2] * vmax[2] >= 0.0f;
}

__device__ __forceinline__ void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ float length3(const float a[3]) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// Vec3::normalize, src/math/Vec3.hpp:57-61
__device__ __forceinline__ void normalize3(const float a[3], float r[3]) {
    const float inv = 1.0f / sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    r[0] = a[0] * inv;
    r[1] = a[1] * inv;
    r[2] = a[2] * inv;
}

// 3-D Morton code, x in the lowest bit of each triple (ascending Morton
// order = the reference's child storage order, src/VoxelOctree.cpp:204-206).
__host__ __device__ __forceinline__ uint64_t spread3(uint64_t v) {
    v &= 0x1FFFFF;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ __forceinline__ uint32_t compact3(uint64_t v) {
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2)) & 0x10C30C30C30C30C3ull;
    v = (v ^ (v >> 4)) & 0x100F00F00F00F00Full;
    v = (v ^ (v >> 8)) & 0x1F0000FF0000FFull;
    v = (v ^ (v >> 16)) & 0x1F00000000FFFFull;
    v = (v ^ (v >> 32)) & 0x1FFFFFull;
    return uint32_t(v);
}

__host__ __device__ __forceinline__ uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return spread3(x) | (spread3(y) << 1) | (spread3(z) << 2);
}

}  // namespace dev
}  // namespace svo
