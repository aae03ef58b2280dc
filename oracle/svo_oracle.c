/*
 * svo_oracle.c — CPU restatement of tunabrain/sparse-voxel-octrees' hot paths.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle). The product never links or
 * calls this file. Each function names the reference code it restates
 * (path:line in the reference repository); the arithmetic follows the
 * reference's operation order exactly, in single precision, with
 * -ffp-contract=off (see oracle/Makefile).
 *
 * Parity of this restatement is pinned by tests/test_oracle_*.py against
 * golden vectors produced by the reference itself (tests/golden/), against
 * the shipped dragon fixture, and against closed forms; see the header of
 * each test for what it pins.
 */
#define _GNU_SOURCE
#include "svo_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* bit casts and helpers                                                      */
/* ------------------------------------------------------------------------ */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* std::min / std::max: return the first argument unless the second compares
 * strictly smaller / larger (keeps -0.0 and NaN behaviour of the reference). */
static inline float smin(float a, float b) { return (b < a) ? b : a; }
static inline float smax(float a, float b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }

/* float -> int32 truncation with x86 cvttss2si semantics: NaN and out of range
 * give INT32_MIN (the reference's (int) casts compile to that instruction). */
static inline int32_t cvt_i32(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}

static inline int popc8(uint32_t v) { return __builtin_popcount(v & 0xFFu); }

static void normalize3(const float a[3], float r[3]);

/* ------------------------------------------------------------------------ */
/* Util.hpp                                                                   */
/* ------------------------------------------------------------------------ */

/* invSqrt, src/Util.hpp:47-58: one Newton step on the 0x5f3759df guess. */
float oracle_inv_sqrt(float x) {
    float half = x * 0.5f;
    float y = u2f(0x5f3759dfu - (f2u(x) >> 1));
    return y * (1.5f - half * y * y);
}

/* compressMaterial, src/Util.hpp:64-84. */
uint32_t oracle_compress_material(const float n[3], float shade) {
    uint32_t face = 0;
    float dom = fabsf(n[0]);
    if (fabsf(n[1]) > dom) { dom = fabsf(n[1]); face = 1; }
    if (fabsf(n[2]) > dom) { dom = fabsf(n[2]); face = 2; }
    uint32_t sign = n[face] < 0.0f;
    float n1 = n[(face + 1) % 3] / dom;
    float n2 = n[(face + 2) % 3] / dom;
    int32_t ui = cvt_i32((n1 * 0.5f + 0.5f) * 2047.0f);
    int32_t vi = cvt_i32((n2 * 0.5f + 0.5f) * 2047.0f);
    int32_t ci = cvt_i32(shade * 127.0f);
    uint32_t u = (uint32_t)(ui < 0x7FF ? ui : 0x7FF);
    uint32_t v = (uint32_t)(vi < 0x7FF ? vi : 0x7FF);
    uint32_t c = (uint32_t)(ci < 0x7F ? ci : 0x7F);
    return (sign << 31) | (face << 29) | (u << 18) | (v << 7) | c;
}

/* decompressMaterial, src/Util.hpp:86-100 (normalised with fastNormalization). */
void oracle_decompress_material(uint32_t w, float n[3], float *shade) {
    uint32_t sign = (w & 0x80000000u) >> 31;
    uint32_t face = (w & 0x60000000u) >> 29;
    uint32_t u = (w & 0x1FFC0000u) >> 18;
    uint32_t v = (w & 0x0003FF80u) >> 7;
    uint32_t c = (w & 0x0000007Fu);
    float a[3] = {0.0f, 0.0f, 0.0f};
    /* face == 3 cannot be produced by compressMaterial; the reference would
     * index past its 3-float array. Treat the slot modulo 3 here. */
    uint32_t f = face % 3;
    a[f] = sign ? -1.0f : 1.0f;
    a[(f + 1) % 3] = (float)u * 4.8852e-4f * 2.0f - 1.0f;
    a[(f + 2) % 3] = (float)v * 4.8852e-4f * 2.0f - 1.0f;
    float s = oracle_inv_sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    n[0] = a[0] * s;
    n[1] = a[1] * s;
    n[2] = a[2] * s;
    *shade = (float)c * 1.0f / 127.0f;
}

/* ------------------------------------------------------------------------ */
/* tribox3.c (Akenine-Moller SAT)                                             */
/* ------------------------------------------------------------------------ */

/* One edge-cross-axis test: separating if the projected interval [lo, hi] of
 * the two vertices lies outside [-rad, rad]. `swap_first` reproduces the two
 * comparison orders the reference uses (tribox3.c:89 vs :121). */
static inline int axis_separates(float pa, float pb, float rad, int b_first) {
    float lo, hi;
    if (b_first) { if (pb < pa) { lo = pb; hi = pa; } else { lo = pa; hi = pb; } }
    else         { if (pa < pb) { lo = pa; hi = pb; } else { lo = pb; hi = pa; } }
    return (lo > rad || hi < -rad);
}

/* triBoxOverlap, src/third-party/tribox3.c:132-206. */
int oracle_tribox(const float c[3], const float h[3], const float tv[9]) {
    float v0[3], v1[3], v2[3], e0[3], e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
        v0[k] = tv[0 + k] - c[k];
        v1[k] = tv[3 + k] - c[k];
        v2[k] = tv[6 + k] - c[k];
    }
    for (int k = 0; k < 3; ++k) {
        e0[k] = v1[k] - v0[k];
        e1[k] = v2[k] - v1[k];
        e2[k] = v0[k] - v2[k];
    }
    float fx, fy, fz, pa, pb, rad;

    /* edge 0: axes X(v0,v2) Y(v0,v2) Z(v1,v2) */
    fx = fabsf(e0[0]); fy = fabsf(e0[1]); fz = fabsf(e0[2]);
    pa = e0[2] * v0[1] - e0[1] * v0[2]; pb = e0[2] * v2[1] - e0[1] * v2[2];
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = -e0[2] * v0[0] + e0[0] * v0[2]; pb = -e0[2] * v2[0] + e0[0] * v2[2];
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = e0[1] * v1[0] - e0[0] * v1[1]; pb = e0[1] * v2[0] - e0[0] * v2[1];
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(pa, pb, rad, 1)) return 0;

    /* edge 1: axes X(v0,v2) Y(v0,v2) Z(v0,v1) */
    fx = fabsf(e1[0]); fy = fabsf(e1[1]); fz = fabsf(e1[2]);
    pa = e1[2] * v0[1] - e1[1] * v0[2]; pb = e1[2] * v2[1] - e1[1] * v2[2];
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = -e1[2] * v0[0] + e1[0] * v0[2]; pb = -e1[2] * v2[0] + e1[0] * v2[2];
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = e1[1] * v0[0] - e1[0] * v0[1]; pb = e1[1] * v1[0] - e1[0] * v1[1];
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(pa, pb, rad, 0)) return 0;

    /* edge 2: axes X(v0,v1) Y(v0,v1) Z(v1,v2) */
    fx = fabsf(e2[0]); fy = fabsf(e2[1]); fz = fabsf(e2[2]);
    pa = e2[2] * v0[1] - e2[1] * v0[2]; pb = e2[2] * v1[1] - e2[1] * v1[2];
    rad = fz * h[1] + fy * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = -e2[2] * v0[0] + e2[0] * v0[2]; pb = -e2[2] * v1[0] + e2[0] * v1[2];
    rad = fz * h[0] + fx * h[2];
    if (axis_separates(pa, pb, rad, 0)) return 0;
    pa = e2[1] * v1[0] - e2[0] * v1[1]; pb = e2[1] * v2[0] - e2[0] * v2[1];
    rad = fy * h[0] + fx * h[1];
    if (axis_separates(pa, pb, rad, 1)) return 0;

    /* triangle AABB against the box (tribox3.c:187-196) */
    for (int k = 0; k < 3; ++k) {
        float lo = v0[k], hi = v0[k];
        if (v1[k] < lo) lo = v1[k];
        if (v1[k] > hi) hi = v1[k];
        if (v2[k] < lo) lo = v2[k];
        if (v2[k] > hi) hi = v2[k];
        if (lo > h[k] || hi < -h[k]) return 0;
    }

    /* plane of the triangle (tribox3.c:201-203, planeBoxOverlap :60-82) */
    float nrm[3];
    nrm[0] = e0[1] * e1[2] - e0[2] * e1[1];
    nrm[1] = e0[2] * e1[0] - e0[0] * e1[2];
    nrm[2] = e0[0] * e1[1] - e0[1] * e1[0];
    float vmin[3], vmax[3];
    for (int k = 0; k < 3; ++k) {
        if (nrm[k] > 0.0f) { vmin[k] = -h[k] - v0[k]; vmax[k] = h[k] - v0[k]; }
        else               { vmin[k] = h[k] - v0[k];  vmax[k] = -h[k] - v0[k]; }
    }
    if (nrm[0] * vmin[0] + nrm[1] * vmin[1] + nrm[2] * vmin[2] > 0.0f) return 0;
    if (nrm[0] * vmax[0] + nrm[1] * vmax[1] + nrm[2] * vmax[2] >= 0.0f) return 1;
    return 0;
}

void oracle_compress_material_batch(int64_t n, const float *normals, const float *shades, uint32_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = oracle_compress_material(normals + 3 * i, shades[i]);
}

void oracle_decompress_material_batch(int64_t n, const uint32_t *w, float *normals, float *shades) {
    for (int64_t i = 0; i < n; ++i) oracle_decompress_material(w[i], normals + 3 * i, shades + i);
}

void oracle_tribox_batch(int64_t n, const float *rec, int32_t *out) {
    for (int64_t i = 0; i < n; ++i) out[i] = oracle_tribox(rec + 15 * i, rec + 15 * i + 3, rec + 15 * i + 6);
}

/* ------------------------------------------------------------------------ */
/* raymarch                                                                   */
/* ------------------------------------------------------------------------ */

/* VoxelOctree::raymarch, src/VoxelOctree.cpp:207-346, with MaxScale = 23
 * (VoxelOctree.hpp:38). Sibling counts use popcount of the same masked bits
 * as the BitCount table (VoxelOctree.cpp:36-53). */
int oracle_raymarch(const uint32_t *tree, uint64_t nwords, const float o[3], const float d[3],
                    float rayScale, uint32_t *normal, float *t, uint64_t *leaf,
                    uint64_t counts[3]) {
    uint64_t stackParent[24];
    float stackMaxT[24];

    float ox = o[0], oy = o[1], oz = o[2];
    float dx = d[0], dy = d[1], dz = d[2];
    if (fabsf(dx) < 1e-4f) dx = 1e-4f;
    if (fabsf(dy) < 1e-4f) dy = 1e-4f;
    if (fabsf(dz) < 1e-4f) dz = 1e-4f;

    float dTx = 1.0f / -fabsf(dx);
    float dTy = 1.0f / -fabsf(dy);
    float dTz = 1.0f / -fabsf(dz);
    float bTx = dTx * ox;
    float bTy = dTy * oy;
    float bTz = dTz * oz;

    int octantMask = 7;
    if (dx > 0.0f) { octantMask ^= 1; bTx = 3.0f * dTx - bTx; }
    if (dy > 0.0f) { octantMask ^= 2; bTy = 3.0f * dTy - bTy; }
    if (dz > 0.0f) { octantMask ^= 4; bTz = 3.0f * dTz - bTz; }

    /* A NaN coefficient (NaN origin or direction, or -0 * inf) makes every
     * corner time NaN; no axis ever steps and the reference's loop never ends.
     * Such rays are reported as misses, as the product's generic ray path
     * reports them. */
    if (dTx != dTx || dTy != dTy || dTz != dTz || bTx != bTx || bTy != bTy || bTz != bTz) return 0;

    float minT = smax(2.0f * dTx - bTx, smax(2.0f * dTy - bTy, 2.0f * dTz - bTz));
    float maxT = smin(dTx - bTx, smin(dTy - bTy, dTz - bTz));
    minT = smax(minT, 0.0f);

    uint32_t current = 0;
    uint64_t parent = 0;
    int idx = 0;
    float posX = 1.0f, posY = 1.0f, posZ = 1.0f;
    int scale = 22;
    float scaleExp2 = 0.5f;

    if (1.5f * dTx - bTx > minT) { idx ^= 1; posX = 1.5f; }
    if (1.5f * dTy - bTy > minT) { idx ^= 2; posY = 1.5f; }
    if (1.5f * dTz - bTz > minT) { idx ^= 4; posZ = 1.5f; }

    while (scale < 23) {
        if (current == 0) {
            if (parent >= nwords) return -1;
            current = tree[parent];
            if (counts) counts[0]++;
        }

        float cornerTX = posX * dTx - bTx;
        float cornerTY = posY * dTy - bTy;
        float cornerTZ = posZ * dTz - bTz;
        float maxTC = smin(cornerTX, smin(cornerTY, cornerTZ));

        int childShift = idx ^ octantMask;
        uint32_t childMasks = current << childShift;

        if ((childMasks & 0x8000u) && minT <= maxT) {
            if (maxTC * rayScale >= scaleExp2) {
                *t = maxTC;
                return 2;
            }
            float maxTV = smin(maxT, maxTC);
            float half = scaleExp2 * 0.5f;
            float centerTX = half * dTx + cornerTX;
            float centerTY = half * dTy + cornerTY;
            float centerTZ = half * dTz + cornerTZ;

            if (minT <= maxTV) {
                uint64_t childOffset = current >> 18;
                if (current & 0x20000u) {
                    if (parent + 1 >= nwords) return -1;
                    childOffset = (childOffset << 32) | (uint64_t)tree[parent + 1];
                    if (counts) counts[1]++;
                }
                if (!(childMasks & 0x80u)) {
                    uint64_t li = childOffset + parent +
                                  (uint64_t)popc8(((childMasks >> (8 + childShift)) << childShift) & 127u);
                    if (li >= nwords) return -1;
                    *normal = tree[li];
                    *leaf = li;
                    if (counts) counts[2]++;
                    break;
                }
                stackParent[scale] = parent;
                stackMaxT[scale] = maxT;

                uint32_t siblingCount = (uint32_t)popc8(childMasks & 127u);
                parent += childOffset + siblingCount;
                if (current & 0x10000u) parent += siblingCount;

                idx = 0;
                scale--;
                scaleExp2 = half;
                if (centerTX > minT) { idx ^= 1; posX += scaleExp2; }
                if (centerTY > minT) { idx ^= 2; posY += scaleExp2; }
                if (centerTZ > minT) { idx ^= 4; posZ += scaleExp2; }
                maxT = maxTV;
                current = 0;
                continue;
            }
        }

        int stepMask = 0;
        if (cornerTX <= maxTC) { stepMask ^= 1; posX -= scaleExp2; }
        if (cornerTY <= maxTC) { stepMask ^= 2; posY -= scaleExp2; }
        if (cornerTZ <= maxTC) { stepMask ^= 4; posZ -= scaleExp2; }

        minT = maxTC;
        idx ^= stepMask;

        if ((idx & stepMask) != 0) {
            uint32_t diff = 0;
            if (stepMask & 1) diff |= f2u(posX) ^ f2u(posX + scaleExp2);
            if (stepMask & 2) diff |= f2u(posY) ^ f2u(posY + scaleExp2);
            if (stepMask & 4) diff |= f2u(posZ) ^ f2u(posZ + scaleExp2);
            scale = (int)(f2u((float)(int32_t)diff) >> 23) - 127;
            scaleExp2 = u2f((uint32_t)(scale - 23 + 127) << 23);
            /* The reference reads rayStack[scale] even when scale >= 23, then
             * leaves the loop; nothing read there reaches the result. */
            if (scale >= 23) break;
            parent = stackParent[scale];
            maxT = stackMaxT[scale];

            int shX = (int)(f2u(posX) >> scale);
            int shY = (int)(f2u(posY) >> scale);
            int shZ = (int)(f2u(posZ) >> scale);
            posX = u2f((uint32_t)shX << scale);
            posY = u2f((uint32_t)shY << scale);
            posZ = u2f((uint32_t)shZ << scale);
            idx = (shX & 1) | ((shY & 1) << 1) | ((shZ & 1) << 2);
            current = 0;
        }
    }

    if (scale >= 23) return 0;
    *t = minT;
    return 1;
}

typedef struct {
    const uint32_t *tree;
    uint64_t nwords;
    int64_t begin, end;
    const float *o, *d, *rs;
    float scale0;
    uint8_t *hit;
    uint32_t *normal;
    float *t;
    uint64_t *leaf;
    uint64_t counts[3];
    int bad;
} rm_job;

static void *rm_worker(void *arg) {
    rm_job *j = (rm_job *)arg;
    for (int64_t i = j->begin; i < j->end; ++i) {
        uint32_t nrm = 0;
        float tt = 0.0f;
        uint64_t li = ORACLE_LEAF_NONE;
        float rs = j->rs ? j->rs[i] : j->scale0;
        int h = oracle_raymarch(j->tree, j->nwords, j->o + 3 * i, j->d + 3 * i, rs, &nrm, &tt, &li,
                                j->counts);
        if (h < 0) { j->bad = 1; h = 0; }
        if (h != 1) { nrm = 0; li = ORACLE_LEAF_NONE; }
        if (h == 0) tt = 0.0f;
        if (j->hit) j->hit[i] = (uint8_t)h;
        if (j->normal) j->normal[i] = nrm;
        if (j->t) j->t[i] = tt;
        if (j->leaf) j->leaf[i] = li;
    }
    return NULL;
}

int oracle_raymarch_batch(const uint32_t *tree, uint64_t nwords, int64_t n, const float *o,
                          const float *d, const float *rayScale, float scale0, uint8_t *hit,
                          uint32_t *normal, float *t, uint64_t *leaf, uint64_t counts[3],
                          int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    rm_job jobs[256];
    pthread_t tid[256];
    for (int k = 0; k < threads; ++k) {
        rm_job *j = &jobs[k];
        memset(j, 0, sizeof(*j));
        j->tree = tree; j->nwords = nwords;
        j->begin = n * k / threads; j->end = n * (k + 1) / threads;
        j->o = o; j->d = d; j->rs = rayScale; j->scale0 = scale0;
        j->hit = hit; j->normal = normal; j->t = t; j->leaf = leaf;
    }
    for (int k = 1; k < threads; ++k) pthread_create(&tid[k], NULL, rm_worker, &jobs[k]);
    rm_worker(&jobs[0]);
    for (int k = 1; k < threads; ++k) pthread_join(tid[k], NULL);
    int bad = 0;
    for (int k = 0; k < threads; ++k) {
        if (counts) for (int c = 0; c < 3; ++c) counts[c] += jobs[k].counts[c];
        bad |= jobs[k].bad;
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* camera (src/math/Mat4.cpp, src/math/MatrixStack.cpp, src/Main.cpp)         */
/* ------------------------------------------------------------------------ */

typedef struct { float a[16]; } m4;

static m4 m4_identity(void) {
    m4 m; memset(&m, 0, sizeof(m));
    m.a[0] = m.a[5] = m.a[10] = m.a[15] = 1.0f;
    return m;
}

/* Mat4::operator*(Mat4), Mat4.cpp:67-78 */
static m4 m4_mul(const m4 *x, const m4 *y) {
    m4 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            r.a[i * 4 + k] = x->a[i * 4 + 0] * y->a[0 * 4 + k] + x->a[i * 4 + 1] * y->a[1 * 4 + k] +
                             x->a[i * 4 + 2] * y->a[2 * 4 + k] + x->a[i * 4 + 3] * y->a[3 * 4 + k];
    return r;
}

/* Mat4::translate, Mat4.cpp:96-103 */
static m4 m4_translate(float x, float y, float z) {
    m4 m = m4_identity();
    m.a[3] = x; m.a[7] = y; m.a[11] = z;
    return m;
}

/* Mat4::rotXYZ, Mat4.cpp:114-125 (degrees; cosf/sinf of the float angles) */
static m4 m4_rot_xyz(float rx, float ry, float rz) {
    const float pi = (float)3.14159265358979323846;
    float r0 = rx * pi / 180.0f, r1 = ry * pi / 180.0f, r2 = rz * pi / 180.0f;
    float c0 = cosf(r0), c1 = cosf(r1), c2 = cosf(r2);
    float s0 = sinf(r0), s1 = sinf(r1), s2 = sinf(r2);
    m4 m;
    m.a[0] = c1 * c2;  m.a[1] = -c0 * s2 + s0 * s1 * c2; m.a[2] = s0 * s2 + c0 * s1 * c2;  m.a[3] = 0.0f;
    m.a[4] = c1 * s2;  m.a[5] = c0 * c2 + s0 * s1 * s2;  m.a[6] = -s0 * c2 + c0 * s1 * s2; m.a[7] = 0.0f;
    m.a[8] = -s1;      m.a[9] = s0 * c1;                 m.a[10] = c0 * c1;                m.a[11] = 0.0f;
    m.a[12] = 0.0f;    m.a[13] = 0.0f;                   m.a[14] = 0.0f;                   m.a[15] = 1.0f;
    return m;
}

/* Mat4::pseudoInvert, Mat4.cpp:59-65 */
static m4 m4_pseudo_invert(const m4 *x) {
    m4 trans = m4_translate(-x->a[3], -x->a[7], -x->a[11]);
    m4 rot;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) rot.a[i * 4 + k] = x->a[k * 4 + i];
    rot.a[12] = rot.a[13] = rot.a[14] = 0.0f;
    return m4_mul(&rot, &trans);
}

void oracle_orbit_camera(float pitch, float yaw, float radius, const float center[3],
                         float tform[16], float pos[3]) {
    m4 rp = m4_rot_xyz(pitch, 0.0f, 0.0f);
    m4 ry = m4_rot_xyz(0.0f, yaw, 0.0f);
    m4 model = m4_mul(&rp, &ry);
    m4 view = m4_translate(0.0f, 0.0f, -radius);
    m4 inv = m4_pseudo_invert(&model);
    m4 tf = m4_mul(&inv, &view); /* MatrixStack::get(INV_MODELVIEW_STACK) */
    memcpy(tform, tf.a, 64);
    /* pos = tform*Vec3() + center + Vec3(1.0)   (Main.cpp:150-152; Mat4.cpp:80-86) */
    for (int r = 0; r < 3; ++r) {
        float p = tf.a[r * 4 + 0] * 0.0f + tf.a[r * 4 + 1] * 0.0f + tf.a[r * 4 + 2] * 0.0f + tf.a[r * 4 + 3];
        pos[r] = (p + center[r]) + 1.0f;
    }
}

void oracle_primary_dirs(const float tf[16], int width, int height, float *dirs) {
    const float pi = (float)3.14159265358979323846;
    float aspect = (float)height / (float)width;         /* Main.cpp:62 */
    float scale = 2.0f / (float)width;                   /* Main.cpp:156 */
    float planeDist = 1.0f / tanf(pi / 6.0f);            /* Main.cpp:158 */
    float zx = planeDist * tf[2], zy = planeDist * tf[6], zz = planeDist * tf[10];
    for (int y = 0; y < height; ++y) {
        float dy = aspect - (float)y * scale;
        for (int x = 0; x < width; ++x) {
            float dx = -1.0f + (float)x * scale;
            float vx = dx * tf[0] + dy * tf[1] + zx;
            float vy = dx * tf[4] + dy * tf[5] + zy;
            float vz = dx * tf[8] + dy * tf[9] + zz;
            float s = oracle_inv_sqrt(vx * vx + vy * vy + vz * vz);
            float *o = dirs + 3 * ((int64_t)y * width + x);
            o[0] = vx * s; o[1] = vy * s; o[2] = vz * s;
        }
    }
}

/* shade, src/Main.cpp:81-90: decoded normal and shade, Lambert + reflection
 * term; the float is widened to a grey Vec3 by the caller. */
static float frame_shade(uint32_t w, const float ray[3], const float light[3]) {
    float n[3], c;
    oracle_decompress_material(w, n, &c);
    float proj = (n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2]) * 2.0f;   /* Vec3::reflect */
    float r[3] = {ray[0] - n[0] * proj, ray[1] - n[1] * proj, ray[2] - n[2] * proj};
    float d = smax(light[0] * r[0] + light[1] * r[1] + light[2] * r[2], 0.0f);
    float specular = d * d;
    return c * 0.9f * fabsf(light[0] * n[0] + light[1] * n[1] + light[2] * n[2]) + specular * 0.2f;
}

/* colour packing of renderTile (non-Apple byte order), src/Main.cpp:130-135 */
static uint32_t frame_pack(float v) {
    uint32_t g = (uint32_t)((double)smin(v, 1.0f) * 255.0);
    return g | (g << 8) | (g << 16) | 0xFF000000u;
}

typedef struct {
    const uint32_t *tree;
    uint64_t nwords;
    float a11, a12, a21, a22, a31, a32, zx, zy, zz;
    float pos[3], light[3];
    float aspect, scale;
    int width;
    uint32_t *rgba;
} frame_ctx;

static void frame_dir(const frame_ctx *f, float dx, float dy, float dir[3]) {
    float vx = dx * f->a11 + dy * f->a12 + f->zx;
    float vy = dx * f->a21 + dy * f->a22 + f->zy;
    float vz = dx * f->a31 + dy * f->a32 + f->zz;
    float s = oracle_inv_sqrt(vx * vx + vy * vy + vz * vz);
    dir[0] = vx * s; dir[1] = vy * s; dir[2] = vz * s;
}

/* renderTile, src/Main.cpp:92-137 */
static void frame_tile(const frame_ctx *f, int x0, int y0, int x1, int y1, int stride, float minT) {
    float dy = f->aspect - (float)y0 * f->scale;
    for (int y = y0; y < y1; ++y, dy -= f->scale) {
        float dx = -1.0f + (float)x0 * f->scale;
        for (int x = x0; x < x1; ++x, dx += f->scale) {
            int cx = x - ((x - x0) % stride), cy = y - ((y - y0) % stride);
            if (cx != x || cy != y) {
                f->rgba[(int64_t)y * f->width + x] = f->rgba[(int64_t)cy * f->width + cx];
                continue;
            }
            float dir[3];
            frame_dir(f, dx, dy, dir);
            float o[3] = {f->pos[0] + dir[0] * minT, f->pos[1] + dir[1] * minT, f->pos[2] + dir[2] * minT};
            uint32_t nrm = 0;
            float t = 0.0f;
            uint64_t leaf = 0;
            float v = 0.0f;
            if (oracle_raymarch(f->tree, f->nwords, o, dir, 0.0f, &nrm, &t, &leaf, NULL) > 0) v = frame_shade(nrm, dir, f->light);
            f->rgba[(int64_t)y * f->width + x] = frame_pack(v);
        }
    }
}

int64_t oracle_frame_depth_count(int width, int height, int stripes) {
    int st = (height - 1) / stripes + 1;
    int64_t n = 0;
    for (int i = 0; i < stripes; ++i) {
        int y0 = i * st, y1 = imin((i + 1) * st, height);
        n += (int64_t)((width - 1) / 8 + 2) * ((y1 - y0 - 1) / 8 + 2);
    }
    return n;
}

void oracle_render_frame(const uint32_t *tree, uint64_t nwords, const float tform[16], const float center[3],
                         int width, int height, int stripes, int stride, uint32_t *rgba, float *depth) {
    const float pi = (float)3.14159265358979323846;
    frame_ctx f;
    f.tree = tree; f.nwords = nwords; f.width = width; f.rgba = rgba;
    /* renderBatch, src/Main.cpp:139-202 */
    for (int r = 0; r < 3; ++r) {
        float p = tform[r * 4 + 0] * 0.0f + tform[r * 4 + 1] * 0.0f + tform[r * 4 + 2] * 0.0f + tform[r * 4 + 3];
        f.pos[r] = (p + center[r]) + 1.0f;
    }
    float scale = 2.0f / (float)width;
    float tileScale = 8.0f * scale;
    float planeDist = 1.0f / tanf(pi / 6.0f);
    f.a11 = tform[0]; f.a12 = tform[1]; f.a21 = tform[4]; f.a22 = tform[5]; f.a31 = tform[8]; f.a32 = tform[9];
    f.zx = planeDist * tform[2]; f.zy = planeDist * tform[6]; f.zz = planeDist * tform[10];
    float coarseScale = 2.0f * 8.0f / (planeDist * (float)height);
    /* light = (tform*Vec3(-1,1,-1)).normalize() with the translation cleared */
    float lv[3];
    for (int r = 0; r < 3; ++r)
        lv[r] = tform[r * 4 + 0] * -1.0f + tform[r * 4 + 1] * 1.0f + tform[r * 4 + 2] * -1.0f + 0.0f;
    normalize3(lv, f.light);
    f.aspect = (float)height / (float)width;
    f.scale = scale;
    const float miss = 1e10f;
    int st = (height - 1) / stripes + 1;
    int64_t dbase = 0;
    for (int i = 0; i < stripes; ++i) {
        int x0 = 0, x1 = width, y0 = i * st, y1 = imin((i + 1) * st, height);
        int tilesX = (x1 - x0 - 1) / 8 + 2, tilesY = (y1 - y0 - 1) / 8 + 2;
        float *db = (float *)malloc(sizeof(float) * (size_t)tilesX * tilesY);
        for (int y = y0; y < y1; ++y) memset(rgba + (int64_t)y * width, 0, sizeof(uint32_t) * (size_t)width);
        float dy = f.aspect - (float)y0 * scale;
        int idx = 0;
        for (int y = 0; y < tilesY; y++, dy -= tileScale) {
            float dx = -1.0f + (float)x0 * scale;
            for (int x = 0; x < tilesX; x++, dx += tileScale, idx++) {
                float dir[3];
                frame_dir(&f, dx, dy, dir);
                uint32_t nrm = 0;
                float t = 0.0f;
                uint64_t leaf = 0;
                db[idx] = oracle_raymarch(tree, nwords, f.pos, dir, coarseScale, &nrm, &t, &leaf, NULL) > 0 ? t : miss;
                if (x > 0 && y > 0) {
                    float minT = smin(smin(db[idx], db[idx - 1]), smin(db[idx - tilesX], db[idx - tilesX - 1]));
                    if (minT != miss) {
                        int tx0 = (x - 1) * 8 + x0, ty0 = (y - 1) * 8 + y0;
                        frame_tile(&f, tx0, ty0, imin(tx0 + 8, x1), imin(ty0 + 8, y1), stride, smax(minT - 0.03f, 0.0f));
                    }
                }
            }
        }
        if (depth) memcpy(depth + dbase, db, sizeof(float) * (size_t)tilesX * tilesY);
        dbase += (int64_t)tilesX * tilesY;
        free(db);
    }
}

/* ------------------------------------------------------------------------ */
/* .oct decode: LZ4 block format, chunks of 64 MiB sharing one output window  */
/* (src/VoxelOctree.cpp:57-90; src/third-party/lz4.c:1384-1400)               */
/* ------------------------------------------------------------------------ */

/* Decodes one LZ4 block from src[0..srcSize) into out[pos..pos+outSize),
 * allowing matches to reach back into out[0..pos). Returns 0 on success. */
static int lz4_block_decode(const uint8_t *src, uint64_t srcSize, uint8_t *out, uint64_t pos,
                            uint64_t outSize) {
    uint64_t ip = 0, op = pos, oend = pos + outSize;
    while (ip < srcSize) {
        uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint32_t b;
            do { if (ip >= srcSize) return -1; b = src[ip++]; lit += b; } while (b == 255);
        }
        if (ip + lit > srcSize || op + lit > oend) return -1;
        memcpy(out + op, src + ip, lit);
        ip += lit; op += lit;
        if (op == oend) return ip == srcSize ? 0 : -1;
        if (ip + 2 > srcSize) return -1;
        uint64_t off = (uint64_t)src[ip] | ((uint64_t)src[ip + 1] << 8);
        ip += 2;
        if (off == 0 || off > op) return -1;
        uint64_t ml = (token & 15u);
        if (ml == 15) {
            uint32_t b;
            do { if (ip >= srcSize) return -1; b = src[ip++]; ml += b; } while (b == 255);
        }
        ml += 4;
        if (op + ml > oend) return -1;
        for (uint64_t k = 0; k < ml; ++k) out[op + k] = out[op + k - off];
        op += ml;
    }
    return op == oend ? 0 : -1;
}

int oracle_oct_decode(const uint8_t *file, uint64_t size, float center[3], uint32_t **words,
                      uint64_t *nwords) {
    if (size < 20) return -1;
    memcpy(center, file, 12);
    uint64_t n;
    memcpy(&n, file + 12, 8);
    if (n > (1ull << 40)) return -1;
    uint64_t bytes = n * 4;
    uint8_t *out = (uint8_t *)malloc(bytes ? bytes : 4);
    if (!out) return -1;
    uint64_t p = 20;
    const uint64_t chunk = 64ull << 20;
    for (uint64_t off = 0; off < bytes; off += chunk) {
        uint64_t cs;
        if (p + 8 > size) { free(out); return -1; }
        memcpy(&cs, file + p, 8);
        p += 8;
        if (p + cs > size) { free(out); return -1; }
        uint64_t os = bytes - off < chunk ? bytes - off : chunk;
        if (lz4_block_decode(file + p, cs, out, off, os) != 0) { free(out); return -1; }
        p += cs;
    }
    *words = (uint32_t *)out;
    *nwords = n;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* structural walk (layout written by VoxelOctree.cpp:199-265, read :312-399) */
/* ------------------------------------------------------------------------ */

static int walk_node(const uint32_t *tree, uint64_t nwords, uint64_t p, uint64_t off, int level,
                     uint8_t *kind, uint64_t *dpl, uint64_t *leaves, uint64_t *farw, int *depth) {
    if (level >= 24) return -1;
    uint32_t w = tree[p];
    if (kind) kind[p] = 1;
    dpl[level]++;
    if (level + 1 > *depth) *depth = level + 1;
    uint32_t valid = (w >> 8) & 0xFFu, nonleaf = w & 0xFFu;
    int large = (w >> 16) & 1;
    uint64_t base = p + off;
    int k = 0;
    for (int m = 0; m < 8; ++m) {
        if (!((valid >> m) & 1u)) continue;
        uint64_t slot = base + (uint64_t)k * (uint64_t)(1 + large);
        k++;
        if (slot >= nwords) return -1;
        if ((nonleaf >> m) & 1u) {
            uint32_t cw = tree[slot];
            uint64_t coff = cw >> 18;
            if (cw & 0x20000u) {
                if (slot + 1 >= nwords) return -1;
                coff = (coff << 32) | tree[slot + 1];
                if (kind) kind[slot + 1] = 2;
                (*farw)++;
            }
            if (walk_node(tree, nwords, slot, coff, level + 1, kind, dpl, leaves, farw, depth)) return -1;
        } else {
            if (kind) kind[slot] = 3;
            (*leaves)++;
            if (level + 2 > *depth) *depth = level + 2;
        }
    }
    return 0;
}

int oracle_octree_walk(const uint32_t *tree, uint64_t nwords, uint8_t *kind,
                       uint64_t desc_per_level[24], uint64_t *leaves, uint64_t *far_words,
                       int *depth) {
    uint64_t dpl[24] = {0}, lv = 0, fw = 0;
    int dep = 0;
    if (nwords == 0) return -1;
    if (kind) memset(kind, 0, nwords);
    int r = walk_node(tree, nwords, 0, tree[0] >> 18, 0, kind, dpl, &lv, &fw, &dep);
    if (desc_per_level) memcpy(desc_per_level, dpl, sizeof(dpl));
    if (leaves) *leaves = lv;
    if (far_words) *far_words = fw;
    if (depth) *depth = dep;
    return r;
}

/* ------------------------------------------------------------------------ */
/* mesh (src/PlyLoader.cpp:40-224)                                            */
/* ------------------------------------------------------------------------ */

typedef struct { float pos[3], nrm[3], col[3]; } o_vertex;
typedef struct { o_vertex v[3]; float lower[3], upper[3]; } o_tri;

struct oracle_mesh {
    o_tri *tris;
    int64_t ntris;
    float lower[3], upper[3];
};

static void cross3(const float a[3], const float b[3], float r[3]) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static float length3(const float a[3]) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
/* Vec3::normalize, src/math/Vec3.hpp:57-61 */
static void normalize3(const float a[3], float r[3]) {
    float inv = 1.0f / sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    r[0] = a[0] * inv; r[1] = a[1] * inv; r[2] = a[2] * inv;
}

oracle_mesh *oracle_mesh_create(const float *verts, int64_t nv, int has_normals,
                                const int32_t *face_counts, const int32_t *face_idx,
                                int64_t nfaces) {
    oracle_mesh *m = (oracle_mesh *)calloc(1, sizeof(*m));
    o_vertex *vs = (o_vertex *)malloc(sizeof(o_vertex) * (size_t)(nv ? nv : 1));
    float lower[3] = {1e30f, 1e30f, 1e30f}, upper[3] = {-1e30f, -1e30f, -1e30f};
    /* readVertices, PlyLoader.cpp:122-165 */
    for (int64_t i = 0; i < nv; ++i) {
        const float *s = verts + 9 * i;
        memcpy(vs[i].pos, s, 12);
        memcpy(vs[i].nrm, s + 3, 12);
        memcpy(vs[i].col, s + 6, 12);
        for (int k = 0; k < 3; ++k) {
            lower[k] = smin(lower[k], s[k]);
            upper[k] = smax(upper[k], s[k]);
        }
    }
    /* rescaleVertices, PlyLoader.cpp:167-181 */
    float diff[3] = {upper[0] - lower[0], upper[1] - lower[1], upper[2] - lower[2]};
    int largest = 2;
    if (diff[0] > diff[1] && diff[0] > diff[2]) largest = 0;
    else if (diff[1] > diff[2]) largest = 1;
    float factor = 1.0f / diff[largest];
    for (int64_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) vs[i].pos[k] = (vs[i].pos[k] - lower[k]) * factor;
    for (int k = 0; k < 3; ++k) {
        m->upper[k] = upper[k] * factor;
        m->lower[k] = lower[k] * factor;
    }
    /* readTriangles, PlyLoader.cpp:183-224: fan triangulation */
    int64_t cap = 0;
    for (int64_t f = 0; f < nfaces; ++f) if (face_counts[f] > 2) cap += face_counts[f] - 2;
    m->tris = (o_tri *)malloc(sizeof(o_tri) * (size_t)(cap ? cap : 1));
    int64_t at = 0, nt = 0;
    for (int64_t f = 0; f < nfaces; ++f) {
        int cnt = face_counts[f];
        const int32_t *e = face_idx + at;
        at += cnt;
        if (cnt < 3) continue;
        int a = e[0], b = e[1];
        for (int k = 2; k < cnt; ++k) {
            int c = e[k];
            o_tri *t = &m->tris[nt++];
            t->v[0] = vs[a]; t->v[1] = vs[b]; t->v[2] = vs[c];
            for (int q = 0; q < 3; ++q) { /* Triangle ctor, PlyLoader.cpp:40-53 */
                t->lower[q] = smin(t->v[0].pos[q], smin(t->v[1].pos[q], t->v[2].pos[q]));
                t->upper[q] = smax(t->v[0].pos[q], smax(t->v[1].pos[q], t->v[2].pos[q]));
            }
            if (!has_normals) {
                float e1[3], e2[3], cr[3], n[3];
                for (int q = 0; q < 3; ++q) { e1[q] = vs[b].pos[q] - vs[a].pos[q]; e2[q] = vs[c].pos[q] - vs[a].pos[q]; }
                cross3(e1, e2, cr);
                normalize3(cr, n);
                for (int q = 0; q < 3; ++q) memcpy(t->v[q].nrm, n, 12);
            }
            b = c;
        }
    }
    m->ntris = nt;
    free(vs);
    return m;
}

void oracle_mesh_free(oracle_mesh *m) {
    if (!m) return;
    free(m->tris);
    free(m);
}

int64_t oracle_mesh_tri_count(const oracle_mesh *m) { return m->ntris; }

void oracle_mesh_tris(const oracle_mesh *m, float *out) {
    for (int64_t i = 0; i < m->ntris; ++i) {
        const o_tri *t = &m->tris[i];
        float *o = out + 33 * i;
        for (int q = 0; q < 3; ++q) {
            memcpy(o + 9 * q + 0, t->v[q].pos, 12);
            memcpy(o + 9 * q + 3, t->v[q].nrm, 12);
            memcpy(o + 9 * q + 6, t->v[q].col, 12);
        }
        memcpy(o + 27, t->lower, 12);
        memcpy(o + 30, t->upper, 12);
    }
}

void oracle_mesh_bounds(const oracle_mesh *m, float lower[3], float upper[3]) {
    memcpy(lower, m->lower, 12);
    memcpy(upper, m->upper, 12);
}

/* ------------------------------------------------------------------------ */
/* voxel source: VoxelData + PlyLoader block processing                       */
/* ------------------------------------------------------------------------ */

/* roundToPow2 / findHighestBit, src/Util.hpp:102-117 (for powers of two the
 * bit trick equals log2). */
static int round_pow2(int x) { int y = 1; while (y < x) y *= 2; return y; }
static int highest_bit(uint32_t v) { int r = 0; while (v >>= 1) r++; return r; }

static uint64_t grid_cells(int levels) {
    uint64_t r = 0, s = 1;
    for (int i = 0; i < levels; ++i) { r += s; s *= 8; }
    return r;
}

typedef struct {
    /* source */
    const oracle_mesh *mesh;
    const uint32_t *grid; /* file source: whole raw volume */
    /* VoxelData */
    int dataW, dataH, dataD;
    int vW, vH, vD, highBit, lowLevels, topLevels;
    int64_t cache;
    uint8_t *topLut, *lowLut;
    uint8_t **topTab, **lowTab;
    uint32_t *buf;
    int bX, bY, bZ, bW, bH, bD;
    /* PlyLoader */
    int side; /* PlyLoader::_sideLength = resolution - 2 */
    int blkW, blkH, blkD, sbW, sbH, sbD, pW, pH, pD, nPart, gW, gH, gD;
    uint32_t *blockOffsets;
    uint32_t *blockLists;
    uint8_t *counts;
    int pbX, pbY, pbZ, pbW, pbH, pbD; /* PlyLoader::_bufferX.. */
    /* emission (ChunkedAllocator) */
    uint32_t *words;
    uint64_t size, cap;
    uint64_t *insIdx;
    uint32_t *insVal;
    uint64_t nIns, insCap;
    uint64_t leaves;
    int err;
} vctx;

static inline uint64_t lut_idx(int l, int x, int y, int z) {
    return (uint64_t)x + ((uint64_t)y << l) + ((uint64_t)z << (2 * l));
}

/* PlyLoader::pointToGrid, PlyLoader.cpp:226-231 */
static void point_to_grid(const vctx *c, const float p[3], int g[3]) {
    for (int k = 0; k < 3; ++k) g[k] = cvt_i32(p[k] * (float)(c->side - 2) + 1.0f);
}

/* PlyLoader::iterateOverlappingBlocks, PlyLoader.cpp:233-270. Calls emit(cell)
 * for every sub-block cell the triangle is binned into. */
typedef void (*cell_fn)(vctx *c, uint64_t cell, uint32_t tri);
static void for_binned_cells(vctx *c, const o_tri *t, uint32_t ti, cell_fn emit) {
    int l[3], u[3];
    point_to_grid(c, t->lower, l);
    point_to_grid(c, t->upper, u);
    int lg[3] = {l[0] / c->sbW, l[1] / c->sbH, l[2] / c->sbD};
    int ug[3] = {(u[0] + 1) / c->sbW, (u[1] + 1) / c->sbH, (u[2] + 1) / c->sbD};
    int maxSide = imax(ug[0] - lg[0], imax(ug[1] - lg[1], ug[2] - lg[2]));
    if (maxSide > 0) {
        float hx = (float)c->sbW / (float)(c->side - 2);
        float hy = (float)c->sbH / (float)(c->side - 2);
        float hz = (float)c->sbD / (float)(c->side - 2);
        float tv[9];
        for (int q = 0; q < 3; ++q) memcpy(tv + 3 * q, t->v[q].pos, 12);
        float half[3] = {0.5f * hx, 0.5f * hy, 0.5f * hz};
        float ctr[3];
        ctr[2] = ((float)lg[2] + 0.5f) * hz;
        for (int z = lg[2]; z <= ug[2]; ++z, ctr[2] += hz) {
            ctr[1] = ((float)lg[1] + 0.5f) * hy;
            for (int y = lg[1]; y <= ug[1]; ++y, ctr[1] += hy) {
                ctr[0] = ((float)lg[0] + 0.5f) * hx;
                for (int x = lg[0]; x <= ug[0]; ++x, ctr[0] += hx)
                    if (oracle_tribox(ctr, half, tv))
                        emit(c, (uint64_t)x + (uint64_t)c->gW * ((uint64_t)y + (uint64_t)c->gH * (uint64_t)z), ti);
            }
        }
    } else {
        emit(c, (uint64_t)lg[0] + (uint64_t)c->gW * ((uint64_t)lg[1] + (uint64_t)c->gH * (uint64_t)lg[2]), ti);
    }
}

static uint64_t n_cells(const vctx *c) { return (uint64_t)c->gW * c->gH * c->gD; }

static void count_cell(vctx *c, uint64_t cell, uint32_t tri) {
    (void)tri;
    if (cell >= n_cells(c)) { c->err = 1; return; }
    c->blockOffsets[1 + cell]++;
}
static void fill_cell(vctx *c, uint64_t cell, uint32_t tri) {
    if (cell >= n_cells(c)) { c->err = 1; return; }
    c->blockLists[c->blockOffsets[cell]++] = tri;
}

/* PlyLoader::writeTriangleCell, PlyLoader.cpp:296-336 */
static void write_cell(vctx *c, uint32_t *data, int x, int y, int z, float cx, float cy, float cz,
                       const o_tri *t) {
    uint64_t idx = (uint64_t)(x - c->pbX) +
                   (uint64_t)c->pbW * ((uint64_t)(y - c->pbY) + (uint64_t)c->pbH * (uint64_t)(z - c->pbZ));
    const float p[3] = {cx, cy, cz};
    /* Triangle::barycentric, PlyLoader.cpp:55-64 */
    float f1[3], f2[3], f3[3], a[3], b[3], cr[3];
    for (int k = 0; k < 3; ++k) {
        f1[k] = t->v[0].pos[k] - p[k];
        f2[k] = t->v[1].pos[k] - p[k];
        f3[k] = t->v[2].pos[k] - p[k];
        a[k] = t->v[0].pos[k] - t->v[1].pos[k];
        b[k] = t->v[0].pos[k] - t->v[2].pos[k];
    }
    cross3(a, b, cr);
    float area = length3(cr);
    cross3(f2, f3, cr);
    float l1 = length3(cr) / area;
    cross3(f3, f1, cr);
    float l2 = length3(cr) / area;
    int inside = l1 >= 0.0f && l2 >= 0.0f && l1 + l2 <= 1.0f;
    if (!inside) {
        l1 = smin(smax(l1, 0.0f), 1.0f);
        l2 = smin(smax(l2, 0.0f), 1.0f);
        float tau = l1 + l2;
        if (tau > 1.0f) { l1 /= tau; l2 /= tau; }
    }
    float l3 = 1.0f - l1 - l2;

    float nb[3], n[3], col[3];
    for (int k = 0; k < 3; ++k) {
        nb[k] = t->v[0].nrm[k] * l1 + t->v[1].nrm[k] * l2 + t->v[2].nrm[k] * l3;
        col[k] = t->v[0].col[k] * l1 + t->v[1].col[k] * l2 + t->v[2].col[k] * l3;
    }
    normalize3(nb, n);
    float shade = (col[0] * 0.2126f + col[1] * 0.7152f + col[2] * 0.0722f) * (1.0f / 256.0f);

    if (data[idx] == 0) {
        c->counts[idx] = 1;
        data[idx] = oracle_compress_material(n, shade);
    } else {
        float cnt = (float)c->counts[idx];
        float curRatio = cnt / (cnt + 1.0f);
        float newRatio = 1.0f - curRatio;
        float curN[3], curShade;
        oracle_decompress_material(data[idx], curN, &curShade);
        float nn[3];
        for (int k = 0; k < 3; ++k) nn[k] = curN[k] * curRatio + n[k] * newRatio;
        float ns = curShade * curRatio + shade * newRatio;
        if (nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2] < 1e-3f) memcpy(nn, curN, 12);
        data[idx] = oracle_compress_material(nn, ns);
        int nc = (int)c->counts[idx] + 1;
        c->counts[idx] = (uint8_t)(nc < 255 ? nc : 255);
    }
}

/* PlyLoader::triangleToVolume, PlyLoader.cpp:338-374 */
static void tri_to_volume(vctx *c, uint32_t *data, const o_tri *t, int offX, int offY, int offZ) {
    int l[3], u[3];
    point_to_grid(c, t->lower, l);
    point_to_grid(c, t->upper, u);
    l[0] = imax(l[0], c->pbX + offX);
    l[1] = imax(l[1], c->pbY + offY);
    l[2] = imax(l[2], c->pbZ + offZ);
    u[0] = imin(u[0], c->pbX + imin(offX + c->sbW, c->pbW) - 1);
    u[1] = imin(u[1], c->pbY + imin(offY + c->sbH, c->pbH) - 1);
    u[2] = imin(u[2], c->pbZ + imin(offZ + c->sbD, c->pbD) - 1);
    if (l[0] > u[0] || l[1] > u[1] || l[2] > u[2]) return;

    float hx = 1.0f / (float)(c->side - 2);
    float tv[9];
    for (int q = 0; q < 3; ++q) memcpy(tv + 3 * q, t->v[q].pos, 12);
    float half[3] = {0.5f * hx, 0.5f * hx, 0.5f * hx};
    float ctr[3];
    ctr[2] = ((float)l[2] - 0.5f) * hx;
    for (int z = l[2]; z <= u[2]; z++, ctr[2] += hx) {
        ctr[1] = ((float)l[1] - 0.5f) * hx;
        for (int y = l[1]; y <= u[1]; y++, ctr[1] += hx) {
            ctr[0] = ((float)l[0] - 0.5f) * hx;
            for (int x = l[0]; x <= u[0]; x++, ctr[0] += hx)
                if (oracle_tribox(ctr, half, tv)) write_cell(c, data, x, y, z, ctr[0], ctr[1], ctr[2], t);
        }
    }
}

/* PlyLoader::processBlock, PlyLoader.cpp:439-474 (partitions in sequence; each
 * writes a disjoint region, so the order does not change the result). */
static void process_block(vctx *c, uint32_t *data, int x, int y, int z, int w, int h, int d) {
    c->pbX = x; c->pbY = y; c->pbZ = z; c->pbW = w; c->pbH = h; c->pbD = d;
    for (int i = 0; i < c->nPart; ++i) {
        int px = i % c->pW;
        int py = (i / c->pW) % c->pH;
        int pz = i / (c->pW * c->pH);
        uint64_t cell = (uint64_t)(x / c->sbW + px) +
                        (uint64_t)c->gW * ((uint64_t)(y / c->sbH + py) + (uint64_t)c->gH * (uint64_t)(z / c->sbD + pz));
        if (cell >= n_cells(c)) { c->err = 1; return; }
        uint32_t s = c->blockOffsets[cell], e = c->blockOffsets[cell + 1];
        for (uint32_t k = s; k < e; ++k)
            tri_to_volume(c, data, &c->mesh->tris[c->blockLists[k]], px * c->sbW, py * c->sbH, pz * c->sbD);
    }
}

/* PlyLoader::isBlockEmpty, PlyLoader.cpp:476-493 */
static int block_empty(vctx *c, int x, int y, int z) {
    for (int pz = 0; pz < c->pD; ++pz)
        for (int py = 0; py < c->pH; ++py)
            for (int px = 0; px < c->pW; ++px) {
                uint64_t cell = (uint64_t)(x / c->sbW + px) +
                                (uint64_t)c->gW * ((uint64_t)(y / c->sbH + py) + (uint64_t)c->gH * (uint64_t)(z / c->sbD + pz));
                if (cell >= n_cells(c)) { c->err = 1; return 1; }
                if (c->blockOffsets[cell] != c->blockOffsets[cell + 1]) return 0;
            }
    return 1;
}

/* findBestBlockPartition, PlyLoader.cpp:381-405 */
static void best_partition(int *w, int *h, int *d, int threads) {
    int used = 1;
    while (used < threads) {
        int *mx, *md, *mn;
        if (*w > *h && *w > *d) mx = w; else if (*h > *d) mx = h; else mx = d;
        {
            int big = imax(*w, imax(*h, *d)), small = imin(*w, imin(*h, *d));
            if (*w != small && *w != big) md = w;
            else if (*h != small && *h != big) md = h;
            else md = d;
        }
        if (*w < *h && *w < *d) mn = w; else if (*h < *d) mn = h; else mn = d;
        if ((*mx % 2) == 0) *mx /= 2;
        else if ((*md % 2) == 0) *md /= 2;
        else if ((*mn % 2) == 0) *mn /= 2;
        else break;
        used *= 2;
    }
}

/* PlyLoader::setupBlockProcessing + buildBlockLists, PlyLoader.cpp:272-294,407-437 */
static int setup_blocks_dims(vctx *c, int resolution, int threads, int bw, int bh, int bd) {
    c->counts = (uint8_t *)calloc((size_t)bw * bh * bd, 1);
    c->side = resolution - 2;
    c->blkW = c->sbW = bw;
    c->blkH = c->sbH = bh;
    c->blkD = c->sbD = bd;
    best_partition(&c->sbW, &c->sbH, &c->sbD, threads);
    c->pW = c->blkW / c->sbW;
    c->pH = c->blkH / c->sbH;
    c->pD = c->blkD / c->sbD;
    c->nPart = c->pW * c->pH * c->pD;
    c->gW = c->pW * (c->dataW + c->blkW - 1) / c->blkW;
    c->gH = c->pH * (c->dataH + c->blkH - 1) / c->blkH;
    c->gD = c->pD * (c->dataD + c->blkD - 1) / c->blkD;
    uint64_t nc = n_cells(c);
    c->blockOffsets = (uint32_t *)calloc(nc + 1, 4);
    for (int64_t i = 0; i < c->mesh->ntris; ++i) for_binned_cells(c, &c->mesh->tris[i], (uint32_t)i, count_cell);
    for (uint64_t i = 1; i <= nc; ++i) c->blockOffsets[i] += c->blockOffsets[i - 1];
    c->blockLists = (uint32_t *)malloc(sizeof(uint32_t) * (c->blockOffsets[nc] ? c->blockOffsets[nc] : 1));
    for (int64_t i = 0; i < c->mesh->ntris; ++i) for_binned_cells(c, &c->mesh->tris[i], (uint32_t)i, fill_cell);
    for (uint64_t i = nc; i >= 1; --i) c->blockOffsets[i] = c->blockOffsets[i - 1];
    c->blockOffsets[0] = 0;
    return c->err ? -1 : 0;
}

/* VoxelData's mesh source: cubic cache blocks of edge `cache` */
static int setup_blocks(vctx *c, int resolution, int threads) {
    int bs = (int)c->cache;
    return setup_blocks_dims(c, resolution, threads, bs, bs, bs);
}

/* VoxelData::init, src/VoxelData.cpp:211-262 */
static int vd_init(vctx *c, uint64_t mem) {
    c->vW = round_pow2(c->dataW);
    c->vH = round_pow2(c->dataH);
    c->vD = round_pow2(c->dataD);
    c->highBit = highest_bit((uint32_t)imax(c->vW, imax(c->vH, c->vD)));
    uint64_t cellCost = 4 + (c->mesh ? 1 : 0);
    int best = -1;
    for (int i = c->highBit; i >= 0; --i) {
        int lowBit = c->highBit - i;
        uint64_t cost = grid_cells(i + 1) + grid_cells(lowBit) + ((uint64_t)1 << (uint64_t)(lowBit * 3)) * cellCost;
        if (cost < mem) best = c->highBit - i;
    }
    if (best < 0) return -2;
    c->lowLevels = best;
    c->topLevels = c->highBit - best + 1;
    c->cache = (int64_t)1 << best;
    c->topLut = (uint8_t *)calloc(grid_cells(c->topLevels) + 1, 1);
    c->lowLut = (uint8_t *)calloc(grid_cells(c->lowLevels) + 1, 1);
    c->buf = (uint32_t *)calloc((size_t)1 << (best * 3), 4);
    c->topTab = (uint8_t **)calloc((size_t)c->topLevels + 1, sizeof(uint8_t *));
    c->lowTab = (uint8_t **)calloc((size_t)c->lowLevels + 1, sizeof(uint8_t *));
    uint64_t off = 0, s = 1;
    for (int i = 0; i < c->topLevels; ++i) { c->topTab[i] = c->topLut + off; off += s; s *= 8; }
    off = 0; s = 1;
    for (int i = 0; i < c->lowLevels; ++i) { c->lowTab[i] = c->lowLut + off; off += s; s *= 8; }
    c->bX = c->bY = c->bZ = c->bW = c->bH = c->bD = 0;
    return 0;
}

/* VoxelData::upsampleLutLevel, VoxelData.cpp:87-118 */
static void upsample(uint8_t **tab, int l) {
    int size = 2 << l;
    for (int z = 0; z < size; z += 2)
        for (int y = 0; y < size; y += 2)
            for (int x = 0; x < size; x += 2) {
                uint8_t *n = tab[l + 1];
                int v = n[lut_idx(l + 1, x, y, z)] | n[lut_idx(l + 1, x + 1, y, z)] |
                        n[lut_idx(l + 1, x, y + 1, z)] | n[lut_idx(l + 1, x + 1, y + 1, z)] |
                        n[lut_idx(l + 1, x, y, z + 1)] | n[lut_idx(l + 1, x + 1, y, z + 1)] |
                        n[lut_idx(l + 1, x, y + 1, z + 1)] | n[lut_idx(l + 1, x + 1, y + 1, z + 1)];
                tab[l][lut_idx(l, x >> 1, y >> 1, z >> 1)] = (v != 0);
            }
}

/* VoxelData::cacheData, VoxelData.cpp:177-209 */
static void cache_data(vctx *c, int x, int y, int z, int w, int h, int d) {
    if (c->mesh) { process_block(c, c->buf, x, y, z, w, h, d); return; }
    for (int vz = 0; vz < d; ++vz)
        for (int vy = 0; vy < h; ++vy)
            memcpy(c->buf + ((size_t)vy + (size_t)vz * h) * w,
                   c->grid + (size_t)x + (size_t)c->dataW * ((size_t)(y + vy) + (size_t)c->dataH * (size_t)(z + vz)),
                   (size_t)w * 4);
}

/* VoxelData::buildTopLutBlock / buildTopLut, VoxelData.cpp:120-158 */
static void build_top_lut(vctx *c) {
    int size = (int)c->cache;
    for (int z = 0; z < c->vD; z += size)
        for (int y = 0; y < c->vH; y += size)
            for (int x = 0; x < c->vW; x += size) {
                int w = imin(size, c->dataW - x), h = imin(size, c->dataH - y), d = imin(size, c->dataD - z);
                if (w <= 0 || h <= 0 || d <= 0) continue;
                int empty;
                if (c->mesh) {
                    empty = block_empty(c, x, y, z);
                } else {
                    cache_data(c, x, y, z, w, h, d);
                    empty = 1;
                    for (size_t i = 0; i < (size_t)w * h * d; ++i) if (c->buf[i]) { empty = 0; break; }
                }
                if (!empty)
                    c->topTab[c->topLevels - 1][lut_idx(c->topLevels - 1, x >> c->lowLevels, y >> c->lowLevels, z >> c->lowLevels)] = 1;
            }
    for (int i = c->topLevels - 2; i >= 0; --i) upsample(c->topTab, i);
}

/* VoxelData::buildLowLut, VoxelData.cpp:160-175: z covers 2*floor(bD/2) slices */
static void build_low_lut(vctx *c) {
    int zEnd = (c->bD / 2) * 2;
    for (int z = 0; z < zEnd; ++z)
        for (int y = 0; y < c->bH; ++y)
            for (int x = 0; x < c->bW; ++x)
                if (c->buf[(size_t)x + (size_t)c->bW * ((size_t)y + (size_t)c->bH * (size_t)z)])
                    c->lowTab[c->lowLevels - 1][lut_idx(c->lowLevels - 1, x / 2, y / 2, z / 2)] = 1;
    for (int i = c->lowLevels - 2; i >= 0; --i) upsample(c->lowTab, i);
}

/* VoxelData::prepareDataAccess, VoxelData.cpp:264-291 */
static void prepare_access(vctx *c, int x, int y, int z, int size) {
    int w = imin(size, c->dataW - x), h = imin(size, c->dataH - y), d = imin(size, c->dataD - z);
    if (w <= 0 || h <= 0 || d <= 0) return;
    if (x >= c->bX && y >= c->bY && z >= c->bZ && x + w <= c->bX + c->bW && y + h <= c->bY + c->bH &&
        z + d <= c->bZ + c->bD)
        return;
    if ((int64_t)size <= c->cache) {
        c->bX = x; c->bY = y; c->bZ = z; c->bW = w; c->bH = h; c->bD = d;
        cache_data(c, x, y, z, w, h, d);
        build_low_lut(c);
    }
}

/* VoxelData::getVoxel / getVoxelDestructive, VoxelData.hpp:106-121 */
static uint32_t get_voxel(vctx *c, int x, int y, int z, int destructive) {
    if (x >= c->dataW || y >= c->dataH || z >= c->dataD) return 0;
    size_t idx = (size_t)(x - c->bX) + (size_t)c->bW * ((size_t)(y - c->bY) + (size_t)c->bH * (size_t)(z - c->bZ));
    uint32_t v = c->buf[idx];
    if (destructive && v) c->buf[idx] = 0;
    return v;
}

/* VoxelData::cubeContainsVoxelsDestructive, VoxelData.hpp:123-138 */
static int cube_contains(vctx *c, int x, int y, int z, int size) {
    if (x >= c->dataW || y >= c->dataH || z >= c->dataD) return 0;
    int bit = highest_bit((uint32_t)size);
    if (size == 1) return get_voxel(c, x, y, z, 0) != 0;
    if (bit < c->lowLevels) {
        int l = c->lowLevels - bit;
        uint8_t *cell = &c->lowTab[l][lut_idx(l, (x - c->bX) >> bit, (y - c->bY) >> bit, (z - c->bZ) >> bit)];
        uint8_t v = *cell;
        if (v) *cell = 0;
        return v != 0;
    }
    int l = c->highBit - bit;
    return c->topTab[l][lut_idx(l, x >> bit, y >> bit, z >> bit)] != 0;
}

/* ChunkedAllocator pushBack / operator[] / insert, src/ChunkedAllocator.hpp:64-74 */
static void push(vctx *c, uint32_t w) {
    if (c->size == c->cap) {
        c->cap = c->cap ? c->cap * 2 : 4096;
        c->words = (uint32_t *)realloc(c->words, c->cap * 4);
    }
    c->words[c->size++] = w;
}
static void insert_word(vctx *c, uint64_t idx, uint32_t w) {
    if (c->nIns == c->insCap) {
        c->insCap = c->insCap ? c->insCap * 2 : 1024;
        c->insIdx = (uint64_t *)realloc(c->insIdx, c->insCap * 8);
        c->insVal = (uint32_t *)realloc(c->insVal, c->insCap * 4);
    }
    c->insIdx[c->nIns] = idx;
    c->insVal[c->nIns] = w;
    c->nIns++;
}

/* VoxelOctree::buildOctree, src/VoxelOctree.cpp:199-265. Child i of the
 * reference sits at +half on x when i is even; children are visited in
 * descending i, i.e. ascending Morton order m = 7 - i. */
static uint64_t build_node(vctx *c, int x, int y, int z, int size, uint64_t desc) {
    prepare_access(c, x, y, z, size);
    int half = size >> 1;
    int px[8] = {x + half, x, x + half, x, x + half, x, x + half, x};
    int py[8] = {y + half, y + half, y, y, y + half, y + half, y, y};
    int pz[8] = {z + half, z + half, z + half, z + half, z, z, z, z};

    uint64_t childOffset = c->size - desc;
    int childCount = 0, childIdx[8];
    uint32_t childMask = 0;
    for (int i = 0; i < 8; ++i)
        if (cube_contains(c, px[i], py[i], pz[i], half)) {
            childMask |= 128u >> i;
            childIdx[childCount++] = i;
        }

    int large = 0;
    uint32_t leafMask;
    if (half == 1) {
        leafMask = 0;
        for (int i = 0; i < childCount; ++i) {
            int k = childIdx[childCount - i - 1];
            push(c, get_voxel(c, px[k], py[k], pz[k], 1));
            c->leaves++;
        }
    } else {
        leafMask = childMask;
        for (int i = 0; i < childCount; ++i) push(c, 0);
        uint64_t gco[8], delta = 0, insBefore = c->nIns;
        for (int i = 0; i < childCount; ++i) {
            int k = childIdx[childCount - i - 1];
            gco[i] = delta + build_node(c, px[k], py[k], pz[k], half, desc + childOffset + (uint64_t)i);
            delta += c->nIns - insBefore;
            insBefore = c->nIns;
            if (gco[i] > 0x3FFF) large = 1;
        }
        for (int i = 0; i < childCount; ++i) {
            uint64_t ci = desc + childOffset + (uint64_t)i;
            uint64_t off = gco[i];
            if (large) {
                off += (uint64_t)(childCount - i);
                insert_word(c, ci + 1, (uint32_t)off);
                c->words[ci] |= 0x20000u;
                off >>= 32;
            }
            c->words[ci] |= (uint32_t)(off << 18);
        }
    }
    c->words[desc] = (childMask << 8) | leafMask;
    if (large) c->words[desc] |= 0x10000u;
    return childOffset;
}

static int cmp_u64_pair(const void *a, const void *b) {
    uint64_t x = ((const uint64_t *)a)[0], y = ((const uint64_t *)b)[0];
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* ChunkedAllocator::finalize, src/ChunkedAllocator.hpp:76-114: insertion at
 * index i lands immediately before original element i. */
static uint32_t *finalize(vctx *c, uint64_t *n) {
    uint64_t *pairs = (uint64_t *)malloc(16 * (c->nIns ? c->nIns : 1));
    for (uint64_t i = 0; i < c->nIns; ++i) { pairs[2 * i] = c->insIdx[i]; pairs[2 * i + 1] = c->insVal[i]; }
    qsort(pairs, c->nIns, 16, cmp_u64_pair);
    uint64_t total = c->size + c->nIns;
    uint32_t *out = (uint32_t *)malloc(4 * (total ? total : 1));
    uint64_t o = 0, k = 0;
    for (uint64_t i = 0; i < c->size; ++i) {
        while (k < c->nIns && pairs[2 * k] == i) out[o++] = (uint32_t)pairs[2 * k++ + 1];
        out[o++] = c->words[i];
    }
    while (k < c->nIns && pairs[2 * k] == c->size) out[o++] = (uint32_t)pairs[2 * k++ + 1];
    free(pairs);
    *n = o;
    return out;
}

static void vctx_free(vctx *c) {
    free(c->topLut); free(c->lowLut); free(c->topTab); free(c->lowTab); free(c->buf);
    free(c->blockOffsets); free(c->blockLists); free(c->counts);
    free(c->words); free(c->insIdx); free(c->insVal);
}

/* VoxelOctree(VoxelData*), src/VoxelOctree.cpp:125-137, plus getCenter
 * (VoxelData.cpp:297-303). */
static int emit_tree(vctx *c, uint32_t **words, uint64_t *nwords, float center[3]) {
    int side = imax(c->vW, imax(c->vH, c->vD));
    push(c, 0);
    build_node(c, 0, 0, 0, side, 0);
    c->words[0] |= 1u << 18;
    if (c->err) return -3;
    *words = finalize(c, nwords);
    center[0] = (float)c->dataW * 0.5f / (float)side;
    center[1] = (float)c->dataH * 0.5f / (float)side;
    center[2] = (float)c->dataD * 0.5f / (float)side;
    return 0;
}

static void fill_info(const vctx *c, oracle_build_info *info) {
    if (!info) return;
    memset(info, 0, sizeof(*info));
    info->data_dims[0] = c->dataW; info->data_dims[1] = c->dataH; info->data_dims[2] = c->dataD;
    info->side_length = imax(c->vW, imax(c->vH, c->vD));
    info->cache_size = (int)c->cache;
    info->sub_block[0] = c->sbW; info->sub_block[1] = c->sbH; info->sub_block[2] = c->sbD;
    info->grid[0] = c->gW; info->grid[1] = c->gH; info->grid[2] = c->gD;
    if (c->blockOffsets) info->bin_entries = c->blockOffsets[n_cells(c)];
    info->leaves = c->leaves;
}

/* PlyLoader::suggestedDimensions, PlyLoader.cpp:499-504 */
static void suggested_dims(const oracle_mesh *m, int side, int dims[3]) {
    for (int k = 0; k < 3; ++k) {
        float s = (m->upper[k] - m->lower[k]) * (float)(side - 2);
        dims[k] = cvt_i32(s) + 2;
    }
}

int oracle_build_mesh(oracle_mesh *m, int resolution, int threads, uint64_t mem,
                      uint32_t **words, uint64_t *nwords, float center[3],
                      oracle_build_info *info) {
    vctx c;
    memset(&c, 0, sizeof(c));
    c.mesh = m;
    int dims[3];
    suggested_dims(m, resolution, dims);
    c.dataW = dims[0]; c.dataH = dims[1]; c.dataD = dims[2];
    int r = vd_init(&c, mem);
    if (r == 0) r = setup_blocks(&c, resolution, threads);
    if (r == 0) { build_top_lut(&c); r = c.err ? -3 : 0; }
    if (r == 0) r = emit_tree(&c, words, nwords, center);
    fill_info(&c, info);
    vctx_free(&c);
    return r;
}

int oracle_build_grid(const uint32_t *grid, int w, int h, int d, uint64_t mem, uint32_t **words,
                      uint64_t *nwords, float center[3], oracle_build_info *info) {
    vctx c;
    /* a one-voxel volume (octree side 1): the reference builder crashes on it */
    if (w < 2 && h < 2 && d < 2) return -2;
    memset(&c, 0, sizeof(c));
    c.grid = grid;
    c.dataW = w; c.dataH = h; c.dataD = d;
    int r = vd_init(&c, mem);
    if (r == 0) { build_top_lut(&c); r = emit_tree(&c, words, nwords, center); }
    if (info) {
        fill_info(&c, info);
        uint64_t present = 0;
        for (size_t i = 0; i < (size_t)w * h * d; ++i) present += grid[i] != 0;
        info->hazard_voxels = present - c.leaves;
    }
    vctx_free(&c);
    return r;
}

int oracle_voxelize_mesh(oracle_mesh *m, int resolution, int threads, uint64_t mem,
                         uint32_t *out, int dims[3]) {
    vctx c;
    memset(&c, 0, sizeof(c));
    c.mesh = m;
    suggested_dims(m, resolution, dims);
    c.dataW = dims[0]; c.dataH = dims[1]; c.dataD = dims[2];
    int r = vd_init(&c, mem);
    if (r == 0) r = setup_blocks(&c, resolution, threads);
    if (r == 0) {
        int size = (int)c.cache;
        for (int z = 0; z < c.dataD; z += size)
            for (int y = 0; y < c.dataH; y += size)
                for (int x = 0; x < c.dataW; x += size) {
                    int w = imin(size, c.dataW - x), h = imin(size, c.dataH - y), d = imin(size, c.dataD - z);
                    memset(c.buf, 0, (size_t)size * size * size * 4);
                    process_block(&c, c.buf, x, y, z, w, h, d);
                    for (int vz = 0; vz < d; ++vz)
                        for (int vy = 0; vy < h; ++vy)
                            memcpy(out + (size_t)x + (size_t)dims[0] * ((size_t)(y + vy) + (size_t)dims[1] * (size_t)(z + vz)),
                                   c.buf + ((size_t)vy + (size_t)vz * h) * w, (size_t)w * 4);
                }
        r = c.err ? -3 : 0;
    }
    vctx_free(&c);
    return r;
}

/* PlyLoader::convertToVolume, PlyLoader.cpp:506-534 */
int oracle_convert_to_volume(oracle_mesh *m, int maxSize, int threads, uint64_t mem, uint32_t **out,
                             int dims[3], int *sliceZ) {
    vctx c;
    memset(&c, 0, sizeof(c));
    c.mesh = m;
    *out = NULL;
    suggested_dims(m, maxSize, dims);
    const int w = dims[0], h = dims[1], d = dims[2];
    c.dataW = w; c.dataH = h; c.dataD = d;
    /* blockMemRequirement(w, h, 1): one count byte per voxel of a slice */
    const uint64_t slices = mem / ((uint64_t)w * (uint64_t)h);
    const int sz = slices < (uint64_t)d ? (int)slices : d;
    *sliceZ = sz;
    if (sz == 0) return -2; /* the reference asserts "Not enough memory available for single slice" */
    const size_t slab = (size_t)w * h * sz, plane = (size_t)w * h;
    int r = setup_blocks_dims(&c, maxSize, threads, w, h, sz);
    /* the slab buffer starts zeroed (a fresh allocation of that size is) and is
     * reused for every slab without clearing, as are the counts: cells a slab
     * does not write keep the previous slab's words, cells it writes average
     * with them */
    uint32_t *data = r == 0 ? (uint32_t *)calloc(slab, 4) : NULL;
    uint32_t *vol = r == 0 ? (uint32_t *)malloc(plane * (size_t)d * 4) : NULL;
    if (r == 0 && (!data || !vol)) r = -4;
    for (int z = 0; r == 0 && z < d; z += sz) {
        process_block(&c, data, 0, 0, z, w, h, sz);
        if (c.err) { r = -3; break; }
        memcpy(vol + plane * (size_t)z, data, plane * (size_t)imin(sz, d - z) * 4);
    }
    free(data);
    vctx_free(&c);
    if (r) { free(vol); return r; }
    *out = vol;
    return 0;
}

void oracle_free(void *p) { free(p); }
