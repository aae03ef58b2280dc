// Octree raymarch on gfx950.
//
// Traversal restates VoxelOctree::raymarch (src/VoxelOctree.cpp:207-346) per
// lane, in FP32 with contraction off and correctly rounded division, so
// results are bit-identical to the reference on the same rays. Design:
//  - one lane per ray; a wave64 owns an 8x8 pixel tile (render path) so its
//    rays walk the same nodes and its node loads coalesce through L2;
//  - the restart stack {parent, maxT} per scale lives in LDS, laid out
//    [slot][lane] so a wave's pushes/pops hit 64 different banks;
//  - node words are fetched with plain 32-bit global loads (the tree is
//    read-only, so they stay cached in L2 / MALL across rays);
//  - word indices are 32-bit when the tree has < 2^32 words, else 64-bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <mutex>

#include "march.hpp"
#include "tile_order.hpp"
#include "svo_internal.hpp"

namespace svo {

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(SVO_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace march;

// Runs one ray to completion. Returns 0 miss, 1 leaf hit, 2 LOD hit. With
// Normal set, a leaf hit also gives the leaf's word index and word.
// With Count, *passes counts the loop passes.
template <bool Lod, bool Finite, typename Idx, bool Normal = true, int Reg = 1, bool Count = false>
__device__ __forceinline__ int march_ray(const uint32_t *__restrict__ tree, float ox, float oy, float oz, float dx,
                                         float dy, float dz, float rayScale, const Stack<Idx> &stk, uint32_t &normal,
                                         float &tOut, Idx &leaf, uint32_t *passes = nullptr) {
    Ray<Idx> r;
    ray_init(r, tree, ox, oy, oz, dx, dy, dz, rayScale);
    if (!Finite && !ray_defined(r)) return kMiss;  // the reference never ends such a ray
    int code;
    do {
        code = ray_step<Lod, Finite, Idx, Normal, Reg>(r, tree, stk, normal, tOut, leaf);
        if (Count) ++*passes;
    } while (code == kRunning);
    return code;
}

template <typename Idx>
__device__ __forceinline__ void store_result(int code, uint32_t normal, float t, Idx leaf, int64_t i,
                                             uint8_t *hit, uint32_t *nrm, float *tt, uint64_t *lf) {
    if (hit) hit[i] = uint8_t(code);
    if (nrm) nrm[i] = code == 1 ? normal : 0u;
    if (tt) tt[i] = code == 0 ? 0.0f : t;
    if (lf) lf[i] = code == 1 ? uint64_t(leaf) : SVO_LEAF_NONE;
}

template <typename Idx>
__global__ void __launch_bounds__(256) rays_kernel(const uint32_t *__restrict__ tree, int64_t n,
                                                   const float *__restrict__ o, const float *__restrict__ d,
                                                   const float *__restrict__ rs, float rs0, int slots,
                                                   uint8_t *hit, uint32_t *nrm, float *tt, uint64_t *lf) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Stack<Idx> stk(lds, int(blockDim.x), slots, int(threadIdx.x));
    uint32_t normal = 0;
    float t = 0.0f;
    Idx leaf = 0;
    const float scale = rs ? rs[i] : rs0;
    // arbitrary caller rays: LOD test kept, exact std::min semantics
    int code = march_ray<true, false, Idx>(tree, o[3 * i], o[3 * i + 1], o[3 * i + 2], d[3 * i], d[3 * i + 1],
                                           d[3 * i + 2], scale, stk, normal, t, leaf);
    store_result<Idx>(code, normal, t, leaf, i, hit, nrm, tt, lf);
}

struct PrimaryParams {
    float a11, a12, a21, a22, a31, a32;  // tform rotation columns used by ray generation
    float zx, zy, zz;                    // planeDist * third column
    float px, py, pz;                    // origin
    float aspect, scale;                 // H/W and 2/W
    int width, height, tilesX;
};

// ray generation: src/Main.cpp:97-113 with closed-form dx, dy
__device__ __forceinline__ void primary_dir(const PrimaryParams &p, int x, int y, float d[3]) {
    const float dy = p.aspect - float(y) * p.scale;
    const float dx = -1.0f + float(x) * p.scale;
    const float vx = dx * p.a11 + dy * p.a12 + p.zx;
    const float vy = dx * p.a21 + dy * p.a22 + p.zy;
    const float vz = dx * p.a31 + dy * p.a32 + p.zz;
    const float s = quake_rsqrt(vx * vx + vy * vy + vz * vz);
    d[0] = vx * s;
    d[1] = vy * s;
    d[2] = vz * s;
}

// 8x8 pixel tile per wave64; lanes raster the tile row by row.
template <typename Idx, int Reg>
__global__ void __launch_bounds__(64) primary_kernel(const uint32_t *__restrict__ tree, PrimaryParams p,
                                                     int slots, uint8_t *hit, uint32_t *nrm, float *tt,
                                                     uint64_t *lf) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tile = int(blockIdx.x);
    const int lane = threadIdx.x;
    const int x = (tile % p.tilesX) * 8 + (lane & 7);
    const int y = (tile / p.tilesX) * 8 + (lane >> 3);
    if (x >= p.width || y >= p.height) return;
    float d[3];
    primary_dir(p, x, y, d);
    Stack<Idx> stk(lds, 64, slots, lane);
    uint32_t normal = 0;
    float t = 0.0f;
    Idx leaf = 0;
    int code = march_ray<false, true, Idx, true, Reg>(tree, p.px, p.py, p.pz, d[0], d[1], d[2], 0.0f, stk, normal, t,
                                                      leaf);
    store_result<Idx>(code, normal, t, leaf, int64_t(y) * p.width + x, hit, nrm, tt, lf);
}

// Cost-ordered form of primary_kernel (tile_order.hpp): workgroup b traces the
// tile the plan's order gives and files it by cost for the next launch.
template <typename Idx, int Reg>
__global__ void __launch_bounds__(64) primary_ordered_kernel(const uint32_t *__restrict__ tree, PrimaryParams p,
                                                             int slots, TileOrder o, uint8_t *hit, uint32_t *nrm,
                                                             float *tt, uint64_t *lf) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const int tile = ordered_tile(o, int(blockIdx.x), lane);
    const int x = (tile % p.tilesX) * 8 + (lane & 7);
    const int y = (tile / p.tilesX) * 8 + (lane >> 3);
    uint32_t passes = 0;
    if (x < p.width && y < p.height) {
        float d[3];
        primary_dir(p, x, y, d);
        Stack<Idx> stk(lds, 64, slots, lane);
        Ray<Idx> r;
        ray_init(r, tree, p.px, p.py, p.pz, d[0], d[1], d[2], 0.0f);
        uint32_t normal = 0;
        float t = 0.0f;
        Idx leaf = 0;
        int code;
        do {
            code = ray_step<false, true, Idx, true, Reg>(r, tree, stk, normal, t, leaf);
            passes++;
        } while (code == kRunning);
        store_result<Idx>(code, normal, t, leaf, int64_t(y) * p.width + x, hit, nrm, tt, lf);
    }
    file_tile(o, int(blockIdx.x), tile, passes, lane);
}

// Primary rays and hard shadows of one 8x8 tile in one wave: every lane's
// primary ray (as primary_kernel), then one shadow ray per primary leaf hit,
// from o + d*(t - eps) (the hit leaf's entry point pulled back along the
// primary ray) toward the light, with no LOD cut-off; shadow = 1 when it hits
// any voxel, 0 when it escapes or the pixel had no primary hit. The two loops
// run one after the other (the wave finishes every primary ray before any
// shadow ray starts, so lanes never mix the two), and an image pays one launch
// tail instead of two (8192^3 terrain at 3840x2160: 1.644 ms against 1.779 ms
// for a primary launch then a shadow launch, each split in two halves over two
// streams).
//
// Planned: the tiles in a tile plan's cost order (tile_order.hpp), each filed
// by its primary plus shadow passes for the plan's next launch.
template <typename Idx, int Reg, bool Planned = false>
__global__ void __launch_bounds__(64) primary_shadow_kernel(const uint32_t *__restrict__ tree, PrimaryParams p,
                                                            float lx, float ly, float lz, float eps, int slots,
                                                            uint8_t *hit, uint32_t *nrm, float *tt, uint64_t *lf,
                                                            uint8_t *__restrict__ shadow, TileOrder o) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const int tile = Planned ? ordered_tile(o, int(blockIdx.x), lane) : int(blockIdx.x);
    const int x = (tile % p.tilesX) * 8 + (lane & 7);
    const int y = (tile / p.tilesX) * 8 + (lane >> 3);
    uint32_t passes = 0;
    if (x < p.width && y < p.height) {
        float d[3];
        primary_dir(p, x, y, d);
        Stack<Idx> stk(lds, 64, slots, lane);
        uint32_t normal = 0;
        float t = 0.0f;
        Idx leaf = 0;
        const int code = march_ray<false, true, Idx, true, Reg, Planned>(tree, p.px, p.py, p.pz, d[0], d[1], d[2], 0.0f,
                                                                         stk, normal, t, leaf, &passes);
        const int64_t i = int64_t(y) * p.width + x;
        store_result<Idx>(code, normal, t, leaf, i, hit, nrm, tt, lf);
        uint8_t sh = 0;
        if (code == 1) {
            const float back = t - eps;
            uint32_t n2 = 0;
            float t2 = 0.0f;
            Idx leaf2 = 0;
            const int c2 = march_ray<false, true, Idx, false, Reg, Planned>(tree, p.px + d[0] * back, p.py + d[1] * back,
                                                                            p.pz + d[2] * back, lx, ly, lz, 0.0f, stk,
                                                                            n2, t2, leaf2, &passes);
            sh = (c2 == 1 || c2 == 2) ? 1 : 0;
        }
        shadow[i] = sh;
    }
    if (Planned) file_tile(o, int(blockIdx.x), tile, passes, lane);
}

int stack_slots(const svo_octree &h, int reg) { return lds_slots(h.depth, reg); }

template <typename Idx>
int walk_one(const svo_octree &h, const float o[3], const float d[3], float rs, uint32_t &normal, float &t,
             uint64_t &leaf) {
    const uint32_t *tree = h.words.data();
    uint64_t buf[48];  // the restart stack: at most 21 entries of (index, maxT)
    const Stack<Idx> stk(buf, 1, stack_slots(h, 1), 0);
    Ray<Idx> r;
    ray_init(r, tree, o[0], o[1], o[2], d[0], d[1], d[2], rs);
    if (!ray_defined(r)) return kMiss;  // the reference never ends such a ray
    Idx lf = 0;
    int code;
    do {
        code = ray_step<true, false, Idx, true, 1>(r, tree, stk, normal, t, lf);
    } while (code == kRunning);
    leaf = uint64_t(lf);
    return code;
}

}  // namespace

// One ray walked on the calling host thread by the same traversal source as
// the kernels (march.hpp): hit code 0/1/2, normal and leaf on a leaf hit, t on
// any hit.
int raymarch_one(const svo_octree &h, const float o[3], const float d[3], float rs, uint32_t &normal, float &t,
                 uint64_t &leaf) {
    if (h.words.size() < (uint64_t(1) << 32)) return walk_one<uint32_t>(h, o, d, rs, normal, t, leaf);
    return walk_one<uint64_t>(h, o, d, rs, normal, t, leaf);
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ensure_device(svo_octree &h) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(h.upload);
    if (h.dwords && h.device == dev) return SVO_OK;
    release_device(h);
    size_t bytes = h.words.size() * sizeof(uint32_t);
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(SVO_ERR_NOMEM, std::string("hipMalloc octree: ") + hipGetErrorString(e));
    e = hipMemcpy(p, h.words.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return fail(SVO_ERR_DEVICE, std::string("upload octree: ") + hipGetErrorString(e));
    }
    h.dwords = static_cast<uint32_t *>(p);
    h.device = dev;
    return SVO_OK;
}

void release_device(svo_octree &h) {
    if (h.dwords) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != h.device) (void)hipSetDevice(h.device);
        (void)hipFree(h.dwords);
        if (cur != h.device) (void)hipSetDevice(cur);
    }
    h.dwords = nullptr;
    h.device = -1;
}

int raymarch_device(svo_octree &h, int64_t n, const float *o, const float *d, const float *rs, float rs0,
                    uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, void *stream) {
    if (n < 0) return fail(SVO_ERR_INVALID, "negative ray count");
    if (n == 0) return SVO_OK;
    if (!o || !d) return fail(SVO_ERR_INVALID, "null ray arrays");
    int r = ensure_device(h);
    if (r) return r;
    const int threads = 256;
    const int slots = stack_slots(h, 1);
    const bool wide = h.words.size() >= (uint64_t(1) << 32);
    const size_t lds = size_t(slots) * threads * ((wide ? 8 : 4) + sizeof(float));
    const int64_t blocks = (n + threads - 1) / threads;
    if (blocks > 0x7FFFFFFF) return fail(SVO_ERR_LIMIT, "too many rays for one launch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!wide)
        hipLaunchKernelGGL(rays_kernel<uint32_t>, dim3(unsigned(blocks)), dim3(threads), lds, s, h.dwords, n, o, d,
                           rs, rs0, slots, hit, normal, t, leaf);
    else
        hipLaunchKernelGGL(rays_kernel<uint64_t>, dim3(unsigned(blocks)), dim3(threads), lds, s, h.dwords, n, o, d,
                           rs, rs0, slots, hit, normal, t, leaf);
    HIP_TRY(hipGetLastError());
    return SVO_OK;
}

namespace {

PrimaryParams primary_params(const svo_camera &cam, int width, int height) {
    PrimaryParams p;
    const float pi = float(3.14159265358979323846);
    const float planeDist = 1.0f / std::tan(pi / 6.0f);  // src/Main.cpp:158
    p.a11 = cam.tform[0]; p.a12 = cam.tform[1];
    p.a21 = cam.tform[4]; p.a22 = cam.tform[5];
    p.a31 = cam.tform[8]; p.a32 = cam.tform[9];
    p.zx = planeDist * cam.tform[2];
    p.zy = planeDist * cam.tform[6];
    p.zz = planeDist * cam.tform[10];
    p.px = cam.pos[0]; p.py = cam.pos[1]; p.pz = cam.pos[2];
    p.aspect = float(height) / float(width);
    p.scale = 2.0f / float(width);
    p.width = width;
    p.height = height;
    p.tilesX = (width + 7) / 8;
    return p;
}

}  // namespace

namespace {

// Launch shape shared by the primary and shadow passes of one octree.
struct TileLaunch {
    int reg, slots;
    bool wide;
    size_t lds;
};

TileLaunch tile_launch(const svo_octree &h) {
    TileLaunch L;
    L.reg = reg_entries(h.depth);
    L.slots = stack_slots(h, L.reg);
    L.wide = h.words.size() >= (uint64_t(1) << 32);
    L.lds = size_t(L.slots) * 64 * ((L.wide ? 8 : 4) + sizeof(float));
    return L;
}

// The primary pass over every tile of the image, one wave per 8x8 tile.
void launch_primary(const svo_octree &h, const TileLaunch &L, const PrimaryParams &p, int tiles, uint8_t *hit,
                    uint32_t *normal, float *t, uint64_t *leaf, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>(tiles)), block(64);
    if (!L.wide && L.reg == 1)
        hipLaunchKernelGGL((primary_kernel<uint32_t, 1>), grid, block, L.lds, s, h.dwords, p, L.slots, hit, normal, t, leaf);
    else if (!L.wide)
        hipLaunchKernelGGL((primary_kernel<uint32_t, 3>), grid, block, L.lds, s, h.dwords, p, L.slots, hit, normal, t, leaf);
    else if (L.reg == 1)
        hipLaunchKernelGGL((primary_kernel<uint64_t, 1>), grid, block, L.lds, s, h.dwords, p, L.slots, hit, normal, t, leaf);
    else
        hipLaunchKernelGGL((primary_kernel<uint64_t, 3>), grid, block, L.lds, s, h.dwords, p, L.slots, hit, normal, t, leaf);
}

}  // namespace

// Tile plans (tile_order.hpp): the cost order of one image size or one batch,
// kept on the device between launches.
int tile_plan_alloc(int64_t tiles, int width, int height, svo_tile_plan **out) {
    if (!out) return fail(SVO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (tiles <= 0) return fail(SVO_ERR_INVALID, "a tile plan needs at least one tile");
    if (tiles > 0x7FFFFFFF) return fail(SVO_ERR_LIMIT, "too many tiles for one launch");
    svo_tile_plan *p = new (std::nothrow) svo_tile_plan();
    if (!p) return fail(SVO_ERR_NOMEM, "out of host memory");
    p->width = width;
    p->height = height;
    p->tiles = int(tiles);
    p->cap = int((tiles + kReplicas - 1) / kReplicas);
    const size_t countBytes = sizeof(uint32_t) * kCountSets * kReplicas * kReplicaStride;
    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipMalloc(&p->counts, countBytes);
    if (e == hipSuccess) e = hipMalloc(&p->ids, sizeof(uint32_t) * kIdSets * kReplicas * kCostBuckets * size_t(p->cap));
    if (e == hipSuccess) e = hipMemset(p->counts, 0, countBytes);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);  // cleared before any stream's launch uses it
    if (e != hipSuccess) {
        tile_plan_free(p);
        return fail(SVO_ERR_NOMEM, std::string("tile plan: ") + hipGetErrorString(e));
    }
    *out = p;
    return SVO_OK;
}

int tile_plan_create(int width, int height, svo_tile_plan **out) {
    if (!out) return fail(SVO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return fail(SVO_ERR_INVALID, "image size out of range");
    return tile_plan_alloc(int64_t((width + 7) / 8) * ((height + 7) / 8), width, height, out);
}

void tile_plan_free(svo_tile_plan *p) {
    if (!p) return;
    if (p->counts) (void)hipFree(p->counts);
    if (p->ids) (void)hipFree(p->ids);
    delete p;
}

namespace {

template <typename Idx, int Reg>
void launch_ordered(const svo_octree &h, const TileLaunch &L, const PrimaryParams &p, const svo_tile_plan &plan,
                    uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, hipStream_t s) {
    const TileOrder o = next_order(plan);
    hipLaunchKernelGGL((primary_ordered_kernel<Idx, Reg>), dim3(unsigned(plan.tiles)), dim3(64), L.lds, s, h.dwords, p,
                       L.slots, o, hit, normal, t, leaf);
}

}  // namespace

int render_primary_planned_device(svo_octree &h, svo_tile_plan *plan, const svo_camera &cam, int width, int height,
                                  uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, void *stream) {
    if (!plan) return fail(SVO_ERR_INVALID, "null tile plan");
    if (plan->kind != svo_tile_plan::kPrimary || width != plan->width || height != plan->height)
        return fail(SVO_ERR_INVALID, "the tile plan was made for another image size or render entry");
    if (!camera_ok(cam)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != plan->device) return fail(SVO_ERR_INVALID, "the tile plan lives on another device");
    int r = ensure_device(h);
    if (r) return r;
    const PrimaryParams p = primary_params(cam, width, height);
    const TileLaunch L = tile_launch(h);
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (!L.wide && L.reg == 1) launch_ordered<uint32_t, 1>(h, L, p, *plan, hit, normal, t, leaf, s);
    else if (!L.wide) launch_ordered<uint32_t, 3>(h, L, p, *plan, hit, normal, t, leaf, s);
    else if (L.reg == 1) launch_ordered<uint64_t, 1>(h, L, p, *plan, hit, normal, t, leaf, s);
    else launch_ordered<uint64_t, 3>(h, L, p, *plan, hit, normal, t, leaf, s);
    HIP_TRY(hipGetLastError());
    plan->renders++;
    return SVO_OK;
}

int render_primary_device(svo_octree &h, const svo_camera &cam, int width, int height, uint8_t *hit,
                          uint32_t *normal, float *t, uint64_t *leaf, void *stream) {
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return fail(SVO_ERR_INVALID, "image size out of range");
    if (!camera_ok(cam)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    int r = ensure_device(h);
    if (r) return r;
    const PrimaryParams p = primary_params(cam, width, height);
    const int tiles = p.tilesX * ((height + 7) / 8);
    launch_primary(h, tile_launch(h), p, tiles, hit, normal, t, leaf, static_cast<hipStream_t>(stream));
    HIP_TRY(hipGetLastError());
    return SVO_OK;
}

// Primary pass and one shadow ray per leaf hit, both in primary_shadow_kernel;
// with `plan` (made by tile_plan_create for the image size) in its cost order.
int render_shadow_device(svo_octree &h, const svo_camera &cam, int width, int height, const float light[3], float eps,
                         uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow, void *stream,
                         svo_tile_plan *plan) {
    if (!hit || !t || !shadow || !light) return fail(SVO_ERR_INVALID, "shadow pass needs hit, t, shadow and light");
    for (int k = 0; k < 3; ++k)
        if (!(std::fabs(light[k]) <= 1e6f)) return fail(SVO_ERR_INVALID, "light direction must be finite");
    if (!(std::fabs(light[0]) + std::fabs(light[1]) + std::fabs(light[2]) > 0.0f))
        return fail(SVO_ERR_INVALID, "light direction must be non-zero");
    // shadow origins o + d*(t - eps) must stay finite (the shadow traversal assumes finite rays)
    if (!(std::fabs(eps) <= 1e6f)) return fail(SVO_ERR_INVALID, "eps must be finite and within +-1e6");
    if (width <= 0 || height <= 0 || width > 65536 || height > 65536)
        return fail(SVO_ERR_INVALID, "image size out of range");
    if (!camera_ok(cam)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    if (plan) {
        if (plan->kind != svo_tile_plan::kPrimary || width != plan->width || height != plan->height)
            return fail(SVO_ERR_INVALID, "the tile plan was made for another image size or render entry");
        int dev = -1;
        HIP_TRY(hipGetDevice(&dev));
        if (dev != plan->device) return fail(SVO_ERR_INVALID, "the tile plan lives on another device");
    }
    int r = ensure_device(h);
    if (r) return r;
    const PrimaryParams p = primary_params(cam, width, height);
    const int tiles = p.tilesX * ((height + 7) / 8);
    const TileLaunch L = tile_launch(h);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const dim3 grid(static_cast<unsigned>(tiles)), block(64);
    const TileOrder o = plan ? next_order(*plan) : TileOrder{};
    if (plan) {
        if (!L.wide && L.reg == 1)
            hipLaunchKernelGGL((primary_shadow_kernel<uint32_t, 1, true>), grid, block, L.lds, s, h.dwords, p, light[0],
                               light[1], light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
        else if (!L.wide)
            hipLaunchKernelGGL((primary_shadow_kernel<uint32_t, 3, true>), grid, block, L.lds, s, h.dwords, p, light[0],
                               light[1], light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
        else if (L.reg == 1)
            hipLaunchKernelGGL((primary_shadow_kernel<uint64_t, 1, true>), grid, block, L.lds, s, h.dwords, p, light[0],
                               light[1], light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
        else
            hipLaunchKernelGGL((primary_shadow_kernel<uint64_t, 3, true>), grid, block, L.lds, s, h.dwords, p, light[0],
                               light[1], light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
    } else if (!L.wide && L.reg == 1) {
        hipLaunchKernelGGL((primary_shadow_kernel<uint32_t, 1>), grid, block, L.lds, s, h.dwords, p, light[0], light[1],
                           light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
    } else if (!L.wide) {
        hipLaunchKernelGGL((primary_shadow_kernel<uint32_t, 3>), grid, block, L.lds, s, h.dwords, p, light[0], light[1],
                           light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
    } else if (L.reg == 1) {
        hipLaunchKernelGGL((primary_shadow_kernel<uint64_t, 1>), grid, block, L.lds, s, h.dwords, p, light[0], light[1],
                           light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
    } else {
        hipLaunchKernelGGL((primary_shadow_kernel<uint64_t, 3>), grid, block, L.lds, s, h.dwords, p, light[0], light[1],
                           light[2], eps, L.slots, hit, normal, t, leaf, shadow, o);
    }
    HIP_TRY(hipGetLastError());
    if (plan) plan->renders++;
    return SVO_OK;
}

}  // namespace svo

namespace svo {

namespace {
__global__ void code_probe_kernel() {}
}  // namespace

// Makes the HIP runtime load this source file's device code on the current
// device now (it loads a code object at the first launch of one of its kernels).
hipError_t load_code_raymarch() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&code_probe_kernel));
}

}  // namespace svo
