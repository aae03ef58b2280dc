// Batched renderer: many views of many resident octrees in one launch, in
// one of two forms.
//
// Grid form (default, tiles_kernel): an ordinary launch with one wave per
// 8x8-pixel tile of every view. A single view's launch spends about half its
// time in a tail where a few slow tiles (rays grazing the surface) run on an
// otherwise idle GPU; with all views in one grid the other views' tiles fill
// that time, and only the last tail remains.
//
// Persistent form (batch_kernel), described below.
//
// Work is a single queue of 8x8-pixel tiles over all views. Tiles of a view
// are numbered in 8x8-tile blocks (64x64 pixels) with a Morton order inside
// each block, so neighbouring queue entries trace neighbouring rays through
// the same octree nodes. The queue is split into 8 contiguous chunks, one per
// XCD (workgroup i runs on XCD i % 8), so each XCD's L2 mostly sees one region;
// a wave whose chunk is exhausted steals from the next chunks in turn and
// exits after all eight are empty, so every wave terminates.
//
// Each wave owns one tile at a time and hands its pixels to idle lanes. A
// lane that finishes a ray becomes idle; once `refill` lanes are idle the
// wave pulls the next pixels (and, when the tile is used up, the next tile)
// for them, so SIMD lanes stay busy while the rays of a tile finish at very
// different depths. Results per ray are those of VoxelOctree::raymarch
// (src/VoxelOctree.cpp:207-346): the traversal code is march.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "march.hpp"
#include "svo_internal.hpp"
#include "tile_order.hpp"
#include "view_desc.hpp"

namespace svo {

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SVO_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace march;

constexpr uint64_t kNoTile = ~uint64_t(0);
constexpr int kTileShapes = 4;                      // grid-form tile shapes the autotuner times
constexpr int kAutotuneShapes = kTileShapes + 9;    // plus 3 x 3 persistent shapes

__device__ __forceinline__ int view_of(const uint64_t *__restrict__ tileBase, int nviews, uint64_t tile) {
    int lo = 0, hi = nviews;  // tileBase[lo] <= tile < tileBase[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tileBase[mid] <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint32_t even_bits(uint32_t m) {  // bits 0,2,4 -> 0,1,2
    return (m & 1u) | ((m >> 1) & 2u) | ((m >> 2) & 4u);
}

// Lod = false when every view's rayScale is 0 (the LOD test can never fire);
// camera rays are finite (cameras are checked on the host). Reg: top stack
// entries held in registers, for the deepest octree of the batch.
template <bool Lod, typename Idx, int Reg>
__global__ void __launch_bounds__(256) batch_kernel(const ViewDesc *__restrict__ views,
                                                    const uint64_t *__restrict__ tileBase, int nviews,
                                                    uint32_t *__restrict__ counters, int slots, int refill) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    Stack<Idx> stk(lds + size_t(wave) * slots * 64 * (sizeof(Idx) + sizeof(float)), 64, slots, lane);
    const uint64_t total = tileBase[nviews];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    int queue = blockIdx.x & (kQueues - 1);
    int tried = 0;
    bool queueDone = false;
    int pix = 64;                 // next pixel of the current tile (wave-uniform)
    int tView = 0, tx = 0, ty = 0;

    bool active = false;
    Ray<Idx> r;
    const uint32_t *tree = nullptr;
    int rView = 0;
    int64_t rPix = 0;

    for (;;) {
        // ---- hand pixels to idle lanes ----
        uint64_t idle = __ballot(!active);
        if (!queueDone && (__popcll(idle) >= refill || idle == ~0ull)) {
            while (idle && !queueDone) {
                if (pix >= 64) {
                    uint64_t tile = kNoTile;
                    while (tried < kQueues) {
                        const uint64_t lo = total * uint64_t(queue) / kQueues;
                        const uint64_t hi = total * uint64_t(queue + 1) / kQueues;
                        uint32_t k = 0;
                        if (lane == 0) k = atomicAdd(&counters[queue * kCounterStride], 1u);
                        k = __shfl(k, 0);
                        if (lo + k < hi) {
                            tile = lo + k;
                            break;
                        }
                        queue = (queue + 1) & (kQueues - 1);
                        tried++;
                    }
                    if (tile == kNoTile) {
                        queueDone = true;
                        break;
                    }
                    tView = view_of(tileBase, nviews, tile);
                    const ViewDesc &vd = views[tView];
                    const uint64_t local = tile - tileBase[tView];
                    const uint32_t m = uint32_t(local & 63);
                    const uint64_t block = local >> 6;
                    tx = int(block % uint64_t(vd.blocksX)) * 8 + int(even_bits(m));
                    ty = int(block / uint64_t(vd.blocksX)) * 8 + int(even_bits(m >> 1));
                    pix = (tx < vd.tilesX && ty < vd.tilesY) ? 0 : 64;  // padding tiles hold no pixels
                    continue;
                }
                const int avail = 64 - pix;
                const int nIdle = __popcll(idle);
                const int take = nIdle < avail ? nIdle : avail;
                const int rank = __popcll(idle & below);
                const bool mine = ((idle >> lane) & 1ull) && rank < take;
                if (mine) {
                    const ViewDesc &vd = views[tView];
                    const int p = pix + rank;
                    const int x = tx * 8 + (p & 7);
                    const int y = ty * 8 + (p >> 3);
                    if (x < vd.width && y < vd.height) {
                        float dx, dy, dz;
                        view_ray_dir(vd, x, y, dx, dy, dz);
                        ray_init(r, vd.tree, vd.px, vd.py, vd.pz, dx, dy, dz, vd.rayScale);
                        tree = vd.tree;
                        rView = tView;
                        rPix = int64_t(y) * vd.width + x;
                        active = true;
                    }
                }
                // the first `take` idle lanes are consumed (edge pixels outside the image included)
                const uint64_t consumed = __ballot(mine);
                idle &= ~consumed;
                pix += take;
            }
        }
        if (!__any(active)) {
            if (queueDone) break;
            continue;
        }
        // ---- walk until enough lanes are idle ----
        for (;;) {
            if (active) {
                uint32_t normal = 0;
                float t = 0.0f;
                Idx leaf = 0;
                int code = ray_step<Lod, true, Idx, true, Reg>(r, tree, stk, normal, t, leaf);
                if (code != kRunning) {
                    view_store<Idx>(views[rView], rPix, code, normal, t, leaf);
                    active = false;
                }
            }
            const uint64_t idleNow = __ballot(!active);
            if (idleNow == ~0ull) break;
            if (!queueDone && __popcll(idleNow) >= refill) break;
        }
    }
}

// Grid form: one ordinary launch with one workgroup (one wave) per 8x8 tile
// of every view, views one after another, tiles in raster order. The hardware dispatcher hands out
// the tiles; a wave traces its tile's 64 rays and exits. A single view's
// launch ends with a tail where a few slow tiles (rays that graze the
// surface) run on an otherwise idle GPU; with many views in one launch that
// tail is paid once. `first` offsets blockIdx for grids split into several
// launches. A tile is 2^twLog pixels wide and 64 / 2^twLog high (8x8 by
// default; 16x4, 4x16 and 32x2 are the autotuner's other shapes); the views'
// tilesX / tilesY are counted in that shape.
//
// Planned: the batch's tiles in a tile plan's cost order (tile_order.hpp),
// each filed by cost for the plan's next launch; one launch (first = 0).
template <bool Lod, typename Idx, int Reg, bool Planned = false>
__global__ void __launch_bounds__(64) tiles_kernel(const ViewDesc *__restrict__ views,
                                                   const uint64_t *__restrict__ tileBase, int nviews, uint64_t first,
                                                   int slots, int twLog, TileOrder o) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x;
    const uint64_t tile = Planned ? uint64_t(ordered_tile(o, int(blockIdx.x), lane)) : first + blockIdx.x;
    const int v = view_of(tileBase, nviews, tile);
    const ViewDesc &vd = views[v];
    const uint64_t local = tile - tileBase[v];
    const int tx = int(local % uint64_t(vd.tilesX));
    const int ty = int(local / uint64_t(vd.tilesX));
    const int x = (tx << twLog) + (lane & ((1 << twLog) - 1));
    const int y = (ty << (6 - twLog)) + (lane >> twLog);
    const bool inside = !(tx >= vd.tilesX || ty >= vd.tilesY || x >= vd.width || y >= vd.height);
    if (!Planned && !inside) return;
    uint32_t passes = 0;
    if (inside) {
        float dx, dy, dz;
        view_ray_dir(vd, x, y, dx, dy, dz);
        Stack<Idx> stk(lds, 64, slots, lane);
        Ray<Idx> r;
        ray_init(r, vd.tree, vd.px, vd.py, vd.pz, dx, dy, dz, vd.rayScale);
        uint32_t normal = 0;
        float t = 0.0f;
        Idx leaf = 0;
        int code;
        do {
            code = ray_step<Lod, true, Idx, true, Reg>(r, vd.tree, stk, normal, t, leaf);
            if (Planned) passes++;
        } while (code == kRunning);
        view_store<Idx>(vd, int64_t(y) * vd.width + x, code, normal, t, leaf);
    }
    if (Planned) file_tile(o, int(blockIdx.x), int(tile), passes, lane);
}

__global__ void zero_counters_kernel(uint32_t *c) {
    if (threadIdx.x < kQueues) c[threadIdx.x * kCounterStride] = 0u;
}

}  // namespace

}  // namespace svo

struct svo_renderer {
    int maxViews = 0;
    svo::ViewDesc *dViews = nullptr;
    uint64_t *dTileBase = nullptr;
    uint32_t *dCounters = nullptr;
    svo::ViewDesc *hViews = nullptr;  // pinned staging
    uint64_t *hTileBase = nullptr;  // pinned staging
    hipEvent_t staged = nullptr;    // the last call's staging copies have been read
    hipEvent_t done = nullptr;      // the last call's launch has finished with the device tables
    int cus = 0;
    std::mutex mu;
};

namespace svo {

int renderer_create(int maxViews, svo_renderer **out) {
    if (maxViews <= 0 || maxViews > (1 << 20)) return fail(SVO_ERR_INVALID, "max_views out of range");
    svo_renderer *r = new (std::nothrow) svo_renderer();
    if (!r) return fail(SVO_ERR_NOMEM, "out of host memory");
    r->maxViews = maxViews;
    hipError_t e = hipMalloc(&r->dViews, sizeof(ViewDesc) * maxViews);
    if (e == hipSuccess) e = hipMalloc(&r->dTileBase, sizeof(uint64_t) * (maxViews + 1));
    if (e == hipSuccess) e = hipMalloc(&r->dCounters, sizeof(uint32_t) * kQueues * kCounterStride);
    if (e == hipSuccess) e = hipHostMalloc(&r->hViews, sizeof(ViewDesc) * maxViews, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc(&r->hTileBase, sizeof(uint64_t) * (maxViews + 1), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->done, hipEventDisableTiming);
    int dev = 0;
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) {
        renderer_free(r);
        return fail(SVO_ERR_DEVICE, std::string("renderer: ") + hipGetErrorString(e));
    }
    r->cus = prop.multiProcessorCount;
    *out = r;
    return SVO_OK;
}

void renderer_free(svo_renderer *r) {
    if (!r) return;
    if (r->done) {
        (void)hipEventSynchronize(r->done);
        (void)hipEventDestroy(r->done);
    }
    if (r->staged) (void)hipEventDestroy(r->staged);
    if (r->dViews) (void)hipFree(r->dViews);
    if (r->dTileBase) (void)hipFree(r->dTileBase);
    if (r->dCounters) (void)hipFree(r->dCounters);
    if (r->hViews) (void)hipHostFree(r->hViews);
    if (r->hTileBase) (void)hipHostFree(r->hTileBase);
    delete r;
}

int fill_view_desc(const svo_view &v, ViewDesc &d) {
    if (!v.tree || !v.tree->dwords) return fail(SVO_ERR_INVALID, "view without a resident octree");
    if (v.width <= 0 || v.height <= 0 || v.width > 65536 || v.height > 65536)
        return fail(SVO_ERR_INVALID, "view size out of range");
    if (!camera_ok(v.camera)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    if (!(std::fabs(v.ray_scale) <= 1e6f)) return fail(SVO_ERR_INVALID, "ray_scale must be finite");
    const float pi = float(3.14159265358979323846);
    const float planeDist = 1.0f / std::tan(pi / 6.0f);  // src/Main.cpp:158
    d.tree = v.tree->dwords;
    d.hit = v.hit;
    d.normal = v.normal;
    d.t = v.t;
    d.leaf = v.leaf;
    const float *tf = v.camera.tform;
    d.a11 = tf[0]; d.a12 = tf[1]; d.a21 = tf[4]; d.a22 = tf[5]; d.a31 = tf[8]; d.a32 = tf[9];
    d.zx = planeDist * tf[2];
    d.zy = planeDist * tf[6];
    d.zz = planeDist * tf[10];
    d.px = v.camera.pos[0]; d.py = v.camera.pos[1]; d.pz = v.camera.pos[2];
    d.aspect = float(v.height) / float(v.width);
    d.scale = 2.0f / float(v.width);
    d.rayScale = v.ray_scale;
    d.width = v.width;
    d.height = v.height;
    d.tilesX = (v.width + 7) / 8;
    d.tilesY = (v.height + 7) / 8;
    d.blocksX = (d.tilesX + 7) / 8;
    d.pad = 0;
    return SVO_OK;
}

namespace {

// log2 of a grid-form tile width (0 = 8); -1 for a width that is not 4, 8, 16 or 32.
int tile_width_log(int w) {
    switch (w) {
        case 0:
        case 8: return 3;
        case 4: return 2;
        case 16: return 4;
        case 32: return 5;
        default: return -1;
    }
}

}  // namespace

// Validates the views and fills the host view table, the tile bases and the
// launch shape. upload: octrees not yet resident are uploaded (otherwise a
// view of an octree not resident on the current device is refused).
int plan_views(const svo_view *views, int nviews, const svo_batch_opts *opts, int cus, ViewDesc *hv, uint64_t *htb,
               LaunchPlan &plan, bool upload) {
    int maxDepth = 1;
    bool wide = false, lod = false;
    uint64_t tiles = 0;
    const bool persistentForm = opts && opts->form == SVO_BATCH_PERSISTENT;
    const int twLog = tile_width_log(opts ? opts->tile_w : 0);
    if (twLog < 0) return fail(SVO_ERR_INVALID, "tile_w must be 4, 8, 16 or 32");
    if (persistentForm && twLog != 3) return fail(SVO_ERR_INVALID, "the persistent form traces 8x8 tiles");
    for (int i = 0; i < nviews; ++i) {
        const svo_view &v = views[i];
        if (!v.tree) return fail(SVO_ERR_INVALID, "view without an octree");
        int r = SVO_OK;
        if (upload) {
            r = ensure_device(*v.tree);
        } else {
            int dev = -1;
            if (hipGetDevice(&dev) != hipSuccess || !v.tree->dwords || v.tree->device != dev)
                r = fail(SVO_ERR_INVALID, "view of an octree not resident on this device (svo_upload it first)");
        }
        if (r) return r;
        ViewDesc &d = hv[i];
        r = fill_view_desc(v, d);
        if (r) return r;
        lod = lod || v.ray_scale != 0.0f;
        maxDepth = std::max(maxDepth, v.tree->depth);
        wide = wide || v.tree->words.size() >= (uint64_t(1) << 32);
        const int blocksY = (d.tilesY + 7) / 8;
        if (!persistentForm) {  // tiles of the chosen shape
            const int tw = 1 << twLog, th = 64 >> twLog;
            d.tilesX = (v.width + tw - 1) / tw;
            d.tilesY = (v.height + th - 1) / th;
        }
        htb[i] = tiles;
        // persistent form: Morton order in 64x64-pixel blocks (padded); grid
        // form: raster order (1% faster there, measured on 8 views at 1080p)
        tiles += persistentForm ? uint64_t(d.blocksX) * uint64_t(blocksY) * 64 : uint64_t(d.tilesX) * uint64_t(d.tilesY);
    }
    plan.twLog = twLog;
    htb[nviews] = tiles;
    if (tiles >= (uint64_t(kQueues) << 32)) return fail(SVO_ERR_LIMIT, "too many tiles for one launch");
    plan.nviews = nviews;
    plan.tiles = tiles;
    plan.reg = reg_entries(maxDepth);
    plan.slots = lds_slots(maxDepth, plan.reg);
    plan.wide = wide;
    plan.lod = lod;
    const size_t ldsPerWave = size_t(plan.slots) * 64 * ((wide ? 8 : 4) + sizeof(float));
    plan.ldsPerWave = ldsPerWave;
    plan.ldsPerBlock = ldsPerWave * 4;
    const int form = opts ? opts->form : SVO_BATCH_GRID;
    if (form != SVO_BATCH_GRID && form != SVO_BATCH_PERSISTENT) return fail(SVO_ERR_INVALID, "unknown batch form");
    plan.persistent = form == SVO_BATCH_PERSISTENT;
    const size_t ldsPerCU = 160 * 1024;
    const int wavesPerCU = (opts && opts->waves_per_cu > 0) ? opts->waves_per_cu : 32;
    const int blocksPerCU = std::max(1, std::min(wavesPerCU / 4, int(ldsPerCU / plan.ldsPerBlock)));
    if (plan.ldsPerBlock > 64 * 1024) return fail(SVO_ERR_LIMIT, "octree too deep for the LDS stack");
    plan.refill = (opts && opts->refill_threshold > 0) ? std::min(opts->refill_threshold, 64) : 16;
    // persistent grid: enough blocks to fill the GPU, but no more than four
    // tiles per wave of work (small batches would otherwise launch idle waves)
    const uint64_t workBlocks = (tiles + 15) / 16;
    plan.grid = unsigned(std::max<uint64_t>(kQueues, std::min<uint64_t>(uint64_t(cus) * blocksPerCU, workBlocks)));
    return SVO_OK;
}

namespace {

// Queue reset and the persistent kernel (no memset: the counters are cleared
// by a kernel, so the pair can be captured in a graph).
template <int Reg>
void launch_batch(const LaunchPlan &p, const ViewDesc *dViews, const uint64_t *dTileBase, uint32_t *dCounters,
                  hipStream_t s) {
    if (!p.wide && !p.lod)
        hipLaunchKernelGGL((batch_kernel<false, uint32_t, Reg>), dim3(p.grid), dim3(256), p.ldsPerBlock, s, dViews,
                           dTileBase, p.nviews, dCounters, p.slots, p.refill);
    else if (!p.wide)
        hipLaunchKernelGGL((batch_kernel<true, uint32_t, Reg>), dim3(p.grid), dim3(256), p.ldsPerBlock, s, dViews,
                           dTileBase, p.nviews, dCounters, p.slots, p.refill);
    else if (!p.lod)
        hipLaunchKernelGGL((batch_kernel<false, uint64_t, Reg>), dim3(p.grid), dim3(256), p.ldsPerBlock, s, dViews,
                           dTileBase, p.nviews, dCounters, p.slots, p.refill);
    else
        hipLaunchKernelGGL((batch_kernel<true, uint64_t, Reg>), dim3(p.grid), dim3(256), p.ldsPerBlock, s, dViews,
                           dTileBase, p.nviews, dCounters, p.slots, p.refill);
}

// Grid form: the tiles in launches of at most 2^31 - 1 workgroups.
template <int Reg>
void launch_tiles(const LaunchPlan &p, const ViewDesc *dViews, const uint64_t *dTileBase, hipStream_t s) {
    const uint64_t maxGrid = 0x7FFFFFFFull;
    const TileOrder none{};
    for (uint64_t first = 0; first < p.tiles; first += maxGrid) {
        const dim3 grid(unsigned(std::min(maxGrid, p.tiles - first)));
        if (!p.wide && !p.lod)
            hipLaunchKernelGGL((tiles_kernel<false, uint32_t, Reg>), grid, dim3(64), p.ldsPerWave, s, dViews,
                               dTileBase, p.nviews, first, p.slots, p.twLog, none);
        else if (!p.wide)
            hipLaunchKernelGGL((tiles_kernel<true, uint32_t, Reg>), grid, dim3(64), p.ldsPerWave, s, dViews,
                               dTileBase, p.nviews, first, p.slots, p.twLog, none);
        else if (!p.lod)
            hipLaunchKernelGGL((tiles_kernel<false, uint64_t, Reg>), grid, dim3(64), p.ldsPerWave, s, dViews,
                               dTileBase, p.nviews, first, p.slots, p.twLog, none);
        else
            hipLaunchKernelGGL((tiles_kernel<true, uint64_t, Reg>), grid, dim3(64), p.ldsPerWave, s, dViews,
                               dTileBase, p.nviews, first, p.slots, p.twLog, none);
    }
}

// Grid form in a tile plan's order: one launch (the plan holds at most 2^31 - 1 tiles).
template <int Reg>
void launch_tiles_planned(const LaunchPlan &p, const ViewDesc *dViews, const uint64_t *dTileBase, const TileOrder &o,
                          hipStream_t s) {
    const dim3 grid(unsigned(p.tiles));
    if (!p.wide && !p.lod)
        hipLaunchKernelGGL((tiles_kernel<false, uint32_t, Reg, true>), grid, dim3(64), p.ldsPerWave, s, dViews,
                           dTileBase, p.nviews, uint64_t(0), p.slots, p.twLog, o);
    else if (!p.wide)
        hipLaunchKernelGGL((tiles_kernel<true, uint32_t, Reg, true>), grid, dim3(64), p.ldsPerWave, s, dViews,
                           dTileBase, p.nviews, uint64_t(0), p.slots, p.twLog, o);
    else if (!p.lod)
        hipLaunchKernelGGL((tiles_kernel<false, uint64_t, Reg, true>), grid, dim3(64), p.ldsPerWave, s, dViews,
                           dTileBase, p.nviews, uint64_t(0), p.slots, p.twLog, o);
    else
        hipLaunchKernelGGL((tiles_kernel<true, uint64_t, Reg, true>), grid, dim3(64), p.ldsPerWave, s, dViews,
                           dTileBase, p.nviews, uint64_t(0), p.slots, p.twLog, o);
}

}  // namespace

void launch_plan(const LaunchPlan &p, const ViewDesc *dViews, const uint64_t *dTileBase, uint32_t *dCounters,
                 hipStream_t s) {
    if (!p.persistent) {
        if (p.reg == 3) launch_tiles<3>(p, dViews, dTileBase, s);
        else launch_tiles<1>(p, dViews, dTileBase, s);
        return;
    }
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, dCounters);
    if (p.reg == 3) launch_batch<3>(p, dViews, dTileBase, dCounters, s);
    else launch_batch<1>(p, dViews, dTileBase, dCounters, s);
}


int render_batch(svo_renderer *rd, const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream) {
    if (!rd || !views) return fail(SVO_ERR_INVALID, "null argument");
    if (nviews <= 0 || nviews > rd->maxViews) return fail(SVO_ERR_LIMIT, "view count exceeds the renderer's max_views");
    std::lock_guard<std::mutex> lock(rd->mu);
    // the host waits only until the previous call's copies have read the
    // pinned staging buffers; the device tables are protected on the device:
    // this call's stream waits for the previous launch (a no-op on one stream)
    HIP_TRY(hipEventSynchronize(rd->staged));
    LaunchPlan plan;
    int r = plan_views(views, nviews, opts, rd->cus, rd->hViews, rd->hTileBase, plan, true);
    if (r) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamWaitEvent(s, rd->done, 0));
    HIP_TRY(hipMemcpyAsync(rd->dViews, rd->hViews, sizeof(ViewDesc) * nviews, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(rd->dTileBase, rd->hTileBase, sizeof(uint64_t) * (nviews + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(rd->staged, s));
    launch_plan(plan, rd->dViews, rd->dTileBase, rd->dCounters, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(rd->done, s));
    return SVO_OK;
}

// The grid form's tile count of a batch (the views' sizes and the tile shape).
int tile_plan_create_batch(const svo_view *views, int nviews, const svo_batch_opts *opts, svo_tile_plan **out) {
    if (!out) return fail(SVO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (!views || nviews <= 0) return fail(SVO_ERR_INVALID, "no views");
    if (opts && opts->form == SVO_BATCH_PERSISTENT) return fail(SVO_ERR_INVALID, "tile plans order the grid form");
    const int twLog = tile_width_log(opts ? opts->tile_w : 0);
    if (twLog < 0) return fail(SVO_ERR_INVALID, "tile_w must be 4, 8, 16 or 32");
    const int tw = 1 << twLog, th = 64 >> twLog;
    int64_t tiles = 0;
    for (int i = 0; i < nviews; ++i) {
        if (views[i].width <= 0 || views[i].height <= 0 || views[i].width > 65536 || views[i].height > 65536)
            return fail(SVO_ERR_INVALID, "image size out of range");
        tiles += int64_t((views[i].width + tw - 1) / tw) * int64_t((views[i].height + th - 1) / th);
    }
    const int r = tile_plan_alloc(tiles, 0, 0, out);
    if (r == SVO_OK) (*out)->kind = svo_tile_plan::kBatch;
    return r;
}

// render_batch in the grid form with the tiles in a tile plan's order.
int render_batch_planned(svo_renderer *rd, const svo_view *views, int nviews, const svo_batch_opts *opts,
                         svo_tile_plan *tp, void *stream) {
    if (!rd || !views || !tp) return fail(SVO_ERR_INVALID, "null argument");
    if (nviews <= 0 || nviews > rd->maxViews) return fail(SVO_ERR_LIMIT, "view count exceeds the renderer's max_views");
    if (opts && opts->form == SVO_BATCH_PERSISTENT) return fail(SVO_ERR_INVALID, "tile plans order the grid form");
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != tp->device) return fail(SVO_ERR_INVALID, "the tile plan lives on another device");
    if (tp->kind != svo_tile_plan::kBatch) return fail(SVO_ERR_INVALID, "the tile plan was not made for a batch");
    std::lock_guard<std::mutex> lock(rd->mu);
    HIP_TRY(hipEventSynchronize(rd->staged));
    LaunchPlan plan;
    int r = plan_views(views, nviews, opts, rd->cus, rd->hViews, rd->hTileBase, plan, true);
    if (r) return r;
    if (plan.tiles != uint64_t(tp->tiles)) return fail(SVO_ERR_INVALID, "the tile plan was made for another batch");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipStreamWaitEvent(s, rd->done, 0));
    HIP_TRY(hipMemcpyAsync(rd->dViews, rd->hViews, sizeof(ViewDesc) * nviews, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(rd->dTileBase, rd->hTileBase, sizeof(uint64_t) * (nviews + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(rd->staged, s));
    const TileOrder o = next_order(*tp);
    if (plan.reg == 3) launch_tiles_planned<3>(plan, rd->dViews, rd->dTileBase, o, s);
    else launch_tiles_planned<1>(plan, rd->dViews, rd->dTileBase, o, s);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(rd->done, s));
    tp->renders++;
    return SVO_OK;
}

int renderer_autotune(svo_renderer *rd, const svo_view *views, int nviews, void *stream, svo_batch_opts *best,
                      float *times) {
    if (!rd || !views || !best) return fail(SVO_ERR_INVALID, "null argument");
    static const int kTileW[kTileShapes] = {8, 16, 4, 32}, kWaves[3] = {16, 24, 32}, kRefill[3] = {16, 32, 64};
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(SVO_ERR_DEVICE, std::string("autotune: ") + hipGetErrorString(he));
    }
    float bestMs = 0.0f;
    int r = SVO_OK;
    for (int i = 0; i < kAutotuneShapes && !r; ++i) {
        svo_batch_opts o{};
        if (i < kTileShapes) {
            o.form = SVO_BATCH_GRID;
            o.tile_w = kTileW[i];
        } else {
            const int k = i - kTileShapes;
            o.form = SVO_BATCH_PERSISTENT;
            o.waves_per_cu = kWaves[k / 3];
            o.refill_threshold = kRefill[k % 3];
        }
        r = render_batch(rd, views, nviews, &o, stream);  // warm-up
        if (!r && (he = hipEventRecord(e0, s)) != hipSuccess) r = fail(SVO_ERR_DEVICE, hipGetErrorString(he));
        for (int k = 0; k < 2 && !r; ++k) r = render_batch(rd, views, nviews, &o, stream);
        if (!r && (he = hipEventRecord(e1, s)) != hipSuccess) r = fail(SVO_ERR_DEVICE, hipGetErrorString(he));
        if (!r && (he = hipEventSynchronize(e1)) != hipSuccess) r = fail(SVO_ERR_DEVICE, hipGetErrorString(he));
        float ms = 0.0f;
        if (!r && (he = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) r = fail(SVO_ERR_DEVICE, hipGetErrorString(he));
        if (r) break;
        ms *= 0.5f;
        if (times) times[i] = ms;
        if (i == 0 || ms < bestMs) {
            bestMs = ms;
            *best = o;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

}  // namespace svo

// One batched launch captured in a HIP graph: its own view table, tile bases
// and queue counters, fixed at capture.
struct svo_batch_graph {
    svo::ViewDesc *dViews = nullptr;
    uint64_t *dTileBase = nullptr;
    uint32_t *dCounters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace svo {

void batch_graph_free(svo_batch_graph *g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->dViews) (void)hipFree(g->dViews);
    if (g->dTileBase) (void)hipFree(g->dTileBase);
    if (g->dCounters) (void)hipFree(g->dCounters);
    delete g;
}

int batch_graph_create(svo_renderer *rd, const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream,
                       svo_batch_graph **out) {
    if (!rd || !views || !out) return fail(SVO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (nviews <= 0) return fail(SVO_ERR_INVALID, "no views");
    std::vector<ViewDesc> hv(static_cast<size_t>(nviews));
    std::vector<uint64_t> htb(static_cast<size_t>(nviews) + 1);
    LaunchPlan plan;
    int r = plan_views(views, nviews, opts, rd->cus, hv.data(), htb.data(), plan, true);
    if (r) return r;
    svo_batch_graph *g = new (std::nothrow) svo_batch_graph();
    if (!g) return fail(SVO_ERR_NOMEM, "out of host memory");
    hipError_t e = hipMalloc(&g->dViews, sizeof(ViewDesc) * nviews);
    if (e == hipSuccess) e = hipMalloc(&g->dTileBase, sizeof(uint64_t) * (nviews + 1));
    if (e == hipSuccess) e = hipMalloc(&g->dCounters, sizeof(uint32_t) * kQueues * kCounterStride);
    if (e == hipSuccess) e = hipMemcpy(g->dViews, hv.data(), sizeof(ViewDesc) * nviews, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->dTileBase, htb.data(), sizeof(uint64_t) * (nviews + 1), hipMemcpyHostToDevice);
    // capture needs a stream other than the null stream
    hipStream_t s = static_cast<hipStream_t>(stream), own = nullptr;
    if (e == hipSuccess && !s) {
        e = hipStreamCreateWithFlags(&own, hipStreamNonBlocking);
        s = own;
    }
    if (e == hipSuccess) e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        launch_plan(plan, g->dViews, g->dTileBase, g->dCounters, s);
        hipError_t le = hipGetLastError();
        e = hipStreamEndCapture(s, &g->graph);
        if (e == hipSuccess) e = le;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(own);
    if (e != hipSuccess) {
        batch_graph_free(g);
        return fail(SVO_ERR_DEVICE, std::string("batch graph: ") + hipGetErrorString(e));
    }
    *out = g;
    return SVO_OK;
}

int batch_graph_launch(svo_batch_graph *g, void *stream) {
    if (!g) return fail(SVO_ERR_INVALID, "null graph");
    HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    return SVO_OK;
}

}  // namespace svo

namespace svo {

namespace {
__global__ void code_probe_kernel() {}
}  // namespace

// Makes the HIP runtime load this source file's device code on the current
// device now (it loads a code object at the first launch of one of its kernels).
hipError_t load_code_render_batch() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&code_probe_kernel));
}

}  // namespace svo
