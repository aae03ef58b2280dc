// The viewer's frame algorithm on the GPU (src/Main.cpp renderBatch,
// renderTile and shade, :81-202, over the per-thread stripes of :351-361).
//
// Per stripe the reference walks a grid of 8x8-pixel tiles with one extra
// column and row of corners: every corner gets one coarse ray with the LOD
// cut-off rayScale = 2*8/(planeDist*H), and every tile whose four corner
// depths are not all "miss" is traced per pixel from pos + dir*max(minT-0.03,0),
// shaded and packed to RGBA; other tiles stay 0. Ray directions use the
// reference's running sums (dx += scale per pixel inside a tile, dx +=
// 8*scale per corner along a stripe), so they are reproduced exactly: every
// lane replays the running sums it needs with the same float steps. Stripe
// bounds are closed-form (every stripe but the last is full), so nothing is
// staged from the host and consecutive frames queue on the stream without a
// host round trip.
//
//  coarse_kernel: one lane per corner of every stripe -> depth table
//  fine_kernel:   one wave per tile; lane = pixel; with stride > 1 only the
//                 stride-corner lanes trace and the others copy their colour
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "build_common.hpp"
#include "march.hpp"
#include "tile_order.hpp"
#include "svo_internal.hpp"

namespace svo {

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SVO_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace march;

constexpr float kTreeMiss = 1e10f;  // TreeMiss, src/Main.cpp:140

struct FrameParams {
    float a11, a12, a21, a22, a31, a32, zx, zy, zz;
    float px, py, pz;
    float lx, ly, lz;  // light direction
    float aspect, scale, coarseScale, tileScale;
    int width, height, tilesX, stripes, stride;
    int sh;    // rows of a full stripe: (H - 1) / stripes + 1
    int full;  // corner rows of a full stripe: (sh - 1) / 8 + 2
};

// Up to N frames of one size traced by one coarse and one fine launch: frame f
// has its own camera and light (p[f]), image and depth table. The grids are
// frame-major, so a frame's tiles are contiguous and the frames' coarse passes
// and fine-pass tails overlap each other's work.
template <int N>
struct FrameSet {
    FrameParams p[N];
    uint32_t *rgba[N];
    float *depth[N];
};

// Stripe s covers rows [s*sh, min((s+1)*sh, H)) (src/Main.cpp:351-361). Every
// stripe but the last is full (the layout check guarantees it), so stripe s
// starts at corner s*full*tilesX of the depth table and at fine tile
// s*(full-1)*(tilesX-1).
__device__ __forceinline__ int stripe_y0(const FrameParams &p, int s) { return s * p.sh; }
__device__ __forceinline__ int stripe_y1(const FrameParams &p, int s) { return min((s + 1) * p.sh, p.height); }


// One ray to completion; the leaf word (normal) is read only when Normal is set.
// The frame kernels keep the whole restart stack in LDS (the fine pass measured
// slower with the root's entry in registers). With Count, *passes counts the
// loop passes.
template <bool Lod, bool Normal, typename Idx, bool Count = false>
__device__ __forceinline__ int trace(const uint32_t *__restrict__ tree, float ox, float oy, float oz, float dx, float dy,
                                     float dz, float rayScale, const Stack<Idx> &stk, uint32_t &normal, float &t,
                                     uint32_t *passes = nullptr) {
    Ray<Idx> r;
    ray_init(r, tree, ox, oy, oz, dx, dy, dz, rayScale);
    Idx leaf = 0;
    int code;
    do {
        code = ray_step<Lod, true, Idx, Normal, 0>(r, tree, stk, normal, t, leaf);  // camera rays are finite
        if (Count) ++*passes;
    } while (code == kRunning);
    return code;
}

__device__ __forceinline__ void frame_dir(const FrameParams &p, float dx, float dy, float d[3]) {
    const float vx = dx * p.a11 + dy * p.a12 + p.zx;
    const float vy = dx * p.a21 + dy * p.a22 + p.zy;
    const float vz = dx * p.a31 + dy * p.a32 + p.zz;
    const float s = quake_rsqrt(vx * vx + vy * vy + vz * vz);
    d[0] = vx * s;
    d[1] = vy * s;
    d[2] = vz * s;
}

template <typename Idx, int N>
__global__ void __launch_bounds__(64) coarse_kernel(const uint32_t *__restrict__ tree, const FrameSet<N> set,
                                                    int64_t corners, unsigned blocksPerFrame, int slots) {
    extern __shared__ __align__(16) unsigned char lds[];
    const unsigned f = N == 1 ? 0u : blockIdx.x / blocksPerFrame;
    const int64_t i = int64_t(blockIdx.x - f * blocksPerFrame) * blockDim.x + threadIdx.x;
    if (i >= corners) return;
    const FrameParams &p = set.p[f];
    float *__restrict__ depth = set.depth[f];
    const int64_t perStripe = int64_t(p.full) * p.tilesX;
    const int s = int(min(i / perStripe, int64_t(p.stripes - 1)));
    const int64_t local = i - s * perStripe;
    const int ty = int(local / p.tilesX), tx = int(local % p.tilesX);
    // the reference's running corner sums (src/Main.cpp:156-176), step by step
    float dx = -1.0f + 0.0f * p.scale;
    for (int k = 0; k < tx; ++k) dx += p.tileScale;
    float dy = p.aspect - float(stripe_y0(p, s)) * p.scale;
    for (int k = 0; k < ty; ++k) dy -= p.tileScale;
    float d[3];
    frame_dir(p, dx, dy, d);
    Stack<Idx> stk(lds, 64, slots, int(threadIdx.x));
    uint32_t normal = 0;
    float t = 0.0f;
    const int code = trace<true, false, Idx>(tree, p.px, p.py, p.pz, d[0], d[1], d[2], p.coarseScale, stk, normal, t);
    depth[i] = code != march::kMiss ? t : kTreeMiss;
}

//
// Planned (one frame): the tiles in a tile plan's cost order (tile_order.hpp),
// each filed by cost for the plan's next frame.
template <typename Idx, int N, bool Planned = false>
__global__ void __launch_bounds__(64) fine_kernel(const uint32_t *__restrict__ tree, const FrameSet<N> set,
                                                  unsigned tilesPerFrame, int slots, TileOrder o) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int64_t g = Planned ? int64_t(ordered_tile(o, int(blockIdx.x), int(threadIdx.x))) : int64_t(blockIdx.x);
    const unsigned f = N == 1 ? 0u : unsigned(g / tilesPerFrame);
    const int64_t tile = g - int64_t(f) * tilesPerFrame;
    const FrameParams &p = set.p[f];
    const float *__restrict__ depth = set.depth[f];
    uint32_t *__restrict__ rgba = set.rgba[f];
    const int64_t perStripe = int64_t(p.full - 1) * (p.tilesX - 1);
    const int s = int(min(tile / perStripe, int64_t(p.stripes - 1)));
    const int64_t local = tile - s * perStripe;
    const int fx = int(local % (p.tilesX - 1)), fy = int(local / (p.tilesX - 1));
    const int tx = fx + 1, ty = fy + 1;  // the reference renders tile (x-1, y-1) at corner (x, y)
    const int64_t idx = int64_t(s) * p.full * p.tilesX + int64_t(ty) * p.tilesX + tx;
    const float minT = dev::fmin_ref(dev::fmin_ref(depth[idx], depth[idx - 1]),
                                dev::fmin_ref(depth[idx - p.tilesX], depth[idx - p.tilesX - 1]));
    const int x0 = fx * 8, y0 = stripe_y0(p, s) + fy * 8;
    const int x1 = min(x0 + 8, p.width), y1 = min(y0 + 8, stripe_y1(p, s));
    const int lane = threadIdx.x;
    const int px = lane & 7, py = lane >> 3;
    const int x = x0 + px, y = y0 + py;
    const bool valid = x < x1 && y < y1;
    uint32_t color = 0, passes = 0;
    if (minT != kTreeMiss) {
        const bool corner = (px % p.stride) == 0 && (py % p.stride) == 0;
        if (valid && corner) {
            float dx = -1.0f + float(x0) * p.scale;
            for (int k = 0; k < px; ++k) dx += p.scale;
            float dy = p.aspect - float(y0) * p.scale;
            for (int k = 0; k < py; ++k) dy -= p.scale;
            float d[3];
            frame_dir(p, dx, dy, d);
            const float m = dev::fmax_ref(minT - 0.03f, 0.0f);
            Stack<Idx> stk(lds, 64, slots, lane);
            uint32_t normal = 0;
            float t = 0.0f;
            float v = 0.0f;
            const int code = trace<false, true, Idx, Planned>(tree, p.px + d[0] * m, p.py + d[1] * m, p.pz + d[2] * m,
                                                              d[0], d[1], d[2], 0.0f, stk, normal, t, &passes);
            if (code != march::kMiss) {
                // shade, src/Main.cpp:81-90
                float n[3], c;
                dev::decompress_material(normal, n, c);
                const float proj = (n[0] * d[0] + n[1] * d[1] + n[2] * d[2]) * 2.0f;
                const float r[3] = {d[0] - n[0] * proj, d[1] - n[1] * proj, d[2] - n[2] * proj};
                const float dd = dev::fmax_ref(p.lx * r[0] + p.ly * r[1] + p.lz * r[2], 0.0f);
                const float specular = dd * dd;
                v = c * 0.9f * fabsf(p.lx * n[0] + p.ly * n[1] + p.lz * n[2]) + specular * 0.2f;
            }
            const uint32_t grey = uint32_t(double(dev::fmin_ref(v, 1.0f)) * 255.0);
            color = grey | (grey << 8) | (grey << 16) | 0xFF000000u;
        }
        // pixels between stride corners copy their corner (renderTile's half-size mode)
        const int src = (py - py % p.stride) * 8 + (px - px % p.stride);
        color = __shfl(color, src);
    }
    if (valid) rgba[int64_t(y) * p.width + x] = color;
    if (Planned) file_tile(o, int(blockIdx.x), int(g), passes, lane);
}

}  // namespace

namespace {

// Image size and stripe split checks shared by the renderer and the depth
// count; stripes 0 selects the viewer's 16 (NumThreads, src/Main.cpp:57).
int frame_layout(int width, int height, int &stripes) {
    if (width <= 0 || height <= 0 || width > 32768 || height > 32768) return fail(SVO_ERR_INVALID, "image size out of range");
    if (stripes <= 0) stripes = 16;
    // the viewer's split (rows per stripe rounded up) must leave every stripe non-empty; where it does
    // not, the reference clears a negative number of rows (src/Main.cpp:165)
    if (int64_t(stripes - 1) * ((height - 1) / stripes + 1) >= height)
        return fail(SVO_ERR_INVALID, "stripe split leaves an empty stripe at this height");
    return SVO_OK;
}

}  // namespace

namespace {

// The viewer's per-frame constants (src/Main.cpp:154-176) for one camera.
FrameParams frame_params(const svo_camera &cam, int width, int height, int stripes, int stride) {
    const float pi = float(3.14159265358979323846);
    FrameParams p;
    const float *tf = cam.tform;
    p.px = cam.pos[0];
    p.py = cam.pos[1];
    p.pz = cam.pos[2];
    p.scale = 2.0f / float(width);
    const float tileScale = 8.0f * p.scale;
    const float planeDist = 1.0f / std::tan(pi / 6.0f);
    p.a11 = tf[0]; p.a12 = tf[1]; p.a21 = tf[4]; p.a22 = tf[5]; p.a31 = tf[8]; p.a32 = tf[9];
    p.zx = planeDist * tf[2];
    p.zy = planeDist * tf[6];
    p.zz = planeDist * tf[10];
    p.coarseScale = 2.0f * 8.0f / (planeDist * float(height));
    {  // light = (tform*Vec3(-1,1,-1)).normalize(), translation cleared (src/Main.cpp:154,162)
        float lv[3];
        for (int k = 0; k < 3; ++k) {
            float a = tf[k * 4 + 0] * -1.0f;
            a = a + tf[k * 4 + 1] * 1.0f;
            a = a + tf[k * 4 + 2] * -1.0f;
            lv[k] = a + 0.0f;
        }
        const float inv = 1.0f / std::sqrt(lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2]);
        p.lx = lv[0] * inv;
        p.ly = lv[1] * inv;
        p.lz = lv[2] * inv;
    }
    p.aspect = float(height) / float(width);
    p.width = width;
    p.height = height;
    p.tilesX = (width - 1) / 8 + 2;
    p.stripes = stripes;
    p.stride = stride;
    p.tileScale = tileScale;
    p.sh = (height - 1) / stripes + 1;
    p.full = (p.sh - 1) / 8 + 2;
    return p;
}

constexpr int kFramesPerLaunch = 16;

// Frames [0, n) of `set` (n <= N), all of the size in set.p[0]: one coarse
// launch over every frame's corners, then one fine launch over every frame's
// tiles. Frames with a null depth table get one from a stream-ordered scratch
// allocation.
// Corner and fine-tile counts of a frame over its stripes (the last may be shorter).
void frame_counts(const FrameParams &p, int64_t &corners, int64_t &tiles) {
    corners = 0;
    tiles = 0;
    for (int i = 0; i < p.stripes; ++i) {
        const int y0 = i * p.sh, y1 = std::min(y0 + p.sh, p.height);
        const int tilesY = (y1 - y0 - 1) / 8 + 2;
        corners += int64_t(tilesY) * p.tilesX;
        tiles += int64_t(tilesY - 1) * (p.tilesX - 1);
    }
}

// With `plan` (made for n frames), the fine pass runs in the plan's cost order.
template <int N>
int launch_frames(svo_octree &h, FrameSet<N> &set, int n, hipStream_t s, svo_tile_plan *plan = nullptr) {
    const FrameParams &p = set.p[0];
    int64_t corners = 0, tiles = 0;
    frame_counts(p, corners, tiles);
    if (plan && (n != plan->frames || plan->kind != svo_tile_plan::kFrame || plan->width != p.width ||
                 plan->height != p.height || plan->stripes != p.stripes || int64_t(plan->tiles) != tiles * n))
        return fail(SVO_ERR_INVALID, "the tile plan was made for another frame size, stripe split or frame count");
    const int64_t cblocks = (corners + 63) / 64;
    if (tiles * n > 0x7FFFFFFF || cblocks * n > 0x7FFFFFFF) return fail(SVO_ERR_LIMIT, "image too large");
    int nulls = 0;
    for (int f = 0; f < n; ++f) nulls += set.depth[f] ? 0 : 1;
    void *scratch = nullptr;
    if (nulls) HIP_TRY(hipMallocAsync(&scratch, sizeof(float) * size_t(corners) * size_t(nulls), s));
    for (int f = 0, k = 0; f < n; ++f)
        if (!set.depth[f]) set.depth[f] = static_cast<float *>(scratch) + size_t(corners) * size_t(k++);
    const int slots = lds_slots(h.depth, 0);  // the whole stack in LDS (Reg = 0)
    const bool wide = h.words.size() >= (uint64_t(1) << 32);
    const size_t lds = size_t(slots) * 64 * ((wide ? 8 : 4) + sizeof(float));
    const dim3 cgrid(static_cast<unsigned>(cblocks * n)), fgrid(static_cast<unsigned>(tiles * n)), block(64);
    const TileOrder none{};
    if (!wide) {
        hipLaunchKernelGGL((coarse_kernel<uint32_t, N>), cgrid, block, lds, s, h.dwords, set, corners, unsigned(cblocks),
                           slots);
        if (tiles && plan)
            hipLaunchKernelGGL((fine_kernel<uint32_t, N, true>), fgrid, block, lds, s, h.dwords, set, unsigned(tiles),
                               slots, next_order(*plan));
        else if (tiles)
            hipLaunchKernelGGL((fine_kernel<uint32_t, N>), fgrid, block, lds, s, h.dwords, set, unsigned(tiles), slots,
                               none);
    } else {
        hipLaunchKernelGGL((coarse_kernel<uint64_t, N>), cgrid, block, lds, s, h.dwords, set, corners, unsigned(cblocks),
                           slots);
        if (tiles && plan)
            hipLaunchKernelGGL((fine_kernel<uint64_t, N, true>), fgrid, block, lds, s, h.dwords, set, unsigned(tiles),
                               slots, next_order(*plan));
        else if (tiles)
            hipLaunchKernelGGL((fine_kernel<uint64_t, N>), fgrid, block, lds, s, h.dwords, set, unsigned(tiles), slots,
                               none);
    }
    hipError_t e = hipGetLastError();
    if (scratch) (void)hipFreeAsync(scratch, s);
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("frame: ") + hipGetErrorString(e));
    if (plan && tiles) plan->renders++;
    return SVO_OK;
}

}  // namespace

int tile_plan_create_frames(int width, int height, int stripes, int nframes, svo_tile_plan **out) {
    if (!out) return fail(SVO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (nframes < 1 || nframes > kFramesPerLaunch)
        return fail(SVO_ERR_INVALID, "a frame tile plan covers 1 to 16 frames (one launch)");
    int r = frame_layout(width, height, stripes);
    if (r) return r;
    svo_camera cam{};
    const FrameParams p = frame_params(cam, width, height, stripes, 1);
    int64_t corners = 0, tiles = 0;
    frame_counts(p, corners, tiles);
    r = tile_plan_alloc(tiles * nframes, width, height, out);
    if (r == SVO_OK) {
        (*out)->kind = svo_tile_plan::kFrame;
        (*out)->stripes = stripes;
        (*out)->frames = nframes;
    }
    return r;
}

int render_frame_device(svo_octree &h, const svo_camera &cam, int width, int height, int stripes, int stride,
                        uint32_t *rgba, float *depthOut, void *stream, svo_tile_plan *plan) {
    if (plan) {
        int dev = -1;
        HIP_TRY(hipGetDevice(&dev));
        if (dev != plan->device) return fail(SVO_ERR_INVALID, "the tile plan lives on another device");
    }
    int r = frame_layout(width, height, stripes);
    if (r) return r;
    if (stride <= 0 || stride > 8) return fail(SVO_ERR_INVALID, "stride must be in [1, 8]");
    if (!rgba) return fail(SVO_ERR_INVALID, "null image");
    if (!camera_ok(cam)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    r = ensure_device(h);
    if (r) return r;
    FrameSet<1> set;
    set.p[0] = frame_params(cam, width, height, stripes, stride);
    set.rgba[0] = rgba;
    set.depth[0] = depthOut;
    return launch_frames(h, set, 1, static_cast<hipStream_t>(stream), plan);
}

// Several frames of one size in one call: each group of up to 16 frames is one
// coarse launch over all their corners and one fine launch over all their
// tiles, so one frame's coarse pass and fine-pass tail (both leave most of the
// GPU idle) overlap the other frames' fine work. Each frame equals
// render_frame_device's.
int render_frames_device(svo_octree &h, const svo_camera *cams, int n, int width, int height, int stripes, int stride,
                         uint32_t *const *rgba, float *const *depth, void *stream, svo_tile_plan *plan) {
    if (!cams || !rgba || n <= 0) return fail(SVO_ERR_INVALID, "null camera or image array, or no frames");
    if (plan) {
        if (n != plan->frames) return fail(SVO_ERR_INVALID, "the tile plan was made for another frame count");
        int dev = -1;
        HIP_TRY(hipGetDevice(&dev));
        if (dev != plan->device) return fail(SVO_ERR_INVALID, "the tile plan lives on another device");
    }
    int r = frame_layout(width, height, stripes);
    if (r) return r;
    if (stride <= 0 || stride > 8) return fail(SVO_ERR_INVALID, "stride must be in [1, 8]");
    for (int f = 0; f < n; ++f) {
        if (!rgba[f]) return fail(SVO_ERR_INVALID, "null image");
        if (!camera_ok(cams[f])) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    }
    r = ensure_device(h);
    if (r) return r;
    hipStream_t s = static_cast<hipStream_t>(stream);
    for (int f0 = 0; f0 < n && !r; f0 += kFramesPerLaunch) {
        const int m = std::min(kFramesPerLaunch, n - f0);
        FrameSet<kFramesPerLaunch> set;
        for (int k = 0; k < kFramesPerLaunch; ++k) {
            const int f = f0 + std::min(k, m - 1);  // unused entries repeat the last frame (never indexed)
            set.p[k] = frame_params(cams[f], width, height, stripes, stride);
            set.rgba[k] = rgba[f];
            set.depth[k] = depth ? depth[f] : nullptr;
        }
        r = launch_frames(h, set, m, s, plan);  // a plan covers one launch group (n <= 16)
    }
    return r;
}

int64_t frame_depth_count(int width, int height, int stripes) {
    if (frame_layout(width, height, stripes)) return -1;
    const int sh = (height - 1) / stripes + 1;
    int64_t n = 0;
    for (int i = 0; i < stripes; ++i) {
        const int y0 = i * sh, y1 = std::min((i + 1) * sh, height);
        n += int64_t((width - 1) / 8 + 2) * ((y1 - y0 - 1) / 8 + 2);
    }
    return n;
}

}  // namespace svo

namespace svo {

namespace {
__global__ void code_probe_kernel() {}
}  // namespace

// Makes the HIP runtime load this source file's device code on the current
// device now (it loads a code object at the first launch of one of its kernels).
hipError_t load_code_frame() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&code_probe_kernel));
}

}  // namespace svo
