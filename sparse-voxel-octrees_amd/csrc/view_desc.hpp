// One view of a batched or served render as the kernels read it: octree,
// output arrays and the camera terms of the primary-ray generation. Shared by
// the batched renderer (render_batch.hip) and the request server (server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "march.hpp"
#include "svo_internal.hpp"

namespace svo {

struct ViewDesc {
    const uint32_t *tree;
    uint8_t *hit;
    uint32_t *normal;
    float *t;
    uint64_t *leaf;
    float a11, a12, a21, a22, a31, a32;  // camera rotation columns used by ray generation
    float zx, zy, zz;                    // planeDist * third column
    float px, py, pz;                    // ray origin
    float aspect, scale, rayScale;       // H/W, 2/W, LOD scale
    int width, height, tilesX, tilesY, blocksX, pad;
};
static_assert(sizeof(ViewDesc) == 128, "ViewDesc is copied as 32 words");

constexpr int kQueues = 8;          // batched persistent form: one tile-queue chunk per XCD
constexpr int kCounterStride = 64;  // 256 B between queue counters

// Everything one batched launch needs besides the device view table.
struct LaunchPlan {
    int nviews = 0;
    uint64_t tiles = 0;
    int slots = 1;
    int reg = 1;  // stack entries held in registers
    bool wide = false, lod = false;
    size_t ldsPerBlock = 0;
    unsigned grid = 0;
    int refill = 16;
    bool persistent = false;  // SVO_BATCH_PERSISTENT; otherwise one workgroup per tile
    size_t ldsPerWave = 0;
    int twLog = 3;  // grid form: log2 of the tile width (tile height 64 >> twLog)
};


// Validates views and fills the host view table (hv), the tile bases (htb,
// nviews + 1) and the launch shape; render_batch.hip.
int plan_views(const svo_view *views, int nviews, const svo_batch_opts *opts, int cus, ViewDesc *hv, uint64_t *htb,
               LaunchPlan &plan, bool upload);
// Enqueues the planned launch over device copies of the view table and tile
// bases (and, in the persistent form, the queue counters, which it clears).
void launch_plan(const LaunchPlan &p, const ViewDesc *dViews, const uint64_t *dTileBase, uint32_t *dCounters,
                 hipStream_t s);

// Checks a view (tree present, size in range, finite camera and LOD scale)
// and fills `d` from it; the octree must already be resident (d.tree is its
// device word array).
int fill_view_desc(const svo_view &v, ViewDesc &d);

// Primary-ray direction of pixel (x, y): src/Main.cpp:97-113 with closed-form
// dx, dy, normalised by the reference's invSqrt (src/Util.hpp:47-58).
__device__ __forceinline__ void view_ray_dir(const ViewDesc &vd, int x, int y, float &ox, float &oy, float &oz) {
    const float dy = vd.aspect - float(y) * vd.scale;
    const float dx = -1.0f + float(x) * vd.scale;
    const float vx = dx * vd.a11 + dy * vd.a12 + vd.zx;
    const float vy = dx * vd.a21 + dy * vd.a22 + vd.zy;
    const float vz = dx * vd.a31 + dy * vd.a32 + vd.zz;
    const float s = march::quake_rsqrt(vx * vx + vy * vy + vz * vz);
    ox = vx * s;
    oy = vy * s;
    oz = vz * s;
}

// Writes one ray's outcome into the view's output arrays (any may be null).
template <typename Idx>
__device__ __forceinline__ void view_store(const ViewDesc &vd, int64_t i, int code, uint32_t normal, float t, Idx leaf) {
    if (vd.hit) vd.hit[i] = uint8_t(code);
    if (vd.normal) vd.normal[i] = code == march::kLeaf ? normal : 0u;
    if (vd.t) vd.t[i] = code == march::kMiss ? 0.0f : t;
    if (vd.leaf) vd.leaf[i] = code == march::kLeaf ? uint64_t(leaf) : SVO_LEAF_NONE;
}

}  // namespace svo
