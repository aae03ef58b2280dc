// Cost-ordered tile dispatch, shared by the plain primary launch
// (raymarch.hip) and the batched grid launch (render_batch.hip).
//
// A tile's cost is the largest loop-pass count of its 64 rays; tiles are
// grouped in 16 buckets by log2(cost). A launch made through a tile plan
// starts the tiles its previous launch found slowest first (rays grazing the
// surface take up to ~500 dependent passes against a median ~40), so they run
// beside the bulk of the tiles instead of in the launch's tail on an
// otherwise idle GPU. The order only decides when a tile runs: every ray and
// its result are those of the unordered launch.
//
// The order is kept by the launches themselves, with no pass between them.
// Workgroup b files its tile for the next launch in replica b % 64, under its
// cost bucket (one atomic on that replica's bucket counter, one store into
// the replica's id region); workgroup b of the next launch takes entry b / 64
// of replica b % 64's tiles in descending bucket order (its 16 counts give
// the bucket). Replica r holds exactly the tiles of the workgroups b with
// b % 64 == r, the count the next launch takes from it, so every tile runs
// once; dispatch in workgroup order starts the costliest tiles of all
// replicas first. A single set of 16 counters would take one device-scope
// atomic per tile on one cache line, which serialises (measured: a 2.07 M-ray
// view at 0.73 ms against 0.33); 64 replicas on lines of their own take ~500
// each.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// A tile plan: counter sets and id regions on one device (opaque in the C ABI).
struct svo_tile_plan {
    enum Kind : int { kPrimary, kBatch, kFrame };
    Kind kind = kPrimary;        // the render entry the plan was made for
    int width = 0, height = 0;   // primary renders and frames: the image size (0 x 0 for a batch plan)
    int stripes = 0;             // frames: the stripe split
    int frames = 0;              // frames: the frames of each launch made through the plan
    int tiles = 0;               // tiles of every launch made through the plan
    int cap = 0;                 // ceil(tiles / 64): the most tiles a replica holds
    int device = -1;
    uint64_t renders = 0;        // launches made through the plan (0: no filing yet)
    uint32_t *counts = nullptr;  // 3 sets of 64 replicas of 16 bucket counters
    uint32_t *ids = nullptr;     // 2 regions of 64 replicas x 16 buckets x cap tile ids
};

namespace svo {

constexpr int kCostBuckets = 16;
constexpr int kReplicas = 64;
constexpr int kReplicaStride = 32;  // counters per replica: 16 used, 128 B apart
constexpr int kCountSets = 3;       // a set is read by launch n, cleared by launch n+1, filled by launch n+2
constexpr int kIdSets = 2;          // a region is read by launch n and filled by launch n+1

// What one launch reads, fills and clears.
struct TileOrder {
    uint32_t *counts;   // [kCountSets][kReplicas][kReplicaStride]
    uint32_t *ids;      // [kIdSets][kReplicas][kCostBuckets][cap]
    int tiles, cap;
    int src, dst, clr;  // counter sets read (< 0: raster order), filled, cleared
    int idSrc, idDst;   // id regions read, filled
};

// Launch n of a plan reads counter set n % 3 and id region n % 2 (filled by
// launch n-1), fills set (n+1) % 3 and region (n+1) % 2, clears set (n+2) % 3.
inline TileOrder next_order(const svo_tile_plan &plan) {
    const uint64_t n = plan.renders;
    TileOrder o;
    o.counts = plan.counts;
    o.ids = plan.ids;
    o.tiles = plan.tiles;
    o.cap = plan.cap;
    o.src = n ? int(n % kCountSets) : -1;
    o.dst = int((n + 1) % kCountSets);
    o.clr = int((n + 2) % kCountSets);
    o.idSrc = int(n % kIdSets);
    o.idDst = int((n + 1) % kIdSets);
    return o;
}

__device__ __forceinline__ int cost_bucket(uint32_t c) { return min(kCostBuckets - 1, 31 - __builtin_clz(c | 1u)); }

__device__ __forceinline__ size_t replica_counts(int set, int rep) {
    return (size_t(set) * kReplicas + rep) * kReplicaStride;
}
__device__ __forceinline__ size_t replica_ids(const TileOrder &o, int set, int rep, int bucket) {
    return ((size_t(set) * kReplicas + rep) * kCostBuckets + bucket) * size_t(o.cap);
}

// The tile of workgroup b (lane k < 16 holds bucket 15 - k's count of replica
// b % 64); workgroups b < 64 also clear replica b of the set the next launch
// fills.
__device__ __forceinline__ int ordered_tile(const TileOrder &o, int b, int lane) {
    if (b < kReplicas && lane < kCostBuckets) o.counts[replica_counts(o.clr, b) + lane] = 0u;
    if (o.src < 0) return b;
    const int rep = b % kReplicas, k = b / kReplicas;
    const uint32_t c = lane < kCostBuckets ? o.counts[replica_counts(o.src, rep) + (kCostBuckets - 1 - lane)] : 0u;
    uint32_t incl = c;  // inclusive prefix over lanes 0..lane
#pragma unroll
    for (int d = 1; d < kCostBuckets; d <<= 1) {
        const uint32_t v = uint32_t(__shfl_up(int(incl), d));
        if (lane >= d) incl += v;
    }
    const uint64_t past = __ballot(lane < kCostBuckets && uint32_t(k) < incl);
    if (!past) return b;  // counts that do not cover the launch (not reached through the plan's checks)
    const int j = __ffsll(static_cast<unsigned long long>(past)) - 1;  // the first bucket reaching past k
    const uint32_t before = uint32_t(__shfl(int(incl - c), j));
    return int(o.ids[replica_ids(o, o.idSrc, rep, kCostBuckets - 1 - j) + (uint32_t(k) - before)]);
}

// Files workgroup b's tile for the next launch under the wave's largest pass
// count (every lane of the wave calls this).
__device__ __forceinline__ void file_tile(const TileOrder &o, int b, int tile, uint32_t passes, int lane) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) passes = max(passes, uint32_t(__shfl_xor(int(passes), d)));
    if (lane == 0) {
        const int bucket = cost_bucket(passes);
        const int rep = b % kReplicas;
        const uint32_t at = atomicAdd(&o.counts[replica_counts(o.dst, rep) + bucket], 1u);
        if (at < uint32_t(o.cap))  // a replica files at most cap tiles per launch
            o.ids[replica_ids(o, o.idDst, rep, bucket) + at] = uint32_t(tile);
    }
}

}  // namespace svo
