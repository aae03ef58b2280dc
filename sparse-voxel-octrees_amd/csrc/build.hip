// GPU octree construction.
//
// Produces the word array the reference builder writes (VoxelOctree(VoxelData*),
// src/VoxelOctree.cpp:125-265, fed by PlyLoader/VoxelData), for the same
// (resolution, memory budget, thread count), without its recursion:
//
//  mesh source (PlyLoader, src/PlyLoader.cpp)
//   1. rescale vertices to the unit cube; fan-triangulate; face normals
//   2. bin triangles into sub-block cells (iterateOverlappingBlocks: SAT of
//      the triangle against each +1-shifted cell box)          -> (cell, tri)
//   3. per (cell, tri): voxel rows of the triangle's grid box clipped to the
//      cell; SAT per voxel with the incrementally stepped centres of
//      triangleToVolume                                          -> keys
//      key = morton(x,y,z) << tribits | tri
//   4. radix sort the keys: each voxel's triangles end up in ascending order
//   5. per voxel, fold writeTriangleCell over its triangles in that order
//      (re-quantised running average)                           -> leaf words
//  file source (VoxelData(path)): non-zero voxels -> morton keys -> sort
//
//  both: the reference's occupancy-table behaviour (VoxelData.cpp:87-209):
//   - voxels on the unscanned last slice of an odd-depth cache block are
//     never emitted (buildLowLut covers 2*floor(d/2) slices);
//   - above the cache-block level occupancy comes from the top table, so a
//     cache block counted non-empty but holding no voxel still becomes an
//     (empty) node.
//  then levels bottom-up (parent = key >> 3, child masks), subtree sizes with
//  the far-pointer rule, offsets top-down, and a scatter of every word into
//  the depth-first layout (see size_kernel / place_kernel).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <vector>

#include "build_common.hpp"
#include "range_alloc.hpp"
#include "svo_internal.hpp"

namespace svo {

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SVO_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace dev;

constexpr int kThreads = 256;

inline unsigned grid_for(uint64_t n) { return unsigned((n + kThreads - 1) / kThreads); }

// Device memory and stream of one build, kept between builds: a workspace
// holds one slab from the device's stream-ordered pool and a stream; a build
// carves its arrays from the slab first-fit on the host. All of a build's
// kernels and copies run on the workspace's stream, so an array released and
// carved again is reused in stream order and no carve waits for the device.
// A request the slab cannot hold comes from the pool itself; after the build
// the slab is regrown to the build's peak (up to the reserve bound) so that the
// next build of that size allocates nothing from the runtime. Concurrent builds
// on one device take separate workspaces.
constexpr uint64_t kCarveAlign = 256;

struct Workspace {
    int device = -1;
    hipStream_t stream = nullptr;
    char *slab = nullptr;
    uint64_t size = 0;
    RangeAlloc fit;
};

struct WorkspaceCache {
    std::mutex mu;
    std::vector<Workspace *> idle;
};

WorkspaceCache &workspaces() {
    static WorkspaceCache *c = new WorkspaceCache();  // never freed: a static destructor may outlive the runtime
    return *c;
}

// What the device keeps for later builds: idle slabs plus the pool's free
// memory, 1/16 of the device and at most 16 GiB (a 2048^3 mesh build peaks
// near 2 GiB): memory newly obtained from the driver costs milliseconds per
// large allocation.
size_t build_reserve() {
    size_t freeB = 0, total = 0;
    if (hipMemGetInfo(&freeB, &total) != hipSuccess) return 0;
    return std::min<size_t>(total / 16, size_t(16) << 30);
}

bool default_pool(hipMemPool_t &pool) {
    int dev = 0;
    return hipGetDevice(&dev) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess;
}

struct Arena {
    Workspace *ws = nullptr;
    hipStream_t stream = nullptr;
    std::chrono::steady_clock::time_point started = std::chrono::steady_clock::now();  // device phase begins
    struct Block {
        void *p;
        uint64_t bytes;
        bool carved;
    };
    std::vector<Block> blocks;
    uint64_t live = 0, peak = 0;

    // takes an idle workspace of the current device (or a new one)
    int open() {
        int dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        {
            WorkspaceCache &c = workspaces();
            std::lock_guard<std::mutex> lock(c.mu);
            for (size_t i = 0; i < c.idle.size(); ++i)
                if (c.idle[i]->device == dev) {
                    ws = c.idle[i];
                    c.idle.erase(c.idle.begin() + long(i));
                    break;
                }
        }
        if (!ws) {
            ws = new Workspace();
            ws->device = dev;
            hipError_t e = hipStreamCreate(&ws->stream);
            if (e != hipSuccess) {
                delete ws;
                ws = nullptr;
                return fail(SVO_ERR_DEVICE, std::string("builder: hipStreamCreate: ") + hipGetErrorString(e));
            }
            hipMemPool_t pool;
            uint64_t keep = ~uint64_t(0);  // the pool keeps what builds free; trimmed after each build
            if (default_pool(pool)) (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
        ws->fit.reset(ws->size);
        stream = ws->stream;
        started = std::chrono::steady_clock::now();
        return SVO_OK;
    }

    ~Arena() {
        if (!ws) return;
        clear();
        (void)hipStreamSynchronize(stream);
        const uint64_t reserve = build_reserve();
        // regrow the slab to this build's peak (with slack for fragmentation)
        const uint64_t want = std::min<uint64_t>((peak + peak / 4 + (uint64_t(64) << 20) - 1) >> 26 << 26, reserve);
        if (want > ws->size) {
            if (ws->slab) (void)hipFreeAsync(ws->slab, stream);
            ws->slab = nullptr;
            ws->size = 0;
            void *p = nullptr;
            if (hipMallocAsync(&p, want, stream) == hipSuccess) {
                ws->slab = static_cast<char *>(p);
                ws->size = want;
            }
            (void)hipStreamSynchronize(stream);
        }
        hipMemPool_t pool;
        WorkspaceCache &c = workspaces();
        std::lock_guard<std::mutex> lock(c.mu);
        uint64_t held = ws->size;
        for (Workspace *w : c.idle)
            if (w->device == ws->device) held += w->size;
        if (held > reserve) {  // more idle slabs than the reserve: drop this one
            if (ws->slab) (void)hipFreeAsync(ws->slab, stream);
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
            delete ws;
        } else {
            c.idle.push_back(ws);
        }
        ws = nullptr;
        if (default_pool(pool)) (void)hipMemPoolTrimTo(pool, reserve > held ? reserve - held : 0);
    }

    template <typename T>
    int alloc(T *&out, uint64_t n) {
        out = nullptr;
        const uint64_t bytes = (std::max<uint64_t>(n, 1) * sizeof(T) + kCarveAlign - 1) / kCarveAlign * kCarveAlign;
        uint64_t off = 0;
        void *p = nullptr;
        bool carved = ws->slab && ws->fit.alloc(bytes, off);
        if (carved) {
            p = ws->slab + off;
        } else {
            hipError_t e = hipMallocAsync(&p, bytes, stream);
            if (e != hipSuccess)
                return fail(SVO_ERR_NOMEM, std::string("builder: hipMallocAsync: ") + hipGetErrorString(e));
        }
        blocks.push_back({p, bytes, carved});
        live += bytes;
        peak = std::max(peak, live);
        out = static_cast<T *>(p);
        return SVO_OK;
    }
    void drop(Block &b) {
        if (b.carved) ws->fit.release(uint64_t(static_cast<char *>(b.p) - ws->slab), b.bytes);
        else (void)hipFreeAsync(b.p, stream);
        live -= b.bytes;
        b.p = nullptr;
    }
    void release(void *p) {
        for (Block &b : blocks)
            if (b.p == p) drop(b);
    }
    void clear() {
        for (Block &b : blocks)
            if (b.p) drop(b);
        blocks.clear();
    }
};

// Parameters the reference derives on the host (VoxelData::init,
// PlyLoader::setupBlockProcessing).
struct Geometry {
    int dataW, dataH, dataD;  // VoxelData::_dataW/H/D
    int side;                 // sideLength(): largest power-of-two virtual dim
    int depth;                // log2(side)
    int cacheLog;             // _lowLutLevels: cache block edge = 1 << cacheLog
    int cache;
    // mesh source only
    int scale;               // PlyLoader::_sideLength - 2 = resolution - 4
    int sbW, sbH, sbD;       // sub-block dims
    int pW, pH, pD;          // partitions per cache block
    int gW, gH, gD;          // flat cell grid
    int blocksX, blocksY, blocksZ;  // cache blocks per virtual axis
};

int round_pow2(int x) {
    int y = 1;
    while (y < x) y *= 2;
    return y;
}

int log2i(int v) {
    int r = 0;
    while (v >>= 1) r++;
    return r;
}

uint64_t grid_cells(int levels) {
    uint64_t r = 0, s = 1;
    for (int i = 0; i < levels; ++i) {
        r += s;
        s *= 8;
    }
    return r;
}

// VoxelData::init, src/VoxelData.cpp:211-262
int choose_cache(Geometry &g, uint64_t mem, bool mesh) {
    int vW = round_pow2(g.dataW), vH = round_pow2(g.dataH), vD = round_pow2(g.dataD);
    g.side = std::max(vW, std::max(vH, vD));
    // side 1 (a one-voxel volume) gives an octree of depth 0, on which the
    // reference builder crashes; there is no output to reproduce
    if (g.side < 2) return fail(SVO_ERR_UNSUPPORTED, "a one-voxel volume has no octree (the reference builder crashes)");
    g.depth = log2i(uint32_t(g.side));
    uint64_t cellCost = 4 + (mesh ? 1 : 0);
    int best = -1;
    uint64_t smallest = UINT64_MAX;
    for (int i = g.depth; i >= 0; --i) {
        int lowBit = g.depth - i;
        uint64_t cost = grid_cells(i + 1) + grid_cells(lowBit) + (uint64_t(1) << uint64_t(lowBit * 3)) * cellCost;
        if (cost < mem) best = g.depth - i;
        smallest = std::min(smallest, cost);
    }
    if (best < 0)
        return fail(SVO_ERR_LIMIT, "memory budget below the smallest cache-block partition (" +
                                       std::to_string(smallest) + " bytes)");
    g.cacheLog = best;
    g.cache = 1 << best;
    g.blocksX = std::max(1, vW / g.cache);
    g.blocksY = std::max(1, vH / g.cache);
    g.blocksZ = std::max(1, vD / g.cache);
    return SVO_OK;
}

// findBestBlockPartition, src/PlyLoader.cpp:381-405
void best_partition(int &w, int &h, int &d, int threads) {
    int used = 1;
    while (used < threads) {
        int *mx, *md, *mn;
        if (w > h && w > d) mx = &w; else if (h > d) mx = &h; else mx = &d;
        int big = std::max(w, std::max(h, d)), small = std::min(w, std::min(h, d));
        if (w != small && w != big) md = &w; else if (h != small && h != big) md = &h; else md = &d;
        if (w < h && w < d) mn = &w; else if (h < d) mn = &h; else mn = &d;
        if ((*mx % 2) == 0) *mx /= 2;
        else if ((*md % 2) == 0) *md /= 2;
        else if ((*mn % 2) == 0) *mn /= 2;
        else break;
        used *= 2;
    }
}

// ---------------------------------------------------------------------------
// mesh kernels
// ---------------------------------------------------------------------------

__global__ void rescale_kernel(float *v9, int64_t nv, float lx, float ly, float lz, float factor) {
    int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    float *p = v9 + 9 * i;
    p[0] = (p[0] - lx) * factor;
    p[1] = (p[1] - ly) * factor;
    p[2] = (p[2] - lz) * factor;
}

// PlyLoader::readTriangles (src/PlyLoader.cpp:183-224) + Triangle ctor (:40-53):
// fan (e0, e[k-1], e[k]); face normal (v1-v0)x(v2-v0) normalised when the PLY
// has no vertex normals. Triangle record: 9 pos, 9 normal, 9 colour floats.
__global__ void triangles_kernel(const float *__restrict__ v9, const int32_t *__restrict__ fcount,
                                 const int64_t *__restrict__ fstart, const int64_t *__restrict__ tstart,
                                 const int32_t *__restrict__ fidx, int64_t nfaces, int hasNormals,
                                 float *__restrict__ pos, float *__restrict__ nrm, float *__restrict__ col) {
    int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= nfaces) return;
    const int c = fcount[f];
    if (c < 3) return;
    const int32_t *e = fidx + fstart[f];
    int64_t t = tstart[f];
    const int a = e[0];
    int b = e[1];
    for (int k = 2; k < c; ++k, ++t) {
        const int cc = e[k];
        const int vv[3] = {a, b, cc};
        for (int q = 0; q < 3; ++q)
            for (int s = 0; s < 3; ++s) {
                pos[9 * t + 3 * q + s] = v9[9 * int64_t(vv[q]) + s];
                nrm[9 * t + 3 * q + s] = v9[9 * int64_t(vv[q]) + 3 + s];
                col[9 * t + 3 * q + s] = v9[9 * int64_t(vv[q]) + 6 + s];
            }
        if (!hasNormals) {
            float e1[3], e2[3], cr[3], n[3];
            for (int s = 0; s < 3; ++s) {
                e1[s] = v9[9 * int64_t(b) + s] - v9[9 * int64_t(a) + s];
                e2[s] = v9[9 * int64_t(cc) + s] - v9[9 * int64_t(a) + s];
            }
            cross3(e1, e2, cr);
            normalize3(cr, n);
            for (int q = 0; q < 3; ++q)
                for (int s = 0; s < 3; ++s) nrm[9 * t + 3 * q + s] = n[s];
        }
        b = cc;
    }
}

// ---- PLY ingest on the device (binary little-endian bodies) ----
// A vertex record has a fixed size; a face record too once every face has the
// first face's vertex count (checked here, else the host parser takes over).
// Values convert as plyfile.c stores them: the file's type -> double -> float
// for vertex properties, -> int for face indices (src/PlyLoader.cpp:81-224).

struct PlyVertexLayout {
    int64_t stride;
    int off[9];
    int type[9];  // PType; 0 = property absent (default value)
};

struct PlyFaceLayout {
    int64_t stride;
    int countOff, countType;
    int idxOff, idxType;
    int c0;  // vertices per face
};

__device__ __forceinline__ double ply_value(const uint8_t *p, int type) {
    uint64_t u = 0;
    const int n = type == P_CHAR || type == P_UCHAR ? 1 : type == P_SHORT || type == P_USHORT ? 2 : type == P_DOUBLE ? 8 : 4;
    for (int k = 0; k < n; ++k) u |= uint64_t(p[k]) << (8 * k);
    switch (type) {
        case P_CHAR: return double(int8_t(uint8_t(u)));
        case P_UCHAR: return double(uint8_t(u));
        case P_SHORT: return double(int16_t(uint16_t(u)));
        case P_USHORT: return double(uint16_t(u));
        case P_INT: return double(int32_t(uint32_t(u)));
        case P_UINT: return double(uint32_t(u));
        case P_FLOAT: return double(__uint_as_float(uint32_t(u)));
        default: return __longlong_as_double(int64_t(u));
    }
}

__global__ void ply_vertices_kernel(const uint8_t *__restrict__ body, int64_t nv, PlyVertexLayout L,
                                    float *__restrict__ v9) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const uint8_t *rec = body + i * L.stride;
#pragma unroll
    for (int t = 0; t < 9; ++t)
        v9[9 * i + t] = L.type[t] ? float(ply_value(rec + L.off[t], L.type[t])) : (t < 6 ? 0.0f : 255.0f);
}

// flags: bit 0 a face with another vertex count, bit 1 an index outside [0, nv)
__global__ void ply_faces_kernel(const uint8_t *__restrict__ body, int64_t nf, PlyFaceLayout L, int64_t nv,
                                 int32_t *__restrict__ fcount, int64_t *__restrict__ fstart, int64_t *__restrict__ tstart,
                                 int32_t *__restrict__ fidx, uint32_t *__restrict__ flags) {
    const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint8_t *rec = body + f * L.stride;
    const double c = ply_value(rec + L.countOff, L.countType);
    if (!(c == double(L.c0))) {
        atomicOr(flags, 1u);
        return;
    }
    const int isz = L.idxType == P_CHAR || L.idxType == P_UCHAR ? 1 : L.idxType == P_SHORT || L.idxType == P_USHORT ? 2
                    : L.idxType == P_DOUBLE ? 8 : 4;
    for (int k = 0; k < L.c0; ++k) {
        const double v = ply_value(rec + L.idxOff + k * isz, L.idxType);
        const int32_t e = (v >= -2147483648.0 && v < 2147483648.0) ? int32_t(v) : INT32_MIN;
        if (e < 0 || int64_t(e) >= nv) atomicOr(flags, 2u);
        fidx[f * L.c0 + k] = e;
    }
    fcount[f] = L.c0;
    fstart[f] = f * L.c0;
    tstart[f] = f * (L.c0 - 2);
}

// Vertex position bounds as PlyLoader::readVertices scans them: per axis the
// smallest / largest value, ties going to the earliest vertex (std::min/max
// keep the first of equal values, which decides the sign of a zero bound);
// flags bit 2 marks a non-finite position. One (value, index) pair per block
// and bound; the host folds the blocks.
struct BoundPart {
    float v[6];
    int64_t i[6];
};

__device__ __forceinline__ bool bound_better(float a, int64_t ia, float b, int64_t ib, bool isMin) {
    // does (a, ia) replace (b, ib)?
    return isMin ? (a < b || (!(b < a) && ia < ib)) : (b < a || (!(a < b) && ia < ib));
}

__global__ void __launch_bounds__(256) ply_bounds_kernel(const float *__restrict__ v9, int64_t nv,
                                                         BoundPart *__restrict__ parts, uint32_t *__restrict__ flags) {
    __shared__ float sv[6][256];
    __shared__ int64_t si[6][256];
    float bv[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
    int64_t bi[6] = {INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX, INT64_MAX};
    bool bad = false;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nv; i += int64_t(gridDim.x) * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float x = v9[9 * i + k];
            if (!isfinite(x)) bad = true;
            if (bound_better(x, i, bv[k], bi[k], true)) { bv[k] = x; bi[k] = i; }
            if (bound_better(x, i, bv[3 + k], bi[3 + k], false)) { bv[3 + k] = x; bi[3 + k] = i; }
        }
    }
    if (bad) atomicOr(flags, 4u);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        sv[k][threadIdx.x] = bv[k];
        si[k][threadIdx.x] = bi[k];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const float a = sv[k][threadIdx.x + w];
                const int64_t ia = si[k][threadIdx.x + w];
                if (bound_better(a, ia, sv[k][threadIdx.x], si[k][threadIdx.x], k < 3)) {
                    sv[k][threadIdx.x] = a;
                    si[k][threadIdx.x] = ia;
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            parts[blockIdx.x].v[k] = sv[k][0];
            parts[blockIdx.x].i[k] = si[k][0];
        }
    }
}

// Destination of a mesh voxelized as PlyLoader::convertToVolume writes it.
struct VolumeSink {
    FILE *fp;  // the .voxel file, empty
};

struct MeshParams {
    int scale;  // resolution - 4
    int sb[3];
    int g[3];
    int p[3];
    int data[3];
    int blk[3];   // processBlock's block: the cubic cache block (VoxelData) or a convertToVolume slab
    float hx;     // 1.0f / scale (voxel edge in the unit cube)
    float hb[3];  // sub-block edge per axis in the unit cube
};

// PlyLoader::pointToGrid, src/PlyLoader.cpp:226-231 over the triangle's AABB
// (Triangle ctor, PlyLoader.cpp:40-53)
__device__ __forceinline__ void tri_grid_box(const float *tv, float scale, int lo[3], int hi[3]) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float l = fmin_ref(tv[k], fmin_ref(tv[3 + k], tv[6 + k]));
        const float u = fmax_ref(tv[k], fmax_ref(tv[3 + k], tv[6 + k]));
        lo[k] = cvt_i32(l * scale + 1.0f);
        hi[k] = cvt_i32(u * scale + 1.0f);
    }
}

__device__ __forceinline__ void load_tri(const float *__restrict__ pos, int64_t t, float tv[9]) {
#pragma unroll
    for (int k = 0; k < 9; ++k) tv[k] = pos[9 * t + k];
}

// PlyLoader::iterateOverlappingBlocks, src/PlyLoader.cpp:233-270. Count pass
// (out == nullptr) writes the number of cells per triangle; fill pass writes
// the flat cell indices at the triangle's offset. Flags bad[0] if a cell index
// falls outside the cell table (the reference would write out of bounds).
__global__ void bin_kernel(const float *__restrict__ pos, int64_t T, MeshParams mp, uint32_t *__restrict__ count,
                           const uint64_t *__restrict__ offset, uint32_t *__restrict__ cells,
                           uint32_t *__restrict__ tris, uint32_t *__restrict__ bad) {
    int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= T) return;
    float tv[9];
    load_tri(pos, t, tv);
    int l[3], u[3];
    tri_grid_box(tv, float(mp.scale), l, u);
    int lg[3], ug[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lg[k] = l[k] / mp.sb[k];
        ug[k] = (u[k] + 1) / mp.sb[k];
    }
    const uint64_t ncells = uint64_t(mp.g[0]) * mp.g[1] * mp.g[2];
    uint64_t at = cells ? offset[t] : 0;
    uint32_t n = 0;
    auto emit = [&](int64_t x, int64_t y, int64_t z) {
        int64_t f = x + int64_t(mp.g[0]) * (y + int64_t(mp.g[1]) * z);
        if (f < 0 || uint64_t(f) >= ncells) {
            atomicOr(bad, 1u);
            return;
        }
        if (cells) {
            cells[at + n] = uint32_t(f);
            tris[at + n] = uint32_t(t);
        }
        n++;
    };
    const int maxSide = max(ug[0] - lg[0], max(ug[1] - lg[1], ug[2] - lg[2]));
    if (maxSide > 0) {
        const float half[3] = {0.5f * mp.hb[0], 0.5f * mp.hb[1], 0.5f * mp.hb[2]};
        float ctr[3];
        ctr[2] = (float(lg[2]) + 0.5f) * mp.hb[2];
        for (int z = lg[2]; z <= ug[2]; ++z, ctr[2] += mp.hb[2]) {
            ctr[1] = (float(lg[1]) + 0.5f) * mp.hb[1];
            for (int y = lg[1]; y <= ug[1]; ++y, ctr[1] += mp.hb[1]) {
                ctr[0] = (float(lg[0]) + 0.5f) * mp.hb[0];
                for (int x = lg[0]; x <= ug[0]; ++x, ctr[0] += mp.hb[0])
                    if (tri_box(ctr, half, tv)) emit(x, y, z);
            }
        }
    } else {
        emit(lg[0], lg[1], lg[2]);
    }
    if (!cells) count[t] = n;
}

// Region of voxels that cell f covers in processBlock (src/PlyLoader.cpp:439-474,
// clip in triangleToVolume :343-352): [cell origin, min(cell end, data end)),
// only if the cell's cache block starts inside the volume.
__device__ __forceinline__ bool cell_region(uint32_t f, const MeshParams &mp, int lo[3], int hi[3]) {
    const int gx = int(f % uint32_t(mp.g[0]));
    const int gy = int((f / uint32_t(mp.g[0])) % uint32_t(mp.g[1]));
    const int gz = int(f / (uint32_t(mp.g[0]) * uint32_t(mp.g[1])));
    const int gg[3] = {gx, gy, gz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int blockOrigin = (gg[k] / mp.p[k]) * mp.blk[k];
        if (blockOrigin >= mp.data[k]) return false;
        lo[k] = gg[k] * mp.sb[k];
        hi[k] = min((gg[k] + 1) * mp.sb[k], mp.data[k]) - 1;
    }
    return true;
}

// Clipped voxel box of (cell, tri) pair i; false if empty.
__device__ __forceinline__ bool pair_box(uint32_t f, const float tv[9], const MeshParams &mp, int l[3], int u[3]) {
    int cl[3], cu[3];
    if (!cell_region(f, mp, cl, cu)) return false;
    tri_grid_box(tv, float(mp.scale), l, u);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        l[k] = max(l[k], cl[k]);
        u[k] = min(u[k], cu[k]);
        if (l[k] > u[k]) return false;
    }
    return true;
}

// A pair's clipped voxel box, kept for its rows.
struct PairBox {
    int l[3], u[3];
};

__global__ void pair_rows_kernel(const uint32_t *__restrict__ cells, const uint32_t *__restrict__ tris, uint64_t P,
                                 const float *__restrict__ pos, MeshParams mp, uint32_t *__restrict__ rows,
                                 PairBox *__restrict__ boxes) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float tv[9];
    load_tri(pos, tris[i], tv);
    PairBox b;
    const bool any = pair_box(cells[i], tv, mp, b.l, b.u);
    rows[i] = any ? uint32_t((b.u[1] - b.l[1] + 1) * (b.u[2] - b.l[2] + 1)) : 0u;
    if (any) boxes[i] = b;
}

// rowPair[r] = the pair of row r (each pair's rows are contiguous from
// rowStart[i]). A wave takes 64 pairs and writes their rows one pair after
// another with all lanes (coalesced stores, no lane waiting on a long pair).
__global__ void __launch_bounds__(kThreads) row_pairs_kernel(const uint64_t *__restrict__ rowStart, uint64_t P,
                                                             uint32_t *__restrict__ rowPair) {
    const int lane = threadIdx.x & 63;
    const uint64_t base = (uint64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6)) * 64;
    if (base >= P) return;
    const uint64_t i = base + uint64_t(lane);
    const uint64_t start = i < P ? rowStart[i] : 0, end = i < P ? rowStart[i + 1] : 0;
    const int n = int(min(uint64_t(64), P - base));
    for (int j = 0; j < n; ++j) {
        const uint64_t s = __shfl(start, j), e = __shfl(end, j);
        for (uint64_t r = s + uint64_t(lane); r < e; r += 64) rowPair[r] = uint32_t(base + uint64_t(j));
    }
}

// Cache blocks with a non-empty cell list (PlyLoader::isBlockEmpty,
// src/PlyLoader.cpp:476-493), indexed by the block's Morton key.
__global__ void block_flags_kernel(const uint32_t *__restrict__ cells, uint64_t P, MeshParams mp,
                                   uint32_t *__restrict__ flags) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const uint32_t f = cells[i];
    const uint32_t gx = f % uint32_t(mp.g[0]);
    const uint32_t gy = (f / uint32_t(mp.g[0])) % uint32_t(mp.g[1]);
    const uint32_t gz = f / (uint32_t(mp.g[0]) * uint32_t(mp.g[1]));
    const uint32_t b[3] = {gx / uint32_t(mp.p[0]), gy / uint32_t(mp.p[1]), gz / uint32_t(mp.p[2])};
    for (int k = 0; k < 3; ++k)
        if (int64_t(b[k]) * mp.blk[k] >= mp.data[k]) return;
    flags[morton3(b[0], b[1], b[2])] = 1u;
}

// A row's voxels for the emit pass, recorded by the count pass: the first
// voxel's x, the row's y and z, the triangle, and which of the first 64
// voxels of the row overlap it (rows longer than 64 voxels are marked
// x0 = kLongRow and tested again by the emit pass).
struct RowHits {
    uint64_t mask;
    uint32_t x0, y, z, t;
};
constexpr uint32_t kLongRow = 0xFFFFFFFFu;

// The voxel loop of PlyLoader::triangleToVolume (src/PlyLoader.cpp:353-373)
// for one (y, z) row: centres start at (l - 0.5)*hx and step by +hx, exactly as
// the reference accumulates them. Count pass (keys == nullptr: the row's hit
// count and its RowHits) or emit pass (keys at hitStart).
__device__ __forceinline__ void row_voxels(uint64_t r, const uint32_t *__restrict__ rowPair,
                                           const PairBox *__restrict__ boxes, const uint32_t *__restrict__ tris,
                                           const uint64_t *__restrict__ rowStart, const float *__restrict__ pos,
                                           const MeshParams &mp, int tribits, uint32_t *__restrict__ hitCount,
                                           RowHits *__restrict__ hits, const uint64_t *__restrict__ hitStart,
                                           uint64_t *__restrict__ keys) {
    const uint32_t i = rowPair[r];
    const uint32_t t = tris[i];
    float tv[9];
    load_tri(pos, t, tv);
    const PairBox b = boxes[i];
    const int *l = b.l, *u = b.u;
    const uint32_t local = uint32_t(r - rowStart[i]);  // below the pair's row count (32 bits)
    const uint32_t ny = uint32_t(u[1] - l[1] + 1);
    const int y = l[1] + int(local % ny);
    const int z = l[2] + int(local / ny);
    const float hx = mp.hx;
    const float half[3] = {0.5f * hx, 0.5f * hx, 0.5f * hx};
    float ctr[3];
    ctr[2] = (float(l[2]) - 0.5f) * hx;
    for (int k = l[2]; k < z; ++k) ctr[2] += hx;
    ctr[1] = (float(l[1]) - 0.5f) * hx;
    for (int k = l[1]; k < y; ++k) ctr[1] += hx;
    ctr[0] = (float(l[0]) - 0.5f) * hx;
    uint32_t n = 0;
    uint64_t mask = 0;
    uint64_t at = keys ? hitStart[r] : 0;
    for (int x = l[0]; x <= u[0]; ++x, ctr[0] += hx) {
        if (tri_box(ctr, half, tv)) {
            if (keys) keys[at + n] = (morton3(uint32_t(x), uint32_t(y), uint32_t(z)) << tribits) | uint64_t(t);
            else if (x - l[0] < 64) mask |= uint64_t(1) << (x - l[0]);
            n++;
        }
    }
    if (!keys) {
        hitCount[r] = n;
        const bool longRow = u[0] - l[0] >= 64;
        hits[r] = RowHits{mask, longRow ? kLongRow : uint32_t(l[0]), uint32_t(y), uint32_t(z), t};
    }
}

__global__ void row_count_kernel(const uint32_t *__restrict__ rowPair, const PairBox *__restrict__ boxes,
                                 const uint32_t *__restrict__ tris, const uint64_t *__restrict__ rowStart, uint64_t R,
                                 const float *__restrict__ pos, MeshParams mp, int tribits,
                                 uint32_t *__restrict__ hitCount, RowHits *__restrict__ hits) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R) return;
    row_voxels(r, rowPair, boxes, tris, rowStart, pos, mp, tribits, hitCount, hits, nullptr, nullptr);
}

// Emit pass: a row of at most 64 voxels writes its keys from the count pass's
// record (the same keys, in the same ascending x order); a longer row runs
// the voxel loop again.
__global__ void row_emit_kernel(const uint32_t *__restrict__ rowPair, const PairBox *__restrict__ boxes,
                                const uint32_t *__restrict__ tris, const uint64_t *__restrict__ rowStart, uint64_t R,
                                const float *__restrict__ pos, MeshParams mp, int tribits,
                                const RowHits *__restrict__ hits, const uint64_t *__restrict__ hitStart,
                                uint64_t *__restrict__ keys) {
    const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const RowHits h = hits[r];
    if (h.x0 == kLongRow) {
        row_voxels(r, rowPair, boxes, tris, rowStart, pos, mp, tribits, nullptr, nullptr, hitStart, keys);
        return;
    }
    uint64_t at = hitStart[r];
    for (uint64_t m = h.mask; m; m &= m - 1) {
        const uint32_t x = h.x0 + uint32_t(__builtin_ctzll(m));
        keys[at++] = (morton3(x, h.y, h.z) << tribits) | uint64_t(h.t);
    }
}

// PlyLoader::writeTriangleCell, src/PlyLoader.cpp:296-336, folded over one
// voxel's triangles in ascending index order, starting from the buffer state
// (word, count) the voxel's cell holds when its block is processed (0 for a
// fresh cell). The voxel centre each triangle used is rebuilt with the same
// +hx steps from that triangle's clipped start.
// One triangle t of voxel vx folded into (word, count).
__device__ __forceinline__ void fold_step(int64_t t, const int vx[3], const float *__restrict__ pos,
                                          const float *__restrict__ nrm, const float *__restrict__ col,
                                          const MeshParams &mp, uint32_t &word, uint32_t &count) {
    const float hx = mp.hx;
    {
        float tv[9];
        load_tri(pos, t, tv);
        int l[3], u[3];
        tri_grid_box(tv, float(mp.scale), l, u);
        float p[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int start = max(l[a], (vx[a] / mp.sb[a]) * mp.sb[a]);
            float c = (float(start) - 0.5f) * hx;
            for (int q = start; q < vx[a]; ++q) c += hx;
            p[a] = c;
        }
        // Triangle::barycentric, src/PlyLoader.cpp:55-64
        float f1[3], f2[3], f3[3], ea[3], eb[3], cr[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            f1[a] = tv[a] - p[a];
            f2[a] = tv[3 + a] - p[a];
            f3[a] = tv[6 + a] - p[a];
            ea[a] = tv[a] - tv[3 + a];
            eb[a] = tv[a] - tv[6 + a];
        }
        cross3(ea, eb, cr);
        const float area = length3(cr);
        cross3(f2, f3, cr);
        float l1 = length3(cr) / area;
        cross3(f3, f1, cr);
        float l2 = length3(cr) / area;
        const bool inside = l1 >= 0.0f && l2 >= 0.0f && l1 + l2 <= 1.0f;
        if (!inside) {
            l1 = fmin_ref(fmax_ref(l1, 0.0f), 1.0f);
            l2 = fmin_ref(fmax_ref(l2, 0.0f), 1.0f);
            const float tau = l1 + l2;
            if (tau > 1.0f) {
                l1 /= tau;
                l2 /= tau;
            }
        }
        const float l3 = 1.0f - l1 - l2;
        float nb[3], n[3], cc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            nb[a] = nrm[9 * t + a] * l1 + nrm[9 * t + 3 + a] * l2 + nrm[9 * t + 6 + a] * l3;
            cc[a] = col[9 * t + a] * l1 + col[9 * t + 3 + a] * l2 + col[9 * t + 6 + a] * l3;
        }
        normalize3(nb, n);
        const float shade = (cc[0] * 0.2126f + cc[1] * 0.7152f + cc[2] * 0.0722f) * (1.0f / 256.0f);
        if (word == 0) {
            count = 1;
            word = compress_material(n, shade);
        } else {
            const float cnt = float(count);
            const float curRatio = cnt / (cnt + 1.0f);
            const float newRatio = 1.0f - curRatio;
            float cn[3], cs;
            decompress_material(word, cn, cs);
            float nn[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) nn[a] = cn[a] * curRatio + n[a] * newRatio;
            const float ns = cs * curRatio + shade * newRatio;
            if (nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2] < 1e-3f) {
                nn[0] = cn[0];
                nn[1] = cn[1];
                nn[2] = cn[2];
            }
            word = compress_material(nn, ns);
            count = min(count + 1u, 255u);
        }
    }
}

__device__ __forceinline__ void fold_voxel(const uint64_t *__restrict__ keys, uint64_t b, uint64_t e, int tribits,
                                           const float *__restrict__ pos, const float *__restrict__ nrm,
                                           const float *__restrict__ col, const MeshParams &mp, uint32_t &word,
                                           uint32_t &count) {
    const uint64_t m = keys[b] >> tribits;
    const int vx[3] = {int(compact3(m)), int(compact3(m >> 1)), int(compact3(m >> 2))};
    const uint64_t triMask = (uint64_t(1) << tribits) - 1;
    for (uint64_t k = b; k < e; ++k) fold_step(int64_t(keys[k] & triMask), vx, pos, nrm, col, mp, word, count);
}

// Voxels per wave in fold_kernel: a lane that has folded its voxel takes the
// chunk's next one, so a wave's lanes stay busy while their voxels' triangle
// counts differ (one lane per voxel would run every lane as long as the
// voxel with the most triangles).
constexpr int kFoldChunk = 512;
constexpr uint64_t kFoldVoxelsPerBlock = uint64_t(kFoldChunk) * (kThreads / 64);

__global__ void __launch_bounds__(kThreads) fold_kernel(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ segStart, uint64_t S,
                            int tribits, const float *__restrict__ pos, const float *__restrict__ nrm,
                            const float *__restrict__ col, MeshParams mp, uint64_t *__restrict__ vkey,
                            uint32_t *__restrict__ vword, uint8_t *__restrict__ vcount) {
    const int lane = threadIdx.x & 63;
    const uint64_t base = (uint64_t(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6)) * kFoldChunk;
    if (base >= S) return;
    const uint64_t end = min(base + uint64_t(kFoldChunk), S);
    const uint64_t triMask = (uint64_t(1) << tribits) - 1;
    uint64_t next = base;  // the chunk's next unassigned voxel (wave-uniform)
    uint64_t v = 0, k = 0, e = 0;
    bool busy = false;
    int vx[3] = {0, 0, 0};
    uint32_t word = 0, count = 0;
    for (;;) {
        // lanes without a voxel take the next ones, in lane order
        const uint64_t idle = __ballot(!busy);
        if (!busy) {
            const uint64_t w = next + uint64_t(__popcll(idle & ((uint64_t(1) << lane) - 1)));
            if (w < end) {
                v = w;
                k = segStart[w];
                e = segStart[w + 1];
                const uint64_t m = keys[k] >> tribits;
                vx[0] = int(compact3(m));
                vx[1] = int(compact3(m >> 1));
                vx[2] = int(compact3(m >> 2));
                word = 0;
                count = 0;
                busy = true;
            }
        }
        next = min(next + uint64_t(__popcll(idle)), end);
        if (!__ballot(busy)) break;
        if (busy) {
            fold_step(int64_t(keys[k] & triMask), vx, pos, nrm, col, mp, word, count);
            if (++k == e) {
                vkey[v] = keys[e - 1] >> tribits;
                vword[v] = word;
                vcount[v] = uint8_t(count);
                busy = false;
            }
        }
    }
}

// One slab of PlyLoader::convertToVolume (src/PlyLoader.cpp:521-530): every
// voxel whose z lies in [z0, z0 + sliceZ) is folded from the state its cell of
// the slab buffer holds (word, count), which carries over from earlier slabs
// because the reference never clears the buffer or its counts.
__global__ void slab_fold_kernel(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ segStart, uint64_t S,
                                 int tribits, const float *__restrict__ pos, const float *__restrict__ nrm,
                                 const float *__restrict__ col, MeshParams mp, int z0, int sliceZ,
                                 uint32_t *__restrict__ word, uint32_t *__restrict__ count,
                                 uint32_t *__restrict__ bad) {
    uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint64_t m = keys[segStart[s]] >> tribits;
    const int x = int(compact3(m)), y = int(compact3(m >> 1)), z = int(compact3(m >> 2));
    if (z < z0 || z >= z0 + sliceZ) return;
    if (x >= mp.data[0] || y >= mp.data[1]) {
        atomicOr(bad, 1u);
        return;
    }
    const uint64_t cell = uint64_t(x) + uint64_t(mp.data[0]) * (uint64_t(y) + uint64_t(mp.data[1]) * uint64_t(z - z0));
    uint32_t w = word[cell], c = count[cell];
    fold_voxel(keys, segStart[s], segStart[s + 1], tribits, pos, nrm, col, mp, w, c);
    word[cell] = w;
    count[cell] = c;
}

// Folds voxel `voxel[i]` again from a carried buffer state (init word/count).
__global__ void refold_kernel(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ segStart,
                              const uint64_t *__restrict__ voxel, const uint32_t *__restrict__ initWord,
                              const uint32_t *__restrict__ initCount, uint64_t n, int tribits,
                              const float *__restrict__ pos, const float *__restrict__ nrm,
                              const float *__restrict__ col, MeshParams mp, uint32_t *__restrict__ outWord,
                              uint32_t *__restrict__ outCount) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t s = voxel[i];
    uint32_t word = initWord[i], count = initCount[i];
    fold_voxel(keys, segStart[s], segStart[s + 1], tribits, pos, nrm, col, mp, word, count);
    outWord[i] = word;
    outCount[i] = count;
}

// ---------------------------------------------------------------------------
// file-source kernels
// ---------------------------------------------------------------------------

__global__ void grid_keys_kernel(const uint32_t *__restrict__ grid, int w, int h, int d, uint32_t *__restrict__ flag) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(w) * h * d) return;
    flag[i] = grid[i] != 0 ? 1u : 0u;
}

__global__ void grid_compact_kernel(const uint32_t *__restrict__ grid, int w, int h, int d,
                                    const uint32_t *__restrict__ flag, const uint64_t *__restrict__ at,
                                    uint64_t *__restrict__ key, uint32_t *__restrict__ word) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= uint64_t(w) * h * d || !flag[i]) return;
    const uint32_t x = uint32_t(i % uint64_t(w));
    const uint32_t y = uint32_t((i / uint64_t(w)) % uint64_t(h));
    const uint32_t z = uint32_t(i / (uint64_t(w) * h));
    key[at[i]] = morton3(x, y, z);
    word[at[i]] = grid[i];
}

// ---------------------------------------------------------------------------
// occupancy-table behaviour shared by both sources
// ---------------------------------------------------------------------------

struct BlockParams {
    int data[3];
    int cache;
    int cacheLog;
};

// Voxel on the last slice of an odd-depth cache block: never scanned by
// VoxelData::buildLowLut (src/VoxelData.cpp:163-166), hence never emitted.
__device__ __forceinline__ bool unscanned(uint32_t z, const BlockParams &bp) {
    const int z0 = int(z >> bp.cacheLog) << bp.cacheLog;
    const int bd = min(bp.cache, bp.data[2] - z0);
    return (bd & 1) && int(z) == z0 + bd - 1;
}

// keep[i] = voxel i is emitted; flags cache blocks holding any voxel (the
// file source's top table, VoxelData::buildTopLutBlock :120-145) and, for the
// mesh source, marks blocks that hold an unscanned voxel.
__global__ void keep_kernel(const uint64_t *__restrict__ key, const uint32_t *__restrict__ word, uint64_t V,
                            BlockParams bp, int markBlocks, uint32_t *__restrict__ blockFlags,
                            uint32_t *__restrict__ hazardBlocks, uint32_t *__restrict__ keep) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const uint64_t m = key[i];
    const uint32_t z = compact3(m >> 2);
    const bool nz = word[i] != 0;
    const uint64_t blk = m >> (3 * bp.cacheLog);
    if (nz && markBlocks) blockFlags[blk] = 1u;
    const bool drop = nz && unscanned(z, bp);
    if (drop) hazardBlocks[blk] = 1u;
    keep[i] = (nz && !drop) ? 1u : 0u;
}

__global__ void compact_kernel(const uint64_t *__restrict__ key, const uint32_t *__restrict__ word, uint64_t V,
                               const uint32_t *__restrict__ keep, const uint64_t *__restrict__ at,
                               uint64_t *__restrict__ okey, uint32_t *__restrict__ oword) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= V || !keep[i]) return;
    okey[at[i]] = key[i];
    oword[at[i]] = word[i];
}

// ---------------------------------------------------------------------------
// levels, sizes, placement
// ---------------------------------------------------------------------------

// Level kernels read the level's node count from a device word (nptr): the
// count comes from the previous level's pass, so no host round trip sizes
// them; grids are sized by a host upper bound.

// Level that also takes the cache blocks the occupancy table marks: parents
// are the flagged cells of a dense Morton-indexed table.
__global__ void dense_mark_kernel(const uint64_t *__restrict__ child, const uint64_t *__restrict__ nptr,
                                  uint32_t *__restrict__ dense) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    dense[child[i] >> 3] = 1u;
}

__global__ void dense_compact_kernel(const uint32_t *__restrict__ dense, uint64_t cells,
                                     const uint64_t *__restrict__ at, uint64_t *__restrict__ pkey,
                                     uint64_t *__restrict__ slot) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    if (dense[i]) pkey[at[i]] = i;
    slot[i] = at[i];
}

__global__ void dense_children_kernel(const uint64_t *__restrict__ child, const uint64_t *__restrict__ nptr,
                                      const uint64_t *__restrict__ slot, uint64_t *__restrict__ pstart,
                                      uint32_t *__restrict__ pmask) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= *nptr) return;
    const uint64_t pk = child[i] >> 3;
    const uint64_t pi = slot[pk];
    if (i == 0 || (child[i - 1] >> 3) != pk) pstart[pi] = i;
    atomicOr(&pmask[pi], 1u << uint32_t(child[i] & 7));
}

// Subtree sizes, bottom-up (VoxelOctree::buildOctree, src/VoxelOctree.cpp:199-265
// and ChunkedAllocator insertions, ChunkedAllocator.hpp:72-107): a node with n
// children occupies n*(1+large) words for its child block plus its children's
// subtrees S. Child i's offset from its own slot, before the node's far words,
// is n - i + sum_{j<i} S(c_j); the node needs far pointers (large) if any
// exceeds 0x3FFF, and then every child slot takes a second word.
__global__ void size_kernel(const uint32_t *__restrict__ mask, const uint64_t *__restrict__ start,
                            const uint64_t *__restrict__ nptr, int leafParents, const uint64_t *__restrict__ childS,
                            uint64_t *__restrict__ S, uint8_t *__restrict__ large) {
    uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= *nptr) return;
    const uint32_t c = __popc(mask[k]);
    if (leafParents || c == 0) {
        S[k] = c;
        large[k] = 0;
        return;
    }
    uint64_t acc = 0, maxOff = 0;
    const uint64_t s0 = start[k];
    for (uint32_t i = 0; i < c; ++i) {
        const uint64_t off = uint64_t(c - i) + acc;
        if (off > maxOff) maxOff = off;
        acc += childS[s0 + i];
    }
    const bool lg = maxOff > 0x3FFF;
    large[k] = lg ? 1 : 0;
    S[k] = uint64_t(c) * (lg ? 2 : 1) + acc;
}

// Placement, top-down: a node at `pos` whose child block starts at `B` puts
// child i at B + i*(1+large); that child's own block starts after the node's
// whole child block and the subtrees of children 0..i-1.
__global__ void place_kernel(const uint32_t *__restrict__ mask, const uint64_t *__restrict__ start,
                             const uint64_t *__restrict__ nptr, const uint64_t *__restrict__ B,
                             const uint8_t *__restrict__ large, const uint64_t *__restrict__ childS, int childLeaves,
                             uint64_t *__restrict__ childPos, uint64_t *__restrict__ childB,
                             uint8_t *__restrict__ childFar) {
    uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= *nptr) return;
    const uint32_t c = __popc(mask[k]);
    if (c == 0) return;
    const uint64_t per = large[k] ? 2 : 1;
    const uint64_t b = B[k], s0 = start[k];
    uint64_t acc = 0;
    for (uint32_t i = 0; i < c; ++i) {
        childPos[s0 + i] = b + i * per;
        if (!childLeaves) {
            childB[s0 + i] = b + uint64_t(c) * per + acc;
            childFar[s0 + i] = large[k];
            acc += childS[s0 + i];
        }
    }
}

// Descriptor word (and far word) of every node of one level.
// Positions are checked against the word count W; a violation (which the
// size/placement arithmetic excludes) is reported instead of written.
__global__ void write_nodes_kernel(const uint32_t *__restrict__ mask, const uint64_t *__restrict__ nptr,
                                   const uint64_t *__restrict__ pos, const uint64_t *__restrict__ B,
                                   const uint8_t *__restrict__ large, const uint8_t *__restrict__ far, int leafParents,
                                   uint32_t *__restrict__ out, uint64_t W, uint32_t *__restrict__ bad) {
    uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= *nptr) return;
    const uint32_t m = mask[k];
    uint32_t w = (m << 8) | (leafParents ? 0u : m);
    if (large[k]) w |= 0x10000u;
    const uint64_t p = pos[k];
    const uint64_t off = B[k] - p;
    // the traversal reads child slot B + (non-leaf children before it) * (1 + large)
    // or B + (children before it) for a leaf: the block must follow the node
    // and end inside the array
    const uint64_t kids = uint64_t(__popc(m)), extent = leafParents ? kids : kids * (large[k] ? 2u : 1u);
    if (p >= W || (far[k] && p + 1 >= W) || (kids && (B[k] <= p || B[k] > W || extent > W - B[k]))) {
        atomicOr(bad, 1u);
        return;
    }
    if (far[k]) {
        w |= 0x20000u | uint32_t((off >> 32) << 18);
        out[p + 1] = uint32_t(off);
    } else {
        w |= uint32_t(off << 18);
    }
    out[p] = w;
}

__global__ void write_leaves_kernel(const uint32_t *__restrict__ word, uint64_t n, const uint64_t *__restrict__ pos,
                                    uint32_t *__restrict__ out, uint64_t W, uint32_t *__restrict__ bad) {
    uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (pos[k] >= W) {
        atomicOr(bad, 1u);
        return;
    }
    out[pos[k]] = word[k];
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// The root level of an empty tree or the root's placement: one node.
__global__ void root_init_kernel(uint64_t *key, uint64_t *start, uint32_t *mask, uint64_t *cnt) {
    *key = 0;
    *start = 0;
    *mask = 0;
    *cnt = 1;
}

__global__ void root_place_kernel(uint64_t *pos, uint64_t *B, uint8_t *far) {
    *pos = 0;
    *B = 1;
    *far = 0;
}

// {root subtree size, root level count} for the host, in one read.
__global__ void root_report_kernel(const uint64_t *S0, const uint64_t *cnt0, uint64_t *out) {
    out[0] = *S0;
    out[1] = *cnt0;
}

// ---------------------------------------------------------------------------
// signed-distance scene source (procedural replacement of the raw .voxel
// route for SDF scenes; occupancy/normal rule of the workload definition)
// ---------------------------------------------------------------------------

struct SdfPrim {
    int kind;  // 0 sphere (radius r1), 1 torus (major r1, minor r2, unit axis)
    float cx, cy, cz, r1, r2, ax, ay, az;
};

constexpr int kMaxSdfPrims = 64;

struct SdfScene {
    SdfPrim p[kMaxSdfPrims];
    int count;
    int n;         // voxels per axis
    float thr;     // |sdf| <= thr is occupied: 0.5*sqrt(3)/n
    float h;       // central-difference step 1/n
    float shade;
};

__device__ __forceinline__ float sdf_eval(const SdfScene &sc, float x, float y, float z) {
    float d = INFINITY;
    for (int i = 0; i < sc.count; ++i) {
        const SdfPrim &q = sc.p[i];
        const float qx = x - q.cx, qy = y - q.cy, qz = z - q.cz;
        const float qq = qx * qx + qy * qy + qz * qz;
        float di;
        if (q.kind == 0) {
            di = sqrtf(qq) - q.r1;
        } else {
            const float hh = qx * q.ax + qy * q.ay + qz * q.az;
            const float rad = sqrtf(fmaxf(qq - hh * hh, 0.0f)) - q.r1;
            di = sqrtf(rad * rad + hh * hh) - q.r2;
        }
        d = fminf(d, di);
    }
    return d;
}

__device__ __forceinline__ float vox_center(int v, int n) { return (float(v) + 0.5f) / float(n); }

// Bricks of 8^3 voxels whose centre lies farther from the surface than the
// brick's half diagonal plus the shell width (and a rounding margin) hold no
// occupied voxel: the union of exact sphere/torus distances is 1-Lipschitz.
__global__ void sdf_bricks_kernel(SdfScene sc, uint64_t nBricks, int bricksPerAxis, uint32_t *__restrict__ list,
                                  uint32_t *__restrict__ count) {
    uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nBricks) return;
    const int bx = int(b % uint64_t(bricksPerAxis));
    const int by = int((b / uint64_t(bricksPerAxis)) % uint64_t(bricksPerAxis));
    const int bz = int(b / (uint64_t(bricksPerAxis) * bricksPerAxis));
    const float inv = 1.0f / float(sc.n);
    const float d = sdf_eval(sc, (bx * 8 + 4) * inv, (by * 8 + 4) * inv, (bz * 8 + 4) * inv);
    const float reach = (4.0f * 1.7320508f + 2.0f) * inv + sc.thr;
    if (fabsf(d) <= reach) list[atomicAdd(count, 1u)] = uint32_t(b);
}

// One wave per candidate brick, 8 voxels per lane. Emits (morton key, word)
// for occupied voxels; normal = normalised central-difference gradient,
// packed with compressMaterial (src/Util.hpp:64-84); a zero word becomes 1.
__global__ void __launch_bounds__(64) sdf_voxels_kernel(SdfScene sc, const uint32_t *__restrict__ list, int bricksPerAxis,
                                                        uint64_t *__restrict__ key, uint32_t *__restrict__ word,
                                                        uint32_t *__restrict__ count, uint64_t cap,
                                                        uint32_t *__restrict__ overflow) {
    const uint32_t b = list[blockIdx.x];
    const int bx = int(b % uint32_t(bricksPerAxis));
    const int by = int((b / uint32_t(bricksPerAxis)) % uint32_t(bricksPerAxis));
    const int bz = int(b / (uint32_t(bricksPerAxis) * bricksPerAxis));
    for (int k = 0; k < 8; ++k) {
        const int local = threadIdx.x + 64 * k;  // 0..511
        const int x = bx * 8 + (local & 7), y = by * 8 + ((local >> 3) & 7), z = bz * 8 + (local >> 6);
        const float px = vox_center(x, sc.n), py = vox_center(y, sc.n), pz = vox_center(z, sc.n);
        const float d = sdf_eval(sc, px, py, pz);
        if (!(fabsf(d) <= sc.thr)) continue;
        float g[3];
        g[0] = sdf_eval(sc, px + sc.h, py, pz) - sdf_eval(sc, px - sc.h, py, pz);
        g[1] = sdf_eval(sc, px, py + sc.h, pz) - sdf_eval(sc, px, py - sc.h, pz);
        g[2] = sdf_eval(sc, px, py, pz + sc.h) - sdf_eval(sc, px, py, pz - sc.h);
        float nrm[3];
        normalize3(g, nrm);
        uint32_t w = compress_material(nrm, sc.shade);
        if (w == 0) w = 1;
        const uint32_t at = atomicAdd(count, 1u);
        if (at >= cap) {
            atomicOr(overflow, 1u);
            continue;
        }
        key[at] = morton3(uint32_t(x), uint32_t(y), uint32_t(z));
        word[at] = w;
    }
}

__global__ void scatter_dense_kernel(const uint64_t *__restrict__ key, const uint32_t *__restrict__ word, uint64_t n,
                                     int side, uint32_t *__restrict__ grid) {
    uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t m = key[i];
    const uint64_t x = compact3(m), y = compact3(m >> 1), z = compact3(m >> 2);
    grid[x + uint64_t(side) * (y + uint64_t(side) * z)] = word[i];
}

// ---------------------------------------------------------------------------
// fBm heightfield terrain source (procedural replacement of the raw .voxel
// route for terrain; every float operation separately rounded, in the order
// of the definition: lattice hash -> smoothstep bilinear value noise -> 8
// octaves, gain 0.5, lacunarity 2, base frequency 4 -> integer heights ->
// exposed shell with the gradient normal)
// ---------------------------------------------------------------------------

constexpr int kTerrainOctaves = 8;

__device__ __forceinline__ float terrain_lattice(int i, int k, uint32_t s) {
    uint32_t h = (uint32_t(i) * 0x8DA6B343u) ^ (uint32_t(k) * 0xD8163841u) ^ (s * 0x9E3779B9u);
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return float(h >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float terrain_smooth(float c) { return (c * c) * (3.0f - 2.0f * c); }

// normalised fBm height in [0, 1) of column (x, z)
__device__ float terrain_h(int x, int z, int n, uint32_t seed) {
    const float u = (float(x) + 0.5f) / float(n), w = (float(z) + 0.5f) / float(n);
    float total = 0.0f, norm = 0.0f, amp = 1.0f, freq = 4.0f;
    for (int o = 0; o < kTerrainOctaves; ++o) {
        const float pu = u * freq, pw = w * freq;
        const float fu = floorf(pu), fw = floorf(pw);
        const float cu = terrain_smooth(pu - fu), cw = terrain_smooth(pw - fw);
        const int iu = int(fu), iw = int(fw);
        const uint32_t s = seed + uint32_t(o);
        const float v00 = terrain_lattice(iu, iw, s), v10 = terrain_lattice(iu + 1, iw, s);
        const float v01 = terrain_lattice(iu, iw + 1, s), v11 = terrain_lattice(iu + 1, iw + 1, s);
        const float a0 = v00 + (v10 - v00) * cu;
        const float a1 = v01 + (v11 - v01) * cu;
        total = total + amp * (a0 + (a1 - a0) * cw);
        norm = norm + amp;
        amp = amp * 0.5f;
        freq = freq * 2.0f;
    }
    return total / norm;
}

// per column (index z*n + x): h and the integer height n/8 + int(h*(n/2))
__global__ void terrain_heights_kernel(int n, uint32_t seed, float *__restrict__ hOut, int32_t *__restrict__ H) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= uint64_t(n) * n) return;
    const int x = int(c % uint64_t(n)), z = int(c / uint64_t(n));
    const float h = terrain_h(x, z, n, seed);
    hOut[c] = h;
    H[c] = n / 8 + int(h * float(n / 2));
}

// lowest exposed voxel of a column: a voxel is exposed when a 4-neighbour
// column inside the grid is lower, or it is the top one
__device__ __forceinline__ int terrain_lo(const int32_t *__restrict__ H, int n, int x, int z) {
    const int32_t top = H[uint64_t(z) * n + x];
    int32_t m = INT32_MAX;
    if (x > 0) m = min(m, H[uint64_t(z) * n + x - 1]);
    if (x + 1 < n) m = min(m, H[uint64_t(z) * n + x + 1]);
    if (z > 0) m = min(m, H[uint64_t(z - 1) * n + x]);
    if (z + 1 < n) m = min(m, H[uint64_t(z + 1) * n + x]);
    return m == INT32_MAX ? top : min(top, m + 1);
}

__global__ void terrain_count_kernel(int n, const int32_t *__restrict__ H, uint32_t *__restrict__ count) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= uint64_t(n) * n) return;
    const int x = int(c % uint64_t(n)), z = int(c / uint64_t(n));
    count[c] = uint32_t(H[c] - terrain_lo(H, n, x, z) + 1);
}

// one column per thread: normal from the clamped central difference of
// hv = h*(n/2), word = compressMaterial(normal, 0.35 + 0.65*h), keys of the
// exposed voxels at the column's scanned offset
// A column's voxels all carry the column's word: the keys go out alone (the
// sort then moves 8 bytes per voxel instead of 12) and the column's word goes
// to colWord, from which terrain_words_kernel fills the sorted words.
__global__ void terrain_emit_kernel(int n, const int32_t *__restrict__ H, const float *__restrict__ hh,
                                    const uint64_t *__restrict__ at, uint64_t *__restrict__ key,
                                    uint32_t *__restrict__ colWord) {
    const uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= uint64_t(n) * n) return;
    const int x = int(c % uint64_t(n)), z = int(c / uint64_t(n));
    const float half = float(n / 2);
    const int xm = max(x - 1, 0), xp = min(x + 1, n - 1), zm = max(z - 1, 0), zp = min(z + 1, n - 1);
    const float gx = (hh[uint64_t(z) * n + xp] * half - hh[uint64_t(z) * n + xm] * half) * 0.5f;
    const float gz = (hh[uint64_t(zp) * n + x] * half - hh[uint64_t(zm) * n + x] * half) * 0.5f;
    const float a[3] = {-gx, 1.0f, -gz};
    float nrm[3];
    normalize3(a, nrm);
    uint32_t w = compress_material(nrm, 0.35f + 0.65f * hh[c]);
    if (w == 0) w = 1;
    colWord[c] = w;
    const int top = H[c], lo = terrain_lo(H, n, x, z);
    uint64_t o = at[c];
    for (int y = lo; y <= top; ++y, ++o) key[o] = morton3(uint32_t(x), uint32_t(y), uint32_t(z));
}

// word[i] = the word of sorted voxel i's column.
__global__ void terrain_words_kernel(const uint64_t *__restrict__ key, uint64_t V, int n,
                                     const uint32_t *__restrict__ colWord, uint32_t *__restrict__ word) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= V) return;
    const uint64_t m = key[i];
    word[i] = colWord[uint64_t(compact3(m >> 2)) * uint64_t(n) + compact3(m)];
}

// ---------------------------------------------------------------------------
// exclusive prefix sums with the element count in device memory
// ---------------------------------------------------------------------------

// Three passes over blocks of kScanTile elements: per-block sums, one
// workgroup scanning the block sums (and writing the grand total), then each
// block's scan plus its offset. The count is read on the device, so a pass
// whose input count was produced by an earlier kernel needs no round trip to
// the host; grids are sized by a host-side upper bound and blocks past the
// count exit at once.
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = uint64_t(kThreads) * kScanItems;

// Exclusive sum of `v` over the Threads threads of the workgroup (lds holds
// Threads / 64 words); *all = total.
template <int Threads = kThreads>
__device__ __forceinline__ uint64_t block_exclusive(uint64_t v, uint64_t *lds, uint64_t *all) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint64_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < Threads / 64; ++w) {
        if (w < wave) before += lds[w];
        total += lds[w];
    }
    __syncthreads();
    *all = total;
    return before + x - v;
}

// Thread items [first, first + kScanItems) of in[0, n) (0 past n); a full
// run of 32-bit items is two 16-byte loads (the arrays are 16-byte aligned).
template <typename In>
__device__ __forceinline__ void load_items(const In *__restrict__ in, uint64_t first, uint64_t n,
                                           uint64_t v[kScanItems]) {
    if constexpr (sizeof(In) == 4 && kScanItems == 8) {
        if (first + kScanItems <= n) {
            const uint4 a = *reinterpret_cast<const uint4 *>(in + first);
            const uint4 b = *reinterpret_cast<const uint4 *>(in + first + 4);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
            return;
        }
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = first + k < n ? uint64_t(in[first + k]) : 0;
}

template <typename In>
__global__ void __launch_bounds__(kThreads) scan_sums_kernel(const In *__restrict__ in, const uint64_t *__restrict__ nptr,
                                                             uint64_t *__restrict__ sums) {
    __shared__ uint64_t lds[kThreads / 64];
    const uint64_t n = *nptr;
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    if (base >= n) {
        if (threadIdx.x == 0) sums[blockIdx.x] = 0;
        return;
    }
    uint64_t v[kScanItems];
    load_items(in, base + uint64_t(threadIdx.x) * kScanItems, n, v);
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    uint64_t total;
    block_exclusive(sum, lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// One workgroup of 1024 threads, 4 sums each per round: exclusive scan of nb
// block sums in place; *total = their sum.
constexpr int kSumThreads = 1024, kSumItems = 4;
__global__ void __launch_bounds__(kSumThreads) scan_block_sums_kernel(uint64_t *__restrict__ sums, uint64_t nb,
                                                                      uint64_t *__restrict__ total) {
    __shared__ uint64_t lds[kSumThreads / 64];
    uint64_t carry = 0;
    for (uint64_t base = 0; base < nb; base += uint64_t(kSumThreads) * kSumItems) {
        const uint64_t first = base + uint64_t(threadIdx.x) * kSumItems;
        uint64_t v[kSumItems];
        uint64_t sum = 0;
#pragma unroll
        for (int k = 0; k < kSumItems; ++k) {
            v[k] = first + k < nb ? sums[first + k] : 0;
            sum += v[k];
        }
        uint64_t all;
        uint64_t run = carry + block_exclusive<kSumThreads>(sum, lds, &all);
#pragma unroll
        for (int k = 0; k < kSumItems; ++k) {
            if (first + k < nb) sums[first + k] = run;
            run += v[k];
        }
        carry += all;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <typename In>
__global__ void __launch_bounds__(kThreads) scan_down_kernel(const In *__restrict__ in, const uint64_t *__restrict__ nptr,
                                                             const uint64_t *__restrict__ offs, uint64_t *__restrict__ out) {
    __shared__ uint64_t lds[kThreads / 64];
    const uint64_t n = *nptr;
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    if (base >= n) return;
    const uint64_t first = base + uint64_t(threadIdx.x) * kScanItems;
    uint64_t v[kScanItems];
    load_items(in, first, n, v);
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) sum += v[k];
    uint64_t total;
    uint64_t run = offs[blockIdx.x] + block_exclusive(sum, lds, &total);
    if (first + kScanItems <= n) {
        // a full run: four 16-byte stores
        uint64_t o[kScanItems];
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            o[k] = run;
            run += v[k];
        }
#pragma unroll
        for (int k = 0; k < kScanItems; k += 2)
            *reinterpret_cast<ulonglong2 *>(out + first + k) = make_ulonglong2(o[k], o[k + 1]);
        return;
    }
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (first + k < n) out[first + k] = run;
        run += v[k];
    }
}

// Heads of runs of equal keys >> shift in sorted keys[0, n): i is a head when
// i == 0 or its key differs from key i - 1. A compaction in the scan's three
// passes with the head flags computed from the keys in the first and third
// (no flag or index arrays): the first counts each tile's heads, the second
// scans the tile counts, the third gives each head its index j among the
// heads and calls emit(j, i, n).
__device__ __forceinline__ uint32_t run_head_flags(const uint64_t *__restrict__ keys, uint64_t first, uint64_t n,
                                                   int shift) {
    uint64_t k[kScanItems];
    if (first + kScanItems <= n) {
#pragma unroll
        for (int q = 0; q < kScanItems; q += 2) {
            const ulonglong2 v = *reinterpret_cast<const ulonglong2 *>(keys + first + q);
            k[q] = v.x;
            k[q + 1] = v.y;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) k[q] = first + q < n ? keys[first + q] : 0;
    }
    uint64_t prev = (first && first <= n) ? keys[first - 1] >> shift : 0;
    uint32_t flags = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const uint64_t cur = k[q] >> shift;
        if (first + q < n && (first + q == 0 || cur != prev)) flags |= 1u << q;
        prev = cur;
    }
    return flags;
}

__global__ void __launch_bounds__(kThreads) run_sums_kernel(const uint64_t *__restrict__ keys,
                                                            const uint64_t *__restrict__ nptr, int shift,
                                                            uint64_t *__restrict__ sums) {
    __shared__ uint64_t lds[kThreads / 64];
    const uint64_t n = *nptr;
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    if (base >= n) {
        if (threadIdx.x == 0) sums[blockIdx.x] = 0;
        return;
    }
    const uint32_t flags = run_head_flags(keys, base + uint64_t(threadIdx.x) * kScanItems, n, shift);
    uint64_t total;
    block_exclusive(uint64_t(__builtin_popcount(flags)), lds, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename Emit>
__global__ void __launch_bounds__(kThreads) run_emit_kernel(const uint64_t *__restrict__ keys,
                                                            const uint64_t *__restrict__ nptr, int shift,
                                                            const uint64_t *__restrict__ offs, Emit emit) {
    __shared__ uint64_t lds[kThreads / 64];
    const uint64_t n = *nptr;
    const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
    if (base >= n) return;
    const uint64_t first = base + uint64_t(threadIdx.x) * kScanItems;
    uint32_t flags = run_head_flags(keys, first, n, shift);
    uint64_t total;
    uint64_t j = offs[blockIdx.x] + block_exclusive(uint64_t(__builtin_popcount(flags)), lds, &total);
    for (; flags; flags &= flags - 1, ++j) emit(j, first + uint64_t(__builtin_ctz(flags)), n);
}

// The fold's segments: segStart[j] = the first candidate of voxel j.
struct SegmentEmit {
    uint64_t *segStart;
    __device__ void operator()(uint64_t j, uint64_t i, uint64_t) const { segStart[j] = i; }
};

// A level's parents from the sorted child keys: parent j's key, first child
// and child mask (its children, at most 8 since keys are unique, follow i).
struct ParentEmit {
    const uint64_t *child;
    uint64_t *pkey, *pstart;
    uint32_t *pmask;
    __device__ void operator()(uint64_t j, uint64_t i, uint64_t n) const {
        const uint64_t parent = child[i] >> 3;
        uint32_t mask = 1u << uint32_t(child[i] & 7);
        for (uint64_t q = i + 1; q < n && q < i + 8; ++q) {
            const uint64_t k = child[q];
            if ((k >> 3) != parent) break;
            mask |= 1u << uint32_t(k & 7);
        }
        pkey[j] = parent;
        pstart[j] = i;
        pmask[j] = mask;
    }
};

__global__ void set_u64_kernel(uint64_t *p, uint64_t v) { *p = v; }

// *out = 1 when a cache block other than the last flagged one (the last
// built) holds an unscanned voxel; one workgroup.
__global__ void __launch_bounds__(kThreads) stale_blocks_kernel(const uint32_t *__restrict__ hazard,
                                                                const uint32_t *__restrict__ flags, uint64_t nBlocks,
                                                                uint64_t *__restrict__ out) {
    __shared__ int last;
    if (threadIdx.x == 0) last = -1;
    __syncthreads();
    for (uint64_t b = threadIdx.x; b < nBlocks; b += blockDim.x)
        if (flags[b]) atomicMax(&last, int(b));
    __syncthreads();
    for (uint64_t b = threadIdx.x; b < nBlocks; b += blockDim.x)
        if (hazard[b] && int(b) != last) *out = 1;
}
__global__ void widen_u32_kernel(const uint32_t *src, uint64_t *dst) { *dst = *src; }

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

// Exclusive prefix sum of in[0, n) into out (64-bit), n read from the device
// word *nptr; `bound` (host) is an upper bound of n that sizes the grids.
// *total (device, may be null) receives the sum of all n elements.
template <typename In>
int exclusive_scan_dev(Arena &ar, const In *in, uint64_t *out, const uint64_t *nptr, uint64_t bound, uint64_t *total,
                       hipStream_t s) {
    const uint64_t nb = std::max<uint64_t>(1, (bound + kScanTile - 1) / kScanTile);
    if (nb > 0x7FFFFFFFull) return fail(SVO_ERR_LIMIT, "builder: scan too large");
    if (reinterpret_cast<uintptr_t>(in) % 16 || reinterpret_cast<uintptr_t>(out) % 16)
        return fail(SVO_ERR_DEVICE, "builder: scan arrays must be 16-byte aligned (internal error)");
    uint64_t *sums, *tot = total;
    int r = ar.alloc(sums, nb);
    if (r) return r;
    if (!tot && (r = ar.alloc(tot, 1))) return r;
    hipLaunchKernelGGL(scan_sums_kernel<In>, dim3(unsigned(nb)), dim3(kThreads), 0, s, in, nptr, sums);
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(1), dim3(kSumThreads), 0, s, sums, nb, tot);
    hipLaunchKernelGGL(scan_down_kernel<In>, dim3(unsigned(nb)), dim3(kThreads), 0, s, in, nptr, sums, out);
    HIP_TRY(hipGetLastError());
    ar.release(sums);
    if (!total) ar.release(tot);
    return SVO_OK;
}

// Run heads of keys[0, *nptr) (sorted; runs of equal keys >> shift): emit(j,
// i, n) for each head i, j its index among the heads; *total (device) = the
// head count. `bound` (host) bounds *nptr and sizes the grids.
template <typename Emit>
int run_heads(Arena &ar, const uint64_t *keys, const uint64_t *nptr, uint64_t bound, int shift, uint64_t *total,
              const Emit &emit, hipStream_t s) {
    const uint64_t nb = std::max<uint64_t>(1, (bound + kScanTile - 1) / kScanTile);
    if (nb > 0x7FFFFFFFull) return fail(SVO_ERR_LIMIT, "builder: scan too large");
    if (reinterpret_cast<uintptr_t>(keys) % 16)
        return fail(SVO_ERR_DEVICE, "builder: key arrays must be 16-byte aligned (internal error)");
    uint64_t *sums;
    int r = ar.alloc(sums, nb);
    if (r) return r;
    hipLaunchKernelGGL(run_sums_kernel, dim3(unsigned(nb)), dim3(kThreads), 0, s, keys, nptr, shift, sums);
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(1), dim3(kSumThreads), 0, s, sums, nb, total);
    hipLaunchKernelGGL(run_emit_kernel<Emit>, dim3(unsigned(nb)), dim3(kThreads), 0, s, keys, nptr, shift, sums, emit);
    HIP_TRY(hipGetLastError());
    ar.release(sums);
    return SVO_OK;
}

// The same with a host-known count; *total (device, may be null) as above.
template <typename In>
int exclusive_scan(Arena &ar, const In *in, uint64_t *out, uint64_t n, hipStream_t s, uint64_t *total = nullptr) {
    uint64_t *nd;
    int r = ar.alloc(nd, 1);
    if (r) return r;
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s, nd, n);
    r = exclusive_scan_dev(ar, in, out, nd, n, total, s);
    ar.release(nd);
    return r;
}

// Reads device words into host values in one copy (one round trip).
int read_words(const uint64_t *src, uint64_t *dst, int count, hipStream_t s) {
    HIP_TRY(hipMemcpyAsync(dst, src, sizeof(uint64_t) * count, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return SVO_OK;
}

// Exclusive scan with a host-known count whose total the host needs next.
template <typename In>
int scan_count(Arena &ar, const In *in, uint64_t *out, uint64_t n, hipStream_t s, uint64_t &total) {
    if (n == 0) {
        total = 0;
        return SVO_OK;
    }
    uint64_t *tot;
    int r = ar.alloc(tot, 1);
    if (r || (r = exclusive_scan(ar, in, out, n, s, tot)) || (r = read_words(tot, &total, 1, s))) return r;
    ar.release(tot);
    return SVO_OK;
}

int fill32(uint32_t *p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
    HIP_TRY(hipGetLastError());
    return SVO_OK;
}

struct Level {
    uint64_t n = 0;             // node count when known on the host (the leaves)
    uint64_t bound = 0;         // host upper bound of the node count (sizes arrays and grids)
    uint64_t *cnt = nullptr;    // node count, device word
    uint64_t *key = nullptr;
    uint32_t *mask = nullptr;   // child mask (nodes above the leaves)
    uint64_t *start = nullptr;  // index of the first child in the level below
    uint64_t *S = nullptr;      // subtree size (words after the node's slot)
    uint8_t *large = nullptr;
    uint64_t *pos = nullptr;    // word index of the node
    uint64_t *B = nullptr;      // word index of the node's child block
    uint8_t *far = nullptr;     // the node's slot carries a far word
    uint32_t *word = nullptr;   // leaves: material word
};

// Large copies between pageable host memory and the device go through pinned
// chunks shared by the process (never freed: a static destructor may run after
// the HIP runtime has gone): the DMA moves one chunk while a few host threads
// copy another between the chunk and the pageable side, each a slice of it (a
// copy straight from or into pageable memory goes through the runtime's own
// staging at a few GB/s).
struct Staging {
    static constexpr uint64_t kChunk = uint64_t(8) << 20;  // bytes
    static constexpr int kChunks = 4;
    std::mutex mu;
    void *buf[kChunks] = {};
    hipEvent_t ev[kChunks] = {};
};

Staging &staging() {
    static Staging *s = new Staging();
    return *s;
}

// toDevice: host -> pinned chunk -> device; else device -> pinned chunk -> host.
int staged_copy(void *dev, void *host, uint64_t bytes, bool toDevice, hipStream_t st) {
    Staging &sg = staging();
    std::lock_guard<std::mutex> lock(sg.mu);
    for (int k = 0; k < Staging::kChunks; ++k) {
        if (!sg.buf[k]) HIP_TRY(hipHostMalloc(&sg.buf[k], Staging::kChunk, hipHostMallocDefault));
        if (!sg.ev[k]) HIP_TRY(hipEventCreateWithFlags(&sg.ev[k], hipEventDisableTiming));
    }
    uint8_t *d8 = static_cast<uint8_t *>(dev), *h8 = static_cast<uint8_t *>(host);
    const uint64_t n = (bytes + Staging::kChunk - 1) / Staging::kChunk;
    auto len_of = [&](uint64_t j) { return std::min(Staging::kChunk, bytes - j * Staging::kChunk); };
    const int P = int(std::min<uint64_t>({4, std::max(1u, std::thread::hardware_concurrency()), n}));
    // chunk j is ready for the host threads once ready > j; slices done per chunk slot
    std::atomic<uint64_t> ready{0};
    std::atomic<int> done[Staging::kChunks];
    for (auto &d : done) d.store(0);
    std::atomic<bool> abort{false};
    auto host_slice = [&](uint64_t j, int k) {
        const uint64_t len = len_of(j), a = len * uint64_t(k) / uint64_t(P), b = len * uint64_t(k + 1) / uint64_t(P);
        uint8_t *chunk = static_cast<uint8_t *>(sg.buf[j % Staging::kChunks]);
        uint8_t *side = h8 + j * Staging::kChunk;
        if (toDevice) std::memcpy(chunk + a, side + a, b - a);
        else std::memcpy(side + a, chunk + a, b - a);
        done[j % Staging::kChunks].fetch_add(1, std::memory_order_acq_rel);
    };
    auto helper = [&](int k) {
        for (uint64_t j = 0; j < n; ++j) {
            while (ready.load(std::memory_order_acquire) <= j) {
                if (abort.load(std::memory_order_relaxed)) return;
                std::this_thread::yield();
            }
            host_slice(j, k);
        }
    };
    std::vector<std::thread> helpers;
    try {
        for (int k = 1; k < P; ++k) helpers.emplace_back(helper, k);
    } catch (...) {
        abort.store(true);
        for (std::thread &t : helpers) t.join();
        return fail(SVO_ERR_NOMEM, "builder: copy threads");
    }
    hipError_t e = hipSuccess;
    auto dma = [&](uint64_t j) {
        const int slot = int(j % Staging::kChunks);
        if (toDevice) e = hipMemcpyAsync(d8 + j * Staging::kChunk, sg.buf[slot], len_of(j), hipMemcpyHostToDevice, st);
        else e = hipMemcpyAsync(sg.buf[slot], d8 + j * Staging::kChunk, len_of(j), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(sg.ev[slot], st);
    };
    if (!toDevice)
        for (uint64_t j = 0; j < std::min<uint64_t>(n, Staging::kChunks) && e == hipSuccess; ++j) dma(j);
    for (uint64_t j = 0; j < n && e == hipSuccess; ++j) {
        const int slot = int(j % Staging::kChunks);
        // download: chunk j has landed; upload: the chunk's previous DMA has left it
        if (!toDevice || j >= uint64_t(Staging::kChunks)) e = hipEventSynchronize(sg.ev[slot]);
        if (e != hipSuccess) break;
        ready.store(j + 1, std::memory_order_release);
        host_slice(j, 0);
        while (done[slot].load(std::memory_order_acquire) < P) std::this_thread::yield();
        done[slot].store(0, std::memory_order_relaxed);
        if (toDevice) dma(j);
        else if (j + Staging::kChunks < n) dma(j + Staging::kChunks);
    }
    if (e != hipSuccess) abort.store(true);
    for (std::thread &t : helpers) t.join();
    // no DMA may still use the chunks once the lock is released
    hipError_t e2 = hipStreamSynchronize(st);
    if (e == hipSuccess) e = e2;
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("builder: staged copy: ") + hipGetErrorString(e));
    return SVO_OK;
}

// Copies W device words into out (sized to W by the caller or here).
int download_words(const uint32_t *src, uint64_t W, WordVec &out, hipStream_t st) {
    if (out.size() != W) {
        try {
            out.resize(W);
        } catch (...) {
            return fail(SVO_ERR_NOMEM, "builder: host word array");
        }
    }
    const uint64_t bytes = W * sizeof(uint32_t);
    if (bytes <= Staging::kChunk) {
        HIP_TRY(hipMemcpyAsync(out.data(), src, bytes, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return SVO_OK;
    }
    return staged_copy(const_cast<uint32_t *>(src), out.data(), bytes, false, st);
}

// Builds the word array from sorted, unique leaf keys/words. blockFlags marks
// the cache blocks the occupancy table counts as non-empty (nullptr when the
// whole volume is one cache block).
int emit_tree(Arena &ar, const Geometry &g, uint64_t *leafKey, uint32_t *leafWord, uint64_t L,
              const uint32_t *blockFlags, hipStream_t st, svo_octree &out) {
    const int D = g.depth;
    std::vector<Level> lv(size_t(D) + 1);
    lv[D].n = L;
    lv[D].bound = L;
    lv[D].key = leafKey;
    lv[D].word = leafWord;
    const int cacheLevel = D - g.cacheLog;  // nodes of edge `cache`
    const bool multi = g.cacheLog < D;
    int r;
    uint64_t *cnts;  // every level's node count, one word each, read together once the levels exist
    if ((r = ar.alloc(cnts, uint64_t(D) + 1))) return r;
    for (int d = 0; d <= D; ++d) lv[size_t(d)].cnt = cnts + d;
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, lv[D].cnt, L);
    // Levels bottom-up. Node counts stay on the device: each level's arrays
    // are sized by a host bound (no more nodes than the level below, nor than
    // 8^d cells), so no level waits for the host. `empty`: the level below is
    // known on the host to hold no node (a tree without leaves).
    bool empty = L == 0;
    for (int d = D - 1; d >= 0; --d) {
        Level &c = lv[d + 1];
        Level &p = lv[d];
        if (multi && d == cacheLevel) {
            const uint64_t cells = uint64_t(1) << (3 * d);
            uint32_t *dense;
            uint64_t *at, *slot;
            if ((r = ar.alloc(dense, cells)) || (r = ar.alloc(at, cells)) || (r = ar.alloc(slot, cells))) return r;
            HIP_TRY(hipMemcpyAsync(dense, blockFlags, cells * 4, hipMemcpyDeviceToDevice, st));
            if (c.bound)
                hipLaunchKernelGGL(dense_mark_kernel, dim3(grid_for(c.bound)), dim3(kThreads), 0, st, c.key, c.cnt, dense);
            if ((r = exclusive_scan(ar, dense, at, cells, st, p.cnt))) return r;
            p.bound = cells;
            if ((r = ar.alloc(p.key, p.bound)) || (r = ar.alloc(p.start, p.bound)) || (r = ar.alloc(p.mask, p.bound)))
                return r;
            if ((r = fill32(p.mask, p.bound, 0, st))) return r;
            hipLaunchKernelGGL(dense_compact_kernel, dim3(grid_for(cells)), dim3(kThreads), 0, st, dense, cells, at,
                               p.key, slot);
            if (c.bound)
                hipLaunchKernelGGL(dense_children_kernel, dim3(grid_for(c.bound)), dim3(kThreads), 0, st, c.key, c.cnt,
                                   slot, p.start, p.mask);
            if (L == 0) {
                // no leaves: the table level alone decides whether nodes exist
                uint64_t nTable = 0;
                if ((r = read_words(p.cnt, &nTable, 1, st))) return r;
                empty = nTable == 0;
            }
        } else {
            // an empty level below: only the root remains (an empty tree)
            if (empty) {
                p.bound = d == 0 ? 1 : 0;
                if ((r = ar.alloc(p.key, 1)) || (r = ar.alloc(p.start, 1)) || (r = ar.alloc(p.mask, 1))) return r;
                if (d == 0) hipLaunchKernelGGL(root_init_kernel, dim3(1), dim3(1), 0, st, p.key, p.start, p.mask, p.cnt);
                else hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, p.cnt, uint64_t(0));
                continue;
            }
            p.bound = d < 21 ? std::min(c.bound, uint64_t(1) << (3 * d)) : c.bound;
            if ((r = ar.alloc(p.key, p.bound)) || (r = ar.alloc(p.start, p.bound)) || (r = ar.alloc(p.mask, p.bound)))
                return r;
            // every parent below the count is written once (by its first child)
            if ((r = run_heads(ar, c.key, c.cnt, c.bound, 3, p.cnt, ParentEmit{c.key, p.key, p.start, p.mask}, st)))
                return r;
        }
        HIP_TRY(hipGetLastError());
    }

    // the level counts (one read): the size, placement and scatter arrays
    // below are allocated and launched at the exact counts
    {
        std::vector<uint64_t> n(size_t(D) + 1);
        if ((r = read_words(cnts, n.data(), D + 1, st))) return r;
        for (int d = 0; d < D; ++d) {
            if (n[size_t(d)] > lv[size_t(d)].bound) return fail(SVO_ERR_DEVICE, "builder: level above its bound (internal error)");
            lv[size_t(d)].bound = n[size_t(d)];
        }
    }
    // sizes, bottom-up
    for (int d = D - 1; d >= 0; --d) {
        Level &p = lv[d];
        if ((r = ar.alloc(p.S, p.bound)) || (r = ar.alloc(p.large, p.bound))) return r;
        if (p.bound)
            hipLaunchKernelGGL(size_kernel, dim3(grid_for(p.bound)), dim3(kThreads), 0, st, p.mask, p.start, p.cnt,
                               int(d == D - 1), lv[d + 1].S, p.S, p.large);
        HIP_TRY(hipGetLastError());
    }
    // the one round trip of the level passes: the word count sizes the array
    uint64_t *rep, report[2] = {0, 0};
    if ((r = ar.alloc(rep, 2))) return r;
    hipLaunchKernelGGL(root_report_kernel, dim3(1), dim3(1), 0, st, lv[0].S, lv[0].cnt, rep);
    if ((r = read_words(rep, report, 2, st))) return r;
    if (report[1] != 1) return fail(SVO_ERR_DEVICE, "builder: root level does not hold one node");
    const uint64_t W = 1 + report[0];
    // the word array, allocated before the host array below: both map memory
    // into the process, and a device allocation waits while other threads
    // populate host pages
    uint32_t *words;
    {
        void *pw = nullptr;
        hipError_t e = hipMalloc(&pw, W * sizeof(uint32_t));
        if (e != hipSuccess) return fail(SVO_ERR_NOMEM, std::string("builder: word array: ") + hipGetErrorString(e));
        words = static_cast<uint32_t *>(pw);
    }
    struct WordsGuard {  // frees the word array on every early return
        uint32_t *&w;
        bool keep = false;
        ~WordsGuard() {
            if (!keep && w) (void)hipFree(w);
        }
    } guard{words};
    // the host array is mapped and populated by other threads while the
    // device places and scatters (the future's destructor waits on every path)
    std::future<bool> hostArray = std::async(std::launch::async, [&out, W]() {
        try {
            out.words.resize(W);
            return true;
        } catch (...) {
            return false;
        }
    });

    // placement, top-down
    {
        Level &root = lv[0];
        if ((r = ar.alloc(root.pos, 1)) || (r = ar.alloc(root.B, 1)) || (r = ar.alloc(root.far, 1))) return r;
        hipLaunchKernelGGL(root_place_kernel, dim3(1), dim3(1), 0, st, root.pos, root.B, root.far);
    }
    for (int d = 0; d < D; ++d) {
        Level &p = lv[d], &c = lv[d + 1];
        const bool childLeaves = d + 1 == D;
        if ((r = ar.alloc(c.pos, c.bound))) return r;
        if (!childLeaves && ((r = ar.alloc(c.B, c.bound)) || (r = ar.alloc(c.far, c.bound)))) return r;
        if (p.bound)
            hipLaunchKernelGGL(place_kernel, dim3(grid_for(p.bound)), dim3(kThreads), 0, st, p.mask, p.start, p.cnt,
                               p.B, p.large, childLeaves ? nullptr : c.S, int(childLeaves), c.pos, c.B, c.far);
        HIP_TRY(hipGetLastError());
    }

    // scatter
    uint32_t *bad;
    if ((r = ar.alloc(bad, 1)) || (r = fill32(bad, 1, 0, st))) return r;
    for (int d = 0; d < D; ++d) {
        Level &p = lv[d];
        if (p.bound)
            hipLaunchKernelGGL(write_nodes_kernel, dim3(grid_for(p.bound)), dim3(kThreads), 0, st, p.mask, p.cnt, p.pos,
                               p.B, p.large, p.far, int(d == D - 1), words, W, bad);
    }
    if (L)
        hipLaunchKernelGGL(write_leaves_kernel, dim3(grid_for(L)), dim3(kThreads), 0, st, leafWord, L, lv[D].pos, words,
                           W, bad);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    out.build.leaves = L;
    out.build.words = W;
    out.build.device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ar.started).count();
    out.build.peak_bytes = ar.peak;
    uint32_t badH = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&badH, bad, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess || badH) {
        if (badH) return fail(SVO_ERR_DEVICE, "builder: placement outside the word array (internal error)");
        return fail(SVO_ERR_DEVICE, std::string("builder: scatter: ") + hipGetErrorString(e));
    }
    if (!hostArray.get()) return fail(SVO_ERR_NOMEM, "builder: host word array");
    if ((r = download_words(words, W, out.words, st))) return r;
    int dev = 0;
    (void)hipGetDevice(&dev);
    guard.keep = true;
    out.dwords = words;
    out.device = dev;
    // with leaves, the deepest descriptors with children are the leaves'
    // parents (level D - 1): the depth octree_validate reports, D + 1. The
    // scatter checked every node's slot and child block against the array.
    out.depth = L ? D + 1 : 0;
    out.center[0] = float(g.dataW) * 0.5f / float(g.side);  // VoxelData::getCenter, VoxelData.cpp:297-303
    out.center[1] = float(g.dataH) * 0.5f / float(g.side);
    out.center[2] = float(g.dataD) * 0.5f / float(g.side);
    return SVO_OK;
}

// Drops unscanned voxels, flags occupied blocks (file source), and compacts
// to the emitted leaves. For the mesh source an unscanned voxel in any cache
// block but the last one processed lingers in the reused block buffer and
// alters later blocks (VoxelData reuses _bufferedData without clearing): then
// *staleCarry is set and the caller replays the buffer (stale_buffer_leaves);
// without staleCarry that case is refused.
int finish_leaves(Arena &ar, const Geometry &g, uint64_t *key, uint32_t *word, uint64_t V, bool mesh,
                  uint32_t *blockFlags, uint64_t nBlocks, hipStream_t st, uint64_t *&lkey, uint32_t *&lword,
                  uint64_t &L, bool *staleCarry = nullptr) {
    int r;
    uint32_t *keep, *hazard;
    uint64_t *at;
    if ((r = ar.alloc(keep, V)) || (r = ar.alloc(at, V)) || (r = ar.alloc(hazard, nBlocks))) return r;
    if ((r = fill32(hazard, nBlocks, 0, st))) return r;
    BlockParams bp{{g.dataW, g.dataH, g.dataD}, g.cache, g.cacheLog};
    if (V)
        hipLaunchKernelGGL(keep_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, key, word, V, bp, int(!mesh),
                           blockFlags, hazard, keep);
    HIP_TRY(hipGetLastError());
    // info[0] = leaves kept, info[1] = an unscanned voxel in a cache block that
    // is not the last one built (mesh source), read together
    uint64_t *info, infoH[2] = {0, 0};
    if ((r = ar.alloc(info, 2)) || (r = fill32(reinterpret_cast<uint32_t *>(info), 4, 0, st))) return r;
    if (mesh && nBlocks > 1)
        hipLaunchKernelGGL(stale_blocks_kernel, dim3(1), dim3(kThreads), 0, st, hazard, blockFlags, nBlocks, info + 1);
    if (V && (r = exclusive_scan(ar, keep, at, V, st, info))) return r;
    if ((r = read_words(info, infoH, 2, st))) return r;
    if (infoH[1]) {
        if (!staleCarry) return fail(SVO_ERR_UNSUPPORTED, "unscanned voxels in a non-final cache block");
        *staleCarry = true;  // the caller replays the block buffer (stale_buffer_leaves)
        return SVO_OK;
    }
    L = infoH[0];
    if ((r = ar.alloc(lkey, L)) || (r = ar.alloc(lword, L))) return r;
    if (V) hipLaunchKernelGGL(compact_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, key, word, V, keep, at, lkey, lword);
    HIP_TRY(hipGetLastError());
    return SVO_OK;
}

// Mesh source, several cache blocks, and an odd-depth block whose last slice
// holds voxels that is not the last block built: the reference then carries
// those voxels over. VoxelData reuses one block buffer (_bufferedData) and one
// count array (PlyLoader::_counts) for every cache block without clearing
// them (src/VoxelData.cpp:177-180, src/PlyLoader.cpp:439-474) and relies on
// destructive reads to empty them; the unscanned slice (VoxelData.cpp:163-166)
// is never read, so its entries stay, at their buffer index, into the next
// block processed (blocks are processed in depth-first = ascending Morton
// order). There a carried entry either is the starting state of the voxel
// written at that index (writeTriangleCell then blends instead of starting
// fresh) or, if no triangle writes it, appears as a voxel of its own.
// This replays that block by block on the host (the carried set is a few
// slices at most) and folds affected voxels again on the GPU from their
// carried state; it yields the final leaf list.
int stale_buffer_leaves(Arena &ar, const Geometry &g, const MeshParams &mp, int tribits, const uint64_t *keys,
                        const uint64_t *segStart, const float *pos, const float *nrm, const float *col,
                        const uint64_t *dvkey, const uint32_t *dvword, const uint8_t *dvcount, uint64_t V,
                        const uint32_t *dBlockFlags, uint64_t nBlocks, hipStream_t st, uint64_t *&lkey,
                        uint32_t *&lword, uint64_t &L) {
    std::vector<uint64_t> vkey(V);
    std::vector<uint32_t> vword(V), flags(nBlocks);
    std::vector<uint8_t> vcnt(V);
    HIP_TRY(hipStreamSynchronize(st));
    if (V) {
        HIP_TRY(hipMemcpy(vkey.data(), dvkey, V * 8, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(vword.data(), dvword, V * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(vcnt.data(), dvcount, V, hipMemcpyDeviceToHost));
    }
    HIP_TRY(hipMemcpy(flags.data(), dBlockFlags, nBlocks * 4, hipMemcpyDeviceToHost));
    struct Entry {
        uint64_t idx;
        uint32_t word, count;
    };
    const int shift = 3 * g.cacheLog;
    const int cache = g.cache;
    std::vector<Entry> carry;
    std::vector<uint8_t> emit(V, 1);
    std::vector<std::pair<uint64_t, uint32_t>> extra;
    bool started = false;
    int r;
    for (uint64_t b = 0; b < nBlocks; ++b) {
        if (!flags[b]) continue;
        const int ox = int(compact3(b)) * cache, oy = int(compact3(b >> 1)) * cache, oz = int(compact3(b >> 2)) * cache;
        const int w = std::min(cache, g.dataW - ox), h = std::min(cache, g.dataH - oy), d = std::min(cache, g.dataD - oz);
        const uint64_t wh = uint64_t(w) * uint64_t(h), whd = wh * uint64_t(d);
        const int dEven = (d / 2) * 2;
        const uint64_t lo = uint64_t(std::lower_bound(vkey.begin(), vkey.end(), b << shift) - vkey.begin());
        const uint64_t hi = uint64_t(std::lower_bound(vkey.begin(), vkey.end(), (b + 1) << shift) - vkey.begin());
        bool hazard = false;
        if (d & 1)
            for (uint64_t v = lo; v < hi && !hazard; ++v)
                hazard = vword[v] != 0 && int(compact3(vkey[v] >> 2)) == oz + d - 1;
        if (!started && !hazard) continue;  // nothing carried into or out of this block
        started = true;
        std::vector<Entry> next, standalone;
        std::vector<uint64_t> rqVox;
        std::vector<uint32_t> rqWord, rqCount;
        for (const Entry &e : carry) {
            if (e.idx >= whd) {  // beyond this block's extent: untouched
                next.push_back(e);
                continue;
            }
            const uint32_t x = uint32_t(ox + int(e.idx % uint64_t(w)));
            const uint32_t y = uint32_t(oy + int((e.idx / uint64_t(w)) % uint64_t(h)));
            const uint32_t z = uint32_t(oz + int(e.idx / wh));
            const uint64_t k = morton3(x, y, z);
            const uint64_t v = uint64_t(std::lower_bound(vkey.begin() + lo, vkey.begin() + hi, k) - vkey.begin());
            if (v < hi && vkey[v] == k) {
                rqVox.push_back(v);
                rqWord.push_back(e.word);
                rqCount.push_back(e.count);
            } else {
                standalone.push_back(e);
            }
        }
        if (!rqVox.empty()) {
            const uint64_t n = rqVox.size();
            uint64_t *dv;
            uint32_t *dw, *dc, *ow, *oc;
            if ((r = ar.alloc(dv, n)) || (r = ar.alloc(dw, n)) || (r = ar.alloc(dc, n)) || (r = ar.alloc(ow, n)) ||
                (r = ar.alloc(oc, n)))
                return r;
            HIP_TRY(hipMemcpy(dv, rqVox.data(), n * 8, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(dw, rqWord.data(), n * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(dc, rqCount.data(), n * 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(refold_kernel, dim3(grid_for(n)), dim3(kThreads), 0, st, keys, segStart, dv, dw, dc, n,
                               tribits, pos, nrm, col, mp, ow, oc);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(st));
            std::vector<uint32_t> hw(n), hc(n);
            HIP_TRY(hipMemcpy(hw.data(), ow, n * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(hc.data(), oc, n * 4, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < n; ++i) {
                vword[rqVox[i]] = hw[i];
                vcnt[rqVox[i]] = uint8_t(hc[i]);
            }
            for (void *p : {(void *)dv, (void *)dw, (void *)dc, (void *)ow, (void *)oc}) ar.release(p);
        }
        for (uint64_t v = lo; v < hi; ++v) {
            if (vword[v] == 0) {
                emit[v] = 0;
                continue;
            }
            const int x = int(compact3(vkey[v])), y = int(compact3(vkey[v] >> 1)), z = int(compact3(vkey[v] >> 2));
            if (z - oz >= dEven) {  // unscanned: stays in the buffer
                emit[v] = 0;
                const uint64_t idx = uint64_t(x - ox) + uint64_t(w) * (uint64_t(y - oy) + uint64_t(h) * uint64_t(z - oz));
                next.push_back({idx, vword[v], vcnt[v]});
            }
        }
        for (const Entry &e : standalone) {
            if (int(e.idx / wh) < dEven) {
                const uint32_t x = uint32_t(ox + int(e.idx % uint64_t(w)));
                const uint32_t y = uint32_t(oy + int((e.idx / uint64_t(w)) % uint64_t(h)));
                const uint32_t z = uint32_t(oz + int(e.idx / wh));
                extra.push_back({morton3(x, y, z), e.word});
            } else {
                next.push_back(e);
            }
        }
        carry.swap(next);
    }
    std::vector<std::pair<uint64_t, uint32_t>> leaves;
    leaves.reserve(V + extra.size());
    for (uint64_t v = 0; v < V; ++v)
        if (emit[v] && vword[v] != 0) leaves.push_back({vkey[v], vword[v]});
    leaves.insert(leaves.end(), extra.begin(), extra.end());
    std::sort(leaves.begin(), leaves.end());
    L = leaves.size();
    std::vector<uint64_t> hk(L);
    std::vector<uint32_t> hw(L);
    for (uint64_t i = 0; i < L; ++i) {
        hk[i] = leaves[i].first;
        hw[i] = leaves[i].second;
    }
    if ((r = ar.alloc(lkey, L)) || (r = ar.alloc(lword, L))) return r;
    if (L) {
        HIP_TRY(hipMemcpy(lkey, hk.data(), L * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(lword, hw.data(), L * 4, hipMemcpyHostToDevice));
    }
    return SVO_OK;
}

int host_threads(const svo_build_opts *opts) {
    if (opts && opts->threads > 0) return opts->threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);  // ThreadUtils::idealThreadCount
    return n > 0 ? int(n) : 4;
}

uint64_t budget(const svo_build_opts *opts) {
    return (opts && opts->memory_budget) ? opts->memory_budget : (uint64_t(1) << 30);
}

int check_blocks(const Geometry &g, uint64_t &nBlocks) {
    const int levels = g.depth - g.cacheLog;
    if (3 * levels > 30) return fail(SVO_ERR_LIMIT, "too many cache blocks for the occupancy table (raise the budget)");
    nBlocks = uint64_t(1) << (3 * levels);
    return SVO_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// entry points
// ---------------------------------------------------------------------------

// convertToVolume's slab loop (src/PlyLoader.cpp:516-530) once every voxel's
// (voxel, triangle) pairs are sorted: the w, h, d header, then slab after slab
// the voxels of the slab are folded from the slab buffer's carried state and
// the slab's first min(sliceZ, d - z) slices are appended.
static int write_volume(Arena &ar, const Geometry &g, const MeshParams &mp, int tribits, const uint64_t *keys,
                        const uint64_t *segStart, uint64_t V, const float *pos, const float *nrm, const float *col,
                        int sliceZ, FILE *fp) {
    hipStream_t st = ar.stream;
    const uint64_t plane = uint64_t(g.dataW) * uint64_t(g.dataH), slab = plane * uint64_t(sliceZ);
    uint32_t *word, *count, *bad;
    int r;
    if ((r = ar.alloc(word, slab)) || (r = ar.alloc(count, slab)) || (r = ar.alloc(bad, 1))) return r;
    // the reference's buffer is a fresh allocation (zero pages) and its counts are
    // only read for cells written before
    if ((r = fill32(word, slab, 0, st)) || (r = fill32(count, slab, 0, st)) || (r = fill32(bad, 1, 0, st))) return r;
    const int32_t hdr[3] = {g.dataW, g.dataH, g.dataD};
    if (std::fwrite(hdr, 4, 3, fp) != 3) return fail(SVO_ERR_IO, "convert: short write");
    WordVec host;
    for (int z0 = 0; z0 < g.dataD; z0 += sliceZ) {
        if (V)
            hipLaunchKernelGGL(slab_fold_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, keys, segStart, V, tribits, pos,
                               nrm, col, mp, z0, sliceZ, word, count, bad);
        HIP_TRY(hipGetLastError());
        const uint64_t n = plane * uint64_t(std::min(sliceZ, g.dataD - z0));
        if ((r = download_words(word, n, host, st))) return r;
        if (std::fwrite(host.data(), 4, n, fp) != n) return fail(SVO_ERR_IO, "convert: short write");
    }
    uint32_t badH = 0;
    HIP_TRY(hipMemcpy(&badH, bad, 4, hipMemcpyDeviceToHost));
    if (badH) return fail(SVO_ERR_DEVICE, "convert: a voxel outside the volume (internal error)");
    return SVO_OK;
}

// The mesh build once the vertices (9 floats each: position, normal, colour;
// not yet rescaled), their bounds and the face table (sizes, first index,
// first triangle, indices) are on the device in `ar`. With a sink the mesh is
// instead voxelized into a dense .voxel file as PlyLoader::convertToVolume
// writes it (out is not touched).
static int mesh_core(Arena &ar, float *dv, int64_t nverts, int has_normals, int32_t *dfc, int64_t *dfs, int64_t *dts,
                     int32_t *dfi, int64_t nfaces, int64_t T, const float lower[3], const float upper[3],
                     int resolution, const svo_build_opts *opts, svo_octree &out, const VolumeSink *sink) {
    const float diff[3] = {upper[0] - lower[0], upper[1] - lower[1], upper[2] - lower[2]};
    int largest = 2;
    if (diff[0] > diff[1] && diff[0] > diff[2]) largest = 0;
    else if (diff[1] > diff[2]) largest = 1;
    if (!(diff[largest] > 0.0f)) return fail(SVO_ERR_UNSUPPORTED, "degenerate mesh extent");
    const float factor = 1.0f / diff[largest];
    float sl[3], su[3];
    for (int k = 0; k < 3; ++k) {
        su[k] = upper[k] * factor;
        sl[k] = lower[k] * factor;
    }
    Geometry g{};
    // PlyLoader::suggestedDimensions, src/PlyLoader.cpp:499-504
    g.dataW = cvt_i32((su[0] - sl[0]) * float(resolution - 2)) + 2;
    g.dataH = cvt_i32((su[1] - sl[1]) * float(resolution - 2)) + 2;
    g.dataD = cvt_i32((su[2] - sl[2]) * float(resolution - 2)) + 2;
    int r;
    int blk[3];  // processBlock's block
    int sliceZ = 0;
    uint64_t nBlocks = 0;
    if (!sink) {
        if ((r = choose_cache(g, budget(opts), true))) return r;
        blk[0] = blk[1] = blk[2] = g.cache;
    } else {
        // convertToVolume (src/PlyLoader.cpp:511-517): slabs of the whole w x h
        // plane, as many slices as one count byte per voxel lets into the budget
        if (g.dataW < 1 || g.dataH < 1 || g.dataD < 1) return fail(SVO_ERR_UNSUPPORTED, "degenerate mesh extent");
        const uint64_t slices = budget(opts) / (uint64_t(g.dataW) * uint64_t(g.dataH));
        sliceZ = int(std::min<uint64_t>(slices, uint64_t(g.dataD)));
        if (sliceZ == 0)
            return fail(SVO_ERR_LIMIT, "convert: the memory budget holds less than one slice (the reference asserts)");
        blk[0] = g.dataW;
        blk[1] = g.dataH;
        blk[2] = sliceZ;
        g.depth = 0;  // Morton bits per axis of the voxel keys
        while ((1 << g.depth) < std::max(g.dataW, std::max(g.dataH, g.dataD))) g.depth++;
    }
    // PlyLoader::setupBlockProcessing, src/PlyLoader.cpp:407-437
    g.scale = resolution - 4;
    g.sbW = blk[0];
    g.sbH = blk[1];
    g.sbD = blk[2];
    best_partition(g.sbW, g.sbH, g.sbD, host_threads(opts));
    g.pW = blk[0] / g.sbW;
    g.pH = blk[1] / g.sbH;
    g.pD = blk[2] / g.sbD;
    // multiplied before the division, as the reference's _gridW expression evaluates
    g.gW = g.pW * (g.dataW + blk[0] - 1) / blk[0];
    g.gH = g.pH * (g.dataH + blk[1] - 1) / blk[1];
    g.gD = g.pD * (g.dataD + blk[2] - 1) / blk[2];
    if (uint64_t(g.gW) * g.gH * g.gD >= (uint64_t(1) << 32)) return fail(SVO_ERR_LIMIT, "sub-block grid too large");
    int tribits = 1;
    while ((int64_t(1) << tribits) < std::max<int64_t>(T, 2)) tribits++;
    if (3 * g.depth + tribits > 64) return fail(SVO_ERR_LIMIT, "voxel key and triangle index exceed 64 bits");
    if (!sink && (r = check_blocks(g, nBlocks))) return r;

    MeshParams mp;
    mp.scale = g.scale;
    mp.sb[0] = g.sbW; mp.sb[1] = g.sbH; mp.sb[2] = g.sbD;
    mp.g[0] = g.gW; mp.g[1] = g.gH; mp.g[2] = g.gD;
    mp.p[0] = g.pW; mp.p[1] = g.pH; mp.p[2] = g.pD;
    mp.data[0] = g.dataW; mp.data[1] = g.dataH; mp.data[2] = g.dataD;
    mp.blk[0] = blk[0];
    mp.blk[1] = blk[1];
    mp.blk[2] = blk[2];
    mp.hx = 1.0f / float(g.scale);
    mp.hb[0] = float(g.sbW) / float(g.scale);
    mp.hb[1] = float(g.sbH) / float(g.scale);
    mp.hb[2] = float(g.sbD) / float(g.scale);

    hipStream_t st = ar.stream;
    float *pos, *nrm, *col;
    if ((r = ar.alloc(pos, uint64_t(T) * 9)) || (r = ar.alloc(nrm, uint64_t(T) * 9)) || (r = ar.alloc(col, uint64_t(T) * 9)))
        return r;
    hipLaunchKernelGGL(rescale_kernel, dim3(grid_for(nverts)), dim3(kThreads), 0, st, dv, nverts, lower[0], lower[1],
                       lower[2], factor);
    if (nfaces)
        hipLaunchKernelGGL(triangles_kernel, dim3(grid_for(nfaces)), dim3(kThreads), 0, st, dv, dfc, dfs, dts, dfi, nfaces,
                           has_normals, pos, nrm, col);
    HIP_TRY(hipGetLastError());
    ar.release(dv);
    ar.release(dfc);
    ar.release(dfi);
    ar.release(dfs);
    ar.release(dts);

    // binning
    uint32_t *binCount, *bad;
    uint64_t *binOff;
    if ((r = ar.alloc(binCount, T)) || (r = ar.alloc(binOff, T)) || (r = ar.alloc(bad, 1))) return r;
    if ((r = fill32(bad, 1, 0, st))) return r;
    if (T)
        hipLaunchKernelGGL(bin_kernel, dim3(grid_for(T)), dim3(kThreads), 0, st, pos, T, mp, binCount, nullptr, nullptr,
                           nullptr, bad);
    HIP_TRY(hipGetLastError());
    // the pair count and the out-of-table flag (the count pass already sets
    // it for every offending cell) in one read
    uint64_t *rb, binInfo[2] = {0, 0};
    if ((r = ar.alloc(rb, 2))) return r;
    if ((r = exclusive_scan(ar, binCount, binOff, uint64_t(T), st, rb))) return r;
    hipLaunchKernelGGL(widen_u32_kernel, dim3(1), dim3(1), 0, st, bad, rb + 1);
    if ((r = read_words(rb, binInfo, 2, st))) return r;
    if (binInfo[1]) return fail(SVO_ERR_UNSUPPORTED, "a triangle bins outside the sub-block table (reference undefined behaviour)");
    const uint64_t P = T ? binInfo[0] : 0;
    uint32_t *cells, *ptri;
    if ((r = ar.alloc(cells, P)) || (r = ar.alloc(ptri, P))) return r;
    if (T)
        hipLaunchKernelGGL(bin_kernel, dim3(grid_for(T)), dim3(kThreads), 0, st, pos, T, mp, binCount, binOff, cells, ptri,
                           bad);
    ar.release(binCount);
    ar.release(binOff);

    uint32_t *blockFlags = nullptr;
    if (!sink) {
        if ((r = ar.alloc(blockFlags, nBlocks)) || (r = fill32(blockFlags, nBlocks, 0, st))) return r;
        if (P) hipLaunchKernelGGL(block_flags_kernel, dim3(grid_for(P)), dim3(kThreads), 0, st, cells, P, mp, blockFlags);
    }

    // voxel rows
    if (P > 0xFFFFFFFFull) return fail(SVO_ERR_LIMIT, "builder: more than 2^32 (cell, triangle) pairs");
    uint32_t *rows, *rowPair;
    uint64_t *rowStart;
    PairBox *boxes;
    if ((r = ar.alloc(rows, P)) || (r = ar.alloc(rowStart, P + 1)) || (r = ar.alloc(boxes, P))) return r;
    if (P)
        hipLaunchKernelGGL(pair_rows_kernel, dim3(grid_for(P)), dim3(kThreads), 0, st, cells, ptri, P, pos, mp, rows,
                           boxes);
    HIP_TRY(hipGetLastError());
    uint64_t R;
    if ((r = scan_count(ar, rows, rowStart, P, st, R))) return r;
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, rowStart + P, R);
    ar.release(rows);
    // each row's pair (the row passes read it instead of searching rowStart)
    if ((r = ar.alloc(rowPair, R))) return r;
    if (P) hipLaunchKernelGGL(row_pairs_kernel, dim3(grid_for(P)), dim3(kThreads), 0, st, rowStart, P, rowPair);  // 64 pairs per wave
    // (voxel, triangle) keys: a count pass sizes each row's output, an emit
    // pass writes the keys at the rows' exclusive offsets (a single pass that
    // reserved space per wave with device-scope atomics on one counter
    // measured 5.7 ms against 1.8 ms for the pair: the ~300 K reservations
    // serialise); the count pass also records each row's hits, so the emit
    // pass repeats the SAT only for rows longer than 64 voxels
    uint32_t *hitCount;
    uint64_t *hitStart;
    RowHits *hits;
    if ((r = ar.alloc(hitCount, R)) || (r = ar.alloc(hitStart, R)) || (r = ar.alloc(hits, R))) return r;
    if (R)
        hipLaunchKernelGGL(row_count_kernel, dim3(grid_for(R)), dim3(kThreads), 0, st, rowPair, boxes, ptri, rowStart, R,
                           pos, mp, tribits, hitCount, hits);
    HIP_TRY(hipGetLastError());
    uint64_t K;
    if ((r = scan_count(ar, hitCount, hitStart, R, st, K))) return r;
    uint64_t *keys, *keys2;
    if ((r = ar.alloc(keys, K)) || (r = ar.alloc(keys2, K))) return r;
    if (R)
        hipLaunchKernelGGL(row_emit_kernel, dim3(grid_for(R)), dim3(kThreads), 0, st, rowPair, boxes, ptri, rowStart, R,
                           pos, mp, tribits, hits, hitStart, keys);
    HIP_TRY(hipGetLastError());
    ar.release(hitCount);
    ar.release(hitStart);
    ar.release(hits);
    ar.release(rowPair);
    ar.release(boxes);
    out.build.triangles = uint64_t(T);
    out.build.candidates = K;
    ar.release(cells);
    ar.release(ptri);
    ar.release(rowStart);

    // sort (voxel, triangle). The keys are written in ascending triangle order
    // (pairs in triangle order from the binning, rows in pair order, each
    // row's keys at its scanned offset), and rocPRIM's radix sort is stable,
    // so sorting the voxel bits alone leaves each voxel's triangles ascending:
    // the same array as sorting the whole key, in 5 passes instead of 7 at
    // 2048^3 (33 of 54 bits)
    if (K) {
        size_t bytes = 0;
        const int endBit = 3 * g.depth + tribits;
        HIP_TRY(rocprim::radix_sort_keys(nullptr, bytes, keys, keys2, K, tribits, endBit, st));
        uint8_t *tmp;
        if ((r = ar.alloc(tmp, bytes))) return r;
        HIP_TRY(rocprim::radix_sort_keys(tmp, bytes, keys, keys2, K, tribits, endBit, st));
        ar.release(tmp);
    }
    ar.release(keys);

    // per-voxel fold
    // each voxel's first candidate (segStart sized by the candidate count, the
    // voxel count's bound), then the voxel count in one read
    uint64_t V = 0;
    uint64_t *segStart, *vkey, *segInfo;
    uint32_t *vword;
    uint8_t *vcount;
    if ((r = ar.alloc(segStart, K + 1)) || (r = ar.alloc(segInfo, 2))) return r;
    if (K) {
        hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, segInfo, K);
        if ((r = run_heads(ar, keys2, segInfo, K, tribits, segInfo + 1, SegmentEmit{segStart}, st))) return r;
        uint64_t info[2];
        if ((r = read_words(segInfo, info, 2, st))) return r;
        V = info[1];
    }
    ar.release(segInfo);
    hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, st, segStart + V, K);
    if (sink) return write_volume(ar, g, mp, tribits, keys2, segStart, V, pos, nrm, col, sliceZ, sink->fp);
    if ((r = ar.alloc(vkey, V)) || (r = ar.alloc(vword, V)) || (r = ar.alloc(vcount, V))) return r;
    if (V)
        hipLaunchKernelGGL(fold_kernel, dim3(unsigned((V + kFoldVoxelsPerBlock - 1) / kFoldVoxelsPerBlock)), dim3(kThreads),
                           0, st, keys2, segStart, V, tribits, pos, nrm,
                           col, mp, vkey, vword, vcount);
    HIP_TRY(hipGetLastError());

    uint64_t *lkey, L;
    uint32_t *lword;
    bool stale = false;
    if ((r = finish_leaves(ar, g, vkey, vword, V, true, blockFlags, nBlocks, st, lkey, lword, L, &stale))) return r;
    if (stale &&
        (r = stale_buffer_leaves(ar, g, mp, tribits, keys2, segStart, pos, nrm, col, vkey, vword, vcount, V, blockFlags,
                                 nBlocks, st, lkey, lword, L)))
        return r;
    ar.release(segStart);
    ar.release(keys2);
    ar.release(pos);
    ar.release(nrm);
    ar.release(col);
    const bool multi = g.cacheLog < g.depth;
    return emit_tree(ar, g, lkey, lword, L, multi ? blockFlags : nullptr, st, out);
}

static int mesh_entry(const float *verts, int64_t nverts, int has_normals, const int32_t *face_counts,
                      const int32_t *face_idx, int64_t nfaces, int resolution, const svo_build_opts *opts,
                      svo_octree &out, const VolumeSink *sink) {
    if (resolution < 8 || resolution > 8192) return fail(SVO_ERR_INVALID, "resolution must be in [8, 8192]");
    // PlyLoader::readVertices bounds and rescaleVertices (src/PlyLoader.cpp:122-181)
    float lower[3] = {1e30f, 1e30f, 1e30f}, upper[3] = {-1e30f, -1e30f, -1e30f};
    for (int64_t i = 0; i < nverts; ++i)
        for (int k = 0; k < 3; ++k) {
            const float v = verts[9 * i + k];
            if (!std::isfinite(v)) return fail(SVO_ERR_UNSUPPORTED, "non-finite vertex position");
            lower[k] = fmin_ref(lower[k], v);
            upper[k] = fmax_ref(upper[k], v);
        }
    // face table
    std::vector<int64_t> fstart(size_t(nfaces) + 1), tstart(size_t(nfaces) + 1);
    int64_t at = 0, T = 0;
    for (int64_t f = 0; f < nfaces; ++f) {
        if (face_counts[f] < 0) return fail(SVO_ERR_FORMAT, "negative face size");
        fstart[f] = at;
        tstart[f] = T;
        at += face_counts[f];
        if (face_counts[f] > 2) T += face_counts[f] - 2;
    }
    for (int64_t i = 0; i < at; ++i)
        if (face_idx[i] < 0 || face_idx[i] >= nverts) return fail(SVO_ERR_FORMAT, "face references a missing vertex");
    if (T >= (int64_t(1) << 32)) return fail(SVO_ERR_LIMIT, "too many triangles");

    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    float *dv;
    int32_t *dfc, *dfi;
    int64_t *dfs, *dts;
    int r;
    if ((r = ar.alloc(dv, uint64_t(nverts) * 9)) || (r = ar.alloc(dfc, nfaces)) || (r = ar.alloc(dfi, at)) ||
        (r = ar.alloc(dfs, nfaces)) || (r = ar.alloc(dts, nfaces)))
        return r;
    HIP_TRY(hipMemcpyAsync(dv, verts, uint64_t(nverts) * 36, hipMemcpyHostToDevice, st));
    if (nfaces) {
        HIP_TRY(hipMemcpyAsync(dfc, face_counts, uint64_t(nfaces) * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dfs, fstart.data(), uint64_t(nfaces) * 8, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(dts, tstart.data(), uint64_t(nfaces) * 8, hipMemcpyHostToDevice, st));
    }
    if (at) HIP_TRY(hipMemcpyAsync(dfi, face_idx, uint64_t(at) * 4, hipMemcpyHostToDevice, st));
    // the host vectors must outlive the copies
    HIP_TRY(hipStreamSynchronize(st));
    return mesh_core(ar, dv, nverts, has_normals, dfc, dfs, dts, dfi, nfaces, T, lower, upper, resolution, opts, out,
                     sink);
}

int build_mesh(const float *verts, int64_t nverts, int has_normals, const int32_t *face_counts,
               const int32_t *face_idx, int64_t nfaces, int resolution, const svo_build_opts *opts, svo_octree &out) {
    return mesh_entry(verts, nverts, has_normals, face_counts, face_idx, nfaces, resolution, opts, out, nullptr);
}

// PLY path: binary little-endian files whose vertex records have a fixed
// size and whose faces all have the first face's vertex count are decoded on
// the device straight from the file bytes; other files (ascii, mixed face
// sizes, list properties on vertices) go through the host parser.
static int ply_entry(const char *path, int resolution, const svo_build_opts *opts, svo_octree &out,
                     const VolumeSink *sink) {
    if (resolution < 8 || resolution > 8192) return fail(SVO_ERR_INVALID, "resolution must be in [8, 8192]");
    PlyLayout L;
    int r = ply_open(path, L);
    if (r) return r;
    auto host_path = [&]() {
        std::vector<float> verts;
        std::vector<int32_t> counts, idx;
        int hn = 0;
        int rr = ply_parse_host(L, verts, hn, counts, idx);
        if (rr) return rr;
        return mesh_entry(verts.data(), int64_t(verts.size() / 9), hn, counts.data(), idx.data(),
                          int64_t(counts.size()), resolution, opts, out, sink);
    };
    if (!L.binary) return host_path();
    // element offsets: every element before the face element needs fixed-size records
    const PlyElem &ve = L.elems[size_t(L.vertex)], &fe = L.elems[size_t(L.face)];
    if (L.face < L.vertex) return host_path();
    uint64_t at = L.body, vbody = 0, fbody = 0;
    PlyVertexLayout vl{};
    PlyFaceLayout fl{};
    for (int e = 0; e <= L.face; ++e) {
        const PlyElem &el = L.elems[size_t(e)];
        int64_t stride = 0;
        bool fixed = true;
        for (const PlyProp &pr : el.props) {
            if (pr.list) {
                if (e != L.face || pr.name != "vertex_indices" || fl.idxType) {
                    fixed = false;
                    break;
                }
                fl.countOff = int(stride);
                fl.countType = pr.count_type;
                fl.idxType = pr.type;
                if (at + uint64_t(stride) + ply_type_size(pr.count_type) > L.buf.size()) return host_path();
                double c0 = 0;
                {
                    const uint8_t *q = L.buf.data() + at + stride;
                    uint64_t u = 0;
                    for (int k = 0; k < ply_type_size(pr.count_type); ++k) u |= uint64_t(q[k]) << (8 * k);
                    switch (pr.count_type) {
                        case P_CHAR: c0 = int8_t(u); break;
                        case P_UCHAR: c0 = uint8_t(u); break;
                        case P_SHORT: c0 = int16_t(u); break;
                        case P_USHORT: c0 = uint16_t(u); break;
                        case P_INT: c0 = int32_t(u); break;
                        case P_UINT: c0 = uint32_t(u); break;
                        default: return host_path();
                    }
                }
                if (c0 < 3 || c0 > 64) return host_path();
                fl.c0 = int(c0);
                fl.idxOff = int(stride + ply_type_size(pr.count_type));
                stride += ply_type_size(pr.count_type) + int64_t(fl.c0) * ply_type_size(pr.type);
            } else {
                if (e == L.vertex)
                    for (int t = 0; t < 9; ++t)
                        if (pr.name == kPlyVertexNames[t]) {
                            vl.off[t] = int(stride);
                            vl.type[t] = pr.type;
                        }
                stride += ply_type_size(pr.type);
            }
        }
        if (!fixed) return host_path();
        if (e == L.vertex) {
            vl.stride = stride;
            vbody = at;
        }
        if (e == L.face) {
            fl.stride = stride;
            fbody = at;
        }
        if (stride && el.count > int64_t((L.buf.size() - at) / uint64_t(stride))) return host_path();
        at += uint64_t(el.count) * uint64_t(stride);
    }
    const int64_t nv = ve.count, nf = fe.count;
    if (nf == 0 || !fl.idxType || fl.idxType == P_FLOAT || fl.idxType == P_DOUBLE) return host_path();
    const int64_t T = nf * (fl.c0 - 2), nidx = nf * fl.c0;
    if (T >= (int64_t(1) << 32)) return fail(SVO_ERR_LIMIT, "too many triangles");

    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint8_t *raw;
    float *dv;
    int32_t *dfc, *dfi;
    int64_t *dfs, *dts;
    uint32_t *flags;
    BoundPart *parts;
    const int nparts = int(std::min<int64_t>(std::max<int64_t>((nv + 255) / 256, 1), 1024));
    const uint64_t span = at - vbody;  // vertex element through face element
    if ((r = ar.alloc(raw, span)) || (r = ar.alloc(dv, uint64_t(nv) * 9)) || (r = ar.alloc(dfc, nf)) ||
        (r = ar.alloc(dfi, nidx)) || (r = ar.alloc(dfs, nf)) || (r = ar.alloc(dts, nf)) || (r = ar.alloc(flags, 1)) ||
        (r = ar.alloc(parts, nparts)))
        return r;
    if ((r = fill32(flags, 1, 0, st))) return r;
    if (span > Staging::kChunk) {
        if ((r = staged_copy(raw, const_cast<uint8_t *>(L.buf.data() + vbody), span, true, st))) return r;
    } else {
        HIP_TRY(hipMemcpyAsync(raw, L.buf.data() + vbody, span, hipMemcpyHostToDevice, st));
    }
    if (nv) hipLaunchKernelGGL(ply_vertices_kernel, dim3(grid_for(nv)), dim3(kThreads), 0, st, raw, nv, vl, dv);
    hipLaunchKernelGGL(ply_faces_kernel, dim3(grid_for(nf)), dim3(kThreads), 0, st, raw + (fbody - vbody), nf, fl, nv, dfc,
                       dfs, dts, dfi, flags);
    hipLaunchKernelGGL(ply_bounds_kernel, dim3(nparts), dim3(256), 0, st, dv, nv, parts, flags);
    HIP_TRY(hipGetLastError());
    uint32_t fh = 0;
    std::vector<BoundPart> hp(static_cast<size_t>(nparts));
    HIP_TRY(hipMemcpyAsync(&fh, flags, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hp.data(), parts, sizeof(BoundPart) * size_t(nparts), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (fh & 1u) {
        ar.clear();  // faces of several sizes
        return host_path();
    }
    if (fh & 2u) return fail(SVO_ERR_FORMAT, "face references a missing vertex");
    if (fh & 4u) return fail(SVO_ERR_UNSUPPORTED, "non-finite vertex position");
    float lower[3] = {1e30f, 1e30f, 1e30f}, upper[3] = {-1e30f, -1e30f, -1e30f};
    int64_t li[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, ui[3] = {INT64_MAX, INT64_MAX, INT64_MAX};
    for (const BoundPart &bp : hp)
        for (int k = 0; k < 3; ++k) {
            if (bp.i[k] != INT64_MAX && (bp.v[k] < lower[k] || (!(lower[k] < bp.v[k]) && bp.i[k] < li[k]))) {
                lower[k] = bp.v[k];
                li[k] = bp.i[k];
            }
            if (bp.i[3 + k] != INT64_MAX && (upper[k] < bp.v[3 + k] || (!(bp.v[3 + k] < upper[k]) && bp.i[3 + k] < ui[k]))) {
                upper[k] = bp.v[3 + k];
                ui[k] = bp.i[3 + k];
            }
        }
    ar.release(raw);
    ar.release(parts);
    ar.release(flags);
    return mesh_core(ar, dv, nv, L.has_normals, dfc, dfs, dts, dfi, nf, T, lower, upper, resolution, opts, out, sink);
}

int build_ply(const char *path, int resolution, const svo_build_opts *opts, svo_octree &out) {
    return ply_entry(path, resolution, opts, out, nullptr);
}

int convert_ply(const char *path, int resolution, const svo_build_opts *opts, const char *voxelPath) {
    FILE *fp = std::fopen(voxelPath, "wb");
    if (!fp) return fail(SVO_ERR_IO, std::string("cannot write ") + voxelPath);
    VolumeSink sink{fp};
    svo_octree unused;
    int r = ply_entry(path, resolution, opts, unused, &sink);
    if (std::fclose(fp) != 0 && r == SVO_OK) r = fail(SVO_ERR_IO, std::string("short write to ") + voxelPath);
    if (r != SVO_OK) std::remove(voxelPath);  // no partial volume is left behind
    return r;
}

int build_grid(const uint32_t *grid, int w, int h, int d, const svo_build_opts *opts, svo_octree &out) {
    Geometry g{};
    g.dataW = w;
    g.dataH = h;
    g.dataD = d;
    int r = choose_cache(g, budget(opts), false);
    if (r) return r;
    if (g.depth > 21) return fail(SVO_ERR_LIMIT, "volume too large");
    uint64_t nBlocks;
    if ((r = check_blocks(g, nBlocks))) return r;
    const uint64_t N = uint64_t(w) * h * d;
    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint32_t *dg, *flag;
    uint64_t *at;
    if ((r = ar.alloc(dg, N)) || (r = ar.alloc(flag, N)) || (r = ar.alloc(at, N))) return r;
    HIP_TRY(hipMemcpyAsync(dg, grid, N * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(grid_keys_kernel, dim3(grid_for(N)), dim3(kThreads), 0, st, dg, w, h, d, flag);
    uint64_t V;
    if ((r = scan_count(ar, flag, at, N, st, V))) return r;
    out.build.candidates = V;
    uint64_t *key, *key2;
    uint32_t *word, *word2;
    if ((r = ar.alloc(key, V)) || (r = ar.alloc(word, V)) || (r = ar.alloc(key2, V)) || (r = ar.alloc(word2, V))) return r;
    hipLaunchKernelGGL(grid_compact_kernel, dim3(grid_for(N)), dim3(kThreads), 0, st, dg, w, h, d, flag, at, key, word);
    HIP_TRY(hipGetLastError());
    ar.release(dg);
    ar.release(flag);
    ar.release(at);
    if (V) {
        size_t bytes = 0;
        HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, key, key2, word, word2, V, 0, 3 * g.depth, st));
        uint8_t *tmp;
        if ((r = ar.alloc(tmp, bytes))) return r;
        HIP_TRY(rocprim::radix_sort_pairs(tmp, bytes, key, key2, word, word2, V, 0, 3 * g.depth, st));
        ar.release(tmp);
    }
    uint32_t *blockFlags;
    if ((r = ar.alloc(blockFlags, nBlocks)) || (r = fill32(blockFlags, nBlocks, 0, st))) return r;
    uint64_t *lkey, L;
    uint32_t *lword;
    if ((r = finish_leaves(ar, g, key2, word2, V, false, blockFlags, nBlocks, st, lkey, lword, L))) return r;
    const bool multi = g.cacheLog < g.depth;
    return emit_tree(ar, g, lkey, lword, L, multi ? blockFlags : nullptr, st, out);
}

// Occupied voxels of an SDF scene as sorted (morton key, word) pairs on the device.
static int sdf_leaves(Arena &ar, const svo_sdf_primitive *prims, int count, int n, float shade, uint64_t *&key,
                      uint32_t *&word, uint64_t &V) {
    if (count < 1 || count > kMaxSdfPrims) return fail(SVO_ERR_INVALID, "1..64 SDF primitives");
    if (n < 8 || n > 8192 || (n & (n - 1))) return fail(SVO_ERR_INVALID, "SDF resolution must be a power of two in [8, 8192]");
    SdfScene sc;
    std::memset(&sc, 0, sizeof(sc));
    for (int i = 0; i < count; ++i) {
        const svo_sdf_primitive &p = prims[i];
        if (p.kind != 0 && p.kind != 1) return fail(SVO_ERR_INVALID, "SDF primitive kind must be 0 (sphere) or 1 (torus)");
        sc.p[i] = SdfPrim{p.kind, p.center[0], p.center[1], p.center[2], p.r1, p.r2, p.axis[0], p.axis[1], p.axis[2]};
    }
    sc.count = count;
    sc.n = n;
    sc.thr = 0.5f * std::sqrt(3.0f) / float(n);
    sc.h = 1.0f / float(n);
    sc.shade = shade;
    hipStream_t st = ar.stream;
    const int bpa = n / 8;
    const uint64_t nBricks = uint64_t(bpa) * bpa * bpa;
    uint32_t *list, *cnt, *overflow;
    int r;
    if ((r = ar.alloc(list, nBricks)) || (r = ar.alloc(cnt, 2)) || (r = ar.alloc(overflow, 1))) return r;
    if ((r = fill32(cnt, 2, 0, st)) || (r = fill32(overflow, 1, 0, st))) return r;
    hipLaunchKernelGGL(sdf_bricks_kernel, dim3(grid_for(nBricks)), dim3(kThreads), 0, st, sc, nBricks, bpa, list, cnt);
    HIP_TRY(hipGetLastError());
    uint32_t nb = 0;
    HIP_TRY(hipMemcpy(&nb, cnt, 4, hipMemcpyDeviceToHost));
    // shell voxels: at most a few per brick column; start from 64 per brick, grow if needed
    uint64_t cap = std::max<uint64_t>(uint64_t(nb) * 64, 1024);
    for (int attempt = 0; attempt < 4; ++attempt) {
        uint64_t *k;
        uint32_t *w;
        if ((r = ar.alloc(k, cap)) || (r = ar.alloc(w, cap))) return r;
        if ((r = fill32(cnt + 1, 1, 0, st)) || (r = fill32(overflow, 1, 0, st))) return r;
        if (nb) hipLaunchKernelGGL(sdf_voxels_kernel, dim3(nb), dim3(64), 0, st, sc, list, bpa, k, w, cnt + 1, cap, overflow);
        HIP_TRY(hipGetLastError());
        uint32_t got = 0, of = 0;
        HIP_TRY(hipMemcpy(&got, cnt + 1, 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&of, overflow, 4, hipMemcpyDeviceToHost));
        if (!of) {
            V = got;
            uint64_t *k2;
            uint32_t *w2;
            if ((r = ar.alloc(k2, V)) || (r = ar.alloc(w2, V))) return r;
            if (V) {
                size_t bytes = 0;
                HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, k, k2, w, w2, V, 0, 3 * log2i(n), st));
                uint8_t *tmp;
                if ((r = ar.alloc(tmp, bytes))) return r;
                HIP_TRY(rocprim::radix_sort_pairs(tmp, bytes, k, k2, w, w2, V, 0, 3 * log2i(n), st));
                ar.release(tmp);
            }
            ar.release(k);
            ar.release(w);
            key = k2;
            word = w2;
            return SVO_OK;
        }
        ar.release(k);
        ar.release(w);
        cap = uint64_t(got) + 1024;
    }
    return fail(SVO_ERR_DEVICE, "SDF voxel count did not converge");
}

static int terrain_check(int n) {
    if (n < 8 || n > 8192 || (n & (n - 1))) return fail(SVO_ERR_INVALID, "terrain resolution must be a power of two in [8, 8192]");
    return SVO_OK;
}

// Per-column heights of the terrain on the device.
static int terrain_columns(Arena &ar, int n, uint32_t seed, float *&hh, int32_t *&H) {
    const uint64_t cols = uint64_t(n) * n;
    int r;
    if ((r = ar.alloc(hh, cols)) || (r = ar.alloc(H, cols))) return r;
    hipLaunchKernelGGL(terrain_heights_kernel, dim3(grid_for(cols)), dim3(kThreads), 0, ar.stream, n, seed, hh, H);
    HIP_TRY(hipGetLastError());
    return SVO_OK;
}

// Exposed-shell voxels of the terrain as sorted (morton key, word) pairs.
static int terrain_leaves(Arena &ar, int n, uint32_t seed, uint64_t *&key, uint32_t *&word, uint64_t &V) {
    hipStream_t st = ar.stream;
    const uint64_t cols = uint64_t(n) * n;
    float *hh;
    int32_t *H;
    uint32_t *cnt;
    uint64_t *at;
    int r;
    if ((r = terrain_columns(ar, n, seed, hh, H))) return r;
    if ((r = ar.alloc(cnt, cols)) || (r = ar.alloc(at, cols))) return r;
    hipLaunchKernelGGL(terrain_count_kernel, dim3(grid_for(cols)), dim3(kThreads), 0, st, n, H, cnt);
    HIP_TRY(hipGetLastError());
    if ((r = scan_count(ar, cnt, at, cols, st, V))) return r;
    ar.release(cnt);
    uint64_t *k;
    uint32_t *colWord;
    if ((r = ar.alloc(k, V)) || (r = ar.alloc(colWord, cols))) return r;
    hipLaunchKernelGGL(terrain_emit_kernel, dim3(grid_for(cols)), dim3(kThreads), 0, st, n, H, hh, at, k, colWord);
    HIP_TRY(hipGetLastError());
    ar.release(at);
    ar.release(hh);
    ar.release(H);
    uint64_t *k2;
    uint32_t *w2;
    if ((r = ar.alloc(k2, V)) || (r = ar.alloc(w2, V))) return r;
    size_t bytes = 0;
    HIP_TRY(rocprim::radix_sort_keys(nullptr, bytes, k, k2, V, 0, 3 * log2i(n), st));
    uint8_t *tmp;
    if ((r = ar.alloc(tmp, bytes))) return r;
    HIP_TRY(rocprim::radix_sort_keys(tmp, bytes, k, k2, V, 0, 3 * log2i(n), st));
    ar.release(tmp);
    ar.release(k);
    if (V) hipLaunchKernelGGL(terrain_words_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, k2, V, n, colWord, w2);
    HIP_TRY(hipGetLastError());
    ar.release(colWord);
    key = k2;
    word = w2;
    return SVO_OK;
}

int build_terrain(int n, uint32_t seed, const svo_build_opts *opts, svo_octree &out) {
    int r = terrain_check(n);
    if (r) return r;
    Geometry g{};
    g.dataW = g.dataH = g.dataD = n;
    if ((r = choose_cache(g, budget(opts), false))) return r;
    uint64_t nBlocks;
    if ((r = check_blocks(g, nBlocks))) return r;
    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint64_t *key, V;
    uint32_t *word;
    if ((r = terrain_leaves(ar, n, seed, key, word, V))) return r;
    out.build.candidates = V;
    uint32_t *blockFlags;
    if ((r = ar.alloc(blockFlags, nBlocks)) || (r = fill32(blockFlags, nBlocks, 0, st))) return r;
    uint64_t *lkey, L;
    uint32_t *lword;
    if ((r = finish_leaves(ar, g, key, word, V, false, blockFlags, nBlocks, st, lkey, lword, L))) return r;
    ar.release(key);
    ar.release(word);
    const bool multi = g.cacheLog < g.depth;
    return emit_tree(ar, g, lkey, lword, L, multi ? blockFlags : nullptr, st, out);
}

int terrain_volume(int n, uint32_t seed, uint32_t *grid) {
    int r = terrain_check(n);
    if (r) return r;
    if (n > 1024) return fail(SVO_ERR_LIMIT, "dense terrain volumes are limited to 1024^3");
    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint64_t *key, V;
    uint32_t *word;
    if ((r = terrain_leaves(ar, n, seed, key, word, V))) return r;
    const uint64_t N = uint64_t(n) * n * n;
    uint32_t *dg;
    if ((r = ar.alloc(dg, N)) || (r = fill32(dg, N, 0, st))) return r;
    if (V) hipLaunchKernelGGL(scatter_dense_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, key, word, V, n, dg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(grid, dg, N * 4, hipMemcpyDeviceToHost));
    return SVO_OK;
}

int terrain_heights(int n, uint32_t seed, int32_t *heights, float *h) {
    int r = terrain_check(n);
    if (r) return r;
    Arena ar;
    if (int ro = ar.open()) return ro;
    float *hh;
    int32_t *H;
    if ((r = terrain_columns(ar, n, seed, hh, H))) return r;
    HIP_TRY(hipStreamSynchronize(ar.stream));
    const uint64_t cols = uint64_t(n) * n;
    if (heights) HIP_TRY(hipMemcpy(heights, H, cols * 4, hipMemcpyDeviceToHost));
    if (h) HIP_TRY(hipMemcpy(h, hh, cols * 4, hipMemcpyDeviceToHost));
    return SVO_OK;
}

int build_sdf(const svo_sdf_primitive *prims, int count, int n, float shade, const svo_build_opts *opts, svo_octree &out) {
    Geometry g{};
    g.dataW = g.dataH = g.dataD = n;
    int r = choose_cache(g, budget(opts), false);
    if (r) return r;
    uint64_t nBlocks;
    if ((r = check_blocks(g, nBlocks))) return r;
    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint64_t *key, V;
    uint32_t *word;
    if ((r = sdf_leaves(ar, prims, count, n, shade, key, word, V))) return r;
    out.build.candidates = V;
    uint32_t *blockFlags;
    if ((r = ar.alloc(blockFlags, nBlocks)) || (r = fill32(blockFlags, nBlocks, 0, st))) return r;
    uint64_t *lkey, L;
    uint32_t *lword;
    if ((r = finish_leaves(ar, g, key, word, V, false, blockFlags, nBlocks, st, lkey, lword, L))) return r;
    const bool multi = g.cacheLog < g.depth;
    return emit_tree(ar, g, lkey, lword, L, multi ? blockFlags : nullptr, st, out);
}

int sdf_volume(const svo_sdf_primitive *prims, int count, int n, float shade, uint32_t *grid) {
    if (n > 1024) return fail(SVO_ERR_LIMIT, "dense SDF volumes are limited to 1024^3");
    Arena ar;
    if (int ro = ar.open()) return ro;
    hipStream_t st = ar.stream;
    uint64_t *key, V;
    uint32_t *word;
    int r;
    if ((r = sdf_leaves(ar, prims, count, n, shade, key, word, V))) return r;
    const uint64_t N = uint64_t(n) * n * n;
    uint32_t *dg;
    if ((r = ar.alloc(dg, N)) || (r = fill32(dg, N, 0, st))) return r;
    if (V) hipLaunchKernelGGL(scatter_dense_kernel, dim3(grid_for(V)), dim3(kThreads), 0, st, key, word, V, n, dg);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(grid, dg, N * 4, hipMemcpyDeviceToHost));
    return SVO_OK;
}

}  // namespace svo

namespace svo {

namespace {
__global__ void code_probe_kernel() {}
}  // namespace

// Makes the HIP runtime load this source file's device code on the current
// device now (it loads a code object at the first launch of one of its kernels).
hipError_t load_code_build() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&code_probe_kernel));
}

}  // namespace svo
