// Internal declarations shared by the host sources of libsvo_hip.so.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime.h>

#include "svo.h"

namespace svo {
// Large host arrays (word arrays of big octrees): mapped directly, with
// transparent huge pages requested and the pages populated in one call, so
// that filling them (a download, a decode) does not take a page fault per
// 4 KiB. Both requests are advisory; without them the pages fault in as usual.
void *map_large(size_t bytes);
void unmap_large(void *p, size_t bytes);
constexpr size_t kLargeArray = size_t(4) << 20;

// A whole file mapped read-only (pages read in by the mapping call): parsing
// and uploads read the page cache directly instead of a copy of the file.
class FileView {
public:
    FileView() = default;
    FileView(const FileView &) = delete;
    FileView &operator=(const FileView &) = delete;
    ~FileView() { close(); }
    bool open(const char *path);  // false if the file cannot be opened or mapped
    void close();
    const uint8_t *data() const { return data_; }
    size_t size() const { return size_; }
    uint8_t operator[](size_t i) const { return data_[i]; }

private:
    const uint8_t *data_ = nullptr;
    size_t size_ = 0;
};

// Allocator that leaves default-constructed elements uninitialised, so that
// sizing a word array before it is overwritten does not first write zeros over
// it, and that maps large arrays with map_large.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    using value_type = T;
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    T *allocate(size_t n) {
        if (n * sizeof(T) < kLargeArray) return std::allocator<T>::allocate(n);
        void *p = map_large(n * sizeof(T));
        if (!p) throw std::bad_alloc();
        return static_cast<T *>(p);
    }
    void deallocate(T *p, size_t n) noexcept {
        if (n * sizeof(T) < kLargeArray) std::allocator<T>::deallocate(p, n);
        else unmap_large(p, n * sizeof(T));
    }
    template <typename U, typename... A>
    void construct(U *p, A &&...a) {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
    template <typename U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
};
using WordVec = std::vector<uint32_t, NoInitAlloc<uint32_t>>;
}  // namespace svo

struct svo_octree {
    svo::WordVec words;           // Laine-Karras word array (src/VoxelOctree.hpp:40-41)
    float center[3] = {0, 0, 0};  // VoxelOctree::_center
    int depth = 0;                // descriptor levels + 1
    int device = -1;              // device holding `dwords`
    uint32_t *dwords = nullptr;   // device copy of `words`
    svo_build_stats build{};      // counters of the GPU build that made this tree (zeros otherwise)
    std::mutex upload;            // serialises the first upload (ensure_device) of concurrent callers
};

struct svo_renderer;     // render_batch.hip
struct svo_batch_graph;  // render_batch.hip
struct svo_server;       // server.hip
struct svo_tile_plan;    // raymarch.hip

namespace svo {

// thread-local error message; returns `code` for convenient `return fail(...)`
int fail(int code, const std::string &msg);
void clear_error();

// ---- .oct codec (lz4_oct.cpp) ----
// Appends the LZ4 block encoding of src[pos, pos+len) to `out`, continuing the
// stream state of earlier calls over the same contiguous buffer `src`.
class Lz4StreamEncoder {
public:
    Lz4StreamEncoder();
    // Compresses src[pos, pos+len) (the slice that follows the previous call's)
    // into dst, which holds at least bound(len) bytes; returns the bytes written.
    uint64_t encode(const uint8_t *src, uint64_t pos, uint32_t len, uint8_t *dst);
    static uint64_t bound(uint32_t len);

private:
    uint32_t table_[4096];
    uint64_t base_;      // absolute position that table entries are relative to
    uint64_t dictSize_;  // bytes of history usable as the match window start
    uint64_t consumed_;  // absolute position of the next chunk
};

// Decodes one LZ4 block into out[pos, pos+outLen); matches may reach back
// into out[0, pos). Returns false on malformed input.
bool lz4_decode_block(const uint8_t *src, uint64_t srcLen, uint8_t *out, uint64_t pos,
                      uint64_t outLen);

int oct_decode(const uint8_t *data, uint64_t size, svo_octree &h);
void oct_encode(const svo_octree &h, std::vector<uint8_t> &out);

// ---- octree structure (octree_host.cpp) ----
// Validates every child reference of the tree reachable from the root and
// returns its depth (descriptor levels + 1), -1 if malformed (a reference
// out of the array, a non-leaf slot with two parents, too deep), or -2 if
// the check's one-bit-per-word bitmap cannot be allocated.
int octree_validate(const uint32_t *words, uint64_t n);

// ---- camera (camera.cpp) ----
void orbit_camera(float pitch, float yaw, float radius, const float center[3], svo_camera &cam);
// Every tform and pos entry finite and within +-1e6: the rays the renderers
// generate from such a camera are finite (the traversal relies on it).
bool camera_ok(const svo_camera &cam);

// ---- device (raymarch.hip) ----
int ensure_device(svo_octree &h);
void release_device(svo_octree &h);
int raymarch_device(svo_octree &h, int64_t n, const float *o, const float *d, const float *rs,
                    float rs0, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, void *stream);
int render_primary_device(svo_octree &h, const svo_camera &cam, int width, int height, uint8_t *hit,
                          uint32_t *normal, float *t, uint64_t *leaf, void *stream);
int tile_plan_alloc(int64_t tiles, int width, int height, svo_tile_plan **out);
int tile_plan_create(int width, int height, svo_tile_plan **out);
void tile_plan_free(svo_tile_plan *p);
int render_primary_planned_device(svo_octree &h, svo_tile_plan *plan, const svo_camera &cam, int width, int height,
                                  uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, void *stream);
int device_count();
int raymarch_one(const svo_octree &h, const float o[3], const float d[3], float rs, uint32_t &normal, float &t,
                 uint64_t &leaf);

int render_shadow_device(svo_octree &h, const svo_camera &cam, int width, int height, const float light[3], float eps,
                         uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow, void *stream,
                         svo_tile_plan *plan = nullptr);

// ---- viewer frame (frame.hip) ----
int render_frame_device(svo_octree &h, const svo_camera &cam, int width, int height, int stripes, int stride,
                        uint32_t *rgba, float *depth, void *stream, svo_tile_plan *plan = nullptr);
int tile_plan_create_frames(int width, int height, int stripes, int nframes, svo_tile_plan **out);
int render_frames_device(svo_octree &h, const svo_camera *cams, int n, int width, int height, int stripes, int stride,
                         uint32_t *const *rgba, float *const *depth, void *stream, svo_tile_plan *plan = nullptr);
int64_t frame_depth_count(int width, int height, int stripes);

// ---- batched persistent renderer (render_batch.hip) ----
int renderer_create(int maxViews, svo_renderer **out);
void renderer_free(svo_renderer *r);
int render_batch(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream);
int tile_plan_create_batch(const svo_view *views, int nviews, const svo_batch_opts *opts, svo_tile_plan **out);
int render_batch_planned(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                         svo_tile_plan *plan, void *stream);
int render_batch_noalloc(const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream);
int renderer_autotune(svo_renderer *r, const svo_view *views, int nviews, void *stream, svo_batch_opts *best,
                      float *times);
int batch_graph_create(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream,
                       svo_batch_graph **out);
int batch_graph_launch(svo_batch_graph *g, void *stream);
void batch_graph_free(svo_batch_graph *g);

// ---- request loop (serve.cpp) ----
int serve_replay(svo_octree *const *trees, int ntrees, const svo_request *reqs, const double *arrival, int n,
                 const svo_server_opts *opts, double *latency, double *makespan, uint32_t *const *normalOut,
                 float *const *tOut, uint64_t *const *leafOut);

// ---- device code loading (one probe per source file) ----
hipError_t load_code_raymarch();
hipError_t load_code_render_batch();
hipError_t load_code_server();
hipError_t load_code_frame();
hipError_t load_code_build();

// ---- live request server (server.hip) ----
int server_create(svo_octree *const *trees, int ntrees, const svo_server_config *cfg, svo_server **out);
void server_destroy(svo_server *s);
int server_submit(svo_server *s, const svo_request *r, uint64_t *ticket);
int server_poll(svo_server *s, uint64_t ticket, int *state, double *latency);
int server_wait(svo_server *s, uint64_t ticket, double timeout, double *latency);
int server_outputs(svo_server *s, uint64_t ticket, uint32_t **normal, float **t, uint64_t **leaf);
int server_copy_outputs(svo_server *s, uint64_t ticket, uint32_t *normal, float *t, uint64_t *leaf);
int server_release(svo_server *s, uint64_t ticket);
int server_stats(svo_server *s, svo_server_stats *out);
int server_run_schedule(svo_server *s, const svo_request *reqs, const double *arrival, int n, uint64_t *tickets,
                        double *latency, double *makespan);

// ---- construction (build.hip) ----
int build_mesh(const float *verts, int64_t nverts, int has_normals, const int32_t *face_counts,
               const int32_t *face_idx, int64_t nfaces, int resolution, const svo_build_opts *opts,
               svo_octree &out);
int build_grid(const uint32_t *grid, int w, int h, int d, const svo_build_opts *opts, svo_octree &out);
int build_sdf(const svo_sdf_primitive *prims, int count, int n, float shade, const svo_build_opts *opts,
              svo_octree &out);
int sdf_volume(const svo_sdf_primitive *prims, int count, int n, float shade, uint32_t *grid);
int build_terrain(int n, uint32_t seed, const svo_build_opts *opts, svo_octree &out);
int terrain_volume(int n, uint32_t seed, uint32_t *grid);
int terrain_heights(int n, uint32_t seed, int32_t *heights, float *h);
// ---- PLY (ply_reader.cpp) ----
enum PType { P_INVALID = 0, P_CHAR, P_SHORT, P_INT, P_UCHAR, P_USHORT, P_UINT, P_FLOAT, P_DOUBLE };
int ply_type_size(PType t);
extern const char *const kPlyVertexNames[9];  // x y z nx ny nz red green blue
struct PlyProp {
    std::string name;
    bool list = false;
    PType type = P_INVALID, count_type = P_INVALID;
};
struct PlyElem {
    std::string name;
    int64_t count = 0;
    std::vector<PlyProp> props;
};
struct PlyLayout {
    FileView buf;              // the whole file
    std::vector<PlyElem> elems;
    size_t body = 0;           // first byte after end_header
    bool binary = false;       // binary_little_endian (else ascii)
    int vertex = -1, face = -1;
    int has_normals = 0;
};
// Reads the file and parses its header (the subset src/third-party/plyfile.c accepts).
int ply_open(const char *path, PlyLayout &layout);
// Parses the body on the host: 9 floats per vertex (defaults 0 / 255), face sizes and indices.
int ply_parse_host(const PlyLayout &layout, std::vector<float> &verts, int &has_normals,
                   std::vector<int32_t> &face_counts, std::vector<int32_t> &face_idx);
int read_ply(const char *path, std::vector<float> &verts, int &has_normals,
             std::vector<int32_t> &face_counts, std::vector<int32_t> &face_idx);
int build_ply(const char *path, int resolution, const svo_build_opts *opts, svo_octree &out);
// PlyLoader::convertToVolume: the mesh voxelized into a dense .voxel file.
int convert_ply(const char *path, int resolution, const svo_build_opts *opts, const char *voxelPath);

}  // namespace svo
