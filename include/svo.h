/*
 * svo.h — C ABI of libsvo_hip.so, the MI355X (gfx950) sparse voxel octree
 * engine: .oct load/save, GPU raymarch, GPU octree construction.
 *
 * The reference (tunabrain/sparse-voxel-octrees) has no FFI; its boundary is
 * the C++ class VoxelOctree (src/VoxelOctree.hpp:37-58) and its build inputs
 * VoxelData (src/VoxelData.hpp:101-144) and PlyLoader (src/PlyLoader.hpp:96-113).
 * Each entry point below names the reference member it replaces. A
 * header-only C++ class with the reference's member names, forwarding to these
 * functions, is in include/VoxelOctree.hpp.
 *
 * Conventions
 *  - Every function returns SVO_OK (0) or an SVO_ERR_* code and never exits
 *    the process; svo_last_error() gives a message for the calling thread.
 *  - Handles are immutable after load/build, so concurrent raymarch calls on
 *    one handle from many host threads are safe. Build calls are independent.
 *  - "_device" entry points take device pointers and a hipStream_t (passed as
 *    void*, NULL = the null stream) and return after enqueueing; the other
 *    raymarch/render entry points take host pointers and are synchronous.
 *  - Raymarch outputs per ray: hit (0 miss, 1 leaf hit, 2 level-of-detail hit),
 *    normal (the leaf word; 0 unless hit == 1), t (entry distance; 0 on a miss),
 *    leaf (word index of the leaf; SVO_LEAF_NONE unless hit == 1). Any output
 *    pointer may be NULL to skip that output.
 *  - Arithmetic is FP32 in the reference's operation order; hit, normal, t and
 *    leaf are bit-identical to VoxelOctree::raymarch on the same rays.
 */
#ifndef SVO_H
#define SVO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SVO_LEAF_NONE UINT64_MAX

enum {
    SVO_OK = 0,
    SVO_ERR_INVALID = 1,     /* bad argument */
    SVO_ERR_IO = 2,          /* file could not be opened, read or written */
    SVO_ERR_FORMAT = 3,      /* corrupt .oct / .voxel / PLY, or malformed octree */
    SVO_ERR_NOMEM = 4,       /* host or device allocation failed */
    SVO_ERR_DEVICE = 5,      /* HIP runtime error (no GPU, launch failure, ...) */
    SVO_ERR_UNSUPPORTED = 6, /* input outside what the library handles (see message) */
    SVO_ERR_LIMIT = 7        /* request exceeds a stated limit (budget, size) */
};

typedef struct svo_octree svo_octree;

/* Message for the last error on this thread ("" if none). */
const char *svo_last_error(void);
/* Library version, major*10000 + minor*100 + patch. */
int svo_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int svo_device_count(void);
/* Loads all of the library's device code onto the current device now. The
 * HIP runtime otherwise loads a source file's kernels at the first launch of
 * one of them, which makes the first build of a process (the builder's radix
 * sorts included) take ~0.1 s longer than the next. Optional. */
int svo_load_device_code(void);

/* ---- .oct files --------------------------------------------------------- */

/* VoxelOctree::VoxelOctree(const char *path), src/VoxelOctree.cpp:57-90. The
 * word array is validated before use: every child reference in range, each
 * non-leaf slot the child of one parent, at most 23 levels (SVO_ERR_FORMAT
 * otherwise). */
int svo_load(const char *path, svo_octree **out);
/* Same, from an in-memory .oct image. */
int svo_load_memory(const void *data, uint64_t size, svo_octree **out);
/* VoxelOctree::save, src/VoxelOctree.cpp:92-123: byte-identical output
 * (LZ4 1.7.1 stream compression in 64 MiB chunks). */
int svo_save(const svo_octree *h, const char *path);
/* The .oct image in memory; free *data with svo_free_buffer. */
int svo_encode(const svo_octree *h, void **data, uint64_t *size);
/* The .oct writer of VoxelOctree::save applied to any word array (no tree
 * validation): header + LZ4 stream slices. Free *data with svo_free_buffer. */
int svo_encode_words(const uint32_t *words, uint64_t count, const float center[3], void **data,
                     uint64_t *size);
void svo_free_buffer(void *p);
/* The .oct reader of VoxelOctree(const char*) (src/VoxelOctree.cpp:57-90)
 * without tree validation, for any word array: decodes an in-memory image
 * (slices in parallel on the host threads) into a malloc'd word array (free
 * with svo_free_buffer) and its center. */
int svo_decode_words(const void *data, uint64_t size, uint32_t **words, uint64_t *count, float center[3]);

/* Wrap an existing word array (copied) with its center; validated. */
int svo_from_words(const uint32_t *words, uint64_t count, const float center[3], svo_octree **out);
/* VoxelOctree::center, src/VoxelOctree.hpp:55-57. */
int svo_center(const svo_octree *h, float out[3]);
int svo_word_count(const svo_octree *h, uint64_t *count);
/* Copies min(count, word count) words into dst. */
int svo_copy_words(const svo_octree *h, uint32_t *dst, uint64_t count);
/* Octree depth: descriptor levels + 1 (log2 of the side length). */
int svo_depth(const svo_octree *h, int *depth);
/* Releases host and device memory of the handle (NULL is a no-op). */
void svo_free(svo_octree *h);

/* Makes the word array resident on the current HIP device (done implicitly
 * by the first raymarch); returns its device address. */
int svo_upload(svo_octree *h, const uint32_t **device_words);

/* ---- raymarch ------------------------------------------------------------ */

/* VoxelOctree::raymarch(o, d, rayScale, normal, t), src/VoxelOctree.cpp:207-346,
 * one ray, synchronous. The ray is walked on the calling thread by the same
 * traversal source the GPU kernels run (a device launch per ray would cost
 * tens of microseconds of upload, launch and download against well under one
 * of walk); the batch, render and frame entry points run on the GPU. Safe
 * from many threads at once. *hit gets the hit code; normal/t are
 * written exactly when the reference writes them (normal on a leaf hit; t on
 * any hit). For every raymarch entry point: a ray with a NaN origin or
 * direction component, or an infinite origin on an axis whose direction is
 * infinite, is a miss (the reference's loop never ends on it); every other
 * ray, infinite and huge components included, gets the reference's result. */
int svo_raymarch(svo_octree *h, const float o[3], const float d[3], float ray_scale,
                 uint32_t *normal, float *t, int *hit);

/* n rays, host arrays (o, d: n x 3 floats). ray_scale may be NULL, then
 * ray_scale0 applies to every ray. Synchronous. */
int svo_raymarch_batch(svo_octree *h, int64_t n, const float *o, const float *d,
                       const float *ray_scale, float ray_scale0, uint8_t *hit, uint32_t *normal,
                       float *t, uint64_t *leaf);
/* Same with device arrays, enqueued on `stream`. */
int svo_raymarch_batch_device(svo_octree *h, int64_t n, const float *o, const float *d,
                              const float *ray_scale, float ray_scale0, uint8_t *hit,
                              uint32_t *normal, float *t, uint64_t *leaf, void *stream);

/* ---- cameras and primary rays (src/Main.cpp renderBatch) ------------------ */

typedef struct {
    float tform[16]; /* INV_MODELVIEW, row-major (Mat4 a11..a44), translation included */
    float pos[3];    /* ray origin: tform*0 + center + 1 (src/Main.cpp:150-152) */
} svo_camera;

/* Orbit camera of the viewer (src/Main.cpp:212,244-250; MatrixStack.cpp:65-66):
 * MODEL = rotXYZ(pitch,0,0)*rotXYZ(0,yaw,0), VIEW = translate(0,0,-radius). */
int svo_orbit_camera(const svo_octree *h, float pitch, float yaw, float radius, svo_camera *out);

/* Primary rays of a width x height image generated on the device: per pixel
 * dir = normalize_quake(dx*a11 + dy*a12 + zx, ...) with dx = -1 + x*(2/W),
 * dy = H/W - y*(2/W) (the per-pixel expression of src/Main.cpp:97-113), origin
 * cam->pos, rayScale 0. Outputs are W*H arrays (row-major, y down). */
int svo_render_primary(svo_octree *h, const svo_camera *cam, int width, int height, uint8_t *hit,
                       uint32_t *normal, float *t, uint64_t *leaf);
int svo_render_primary_device(svo_octree *h, const svo_camera *cam, int width, int height,
                              uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf,
                              void *stream);

/* Cost-ordered primary renders (no counterpart in the reference): a tile plan
 * keeps, on the device, the 8x8 tiles of the last render of a view size
 * grouped by cost (log2 of the tile's largest loop-pass count, 16 groups).
 * svo_render_primary_planned_device traces every tile as
 * svo_render_primary_device does (same rays, same results) but starts the
 * tiles in the plan's order, most expensive group first, and files each tile
 * by its cost in this render for the next one (inside the same launch), so a
 * view rendered again (the next frame) starts its slow tiles first instead of
 * leaving them to run alone at the end of the launch. The first render of a
 * plan runs in raster order. A plan holds 128 bytes of device memory per tile
 * (4 MiB at 1920x1080). One plan per image size and stream: renders on one
 * plan must be ordered (one stream). Each render entry takes only plans made
 * by its own create call (SVO_ERR_INVALID otherwise). */
typedef struct svo_tile_plan svo_tile_plan;
int svo_tile_plan_create(int width, int height, svo_tile_plan **out);
void svo_tile_plan_free(svo_tile_plan *p);
int svo_render_primary_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width,
                                      int height, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf,
                                      void *stream);

/* Primary rays (as svo_render_primary) plus one hard-shadow ray per primary
 * leaf hit: origin pos + d*(t - eps) (the hit leaf's entry point pulled back
 * along the primary ray d), direction `light` (unit vector toward a
 * directional light), rayScale 0. shadow[i] = 1 when that ray hits a voxel,
 * 0 when it escapes or pixel i had no primary hit. No reference counterpart:
 * a second, incoherent ray per pixel for the large-terrain workload, traced by
 * the same raymarch. The device form needs hit, t and shadow (device arrays). */
int svo_render_shadow(svo_octree *h, const svo_camera *cam, int width, int height, const float light[3], float eps,
                      uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow);
int svo_render_shadow_device(svo_octree *h, const svo_camera *cam, int width, int height, const float light[3],
                             float eps, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow,
                             void *stream);
/* svo_render_shadow_device with the tiles in a tile plan's cost order (no
 * counterpart in the reference): the plan comes from svo_tile_plan_create for
 * the image size (SVO_ERR_INVALID otherwise) and each launch files a tile by
 * its primary plus shadow passes, so keep a plan for shadow renders apart
 * from one used by svo_render_primary_planned_device. Same outputs. */
int svo_render_shadow_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width, int height,
                                     const float light[3], float eps, uint8_t *hit, uint32_t *normal, float *t,
                                     uint64_t *leaf, uint8_t *shadow, void *stream);

/* The viewer's frame (src/Main.cpp renderBatch/renderTile/shade, :81-202),
 * bit-identical: the image is split into `stripes` horizontal stripes as the
 * viewer's threads split it (0 = 16, src/Main.cpp:57,351-361); per stripe one
 * coarse ray with the LOD cut-off per 8x8 tile corner, then per-pixel rays
 * from pos + dir*max(minT-0.03, 0) for tiles with a finite corner depth,
 * shaded and packed to RGBA (0xFF000000 | grey*0x010101; pixels of skipped
 * tiles are 0). stride 1, or 3 as while the viewer's camera moves (pixels
 * between stride corners copy their corner). depth (may be NULL) receives
 * svo_frame_depth_count() coarse depths (1e10 = miss). rgba is width*height. */
int svo_render_frame(svo_octree *h, const svo_camera *cam, int width, int height, int stripes, int stride,
                     uint32_t *rgba, float *depth);
int svo_render_frame_device(svo_octree *h, const svo_camera *cam, int width, int height, int stripes,
                            int stride, uint32_t *rgba, float *depth, void *stream);
/* Several viewer frames of one octree in one call (no counterpart in the
 * reference, whose viewer renders one frame at a time): frame f from cams[f]
 * into rgba[f] (device pointers, width*height each) and, when depth is not
 * NULL, depth[f] (may be NULL per frame). Each frame equals
 * svo_render_frame_device's; every 16 frames are one coarse launch over all
 * their tile corners and one fine launch over all their tiles, so one frame's
 * coarse pass and tail overlap the other frames' work. Enqueued on `stream`.
 * Every camera and image pointer is checked before anything is enqueued. */
int svo_render_frames_device(svo_octree *h, const svo_camera *cams, int nframes, int width, int height, int stripes,
                             int stride, uint32_t *const *rgba, float *const *depth, void *stream);
/* svo_render_frame_device with the fine pass's tiles in a tile plan's cost
 * order (no counterpart in the reference): a viewer's next frame starts the
 * tiles its previous frame found slowest first; the same pixels and depths.
 * The plan is made for the frame's size and stripe split
 * (svo_tile_plan_create_frame; SVO_ERR_INVALID otherwise, before anything is
 * enqueued); frames through one plan must be ordered (one stream). */
int svo_tile_plan_create_frame(int width, int height, int stripes, svo_tile_plan **out);
int svo_render_frame_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width, int height,
                                    int stripes, int stride, uint32_t *rgba, float *depth, void *stream);
/* svo_render_frames_device through a tile plan (no counterpart in the
 * reference): the fine pass over all nframes frames' tiles runs in the cost
 * order the plan's previous call filed. The plan is made for the frame size,
 * stripe split and frame count (svo_tile_plan_create_frames, 1 to 16 frames:
 * one launch group; SVO_ERR_INVALID otherwise, before anything is enqueued).
 * The same pixels and depths as svo_render_frames_device. */
int svo_tile_plan_create_frames(int width, int height, int stripes, int nframes, svo_tile_plan **out);
int svo_render_frames_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cams, int nframes,
                                     int width, int height, int stripes, int stride, uint32_t *const *rgba,
                                     float *const *depth, void *stream);
/* Coarse depth entries of a frame, or -1 (message in svo_last_error) when the
 * size or stripe split is refused. */
int64_t svo_frame_depth_count(int width, int height, int stripes);

/* Counters of the GPU build that produced a tree (all zero for a loaded tree),
 * for the construction roofline: algorithmic bytes B = 36*triangles (vertex
 * reads) + 24*candidates (each (voxel, triangle) key or source voxel written
 * and read once) + 4*words (output). device_ms is the wall time from the
 * first upload to the last scatter kernel (file reading and the download of
 * the words excluded). peak_bytes is the most device memory the build's
 * temporary arrays held at once (the word array excluded). */
typedef struct {
    uint64_t triangles;
    uint64_t candidates;
    uint64_t leaves;
    uint64_t words;
    double device_ms;
    uint64_t peak_bytes;
    uint64_t reserved[2];
} svo_build_stats;
int svo_build_stats_get(const svo_octree *h, svo_build_stats *out);

/* ---- batched rendering: many views of many octrees in one launch ----------- */

/* One view: an octree, a camera and its device output arrays (width*height
 * each, row-major; any may be NULL). Rays are the primary rays of
 * svo_render_primary with the given LOD scale (0 = none). */
typedef struct {
    svo_octree *tree;
    svo_camera camera;
    int width, height;
    float ray_scale;
    uint8_t *hit;
    uint32_t *normal;
    float *t;
    uint64_t *leaf;
} svo_view;

/* Launch forms of a batched render (svo_batch_opts.form):
 *  SVO_BATCH_GRID: one ordinary launch with one wave per 8x8 tile of every
 *    view; a view's slowest tiles (rays grazing the surface) then overlap the
 *    other views' tiles instead of finishing on an idle GPU.
 *  SVO_BATCH_PERSISTENT: persistent waves over a tile queue, handing a wave's
 *    idle lanes new pixels once refill_threshold of them are idle. */
enum { SVO_BATCH_GRID = 0, SVO_BATCH_PERSISTENT = 1 };

typedef struct {
    int waves_per_cu;     /* persistent form: resident waves per CU; 0 = 32 (LDS permitting) */
    int refill_threshold; /* persistent form: idle lanes of a wave (1..64) that trigger new pixels; 0 = 16 */
    int form;             /* SVO_BATCH_GRID (default) or SVO_BATCH_PERSISTENT */
    int tile_w;           /* grid form: pixels per tile row, 4, 8, 16 or 32 (tile = tile_w x 64/tile_w;
                             0 = 8, the reference's TileSize, src/Main.cpp:63); the persistent form uses 8 */
    int reserved[4];
} svo_batch_opts;

/* A renderer owns the device view table and tile-queue counters of batched
 * launches. Launches on one renderer are ordered: a call waits on the host
 * only until the previous call's small staging copies have been read, and its
 * launch waits on the device for the previous launch of the same renderer
 * (which may be on another stream). Use one renderer per concurrent stream to
 * let launches overlap. No counterpart in the reference (its viewer
 * renders one frame with a CPU thread pool, src/Main.cpp:204-262,342-367). */
typedef struct svo_renderer svo_renderer;
int svo_renderer_create(int max_views, svo_renderer **out);
void svo_renderer_free(svo_renderer *r);

/* Traces every view in one call over a single sequence of 8x8 pixel tiles
 * (grid form: raster order per view; persistent form: Morton order within
 * 64x64-pixel blocks), in the launch form of `opts` (NULL: SVO_BATCH_GRID).
 * Per-ray results equal svo_render_primary's. Enqueued on `stream`. */
int svo_render_batch_device(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                            void *stream);

/* The grid form of svo_render_batch_device with every tile of the batch in a
 * tile plan's cost order (no counterpart in the reference): the first launch
 * through a plan runs in the batch's tile order, each later one starts the
 * tiles the previous launch found slowest (over all views) first, so the slow
 * tiles of the last views no longer form the launch's tail. Same rays and
 * results. The plan is made for the batch's tile count
 * (svo_tile_plan_create_batch with the same views' sizes and tile shape;
 * SVO_ERR_INVALID otherwise, and for the persistent form); renders through
 * one plan must be ordered (one stream). Not for graph capture: a launch
 * reads the plan's order from where its predecessor filed it, chosen on the
 * host per call. */
int svo_tile_plan_create_batch(const svo_view *views, int nviews, const svo_batch_opts *opts, svo_tile_plan **out);
int svo_render_batch_planned_device(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                                    svo_tile_plan *plan, void *stream);

/* svo_render_batch_device that allocates nothing. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * The view table, tile bases and queue counters live in a fixed device array
 * of the library (about 72 KiB per device, held until the process ends), so
 * a call takes at most 512 views (SVO_ERR_LIMIT beyond) and every octree must
 * already be resident on the current device (svo_upload; SVO_ERR_INVALID
 * otherwise). Outputs and launch forms are those of svo_render_batch_device.
 * All calls on one device share that array: calls on one stream are ordered;
 * calls that overlap on one device (several streams or threads) race. */
int svo_render_batch_noalloc(const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream);

/* Launch-shape autotuner: times the batched launch of `views` on `stream` in
 * the grid form for tiles 8x8, 16x4, 4x16 and 32x2 and in the persistent form
 * for waves per CU {16, 24, 32} x refill threshold {16, 32, 64} (one warm-up
 * and two timed launches each; the launches write the views' outputs) and
 * stores the fastest in *best. times (may be NULL) receives the 13 averages in
 * ms: the four grid tile shapes in that order, then the persistent shapes
 * waves-major. */
int svo_renderer_autotune(svo_renderer *r, const svo_view *views, int nviews, void *stream, svo_batch_opts *best,
                          float *times);

/* One batched launch of fixed views (cameras, sizes, output pointers) captured
 * in a HIP graph: the graph holds its own view table and queue counters, so
 * every replay is the grid-form launch (or, in the persistent form, the queue
 * reset kernel plus the persistent kernel), launched with one call. Replays of one graph must not overlap (launch them on one
 * stream); the octrees and output arrays must outlive the graph, and the
 * octrees must stay on the graph's device (using one on another device moves
 * its word array there). */
typedef struct svo_batch_graph svo_batch_graph;
int svo_batch_graph_create(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                           void *stream, svo_batch_graph **out);
int svo_batch_graph_launch(svo_batch_graph *g, void *stream);
void svo_batch_graph_free(svo_batch_graph *g);

/* ---- serving: continuous batching of render requests ------------------------ */

/* One render request: primary rays of an orbit/any camera over octree
 * trees[tree]. Outputs per ray: normal, t, leaf (as svo_render_primary). */
typedef struct {
    int tree;
    svo_camera camera;
    int width, height;
} svo_request;

typedef struct {
    int streams;            /* launch slots, each its own stream (1..16; 0 = 4) */
    uint64_t output_budget; /* device bytes for in-flight outputs, split evenly
                               between slots (0 = 1 GiB); a request is admitted
                               when its 16 B/ray outputs fit its slot's share */
    int max_batch;          /* requests per launch (0 = 256) */
    int one_per_launch;     /* baseline: one request per launch, plain kernel */
    int waves_per_cu;       /* batched launch parameters (svo_batch_opts) */
    int refill_threshold;
    int batch_form;         /* SVO_BATCH_GRID (default) or SVO_BATCH_PERSISTENT */
    int device_outputs;     /* 1: outputs stay in device memory (no copy to the host;
                               latency ends when the launch completes) */
    int reserved[4];
} svo_server_opts;

/* Runs the request loop over a known arrival schedule (arrival_s ascending,
 * seconds from the start) with the octrees resident: requests are admitted
 * by output footprint and folded into the next launch as slots free up, with
 * launches of different slots overlapping on the device. latency_s[i] gets
 * scheduled arrival -> outputs visible in host memory (with device_outputs:
 * -> the launch holding it has completed); makespan_s the time
 * from the first arrival to the last completion. Optional per-request host
 * destinations (arrays of n pointers, entries may be NULL) receive the
 * outputs. Synchronous; spins one host thread. */
int svo_serve_replay(svo_octree *const *trees, int ntrees, const svo_request *requests,
                     const double *arrival_s, int n, const svo_server_opts *opts, double *latency_s,
                     double *makespan_s, uint32_t *const *normal_out, float *const *t_out,
                     uint64_t *const *leaf_out);

/* ---- live server: requests traced as they arrive -------------------------- */

/* A request server keeps octrees resident and traces requests submitted at
 * any time, from any thread. By default a request runs as its own ordinary
 * launch while fewer than four launches are in flight (one of four streams);
 * requests that arrive meanwhile wait, and the next free stream takes them
 * (up to max_batch, while their outputs fit the arena) in one batched
 * grid-form launch. With persistent_launch, the requests beyond four in
 * flight go instead to one persistent launch that runs while work is in
 * flight: a leader wave copies newly published requests from a ring in
 * pinned host memory into device tables, the other waves claim 8x8 tiles of
 * every published request from 8 queues, and a request that arrives while
 * others are being traced is picked up by the next waves that free up
 * (persistent_only sends every request there). Measured on the 64-octree
 * request stream: 14.8 against 12.1 Grays/s under a 1000-request burst and
 * p99 0.38 against 0.58 ms at 2000 req/s for the batched launches.
 * Requests are admitted in submission order while their outputs (16 B per
 * ray: normal u32, t f32, leaf u64) fit the output arena and a ring slot is
 * free; the others wait in the server. Results equal svo_render_primary's for
 * the same camera and size, alone or in any mix. No counterpart in the
 * reference (its viewer renders one frame at a time with CPU threads,
 * src/Main.cpp:204-262,342-367). A persistent launch ends after idle_stop_us
 * with nothing in flight (requests waiting for arena memory do not keep it
 * running); the next admission starts a new one. While it runs it holds
 * waves_per_cu wave slots per CU, and a device-wide synchronisation
 * (hipDeviceSynchronize, hipFree) waits until it has ended. */
typedef struct svo_server svo_server;

typedef struct {
    uint64_t output_budget; /* bytes of the output arena (0 = 1 GiB) */
    int ring_slots;         /* requests in flight at most (0 = 1024, at most 2^20) */
    int waves_per_cu;       /* resident waves of the persistent launch per CU, 4..32 (0 = 24:
                               the other slots stay free for the caller's kernels) */
    int host_outputs;       /* 0: outputs stay in device memory (svo_server_outputs gives
                               device pointers); 1: pinned host memory written by the kernel */
    int idle_stop_us;       /* a persistent launch ends after this long with nothing in flight
                               (0 = 2000) */
    int persistent_only;    /* 1: every request goes through the persistent launch */
    int persistent_launch;  /* 1: requests beyond four in flight go to the persistent launch
                               instead of batched launches */
    int max_batch;          /* requests per batched launch at most, 1..256 (0 = 32) */
    int reserved[3];
} svo_server_config;

enum { SVO_REQ_WAITING = 0, SVO_REQ_IN_FLIGHT = 1, SVO_REQ_DONE = 2 };

typedef struct {
    uint64_t launches;  /* persistent launches started */
    uint64_t published; /* requests handed to the device */
    uint64_t in_flight; /* published, not yet complete */
    uint64_t waiting;   /* submitted, not yet admitted (arena or ring full) */
    int live;           /* a launch is running */
    int reserved0;
    uint64_t direct;    /* requests run as their own launch (few others in flight) */
    uint64_t batched;   /* requests run in batched launches */
    int reserved[2];
} svo_server_stats;

/* Makes the octrees resident on the current device and starts the request
 * loop thread. The octrees must outlive the server. */
int svo_server_create(svo_octree *const *trees, int ntrees, const svo_server_config *cfg, svo_server **out);
/* Queues one request (octree index into the create call's array) and returns
 * its ticket at once. SVO_ERR_LIMIT when its outputs exceed the whole arena. */
int svo_server_submit(svo_server *s, const svo_request *req, uint64_t *ticket);
/* State (SVO_REQ_*) of a ticket; latency_s (may be NULL) = submit -> completion
 * seen by the request loop, once done. */
int svo_server_poll(svo_server *s, uint64_t ticket, int *state, double *latency_s);
/* Blocks until the ticket is done (timeout_s <= 0: no limit; SVO_ERR_LIMIT on timeout). */
int svo_server_wait(svo_server *s, uint64_t ticket, double timeout_s, double *latency_s);
/* Output arrays of a completed request (width*height each, row-major): device
 * pointers, or host pointers with host_outputs. Valid until svo_server_release. */
int svo_server_outputs(svo_server *s, uint64_t ticket, uint32_t **normal, float **t, uint64_t **leaf);
/* Copies a completed request's outputs (any pointer may be NULL) into host
 * arrays of width*height entries; synchronous. */
int svo_server_copy_outputs(svo_server *s, uint64_t ticket, uint32_t *normal, float *t, uint64_t *leaf);
/* Frees a completed (or still waiting) request's outputs and its ticket. */
int svo_server_release(svo_server *s, uint64_t ticket);
int svo_server_stats_get(svo_server *s, svo_server_stats *out);
/* Drives a server with a request schedule: request i is submitted at
 * arrival_s[i] seconds after the call starts (ascending), completions are
 * recorded as they come; latency_s[i] = completion seen - scheduled arrival.
 * With tickets non-NULL the requests are kept (tickets[i] is request i's; the
 * caller reads and releases them, and the arena must hold every output);
 * otherwise each request is released once complete. makespan_s (may be NULL)
 * = first arrival -> last completion. The calling thread polls from 200 us
 * before each arrival (a timed sleep can wake late), so it keeps a core busy
 * while arrivals are close together. */
int svo_server_run_schedule(svo_server *s, const svo_request *requests, const double *arrival_s, int n,
                            uint64_t *tickets, double *latency_s, double *makespan_s);
/* Ends the launch (finishing requests in flight) and frees the server. */
void svo_server_destroy(svo_server *s);

/* ---- construction --------------------------------------------------------- */

typedef struct {
    /* VoxelData memory budget in bytes; 0 = 1 GiB (src/Main.cpp:269). It fixes
     * the cache-block size (src/VoxelData.cpp:211-262), which shapes the tree. */
    uint64_t memory_budget;
    /* Thread count the reference would have used: it fixes the sub-block
     * partition (src/PlyLoader.cpp:381-405,418) and therefore the output bytes.
     * 0 = this host's online processor count, as ThreadUtils::idealThreadCount. */
    int threads;
    int reserved[7];
} svo_build_opts;

/* PlyLoader(path).convertToVolume(voxel_path, resolution, budget),
 * src/PlyLoader.cpp:506-534 (the builder's on-disk mode, src/Main.cpp:315-320):
 * the mesh voxelized on the GPU into a dense .voxel file (int32 w, h, d, then
 * w*h*d uint32 words, x fastest), byte-identical to the voxels the reference's
 * slab loop produces for the same (resolution, budget, threads): slabs of
 * budget / (w*h) slices, each one processBlock over a w x h x slices block,
 * with the slab buffer and counts carried from slab to slab uncleared, as
 * there. The reference's function itself never writes a slab (its progress
 * line divides by a zero block count). SVO_ERR_LIMIT when the budget holds
 * less than one slice; on failure no file is left. */
int svo_convert_to_volume(const char *ply_path, int resolution, const svo_build_opts *opts, const char *voxel_path);

/* PlyLoader(path) -> VoxelData(loader, resolution, budget) -> VoxelOctree,
 * src/Main.cpp:322-327, built on the GPU; .oct bytes identical to the
 * reference's for the same (resolution, budget, threads). */
int svo_build_from_ply(const char *path, int resolution, const svo_build_opts *opts,
                       svo_octree **out);
/* Same with the PLY already parsed: verts n x {x,y,z,nx,ny,nz,r,g,b} as
 * PlyLoader::readVertices stores them (absent normals 0, absent colours 255),
 * faces as CSR (face_counts[nfaces], face_idx flat; fans triangulated). */
int svo_build_from_mesh(const float *verts, int64_t nverts, int has_normals,
                        const int32_t *face_counts, const int32_t *face_idx, int64_t nfaces,
                        int resolution, const svo_build_opts *opts, svo_octree **out);
/* VoxelData(path, budget) -> VoxelOctree (raw .voxel file), src/Main.cpp:315-321. */
int svo_build_from_voxels(const char *path, const svo_build_opts *opts, svo_octree **out);
/* Same with the volume in host memory: w*h*d uint32, x fastest. */
int svo_build_from_grid(const uint32_t *grid, int w, int h, int d, const svo_build_opts *opts,
                        svo_octree **out);

/* ---- procedural sources ----------------------------------------------------
 * The reference builds non-mesh scenes only through a raw .voxel file
 * (VoxelData(path), src/VoxelData.cpp:37-48,177-209). These sources generate
 * the same kind of volume on the GPU and build it with svo_build_from_grid's
 * semantics, without a host volume. */

/* Sphere (kind 0: center, radius r1) or torus (kind 1: center, major radius
 * r1, minor radius r2, unit axis), in the unit cube. */
typedef struct {
    int kind;
    float center[3];
    float r1, r2;
    float axis[3];
} svo_sdf_primitive;

/* Solid-shell voxelization of the union of the primitives on an n^3 grid
 * (n a power of two): voxel (x,y,z) with centre c = (v + 0.5)/n is occupied
 * when |sdf(c)| <= 0.5*sqrt(3)/n; its word is compressMaterial(normalised
 * central-difference gradient (step 1/n), shade), 0 replaced by 1. FP32. */
int svo_build_from_sdf(const svo_sdf_primitive *prims, int count, int resolution, float shade,
                       const svo_build_opts *opts, svo_octree **out);
/* The same voxels as a dense host volume (n^3 uint32, x fastest; n <= 1024). */
int svo_sdf_volume(const svo_sdf_primitive *prims, int count, int resolution, float shade, uint32_t *grid);

/* Procedural fBm heightfield terrain on an n^3 grid (n a power of two, <= 8192),
 * the procedural counterpart of a file-source VoxelData
 * (src/VoxelData.cpp:37-48,177-209), generated column by column on the device
 * with no dense volume. FP32, each operation separately rounded:
 *   h(x,z)  = 8 octaves of smoothstep value noise (gain 0.5, lacunarity 2,
 *             base frequency 4, lattice hash of (cell, seed + octave)) at
 *             ((x+0.5)/n, (z+0.5)/n), normalised to [0,1)
 *   H(x,z)  = n/8 + int(h * n/2); solid for y <= H (y is up)
 *   leaves  = the exposed shell: y from min(H, lowest in-grid 4-neighbour + 1) to H
 *   word    = compressMaterial(normalize(-gx, 1, -gz), 0.35 + 0.65*h), 0 -> 1,
 *             g = clamped central difference of h*n/2
 * The octree is the one the reference builds from the same voxels as a
 * .voxel file with the given memory budget. */
int svo_build_from_terrain(int resolution, uint32_t seed, const svo_build_opts *opts, svo_octree **out);
/* The same voxels as a dense host volume (n^3 uint32, x fastest; n <= 1024). */
int svo_terrain_volume(int resolution, uint32_t seed, uint32_t *grid);
/* Per-column H and h (n*n each, x fastest; either may be NULL). */
int svo_terrain_heights(int resolution, uint32_t seed, int32_t *heights, float *h);

#ifdef __cplusplus
}
#endif
#endif /* SVO_H */
