// C ABI of libsvo_hip.so (declared in include/svo.h). Argument checking,
// handle ownership and error reporting live here; the work is in
// lz4_oct.cpp (files), raymarch.hip (GPU traversal) and build.hip (GPU
// construction).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>

#include "svo_internal.hpp"

namespace svo {

static thread_local std::string g_error;

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

void clear_error() { g_error.clear(); }

// Trees from files and caller arrays are walked on the host (octree_validate)
// before anything traverses them. The GPU builder sets the depth of a tree
// with leaves itself, after its scatter checked every node's slot and child
// block against the array; its other trees are walked like the rest.
static int finish_handle(std::unique_ptr<svo_octree> &h, svo_octree **out) {
    int depth = h->depth > 0 ? h->depth : octree_validate(h->words.data(), h->words.size());
    if (depth < 0) {
        release_device(*h);
        if (depth == -2) return fail(SVO_ERR_NOMEM, "out of host memory (octree check)");
        return fail(SVO_ERR_FORMAT,
                    "malformed octree: a child reference leaves the word array, a subtree has two parents, "
                    "or the tree is deeper than the traversal follows");
    }
    h->depth = depth;
    *out = h.release();
    return SVO_OK;
}

}  // namespace svo

using namespace svo;

// ---- host<->device staging for the synchronous entry points ----

namespace {

template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t n) {
        if (n == 0) return SVO_OK;
        hipError_t e = hipMalloc(&p, n * sizeof(T));
        return e == hipSuccess ? SVO_OK : fail(SVO_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
};

template <typename T>
int to_device(DevBuf<T> &b, const T *src, size_t n) {
    if (!src) return SVO_OK;
    int r = b.alloc(n);
    if (r) return r;
    hipError_t e = hipMemcpy(b.p, src, n * sizeof(T), hipMemcpyHostToDevice);
    return e == hipSuccess ? SVO_OK : fail(SVO_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
}

template <typename T>
int out_buf(DevBuf<T> &b, const T *dst, size_t n) {
    return dst ? b.alloc(n) : SVO_OK;
}

template <typename T>
int from_device(T *dst, const DevBuf<T> &b, size_t n) {
    if (!dst || n == 0) return SVO_OK;
    hipError_t e = hipMemcpy(dst, b.p, n * sizeof(T), hipMemcpyDeviceToHost);
    return e == hipSuccess ? SVO_OK : fail(SVO_ERR_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
}

// Runs an entry point's body; a C++ exception (host allocation failure) becomes
// an error code instead of crossing the C boundary.
template <typename F>
int guarded(F &&body) {
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return fail(SVO_ERR_NOMEM, "out of host memory");
    } catch (const std::length_error &) {
        return fail(SVO_ERR_NOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(SVO_ERR_INVALID, std::string("internal error: ") + e.what());
    } catch (...) {
        return fail(SVO_ERR_INVALID, "internal error");
    }
}

}  // namespace

extern "C" {

const char *svo_last_error(void) { return g_error.c_str(); }

int svo_version(void) { return 100; }

int svo_device_count(void) { return device_count(); }

int svo_load_device_code(void) {
    return guarded([&]() -> int {
        clear_error();
        hipError_t (*const loads[])() = {load_code_raymarch, load_code_render_batch, load_code_server,
                                         load_code_frame, load_code_build};
        for (auto load : loads) {
            const hipError_t e = load();
            if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("loading device code: ") + hipGetErrorString(e));
        }
        return SVO_OK;
    });
}

int svo_load_memory(const void *data, uint64_t size, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!data || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        std::unique_ptr<svo_octree> h(new (std::nothrow) svo_octree());
        if (!h) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = oct_decode(static_cast<const uint8_t *>(data), size, *h);
        if (r) return r;
        return finish_handle(h, out);
    });
}

int svo_load(const char *path, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!path || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        FileView file;
        if (!file.open(path)) return fail(SVO_ERR_IO, std::string("cannot read ") + path);
        static const uint8_t kNone = 0;  // an empty file decodes (and fails) like any short stream
        return svo_load_memory(file.size() ? file.data() : &kNone, file.size(), out);
    });
}

int svo_encode(const svo_octree *h, void **data, uint64_t *size) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !data || !size) return fail(SVO_ERR_INVALID, "null argument");
        std::vector<uint8_t> buf;
        oct_encode(*h, buf);
        void *p = std::malloc(buf.size());
        if (!p) return fail(SVO_ERR_NOMEM, "out of host memory");
        std::memcpy(p, buf.data(), buf.size());
        *data = p;
        *size = buf.size();
        return SVO_OK;
    });
}

int svo_encode_words(const uint32_t *words, uint64_t count, const float center[3], void **data, uint64_t *size) {
    return guarded([&]() -> int {
        clear_error();
        if (!words || !center || !data || !size || count == 0) return fail(SVO_ERR_INVALID, "null or empty argument");
        svo_octree tmp;
        try {
            tmp.words.assign(words, words + count);
        } catch (...) {
            return fail(SVO_ERR_NOMEM, "out of host memory");
        }
        std::memcpy(tmp.center, center, sizeof(tmp.center));
        return svo_encode(&tmp, data, size);
    });
}

void svo_free_buffer(void *p) { std::free(p); }

int svo_decode_words(const void *data, uint64_t size, uint32_t **words, uint64_t *count, float center[3]) {
    return guarded([&]() -> int {
        clear_error();
        if (!data || !words || !count || !center) return fail(SVO_ERR_INVALID, "null argument");
        *words = nullptr;
        *count = 0;
        svo_octree tmp;
        int r = oct_decode(static_cast<const uint8_t *>(data), size, tmp);
        if (r) return r;
        uint32_t *p = static_cast<uint32_t *>(std::malloc(tmp.words.size() * sizeof(uint32_t)));
        if (!p) return fail(SVO_ERR_NOMEM, "out of host memory");
        std::memcpy(p, tmp.words.data(), tmp.words.size() * sizeof(uint32_t));
        std::memcpy(center, tmp.center, sizeof(tmp.center));
        *words = p;
        *count = tmp.words.size();
        return SVO_OK;
    });
}

int svo_save(const svo_octree *h, const char *path) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !path) return fail(SVO_ERR_INVALID, "null argument");
        std::vector<uint8_t> buf;
        oct_encode(*h, buf);
        FILE *fp = std::fopen(path, "wb");
        if (!fp) return fail(SVO_ERR_IO, std::string("cannot write ") + path);
        bool ok = std::fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
        ok = (std::fclose(fp) == 0) && ok;
        return ok ? SVO_OK : fail(SVO_ERR_IO, std::string("short write to ") + path);
    });
}

int svo_from_words(const uint32_t *words, uint64_t count, const float center[3], svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!words || !center || !out || count == 0) return fail(SVO_ERR_INVALID, "null or empty argument");
        *out = nullptr;
        std::unique_ptr<svo_octree> h(new (std::nothrow) svo_octree());
        if (!h) return fail(SVO_ERR_NOMEM, "out of host memory");
        try {
            h->words.assign(words, words + count);
        } catch (...) {
            return fail(SVO_ERR_NOMEM, "out of host memory");
        }
        std::memcpy(h->center, center, sizeof(h->center));
        return finish_handle(h, out);
    });
}

int svo_center(const svo_octree *h, float out[3]) {
    return guarded([&]() -> int {
        if (!h || !out) return fail(SVO_ERR_INVALID, "null argument");
        std::memcpy(out, h->center, sizeof(h->center));
        return SVO_OK;
    });
}

int svo_word_count(const svo_octree *h, uint64_t *count) {
    return guarded([&]() -> int {
        if (!h || !count) return fail(SVO_ERR_INVALID, "null argument");
        *count = h->words.size();
        return SVO_OK;
    });
}

int svo_copy_words(const svo_octree *h, uint32_t *dst, uint64_t count) {
    return guarded([&]() -> int {
        if (!h || !dst) return fail(SVO_ERR_INVALID, "null argument");
        uint64_t n = count < h->words.size() ? count : h->words.size();
        std::memcpy(dst, h->words.data(), n * sizeof(uint32_t));
        return SVO_OK;
    });
}

int svo_build_stats_get(const svo_octree *h, svo_build_stats *out) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = h->build;
        return SVO_OK;
    });
}

int svo_depth(const svo_octree *h, int *depth) {
    return guarded([&]() -> int {
        if (!h || !depth) return fail(SVO_ERR_INVALID, "null argument");
        *depth = h->depth;
        return SVO_OK;
    });
}

void svo_free(svo_octree *h) {
    if (!h) return;
    release_device(*h);
    delete h;
}

int svo_upload(svo_octree *h, const uint32_t **device_words) {
    return guarded([&]() -> int {
        clear_error();
        if (!h) return fail(SVO_ERR_INVALID, "null handle");
        int r = ensure_device(*h);
        if (r) return r;
        if (device_words) *device_words = h->dwords;
        return SVO_OK;
    });
}

// ---- raymarch ----

int svo_raymarch_batch_device(svo_octree *h, int64_t n, const float *o, const float *d, const float *ray_scale,
                              float ray_scale0, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf,
                              void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h) return fail(SVO_ERR_INVALID, "null handle");
        return raymarch_device(*h, n, o, d, ray_scale, ray_scale0, hit, normal, t, leaf, stream);
    });
}

int svo_raymarch_batch(svo_octree *h, int64_t n, const float *o, const float *d, const float *ray_scale,
                       float ray_scale0, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf) {
    return guarded([&]() -> int {
        clear_error();
        if (!h) return fail(SVO_ERR_INVALID, "null handle");
        if (n < 0) return fail(SVO_ERR_INVALID, "negative ray count");
        if (n == 0) return SVO_OK;
        if (!o || !d) return fail(SVO_ERR_INVALID, "null ray arrays");
        size_t m = size_t(n);
        DevBuf<float> dO, dD, dS, dT;
        DevBuf<uint8_t> dH;
        DevBuf<uint32_t> dN;
        DevBuf<uint64_t> dL;
        int r;
        if ((r = to_device(dO, o, 3 * m)) || (r = to_device(dD, d, 3 * m)) || (r = to_device(dS, ray_scale, m))) return r;
        if ((r = out_buf(dH, hit, m)) || (r = out_buf(dN, normal, m)) || (r = out_buf(dT, t, m)) || (r = out_buf(dL, leaf, m)))
            return r;
        r = raymarch_device(*h, n, dO.p, dD.p, ray_scale ? dS.p : nullptr, ray_scale0, dH.p, dN.p, dT.p, dL.p, nullptr);
        if (r) return r;
        hipError_t e = hipStreamSynchronize(nullptr);  // the null stream the launch went to
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("raymarch: ") + hipGetErrorString(e));
        if ((r = from_device(hit, dH, m)) || (r = from_device(normal, dN, m)) || (r = from_device(t, dT, m)) ||
            (r = from_device(leaf, dL, m)))
            return r;
        return SVO_OK;
    });
}

int svo_raymarch(svo_octree *h, const float o[3], const float d[3], float ray_scale, uint32_t *normal, float *t,
                 int *hit) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !o || !d) return fail(SVO_ERR_INVALID, "null argument");
        // One ray per call (the reference's per-pixel interface, src/Main.cpp:118,181):
        // walked on the calling thread by the kernels' own traversal source
        // (march.hpp compiled for the host). A device launch per ray costs about
        // 30 us of upload, launch and download against well under 1 us of walk;
        // tiles, frames and ray batches go to the GPU entry points.
        uint32_t nrm = 0;
        float tt = 0.0f;
        uint64_t leaf = 0;
        const int code = raymarch_one(*h, o, d, ray_scale, nrm, tt, leaf);
        if (hit) *hit = code;
        if (code == 1 && normal) *normal = nrm;
        if (code != 0 && t) *t = tt;
        return SVO_OK;
    });
}

int svo_orbit_camera(const svo_octree *h, float pitch, float yaw, float radius, svo_camera *out) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !out) return fail(SVO_ERR_INVALID, "null argument");
        orbit_camera(pitch, yaw, radius, h->center, *out);
        return SVO_OK;
    });
}

int svo_render_primary_device(svo_octree *h, const svo_camera *cam, int width, int height, uint8_t *hit,
                              uint32_t *normal, float *t, uint64_t *leaf, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam) return fail(SVO_ERR_INVALID, "null argument");
        return render_primary_device(*h, *cam, width, height, hit, normal, t, leaf, stream);
    });
}

int svo_tile_plan_create(int width, int height, svo_tile_plan **out) {
    return guarded([&]() -> int {
        clear_error();
        return tile_plan_create(width, height, out);
    });
}

int svo_tile_plan_create_batch(const svo_view *views, int nviews, const svo_batch_opts *opts, svo_tile_plan **out) {
    return guarded([&]() -> int {
        clear_error();
        return tile_plan_create_batch(views, nviews, opts, out);
    });
}

void svo_tile_plan_free(svo_tile_plan *p) { tile_plan_free(p); }

int svo_render_primary_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width,
                                      int height, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf,
                                      void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam) return fail(SVO_ERR_INVALID, "null argument");
        return render_primary_planned_device(*h, plan, *cam, width, height, hit, normal, t, leaf, stream);
    });
}

int svo_render_primary(svo_octree *h, const svo_camera *cam, int width, int height, uint8_t *hit,
                       uint32_t *normal, float *t, uint64_t *leaf) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam) return fail(SVO_ERR_INVALID, "null argument");
        if (width <= 0 || height <= 0) return fail(SVO_ERR_INVALID, "image size out of range");
        size_t m = size_t(width) * size_t(height);
        DevBuf<uint8_t> dH;
        DevBuf<uint32_t> dN;
        DevBuf<float> dT;
        DevBuf<uint64_t> dL;
        int r;
        if ((r = out_buf(dH, hit, m)) || (r = out_buf(dN, normal, m)) || (r = out_buf(dT, t, m)) || (r = out_buf(dL, leaf, m)))
            return r;
        r = render_primary_device(*h, *cam, width, height, dH.p, dN.p, dT.p, dL.p, nullptr);
        if (r) return r;
        hipError_t e = hipStreamSynchronize(nullptr);  // the null stream the launch went to
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("render: ") + hipGetErrorString(e));
        if ((r = from_device(hit, dH, m)) || (r = from_device(normal, dN, m)) || (r = from_device(t, dT, m)) ||
            (r = from_device(leaf, dL, m)))
            return r;
        return SVO_OK;
    });
}

int svo_render_shadow_device(svo_octree *h, const svo_camera *cam, int width, int height, const float light[3],
                             float eps, uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow,
                             void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam) return fail(SVO_ERR_INVALID, "null argument");
        return render_shadow_device(*h, *cam, width, height, light, eps, hit, normal, t, leaf, shadow, stream);
    });
}

int svo_render_shadow_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width, int height,
                                     const float light[3], float eps, uint8_t *hit, uint32_t *normal, float *t,
                                     uint64_t *leaf, uint8_t *shadow, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam || !plan) return fail(SVO_ERR_INVALID, "null argument");
        return render_shadow_device(*h, *cam, width, height, light, eps, hit, normal, t, leaf, shadow, stream, plan);
    });
}

int svo_render_shadow(svo_octree *h, const svo_camera *cam, int width, int height, const float light[3], float eps,
                      uint8_t *hit, uint32_t *normal, float *t, uint64_t *leaf, uint8_t *shadow) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam || !light || !shadow) return fail(SVO_ERR_INVALID, "null argument");
        if (width <= 0 || height <= 0) return fail(SVO_ERR_INVALID, "image size out of range");
        size_t m = size_t(width) * size_t(height);
        DevBuf<uint8_t> dH, dS;
        DevBuf<uint32_t> dN;
        DevBuf<float> dT;
        DevBuf<uint64_t> dL;
        int r;
        if ((r = dH.alloc(m)) || (r = dS.alloc(m)) || (r = dT.alloc(m)) || (r = out_buf(dN, normal, m)) ||
            (r = out_buf(dL, leaf, m)))
            return r;
        r = render_shadow_device(*h, *cam, width, height, light, eps, dH.p, dN.p, dT.p, dL.p, dS.p, nullptr);
        if (r) return r;
        hipError_t e = hipStreamSynchronize(nullptr);  // the null stream the launch went to
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("render: ") + hipGetErrorString(e));
        if ((r = from_device(hit, dH, m)) || (r = from_device(normal, dN, m)) || (r = from_device(t, dT, m)) ||
            (r = from_device(leaf, dL, m)) || (r = from_device(shadow, dS, m)))
            return r;
        return SVO_OK;
    });
}

int svo_render_frame_device(svo_octree *h, const svo_camera *cam, int width, int height, int stripes, int stride,
                            uint32_t *rgba, float *depth, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam) return fail(SVO_ERR_INVALID, "null argument");
        return render_frame_device(*h, *cam, width, height, stripes, stride, rgba, depth, stream);
    });
}

int svo_tile_plan_create_frame(int width, int height, int stripes, svo_tile_plan **out) {
    return guarded([&]() -> int {
        clear_error();
        return tile_plan_create_frames(width, height, stripes, 1, out);
    });
}

int svo_tile_plan_create_frames(int width, int height, int stripes, int nframes, svo_tile_plan **out) {
    return guarded([&]() -> int {
        clear_error();
        return tile_plan_create_frames(width, height, stripes, nframes, out);
    });
}

int svo_render_frame_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cam, int width, int height,
                                    int stripes, int stride, uint32_t *rgba, float *depth, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam || !plan) return fail(SVO_ERR_INVALID, "null argument");
        return render_frame_device(*h, *cam, width, height, stripes, stride, rgba, depth, stream, plan);
    });
}

int svo_render_frames_device(svo_octree *h, const svo_camera *cams, int nframes, int width, int height, int stripes,
                             int stride, uint32_t *const *rgba, float *const *depth, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h) return fail(SVO_ERR_INVALID, "null handle");
        return render_frames_device(*h, cams, nframes, width, height, stripes, stride, rgba, depth, stream);
    });
}

int svo_render_frames_planned_device(svo_octree *h, svo_tile_plan *plan, const svo_camera *cams, int nframes,
                                     int width, int height, int stripes, int stride, uint32_t *const *rgba,
                                     float *const *depth, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !plan) return fail(SVO_ERR_INVALID, "null argument");
        return render_frames_device(*h, cams, nframes, width, height, stripes, stride, rgba, depth, stream, plan);
    });
}

int64_t svo_frame_depth_count(int width, int height, int stripes) {
    clear_error();
    return frame_depth_count(width, height, stripes);
}

int svo_render_frame(svo_octree *h, const svo_camera *cam, int width, int height, int stripes, int stride, uint32_t *rgba,
                     float *depth) {
    return guarded([&]() -> int {
        clear_error();
        if (!h || !cam || !rgba) return fail(SVO_ERR_INVALID, "null argument");
        if (width <= 0 || height <= 0) return fail(SVO_ERR_INVALID, "image size out of range");
        const size_t m = size_t(width) * size_t(height);
        const int64_t nd = frame_depth_count(width, height, stripes);
        if (nd < 0) return SVO_ERR_INVALID;  // the size or stripe split is refused (message set)
        DevBuf<uint32_t> dR;
        DevBuf<float> dD;
        int r;
        if ((r = dR.alloc(m)) || (r = out_buf(dD, depth, size_t(nd)))) return r;
        r = render_frame_device(*h, *cam, width, height, stripes, stride, dR.p, dD.p, nullptr);
        if (r) return r;
        hipError_t e = hipStreamSynchronize(nullptr);  // the null stream the launch went to
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("frame: ") + hipGetErrorString(e));
        if ((r = from_device(rgba, dR, m)) || (r = from_device(depth, dD, size_t(nd)))) return r;
        return SVO_OK;
    });
}

// ---- batched rendering ----------------------------------------------------------

int svo_renderer_create(int max_views, svo_renderer **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        return renderer_create(max_views, out);
    });
}

void svo_renderer_free(svo_renderer *r) { renderer_free(r); }

int svo_render_batch_device(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                            void *stream) {
    return guarded([&]() -> int {
        clear_error();
        return render_batch(r, views, nviews, opts, stream);
    });
}

int svo_render_batch_planned_device(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                                    svo_tile_plan *plan, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        return render_batch_planned(r, views, nviews, opts, plan, stream);
    });
}

int svo_render_batch_noalloc(const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        return render_batch_noalloc(views, nviews, opts, stream);
    });
}

int svo_renderer_autotune(svo_renderer *r, const svo_view *views, int nviews, void *stream, svo_batch_opts *best,
                          float *times) {
    return guarded([&]() -> int {
        clear_error();
        return renderer_autotune(r, views, nviews, stream, best, times);
    });
}

int svo_batch_graph_create(svo_renderer *r, const svo_view *views, int nviews, const svo_batch_opts *opts,
                           void *stream, svo_batch_graph **out) {
    return guarded([&]() -> int {
        clear_error();
        return batch_graph_create(r, views, nviews, opts, stream, out);
    });
}

int svo_batch_graph_launch(svo_batch_graph *g, void *stream) {
    return guarded([&]() -> int {
        clear_error();
        return batch_graph_launch(g, stream);
    });
}

void svo_batch_graph_free(svo_batch_graph *g) { batch_graph_free(g); }

int svo_serve_replay(svo_octree *const *trees, int ntrees, const svo_request *requests, const double *arrival_s, int n,
                     const svo_server_opts *opts, double *latency_s, double *makespan_s, uint32_t *const *normal_out,
                     float *const *t_out, uint64_t *const *leaf_out) {
    return guarded([&]() -> int {
        clear_error();
        return serve_replay(trees, ntrees, requests, arrival_s, n, opts, latency_s, makespan_s, normal_out, t_out, leaf_out);
    });
}

// ---- live request server ---------------------------------------------------------

int svo_server_create(svo_octree *const *trees, int ntrees, const svo_server_config *cfg, svo_server **out) {
    return guarded([&]() -> int {
        clear_error();
        return server_create(trees, ntrees, cfg, out);
    });
}

int svo_server_submit(svo_server *s, const svo_request *req, uint64_t *ticket) {
    return guarded([&]() -> int {
        clear_error();
        return server_submit(s, req, ticket);
    });
}

int svo_server_poll(svo_server *s, uint64_t ticket, int *state, double *latency_s) {
    return guarded([&]() -> int {
        clear_error();
        return server_poll(s, ticket, state, latency_s);
    });
}

int svo_server_wait(svo_server *s, uint64_t ticket, double timeout_s, double *latency_s) {
    return guarded([&]() -> int {
        clear_error();
        return server_wait(s, ticket, timeout_s, latency_s);
    });
}

int svo_server_outputs(svo_server *s, uint64_t ticket, uint32_t **normal, float **t, uint64_t **leaf) {
    return guarded([&]() -> int {
        clear_error();
        return server_outputs(s, ticket, normal, t, leaf);
    });
}

int svo_server_copy_outputs(svo_server *s, uint64_t ticket, uint32_t *normal, float *t, uint64_t *leaf) {
    return guarded([&]() -> int {
        clear_error();
        return server_copy_outputs(s, ticket, normal, t, leaf);
    });
}

int svo_server_release(svo_server *s, uint64_t ticket) {
    return guarded([&]() -> int {
        clear_error();
        return server_release(s, ticket);
    });
}

int svo_server_stats_get(svo_server *s, svo_server_stats *out) {
    return guarded([&]() -> int {
        clear_error();
        return server_stats(s, out);
    });
}

int svo_server_run_schedule(svo_server *s, const svo_request *requests, const double *arrival_s, int n,
                            uint64_t *tickets, double *latency_s, double *makespan_s) {
    return guarded([&]() -> int {
        clear_error();
        return server_run_schedule(s, requests, arrival_s, n, tickets, latency_s, makespan_s);
    });
}

void svo_server_destroy(svo_server *s) { server_destroy(s); }

// ---- construction -------------------------------------------------------------

int svo_build_from_mesh(const float *verts, int64_t nverts, int has_normals, const int32_t *face_counts,
                        const int32_t *face_idx, int64_t nfaces, int resolution, const svo_build_opts *opts,
                        svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!verts || !face_counts || !face_idx || !out || nverts <= 0 || nfaces < 0)
            return fail(SVO_ERR_INVALID, "null or empty mesh");
        *out = nullptr;
        std::unique_ptr<svo_octree> h(new (std::nothrow) svo_octree());
        if (!h) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = build_mesh(verts, nverts, has_normals, face_counts, face_idx, nfaces, resolution, opts, *h);
        if (r) return r;
        return finish_handle(h, out);
    });
}

int svo_build_from_ply(const char *path, int resolution, const svo_build_opts *opts, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!path || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        std::unique_ptr<svo_octree> t(new (std::nothrow) svo_octree());
        if (!t) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = build_ply(path, resolution, opts, *t);
        if (r) return r;
        return finish_handle(t, out);
    });
}

int svo_convert_to_volume(const char *ply_path, int resolution, const svo_build_opts *opts, const char *voxel_path) {
    return guarded([&]() -> int {
        clear_error();
        if (!ply_path || !voxel_path) return fail(SVO_ERR_INVALID, "null argument");
        return convert_ply(ply_path, resolution, opts, voxel_path);
    });
}

int svo_build_from_grid(const uint32_t *grid, int w, int h, int d, const svo_build_opts *opts, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!grid || !out || w <= 0 || h <= 0 || d <= 0) return fail(SVO_ERR_INVALID, "null or empty volume");
        *out = nullptr;
        std::unique_ptr<svo_octree> t(new (std::nothrow) svo_octree());
        if (!t) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = build_grid(grid, w, h, d, opts, *t);
        if (r) return r;
        return finish_handle(t, out);
    });
}

int svo_build_from_sdf(const svo_sdf_primitive *prims, int count, int resolution, float shade,
                       const svo_build_opts *opts, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!prims || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        std::unique_ptr<svo_octree> t(new (std::nothrow) svo_octree());
        if (!t) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = build_sdf(prims, count, resolution, shade, opts, *t);
        if (r) return r;
        return finish_handle(t, out);
    });
}

int svo_sdf_volume(const svo_sdf_primitive *prims, int count, int resolution, float shade, uint32_t *grid) {
    return guarded([&]() -> int {
        clear_error();
        if (!prims || !grid) return fail(SVO_ERR_INVALID, "null argument");
        return sdf_volume(prims, count, resolution, shade, grid);
    });
}

int svo_build_from_terrain(int resolution, uint32_t seed, const svo_build_opts *opts, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        std::unique_ptr<svo_octree> t(new (std::nothrow) svo_octree());
        if (!t) return fail(SVO_ERR_NOMEM, "out of host memory");
        int r = build_terrain(resolution, seed, opts, *t);
        if (r) return r;
        return finish_handle(t, out);
    });
}

int svo_terrain_volume(int resolution, uint32_t seed, uint32_t *grid) {
    return guarded([&]() -> int {
        clear_error();
        if (!grid) return fail(SVO_ERR_INVALID, "null argument");
        return terrain_volume(resolution, seed, grid);
    });
}

int svo_terrain_heights(int resolution, uint32_t seed, int32_t *heights, float *h) {
    return guarded([&]() -> int {
        clear_error();
        return terrain_heights(resolution, seed, heights, h);
    });
}

int svo_build_from_voxels(const char *path, const svo_build_opts *opts, svo_octree **out) {
    return guarded([&]() -> int {
        clear_error();
        if (!path || !out) return fail(SVO_ERR_INVALID, "null argument");
        *out = nullptr;
        FILE *fp = std::fopen(path, "rb");
        if (!fp) return fail(SVO_ERR_IO, std::string("cannot read ") + path);
        int32_t dims[3];
        if (std::fread(dims, 4, 3, fp) != 3 || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) {
            std::fclose(fp);
            return fail(SVO_ERR_FORMAT, ".voxel: bad header");
        }
        // the volume's word count, checked against the words the file holds
        // (w*h*d of them follow the header) before anything is sized by it
        uint64_t avail = 0;
        if (std::fseek(fp, 0, SEEK_END) == 0) {
            const long long end = ftello(fp);
            if (end >= 12) avail = uint64_t(end - 12) / 4;
        }
        if (std::fseek(fp, 12, SEEK_SET) != 0) {
            std::fclose(fp);
            return fail(SVO_ERR_IO, std::string("cannot read ") + path);
        }
        uint64_t n = uint64_t(dims[0]);
        bool fits = n <= avail && uint64_t(dims[1]) <= avail / n;
        if (fits) {
            n *= uint64_t(dims[1]);
            fits = uint64_t(dims[2]) <= avail / n;
        }
        if (!fits) {
            std::fclose(fp);
            return fail(SVO_ERR_FORMAT, ".voxel: truncated (the header's volume exceeds the file)");
        }
        n *= uint64_t(dims[2]);
        std::vector<uint32_t> vol;
        try {
            vol.resize(n);
        } catch (...) {
            std::fclose(fp);
            return fail(SVO_ERR_NOMEM, ".voxel: volume does not fit in host memory");
        }
        size_t got = std::fread(vol.data(), 4, n, fp);
        std::fclose(fp);
        if (got != n) return fail(SVO_ERR_FORMAT, ".voxel: truncated");
        return svo_build_from_grid(vol.data(), dims[0], dims[1], dims[2], opts, out);
    });
}

}  // extern "C"
