// Request loop for serving renders of resident octrees (continuous batching).
//
// Requests (octree, camera, resolution) arrive over time. The loop keeps a
// FIFO of arrived requests and a few launch slots, each with its own stream,
// renderer, a slab of device output memory and a pinned host slab. Whenever a
// slot is free and requests are waiting, it admits waiting requests in
// arrival order while their output footprint (16 B per ray: normal, t, leaf)
// fits the slot's slab, traces them in one batched launch, and
// copies the outputs to the host slab. Launches of different slots overlap on
// the device, so requests arriving while others run start as soon as a slot
// and memory free up. A request is complete when its outputs are on the host;
// latency is measured from its scheduled arrival to that moment.
//
// A launch that admits a single request uses the plain per-view kernel (no
// view table to stage): at low request rates that is the
// shorter path; results are identical either way. Every slot's stream and both
// kernels are exercised once before the clock starts.
//
// With one_per_launch set, every launch holds exactly one request traced by
// the plain per-view kernel: the baseline the batched loop is compared with.
//
// No counterpart in the reference: its viewer renders one frame at a time on
// CPU threads (src/Main.cpp:204-262).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <vector>

#include "svo_internal.hpp"

namespace svo {

namespace {

constexpr uint64_t kAlign = 256;

uint64_t footprint(const svo_request &r) {
    const uint64_t n = uint64_t(r.width) * uint64_t(r.height);
    auto up = [](uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
    return up(n * 4) + up(n * 4) + up(n * 8);
}

struct Slot {
    hipStream_t stream = nullptr;
    svo_renderer *renderer = nullptr;
    uint8_t *dev = nullptr;
    uint8_t *host = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<int> reqs;
    std::vector<uint64_t> offs;
    uint64_t used = 0;
};

struct Loop {
    std::vector<Slot> slots;
    ~Loop() {
        for (Slot &s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.renderer) renderer_free(s.renderer);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.dev) (void)hipFree(s.dev);
            if (s.host) (void)hipHostFree(s.host);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
    }
};

}  // namespace

int serve_replay(svo_octree *const *trees, int ntrees, const svo_request *reqs, const double *arrival, int n,
                 const svo_server_opts *opts, double *latency, double *makespan, uint32_t *const *normalOut,
                 float *const *tOut, uint64_t *const *leafOut) {
    if (!trees || ntrees <= 0 || !reqs || !arrival || n <= 0 || !latency)
        return fail(SVO_ERR_INVALID, "serve: null or empty argument");
    const int nslots = (opts && opts->streams > 0) ? std::min(opts->streams, 16) : 4;
    const uint64_t budget = (opts && opts->output_budget) ? opts->output_budget : (uint64_t(1) << 30);
    const int maxBatch = (opts && opts->max_batch > 0) ? opts->max_batch : 256;
    const bool single = opts && opts->one_per_launch;
    const bool onDevice = opts && opts->device_outputs;
    svo_batch_opts bo{};
    if (opts) {
        bo.waves_per_cu = opts->waves_per_cu;
        bo.refill_threshold = opts->refill_threshold;
        bo.form = opts->batch_form;
    }
    const uint64_t slab = budget / uint64_t(nslots) / kAlign * kAlign;
    for (int i = 0; i < n; ++i) {
        const svo_request &r = reqs[i];
        if (r.tree < 0 || r.tree >= ntrees || !trees[r.tree]) return fail(SVO_ERR_INVALID, "serve: bad octree index");
        if (r.width <= 0 || r.height <= 0 || r.width > 65536 || r.height > 65536)
            return fail(SVO_ERR_INVALID, "serve: bad request size");
        if (footprint(r) > slab)
            return fail(SVO_ERR_LIMIT, "serve: a request's outputs exceed one slot's share of the output budget");
        if (i && arrival[i] < arrival[i - 1]) return fail(SVO_ERR_INVALID, "serve: arrivals must be sorted");
    }
    for (int t = 0; t < ntrees; ++t) {
        int r = ensure_device(*trees[t]);  // octrees stay resident for the whole loop
        if (r) return r;
    }
    Loop loop;
    loop.slots.resize(size_t(nslots));
    for (Slot &s : loop.slots) {
        hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMalloc(&s.dev, slab);
        if (e == hipSuccess) e = hipHostMalloc(&s.host, slab, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e != hipSuccess) return fail(SVO_ERR_NOMEM, std::string("serve: slot setup: ") + hipGetErrorString(e));
        int r = renderer_create(maxBatch, &s.renderer);
        if (r) return r;
    }

    // warm-up: first use of each stream and kernel, outside the timed loop
    for (Slot &s : loop.slots) {
        const svo_request &r0 = reqs[0];
        svo_view v{};
        v.tree = trees[r0.tree];
        v.camera = r0.camera;
        // no larger than the first request, so it fits the slot's slab
        svo_request w = r0;
        w.width = std::min(r0.width, 8);
        w.height = std::min(r0.height, 8);
        const uint64_t m = uint64_t(w.width) * w.height;
        const uint64_t a = (m * 4 + kAlign - 1) / kAlign * kAlign;
        v.width = w.width;
        v.height = w.height;
        v.normal = reinterpret_cast<uint32_t *>(s.dev);
        v.t = reinterpret_cast<float *>(s.dev + a);
        v.leaf = reinterpret_cast<uint64_t *>(s.dev + 2 * a);
        int r = render_primary_device(*v.tree, v.camera, v.width, v.height, nullptr, v.normal, v.t, v.leaf, s.stream);
        if (!r && !single) r = render_batch(s.renderer, &v, 1, &bo, s.stream);
        if (r) return r;
        // one copy of the whole slab maps every page of both slabs for the copy engine up front
        hipError_t e = hipMemcpyAsync(s.host, s.dev, slab, hipMemcpyDeviceToHost, s.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("serve: warm-up: ") + hipGetErrorString(e));
    }

    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
    std::deque<int> waiting;
    int nextArrival = 0, completed = 0;
    std::vector<svo_view> views;
    while (completed < n) {
        const double tnow = now();
        while (nextArrival < n && arrival[nextArrival] <= tnow) waiting.push_back(nextArrival++);
        // completions
        for (Slot &s : loop.slots) {
            if (!s.busy) continue;
            hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("serve: ") + hipGetErrorString(q));
            const double tdone = now();
            for (size_t k = 0; k < s.reqs.size(); ++k) {
                const int id = s.reqs[k];
                latency[id] = tdone - arrival[id];
                const uint64_t m = uint64_t(reqs[id].width) * reqs[id].height;
                const uint8_t *base = s.host + s.offs[k];
                auto up = [](uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
                if (onDevice) continue;
                if (normalOut && normalOut[id]) std::memcpy(normalOut[id], base, m * 4);
                if (tOut && tOut[id]) std::memcpy(tOut[id], base + up(m * 4), m * 4);
                if (leafOut && leafOut[id]) std::memcpy(leafOut[id], base + 2 * up(m * 4), m * 8);
            }
            completed += int(s.reqs.size());
            s.busy = false;
            s.reqs.clear();
            s.offs.clear();
            s.used = 0;
        }
        // admissions: fill free slots from the FIFO
        for (Slot &s : loop.slots) {
            if (s.busy || waiting.empty()) continue;
            while (!waiting.empty() && int(s.reqs.size()) < (single ? 1 : maxBatch)) {
                const int id = waiting.front();
                const uint64_t fp = footprint(reqs[id]);
                if (s.used + fp > slab) break;
                waiting.pop_front();
                s.reqs.push_back(id);
                s.offs.push_back(s.used);
                s.used += fp;
            }
            if (s.reqs.empty()) continue;
            auto up = [](uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; };
            views.clear();
            for (size_t k = 0; k < s.reqs.size(); ++k) {
                const svo_request &r = reqs[s.reqs[k]];
                const uint64_t m = uint64_t(r.width) * r.height;
                svo_view v{};
                v.tree = trees[r.tree];
                v.camera = r.camera;
                v.width = r.width;
                v.height = r.height;
                v.ray_scale = 0.0f;
                v.normal = reinterpret_cast<uint32_t *>(s.dev + s.offs[k]);
                v.t = reinterpret_cast<float *>(s.dev + s.offs[k] + up(m * 4));
                v.leaf = reinterpret_cast<uint64_t *>(s.dev + s.offs[k] + 2 * up(m * 4));
                views.push_back(v);
            }
            int r;
            if (single || views.size() == 1)
                r = render_primary_device(*views[0].tree, views[0].camera, views[0].width, views[0].height, nullptr,
                                          views[0].normal, views[0].t, views[0].leaf, s.stream);
            else
                r = render_batch(s.renderer, views.data(), int(views.size()), &bo, s.stream);
            if (r) return r;
            hipError_t e = onDevice ? hipSuccess : hipMemcpyAsync(s.host, s.dev, s.used, hipMemcpyDeviceToHost, s.stream);
            if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
            if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("serve: ") + hipGetErrorString(e));
            s.busy = true;
        }
    }
    if (makespan) *makespan = now() - arrival[0];
    return SVO_OK;
}

}  // namespace svo
