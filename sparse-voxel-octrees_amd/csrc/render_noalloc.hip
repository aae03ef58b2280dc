// Batched render that allocates nothing: svo_render_batch_noalloc.
//
// Same launch as svo_render_batch_device (render_batch.hip: the grid or the
// persistent form over one tile sequence of every view), for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid it). Its device tables (view table, tile
// bases, persistent-form queue counters) live in `scratch`, a file-scope
// device array of fixed size: at most kMaxViews views per call. Each device
// holds its own copy of the array; its address on a device is looked up once
// (hipGetSymbolAddress) with that device current and kept. Every call on a
// device shares that device's array: calls enqueued on one stream are
// ordered, calls that overlap on one device (other streams, other threads)
// race on it. The queue counters are cleared by a kernel before each
// persistent-form launch (launch_plan). The octrees must already be resident
// on the device (svo_upload).
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "svo_internal.hpp"
#include "view_desc.hpp"

namespace svo {

namespace {

constexpr int kMaxViews = 512;
constexpr int kMaxDevices = 64;

// scratch layout: view table, tile bases (kMaxViews + 1, padded to 256 B),
// queue counters (kQueues, kCounterStride apart)
constexpr size_t kViewBytes = sizeof(ViewDesc) * kMaxViews;
constexpr size_t kBaseBytes = (sizeof(uint64_t) * (kMaxViews + 1) + 255) / 256 * 256;
constexpr size_t kCounterBytes = sizeof(uint32_t) * kQueues * kCounterStride;
constexpr size_t kScratchBytes = kViewBytes + kBaseBytes + kCounterBytes;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

struct DeviceScratch {
    std::mutex mu;
    unsigned char *ptr[kMaxDevices] = {};
    ViewDesc hv[kMaxViews];        // host staging of the view table
    uint64_t htb[kMaxViews + 1];   // host staging of the tile bases
    int cus[kMaxDevices] = {};
};

DeviceScratch &state() {
    static DeviceScratch *s = new DeviceScratch();  // never freed: a static destructor may outlive the runtime
    return *s;
}

}  // namespace

int render_batch_noalloc(const svo_view *views, int nviews, const svo_batch_opts *opts, void *stream) {
    if (!views) return fail(SVO_ERR_INVALID, "null argument");
    if (nviews <= 0) return fail(SVO_ERR_INVALID, "no views");
    if (nviews > kMaxViews) return fail(SVO_ERR_LIMIT, "svo_render_batch_noalloc takes at most 512 views per call");
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("noalloc render: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(SVO_ERR_LIMIT, "device index above 63");
    DeviceScratch &st = state();
    std::lock_guard<std::mutex> lock(st.mu);
    if (!st.ptr[dev]) {
        // this device's copy of the array (the host-side name is not a device address)
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SVO_ERR_DEVICE, std::string("noalloc render: scratch address: ") + hipGetErrorString(e));
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("noalloc render: ") + hipGetErrorString(e));
        st.cus[dev] = prop.multiProcessorCount;
        st.ptr[dev] = static_cast<unsigned char *>(p);
    }
    LaunchPlan plan;
    int r = plan_views(views, nviews, opts, st.cus[dev], st.hv, st.htb, plan, false);
    if (r) return r;
    unsigned char *base = st.ptr[dev];
    ViewDesc *dViews = reinterpret_cast<ViewDesc *>(base);
    uint64_t *dTileBase = reinterpret_cast<uint64_t *>(base + kViewBytes);
    uint32_t *dCounters = reinterpret_cast<uint32_t *>(base + kViewBytes + kBaseBytes);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // pageable sources: each copy has read its source when it returns, so the
    // staging arrays may be refilled by the next call
    e = hipMemcpyAsync(dViews, st.hv, sizeof(ViewDesc) * nviews, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dTileBase, st.htb, sizeof(uint64_t) * (nviews + 1), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("noalloc render: ") + hipGetErrorString(e));
    launch_plan(plan, dViews, dTileBase, dCounters, s);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("noalloc render: ") + hipGetErrorString(e));
    return SVO_OK;
}

}  // namespace svo
