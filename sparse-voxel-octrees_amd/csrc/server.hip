// Live request server: continuous batching of render requests.
//
// A caller submits render requests (octree id, camera, resolution) at any
// time from any thread. The server keeps the octrees resident and admits
// requests in arrival order while their outputs (normal 4 B, t 4 B, leaf 8 B
// per ray) fit the output arena (admission by HBM footprint). A request
// admitted while fewer than four launches are in flight runs as an ordinary
// primary launch on a stream of its own (one of four); requests arriving
// while four are in flight wait, and the next stream that frees up takes the
// waiting requests (up to max_batch, 32 by default, while their outputs fit)
// in one batched grid-form launch (render_batch). A launch's completion is an
// event.
//
// With persistent_launch, the requests beyond four in flight go instead to
// one running persistent launch, published through a ring in pinned,
// coherent host memory:
//
//  - a leader wave (wave 0 of workgroup 0) polls the ring's tail, copies each
//    new request's view descriptor into device memory, extends the launch's
//    table of first-tile indices (one entry per request), and publishes the
//    new (request count, tile count) pair with a release store;
//  - every other wave claims one global tile index at a time from one of 8
//    device counters; indices beyond the published tiles are held until published
//    (the wave polls the published pair), so a request that arrives while
//    the kernel is tracing others is traced by the next waves that free up;
//  - a wave traces its 8x8 tile (march.hpp, the reference's raymarch per
//    lane: src/VoxelOctree.cpp:207-346), stores its outputs coherently (agent
//    scope in device memory, system scope in host memory), waits for them
//    and counts the tile done; the wave that completes a request's last tile
//    writes the request's sequence number into a completion word in host
//    memory.
//
// Measured on the 64-octree request stream, the batched launches beat the
// persistent one (14.8 against 12.1 Grays/s under a 1000-request burst, p99
// 0.38 against 0.58 ms at 2000 req/s): hardware dispatch of the tiles, all
// wave slots, no claims or done counters, and outputs in the caches until the
// launch ends.
//
// The host side is a request loop on its own thread: it records completions
// (latency = submit -> completion seen on the host), admits waiting requests
// as memory, streams and ring slots free up, keeps a heartbeat the persistent
// kernel checks, and ends a persistent launch when no request has been in
// flight for `idle_stop_us` (the next admission starts a new one). Work the caller synchronises with
// device-wide (hipDeviceSynchronize, hipFree) therefore waits at most until
// the requests in flight are done and the launch has ended.
//
// Every wait inside the kernel is bounded: waves leave once the host asks the
// launch to stop (only when nothing is in flight) or when the host's
// heartbeat has not moved for 2 s; a wave holding a published tile always
// traces it first, so every published tile is traced before the launch ends.
//
// Outputs live in the server's arena: device memory (results stay in HBM, the
// caller reads them with its own kernels or copies) or pinned host memory
// (the kernel writes them straight over the host link). No counterpart in the
// reference, whose viewer renders one frame at a time on CPU threads
// (src/Main.cpp:204-262,342-367).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "march.hpp"
#include "range_alloc.hpp"
#include "svo_internal.hpp"
#include "view_desc.hpp"

namespace svo {

namespace {

using namespace march;

constexpr uint64_t kAlign = 256;
constexpr int kCountShift = 44;                     // published word: requests << 44 | tiles
constexpr uint64_t kTileMask = (uint64_t(1) << kCountShift) - 1;
constexpr uint64_t kHeartbeatTicks = 200000000ull;  // 2 s at the 100 MHz wall clock

// Pinned host memory, shared with the kernel (coherent).
struct ServeCtrl {
    uint64_t tail;       // requests published (global sequence numbers < tail)
    uint64_t stop;       // host: end the launch (only set with nothing in flight)
    uint64_t heartbeat;  // host loop counter
    uint64_t mirrored;   // kernel: requests of this launch copied to the device tables (global)
    uint64_t pad[4];
};
struct ServeSlot {       // one ring entry, written by the host before it publishes
    ViewDesc desc;
    uint64_t seq;
    uint32_t tiles;
    uint32_t pad[13];
};
static_assert(sizeof(ServeSlot) == 192, "ServeSlot layout");

// Tiles are claimed from kCursors counters, each on a 128-byte line of its
// own: workgroup b claims from counter b % kCursors, and that counter's i-th
// claim is tile i * kCursors + b % kCursors. One counter for every wave would
// take one device-scope atomic per tile on one address, and those serialise
// (about 12 ns each: a 9,500-tile request spends ~110 us in them). A request's
// finished tiles are counted the same way, per class of tile index modulo
// kCursors, and the class counters that reach their share count the request.
constexpr int kCursors = 8;
constexpr int kLineWords = 32;  // uint32 per 128-byte line

// Device memory: launch state (cleared before each launch).
struct ServeState {
    struct {
        unsigned long long next;   // claims made from this counter
        uint64_t pad[15];
    } cursor[kCursors];
    uint64_t published;            // (requests << 44) | tiles, stored by the leader after the tables
    unsigned int stop;             // set by the leader: waiting waves leave
    unsigned int pad[29];
};

// done[]: per ring slot, kCursors class counters and one count of finished
// classes, each on its own line.
__device__ __forceinline__ uint32_t *done_word(uint32_t *done, int slot, int k) {
    return done + (size_t(slot) * (kCursors + 1) + size_t(k)) * kLineWords;
}
// tiles t in [a, b) with t % kCursors == c
__device__ __forceinline__ uint32_t class_tiles(uint64_t a, uint64_t b, int c) {
    const auto below = [c](uint64_t x) { return (x + uint64_t(kCursors - 1 - c)) / kCursors; };
    return uint32_t(below(b) - below(a));
}

// Memory ordering without cache maintenance. A release or acquire at agent
// or system scope writes back or invalidates the whole L2 of the wave's XCD
// (buffer_wbl2 / buffer_inv), which, once per tile, costs far more than the
// tile and evicts the octree. Instead every word shared between waves of
// different XCDs or with the host is read and written with scoped relaxed
// atomics (coherent at that scope, so never a stale cached copy), and a
// writer waits for its stores to be performed (s_waitcnt vmcnt(0)) before
// the store or atomic that publishes them.
template <typename T>
__device__ __forceinline__ T ld_sys(T *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
template <typename T>
__device__ __forceinline__ T ld_dev(T *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <typename T>
__device__ __forceinline__ void st_sys(T *p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM); }
template <typename T>
__device__ __forceinline__ void st_dev(T *p, T v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
// s_waitcnt vmcnt(0): this wave's vector memory operations are performed
__device__ __forceinline__ void wait_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Lane 0's value in a wave-uniform (scalar) register.
__device__ __forceinline__ uint64_t lane0(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), 0), hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), 0);
    return (uint64_t(hi) << 32) | lo;
}

// A request's view descriptor, read coherently: lane k < 32 reads word k,
// every field is then taken from its lane into a wave-uniform register.
__device__ __forceinline__ ViewDesc load_desc(const ViewDesc *p, int lane) {
    uint32_t w = 0;
    if (lane < 32) w = ld_dev(const_cast<uint32_t *>(reinterpret_cast<const uint32_t *>(p)) + lane);
    uint32_t words[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) words[k] = __builtin_amdgcn_readlane(w, k);
    ViewDesc d;
    __builtin_memcpy(&d, words, sizeof(d));
    return d;
}

// One ray's outputs as scoped stores, visible once the wave has waited for
// them: system scope for outputs in host memory (read by the host), agent
// scope for outputs in device memory (read by later kernels and copies, which
// start after the host has seen the request complete). Requests carry no hit
// array.
template <typename T>
__device__ __forceinline__ void st_out(T *p, T v, bool host) {
    if (host) st_sys(p, v);
    else st_dev(p, v);
}
template <typename Idx>
__device__ __forceinline__ void store_out(const ViewDesc &vd, int64_t i, int code, uint32_t normal, float t, Idx leaf,
                                          bool host) {
    if (vd.normal) st_out(vd.normal + i, code == kLeaf ? normal : 0u, host);
    if (vd.t) st_out(reinterpret_cast<uint32_t *>(vd.t) + i, code == kMiss ? 0u : __float_as_uint(t), host);
    if (vd.leaf) st_out(reinterpret_cast<unsigned long long *>(vd.leaf) + i,
                        (unsigned long long)(code == kLeaf ? uint64_t(leaf) : SVO_LEAF_NONE), host);
}

// Persistent serving kernel. Workgroups of 256 threads (4 waves); wave 0 of
// workgroup 0 is the leader, every other wave traces tiles.
template <typename Idx, int Reg>
__global__ void __launch_bounds__(256) serve_kernel(ServeCtrl *ctrl, const ServeSlot *hslots, ServeState *st,
                                                    ViewDesc *dslots, uint64_t *dbase,
                                                    uint32_t *done, uint64_t *complete, int ring, uint64_t seq0,
                                                    uint32_t maxReqs, int slots, int hostOut) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    if (blockIdx.x == 0 && wave == 0) {
        // ---- leader: mirror the host ring into the device tables ----
        uint32_t mirrored = 0;
        uint64_t total = 0;
        uint64_t hb = ld_sys(&ctrl->heartbeat);
        uint64_t hbTime = wall_clock64();
        for (;;) {
            // the host writes a slot before it moves the tail; the slot reads
            // below depend on the tail read, so they follow it
            const uint64_t tail = ld_sys(&ctrl->tail);
            uint64_t avail = tail > seq0 ? tail - seq0 : 0;
            if (avail > maxReqs) avail = maxReqs;
            if (avail > mirrored) {
                // two requests per pass: lanes 0-31 copy the first's 32
                // descriptor words, lanes 32-63 the second's
                for (uint32_t i = mirrored; i < uint32_t(avail); i += 2) {
                    const uint32_t k = i + uint32_t(lane >> 5);
                    const bool mine = k < uint32_t(avail);
                    uint32_t tiles = 0;
                    if (mine) {
                        const int s = int((seq0 + k) % uint64_t(ring));
                        uint32_t *src = const_cast<uint32_t *>(reinterpret_cast<const uint32_t *>(&hslots[s].desc)) + (lane & 31);
                        st_dev(reinterpret_cast<uint32_t *>(&dslots[s]) + (lane & 31), ld_sys(src));
                        tiles = ld_sys(const_cast<uint32_t *>(&hslots[s].tiles));
                    }
                    const uint32_t t0 = __shfl(tiles, 0), t1 = __shfl(tiles, 32);
                    if (lane == 0) st_dev(&dbase[i], total);
                    total += t0;
                    if (i + 1 < uint32_t(avail)) {
                        if (lane == 0) st_dev(&dbase[i + 1], total);
                        total += t1;
                    }
                }
                mirrored = uint32_t(avail);
                if (lane == 0) st_dev(&dbase[mirrored], total);
                // every lane's table stores performed before the publishing store
                wait_stores();
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) {
                    st_dev(&st->published, (uint64_t(mirrored) << kCountShift) | total);
                    st_sys(&ctrl->mirrored, seq0 + mirrored);
                }
            }
            bool quit = ld_sys(&ctrl->stop) != 0;
            const uint64_t h = ld_sys(&ctrl->heartbeat);
            const uint64_t now = wall_clock64();
            if (h != hb) {
                hb = h;
                hbTime = now;
            } else if (now - hbTime > kHeartbeatTicks) {
                quit = true;  // the host loop is gone: end the launch
            }
            if (quit) {
                if (lane == 0) st_dev(&st->stop, 1u);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }

    // ---- tracing waves ----
    Stack<Idx> stk(lds + size_t(wave) * slots * 64 * (sizeof(Idx) + sizeof(float)), 64, slots, lane);
    uint64_t known = 0;          // last published pair seen
    uint32_t q = 0;              // launch-relative request of the last tile
    uint64_t qa = 0, qb = 0;     // its tiles [qa, qb); qb = 0 until the first tile
    int s = 0;                   // its ring slot
    ViewDesc vd{};               // and view, kept while the wave's tiles stay in it
    const int cls = int(blockIdx.x % kCursors);
    unsigned long long g = 0;
    if (lane == 0) g = atomicAdd(&st->cursor[cls].next, 1ull) * kCursors + uint64_t(cls);
    g = lane0(g);
    for (;;) {
        // hold the tile index until it is published (or the launch ends);
        // each poll is a device-scope load of one word by every waiting wave,
        // so a wave that keeps finding nothing polls less often (up to ~8 K
        // cycles apart)
        for (uint32_t naps = 0; g >= (known & kTileMask);) {
            uint64_t p = 0;
            if (lane == 0) p = ld_dev(&st->published);
            known = lane0(p);
            if (g < (known & kTileMask)) break;
            uint32_t stop = 0;
            if (lane == 0) stop = ld_dev(&st->stop);
            if (__builtin_amdgcn_readlane(stop, 0)) return;
            for (uint32_t k = 0; k < (1u << naps); ++k) __builtin_amdgcn_s_sleep(2);
            if (naps < 6) ++naps;
        }
        if (g >= qb) {
            // a later request: dbase[lo] <= g < dbase[hi] (g only grows per
            // wave; the tables of published requests do not change)
            uint32_t lo = qb ? q + 1 : 0, hi = uint32_t(known >> kCountShift);
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ld_dev(&dbase[mid]) <= g) lo = mid; else hi = mid;
            }
            q = lo;
            s = int((seq0 + q) % uint64_t(ring));
            vd = load_desc(&dslots[s], lane);
            qa = ld_dev(&dbase[q]);
            qb = ld_dev(&dbase[q + 1]);
        }
        const uint64_t local = g - qa;
        const int tx = int(local % uint64_t(vd.tilesX));
        const int ty = int(local / uint64_t(vd.tilesX));
        const int x = tx * 8 + (lane & 7);
        const int y = ty * 8 + (lane >> 3);
        if (x < vd.width && y < vd.height) {
            float dx, dy, dz;
            view_ray_dir(vd, x, y, dx, dy, dz);
            Ray<Idx> r;
            ray_init(r, vd.tree, vd.px, vd.py, vd.pz, dx, dy, dz, 0.0f);
            uint32_t normal = 0;
            float t = 0.0f;
            Idx leaf = 0;
            int code;
            do {
                code = ray_step<false, true, Idx, true, Reg>(r, vd.tree, stk, normal, t, leaf);
            } while (code == kRunning);
            store_out<Idx>(vd, int64_t(y) * vd.width + x, code, normal, t, leaf, hostOut != 0);
        }
        // this tile's outputs performed, then counted in its class; the next
        // claim travels with the count (one round trip for both)
        wait_stores();
        __builtin_amdgcn_wave_barrier();
        unsigned long long next = 0;
        if (lane == 0) {
            uint32_t *cnt = done_word(done, s, cls);
            const uint32_t prior = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            next = atomicAdd(&st->cursor[cls].next, 1ull) * kCursors + uint64_t(cls);
            if (prior + 1 == class_tiles(qa, qb, cls)) {
                // the class's last tile: reset its counter for the slot's next
                // request (performed before the class is counted), count the class
                st_dev(cnt, 0u);
                wait_stores();
                uint32_t *all = done_word(done, s, kCursors);
                const uint32_t classes = uint32_t(qb - qa < uint64_t(kCursors) ? qb - qa : uint64_t(kCursors));
                if (__hip_atomic_fetch_add(all, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == classes) {
                    // the request's last class: every tile's outputs were
                    // performed before its count; reset, then publish completion
                    st_dev(all, 0u);
                    wait_stores();
                    st_sys(&complete[s], seq0 + q + 1);
                }
            }
        }
        g = lane0(next);
    }
}

struct Ticket {
    svo_request req;
    uint64_t off = 0, bytes = 0;  // arena range
    uint64_t seq = 0;             // global sequence number once admitted
    uint32_t tiles = 0;
    int state = 0;                // 0 waiting, 1 in flight, 2 done
    double submitted = 0.0, completed = 0.0;
};

uint64_t round_up(uint64_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

uint64_t request_bytes(const svo_request &r) {
    const uint64_t n = uint64_t(r.width) * uint64_t(r.height);
    return round_up(n * 4) * 2 + round_up(n * 8);
}

using Clock = std::chrono::steady_clock;

}  // namespace

}  // namespace svo

struct svo_server {
    std::vector<svo_octree *> trees;
    int ring = 0;
    uint32_t maxReqs = 0;
    bool hostOutputs = false;
    uint64_t budget = 0;
    double idleStop = 0.0;
    int device = 0, cus = 0, wavesPerCU = 24;
    int reg = 1, stackSlots = 1;
    bool wide = false;
    // memory
    svo::ServeCtrl *ctrl = nullptr;   // pinned
    svo::ServeSlot *hslots = nullptr; // pinned
    uint64_t *complete = nullptr;     // pinned, one word per ring slot
    svo::ServeState *st = nullptr;    // device
    svo::ViewDesc *dslots = nullptr;  // device
    uint64_t *dbase = nullptr;        // device
    uint32_t *done = nullptr;         // device
    uint8_t *arena = nullptr;         // device or pinned
    hipStream_t stream = nullptr;
    hipStream_t copyStream = nullptr;  // copies of completed outputs to the host
    hipStream_t directStreams[4] = {};   // requests run as their own launch, in turn
    hipEvent_t ended = nullptr;
    bool persistentOnly = false;
    svo::RangeAlloc alloc;  // first fit over the output arena (offsets in bytes, kAlign units)
    // request state
    std::mutex mu;
    std::condition_variable cv;
    std::condition_variable loopCv;  // wakes the request loop when a launch starts
    std::unordered_map<uint64_t, svo::Ticket> tickets;
    std::deque<uint64_t> waiting;        // submitted, not yet admitted
    std::deque<uint64_t> inflight;       // admitted to the persistent launch, in sequence order
    std::deque<std::pair<uint64_t, hipEvent_t>> direct;  // running as their own launch
    std::deque<std::pair<std::vector<uint64_t>, hipEvent_t>> groups;  // running in one batched launch each
    std::vector<hipEvent_t> events;      // free completion events of direct and batched launches
    uint64_t directLaunches = 0, groupLaunches = 0, grouped = 0;
    static constexpr size_t kMaxDirect = 4;  // direct and batched launches in flight at most
    static constexpr int kMaxGroup = 256;    // the largest max_batch (the renderers' view tables)
    bool batchedLaunches = false;
    int maxGroup = 32;                   // requests per batched launch at most (burst of 1000 over
                                         // the 64 octrees: p50 22.8 ms at 32, 24-30 ms at 64-256,
                                         // the same 14.5-14.7 Grays/s)
    svo_renderer *renderers[4] = {};     // batched launches, one per stream
    std::vector<uint64_t> slotTicket;    // ring slot -> ticket
    uint64_t nextTicket = 1, nextSeq = 0;
    bool live = false;
    uint64_t seq0 = 0;                   // first sequence number of the running launch
    svo::Clock::time_point t0, lastBusy;
    int error = SVO_OK;
    std::string errorMsg;
    // loop thread
    std::thread loop;
    bool quit = false;
    uint64_t launches = 0, published = 0;
};

namespace svo {

namespace {

double now_s(const svo_server &s) { return std::chrono::duration<double>(Clock::now() - s.t0).count(); }

void set_error(svo_server &s, const std::string &msg) {
    if (s.error == SVO_OK) {
        s.error = SVO_ERR_DEVICE;
        s.errorMsg = msg;
    }
}

// Makes the server's device current for the scope (admission runs on the
// submitting thread).
struct OnDevice {
    int prev = -1, dev;
    explicit OnDevice(int d) : dev(d) {
        if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    }
    ~OnDevice() {
        if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    }
};

// Starts a launch whose first request is sequence number s.nextSeq (called
// with the lock held, no launch running).
bool start_launch(svo_server &s) {
    OnDevice on(s.device);
    s.seq0 = s.nextSeq;
    __atomic_store_n(&s.ctrl->stop, 0ull, __ATOMIC_RELEASE);
    __atomic_store_n(&s.ctrl->mirrored, s.seq0, __ATOMIC_RELEASE);
    hipError_t e = hipMemsetAsync(s.st, 0, sizeof(ServeState), s.stream);
    if (e != hipSuccess) {
        set_error(s, std::string("server launch: ") + hipGetErrorString(e));
        return false;
    }
    const size_t lds = size_t(4) * s.stackSlots * 64 * ((s.wide ? 8 : 4) + sizeof(float));
    const int blocksPerCU = std::max(1, std::min(s.wavesPerCU / 4, int((160 * 1024) / lds)));
    const dim3 grid(unsigned(s.cus * blocksPerCU)), block(256);
    if (!s.wide && s.reg == 1)
        hipLaunchKernelGGL((serve_kernel<uint32_t, 1>), grid, block, lds, s.stream, s.ctrl, s.hslots, s.st, s.dslots,
                           s.dbase, s.done, s.complete, s.ring, s.seq0, s.maxReqs, s.stackSlots, int(s.hostOutputs));
    else if (!s.wide)
        hipLaunchKernelGGL((serve_kernel<uint32_t, 3>), grid, block, lds, s.stream, s.ctrl, s.hslots, s.st, s.dslots,
                           s.dbase, s.done, s.complete, s.ring, s.seq0, s.maxReqs, s.stackSlots, int(s.hostOutputs));
    else if (s.reg == 1)
        hipLaunchKernelGGL((serve_kernel<uint64_t, 1>), grid, block, lds, s.stream, s.ctrl, s.hslots, s.st, s.dslots,
                           s.dbase, s.done, s.complete, s.ring, s.seq0, s.maxReqs, s.stackSlots, int(s.hostOutputs));
    else
        hipLaunchKernelGGL((serve_kernel<uint64_t, 3>), grid, block, lds, s.stream, s.ctrl, s.hslots, s.st, s.dslots,
                           s.dbase, s.done, s.complete, s.ring, s.seq0, s.maxReqs, s.stackSlots, int(s.hostOutputs));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(s.ended, s.stream);
    if (e != hipSuccess) {
        set_error(s, std::string("server launch: ") + hipGetErrorString(e));
        return false;
    }
    s.live = true;
    s.launches++;
    s.loopCv.notify_one();  // the loop watches the launch from now on
    return true;
}

// Asks the running launch to end and waits for it (lock held; nothing in flight).
void stop_launch(svo_server &s) {
    if (!s.live) return;
    __atomic_store_n(&s.ctrl->stop, 1ull, __ATOMIC_RELEASE);
    hipError_t e = hipEventSynchronize(s.ended);
    if (e != hipSuccess) set_error(s, std::string("server stop: ") + hipGetErrorString(e));
    s.live = false;
}

// Output pointers of a request placed at arena offset `off`.
void request_outputs(const svo_server &s, const Ticket &tk, uint32_t *&normal, float *&t, uint64_t *&leaf) {
    const uint64_t n = uint64_t(tk.req.width) * uint64_t(tk.req.height);
    normal = reinterpret_cast<uint32_t *>(s.arena + tk.off);
    t = reinterpret_cast<float *>(s.arena + tk.off + round_up(n * 4));
    leaf = reinterpret_cast<uint64_t *>(s.arena + tk.off + 2 * round_up(n * 4));
}

// While no persistent launch runs and fewer than kMaxDirect requests are in
// flight, a request runs as an ordinary primary launch (the kernel
// svo_render_primary_device uses) on one of kMaxDirect streams in turn: on a
// lightly loaded GPU that ends sooner than folding it into the persistent
// launch, which pays off once more requests overlap. Returns false when the
// arena cannot hold it yet or on an error (lock held).
bool launch_direct(svo_server &s, uint64_t id) {
    Ticket &tk = s.tickets[id];
    uint64_t off = 0;
    if (!s.alloc.alloc(tk.bytes, off)) return false;
    tk.off = off;
    hipEvent_t ev = nullptr;
    if (!s.events.empty()) {
        ev = s.events.back();
        s.events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        s.alloc.release(off, tk.bytes);
        set_error(s, "server: event for a direct launch");
        return false;
    }
    uint32_t *normal;
    float *t;
    uint64_t *leaf;
    request_outputs(s, tk, normal, t, leaf);
    OnDevice on(s.device);
    hipStream_t st = s.directStreams[(s.directLaunches + s.groupLaunches) % svo_server::kMaxDirect];
    int r = render_primary_device(*s.trees[size_t(tk.req.tree)], tk.req.camera, tk.req.width, tk.req.height, nullptr,
                                  normal, t, leaf, st);
    hipError_t e = r ? hipSuccess : hipEventRecord(ev, st);
    if (r || e != hipSuccess) {
        s.alloc.release(off, tk.bytes);
        s.events.push_back(ev);
        set_error(s, r ? std::string("server: direct launch: ") + svo_last_error()
                       : std::string("server: direct launch: ") + hipGetErrorString(e));
        return false;
    }
    tk.state = 1;
    s.direct.emplace_back(id, ev);
    s.published++;
    s.directLaunches++;
    s.lastBusy = Clock::now();
    s.loopCv.notify_one();  // the loop watches the launch
    return true;
}

// With batched launches: the waiting requests at the front of the queue
// (up to maxGroup, while their outputs fit the arena) run in one batched
// grid-form launch (render_batch) on the next stream; its completion is an
// event. Returns false when nothing could be admitted or on an error (lock
// held).
bool launch_group(svo_server &s) {
    std::vector<uint64_t> ids;
    std::vector<svo_view> views;
    const size_t most = std::min<size_t>(s.waiting.size(), size_t(s.maxGroup));
    for (size_t k = 0; k < most; ++k) {
        Ticket &tk = s.tickets[s.waiting[k]];
        uint64_t off = 0;
        if (!s.alloc.alloc(tk.bytes, off)) break;  // admission by output footprint, in arrival order
        tk.off = off;
        ids.push_back(s.waiting[k]);
        svo_view v{};
        v.tree = s.trees[size_t(tk.req.tree)];
        v.camera = tk.req.camera;
        v.width = tk.req.width;
        v.height = tk.req.height;
        request_outputs(s, tk, v.normal, v.t, v.leaf);
        views.push_back(v);
    }
    if (ids.empty()) return false;
    auto release = [&]() {
        for (uint64_t id : ids) s.alloc.release(s.tickets[id].off, s.tickets[id].bytes);
    };
    if (ids.size() == 1) {  // one request: the plain kernel
        release();
        if (!launch_direct(s, ids[0])) return false;
        s.waiting.pop_front();
        return true;
    }
    hipEvent_t ev = nullptr;
    if (!s.events.empty()) {
        ev = s.events.back();
        s.events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        release();
        set_error(s, "server: event for a batched launch");
        return false;
    }
    OnDevice on(s.device);
    const size_t k = (s.directLaunches + s.groupLaunches) % svo_server::kMaxDirect;
    hipStream_t st = s.directStreams[k];
    int r = render_batch(s.renderers[k], views.data(), int(views.size()), nullptr, st);
    hipError_t e = r ? hipSuccess : hipEventRecord(ev, st);
    if (r || e != hipSuccess) {
        release();
        s.events.push_back(ev);
        set_error(s, r ? std::string("server: batched launch: ") + svo_last_error()
                       : std::string("server: batched launch: ") + hipGetErrorString(e));
        return false;
    }
    for (uint64_t id : ids) {
        s.tickets[id].state = 1;
        s.waiting.pop_front();  // ids are the queue's front, in order
    }
    s.published += ids.size();
    s.grouped += ids.size();
    s.groupLaunches++;
    s.groups.emplace_back(std::move(ids), ev);
    s.lastBusy = Clock::now();
    s.loopCv.notify_one();  // the loop watches the launch
    return true;
}

// Admits waiting requests in arrival order while arena memory, a ring slot
// and room in the launch's request table are free, and publishes them (to a
// direct launch when nothing is in flight; with batched launches, to direct
// and batched launches only).
void admit(svo_server &s) {
    if (s.error) return;
    while (!s.waiting.empty()) {
        Ticket &tk = s.tickets[s.waiting.front()];
        if (!s.persistentOnly && !s.live && s.inflight.empty() &&
            s.direct.size() + s.groups.size() < svo_server::kMaxDirect) {
            // no persistent launch and few launches in flight: the request
            // runs as its own launch (with batched launches, the waiting
            // requests together)
            if (s.batchedLaunches && s.waiting.size() > 1) {
                if (!launch_group(s)) return;
                continue;
            }
            if (!launch_direct(s, s.waiting.front())) return;
            s.waiting.pop_front();
            continue;
        }
        if (s.batchedLaunches && !s.persistentOnly) return;  // the next free stream takes them
        // ring window: the oldest request in flight holds its slot
        const uint64_t oldest = s.inflight.empty() ? s.nextSeq : s.tickets[s.inflight.front()].seq;
        if (s.nextSeq - oldest >= uint64_t(s.ring)) return;
        if (s.live && s.nextSeq - s.seq0 >= s.maxReqs) {
            // this launch's request table is full: start a new launch once it drains
            if (!s.inflight.empty()) return;
            stop_launch(s);
            if (s.error) return;
        }
        uint64_t off = 0;
        if (!s.alloc.alloc(tk.bytes, off)) return;  // admission by output footprint
        if (!s.live && !start_launch(s)) {
            s.alloc.release(off, tk.bytes);
            return;
        }
        tk.off = off;
        const uint64_t id = s.waiting.front();
        s.waiting.pop_front();
        tk.seq = s.nextSeq++;
        const int slot = int(tk.seq % uint64_t(s.ring));
        s.slotTicket[size_t(slot)] = id;
        ServeSlot &hs = s.hslots[slot];
        const svo_request &r = tk.req;
        svo_view v{};
        v.tree = s.trees[size_t(r.tree)];
        v.camera = r.camera;
        v.width = r.width;
        v.height = r.height;
        request_outputs(s, tk, v.normal, v.t, v.leaf);
        ViewDesc d;
        (void)fill_view_desc(v, d);  // checked at submit
        std::memcpy(&hs.desc, &d, sizeof(d));
        hs.seq = tk.seq;
        hs.tiles = tk.tiles;
        tk.state = 1;
        s.inflight.push_back(id);
        __atomic_store_n(&s.ctrl->tail, tk.seq + 1, __ATOMIC_RELEASE);  // publish
        s.published++;
        s.lastBusy = Clock::now();
    }
}

// Records completions of requests in flight (lock held).
bool reap(svo_server &s) {
    bool any = false;
    for (auto it = s.direct.begin(); it != s.direct.end();) {
        const hipError_t e = hipEventQuery(it->second);
        if (e == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (e != hipSuccess) set_error(s, std::string("server: direct launch: ") + hipGetErrorString(e));
        Ticket &tk = s.tickets[it->first];
        tk.state = 2;
        tk.completed = now_s(s);
        s.events.push_back(it->second);
        it = s.direct.erase(it);
        any = true;
    }
    for (auto it = s.groups.begin(); it != s.groups.end();) {
        const hipError_t e = hipEventQuery(it->second);
        if (e == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (e != hipSuccess) set_error(s, std::string("server: batched launch: ") + hipGetErrorString(e));
        const double now = now_s(s);
        for (uint64_t id : it->first) {
            Ticket &tk = s.tickets[id];
            tk.state = 2;
            tk.completed = now;
        }
        s.events.push_back(it->second);
        it = s.groups.erase(it);
        any = true;
    }
    for (auto it = s.inflight.begin(); it != s.inflight.end();) {
        Ticket &tk = s.tickets[*it];
        const int slot = int(tk.seq % uint64_t(s.ring));
        if (__atomic_load_n(&s.complete[slot], __ATOMIC_ACQUIRE) == tk.seq + 1) {
            tk.state = 2;
            tk.completed = now_s(s);
            it = s.inflight.erase(it);
            any = true;
        } else {
            ++it;
        }
    }
    return any;
}

// Drops a ticket that is waiting or done (lock held): its arena range goes
// back to the allocator and waiting requests may now fit.
void alloc_release(svo_server &s, std::unordered_map<uint64_t, Ticket>::iterator it) {
    if (it->second.state == 0) {
        auto w = std::find(s.waiting.begin(), s.waiting.end(), it->first);
        if (w != s.waiting.end()) s.waiting.erase(w);
        s.tickets.erase(it);
        return;
    }
    s.alloc.release(it->second.off, it->second.bytes);
    s.tickets.erase(it);
    admit(s);
}

void loop_main(svo_server *sp) {
    svo_server &s = *sp;
    (void)hipSetDevice(s.device);
    std::unique_lock<std::mutex> lock(s.mu);
    while (!s.quit) {
        __atomic_fetch_add(&s.ctrl->heartbeat, 1ull, __ATOMIC_RELEASE);
        if (reap(s)) {
            s.lastBusy = Clock::now();
            admit(s);
            s.cv.notify_all();
        }
        if (s.live) {
            if (hipEventQuery(s.ended) == hipSuccess) {
                // the launch ended by itself (heartbeat lost): requests it did
                // not take over are published again in a new launch
                s.live = false;
                if (!s.inflight.empty()) {
                    const uint64_t mirrored = __atomic_load_n(&s.ctrl->mirrored, __ATOMIC_ACQUIRE);
                    reap(s);
                    if (!s.inflight.empty()) {
                        const uint64_t first = s.tickets[s.inflight.front()].seq;
                        if (mirrored <= first) {
                            // not yet taken over: republish from `first`
                            s.nextSeq = first;
                            for (auto it = s.inflight.rbegin(); it != s.inflight.rend(); ++it) {
                                Ticket &tk = s.tickets[*it];
                                tk.state = 0;
                                s.waiting.push_front(*it);
                            }
                            s.inflight.clear();
                        } else {
                            set_error(s, "server: a launch ended with requests in flight");
                        }
                    }
                }
                admit(s);
                s.cv.notify_all();
            } else if (s.inflight.empty() &&
                       std::chrono::duration<double>(Clock::now() - s.lastBusy).count() > s.idleStop) {
                // nothing in flight: requests still waiting for arena memory
                // cannot start before the caller releases outputs, so the
                // launch ends (a device-wide synchronisation by the caller
                // would otherwise wait on it); admission starts a new one
                stop_launch(s);
                s.cv.notify_all();
            }
        } else if (!s.waiting.empty()) {
            admit(s);
        }
        if (s.error) s.cv.notify_all();
        if (s.live || !s.direct.empty() || !s.groups.empty()) {
            // spin briefly without the lock, so submitters and pollers get in
            lock.unlock();
            for (int k = 0; k < 64; ++k) __builtin_ia32_pause();
            lock.lock();
        } else {
            // nothing running: sleep until a launch starts (start_launch and
            // launch_direct signal) or a waiting request may be admitted
            s.loopCv.wait_for(lock, std::chrono::microseconds(200));
        }
    }
    if (!s.inflight.empty()) {
        // destroy with requests in flight: let the launch finish them
        const auto deadline = Clock::now() + std::chrono::seconds(10);
        while (!s.inflight.empty() && Clock::now() < deadline) {
            __atomic_fetch_add(&s.ctrl->heartbeat, 1ull, __ATOMIC_RELEASE);
            reap(s);
        }
    }
    stop_launch(s);
}

void free_server(svo_server *s) {
    if (!s) return;
    if (s->loop.joinable()) {
        {
            std::lock_guard<std::mutex> lock(s->mu);
            s->quit = true;
        }
        s->loopCv.notify_one();
        s->loop.join();
    }
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (hipStream_t ds : s->directStreams)
        if (ds) (void)hipStreamSynchronize(ds);
    for (auto &d : s->direct) s->events.push_back(d.second);
    for (auto &g : s->groups) s->events.push_back(g.second);
    for (svo_renderer *rd : s->renderers)
        if (rd) renderer_free(rd);
    for (hipEvent_t ev : s->events) (void)hipEventDestroy(ev);
    for (hipStream_t ds : s->directStreams)
        if (ds) (void)hipStreamDestroy(ds);
    if (s->ended) (void)hipEventDestroy(s->ended);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->copyStream) (void)hipStreamDestroy(s->copyStream);
    if (s->arena) {
        if (s->hostOutputs) (void)hipHostFree(s->arena);
        else (void)hipFree(s->arena);
    }
    if (s->st) (void)hipFree(s->st);
    if (s->dslots) (void)hipFree(s->dslots);
    if (s->dbase) (void)hipFree(s->dbase);
    if (s->done) (void)hipFree(s->done);
    if (s->ctrl) (void)hipHostFree(s->ctrl);
    if (s->hslots) (void)hipHostFree(s->hslots);
    if (s->complete) (void)hipHostFree(s->complete);
    delete s;
}

}  // namespace

int server_create(svo_octree *const *trees, int ntrees, const svo_server_config *cfg, svo_server **out) {
    if (!out) return fail(SVO_ERR_INVALID, "server: null output");
    *out = nullptr;
    if (!trees || ntrees <= 0) return fail(SVO_ERR_INVALID, "server: no octrees");
    for (int i = 0; i < ntrees; ++i)
        if (!trees[i]) return fail(SVO_ERR_INVALID, "server: null octree");
    const int ring = (cfg && cfg->ring_slots > 0) ? cfg->ring_slots : 1024;
    if (ring > (1 << 20)) return fail(SVO_ERR_INVALID, "server: ring_slots above 2^20");
    const uint64_t budget = (cfg && cfg->output_budget) ? cfg->output_budget : (uint64_t(1) << 30);
    const int waves = (cfg && cfg->waves_per_cu > 0) ? cfg->waves_per_cu : 24;
    if (waves < 4 || waves > 32) return fail(SVO_ERR_INVALID, "server: waves_per_cu must be 4..32");
    if (cfg && (cfg->max_batch < 0 || cfg->max_batch > svo_server::kMaxGroup))
        return fail(SVO_ERR_INVALID, "server: max_batch must be 0..256");
    const double idle = (cfg && cfg->idle_stop_us > 0) ? cfg->idle_stop_us * 1e-6 : 2e-3;
    if (device_count() < 1) return fail(SVO_ERR_DEVICE, "server: no HIP device");
    svo_server *s = new (std::nothrow) svo_server();
    if (!s) return fail(SVO_ERR_NOMEM, "out of host memory");
    s->trees.assign(trees, trees + ntrees);
    s->ring = ring;
    s->maxReqs = 1u << 20;
    s->hostOutputs = cfg && cfg->host_outputs;
    s->budget = budget / kAlign * kAlign;
    s->idleStop = idle;
    s->wavesPerCU = waves;
    s->persistentOnly = cfg && cfg->persistent_only;
    s->batchedLaunches = !(cfg && (cfg->persistent_launch || cfg->persistent_only));
    if (cfg && cfg->max_batch > 0) s->maxGroup = cfg->max_batch;
    int maxDepth = 1;
    for (svo_octree *t : s->trees) {
        int r = ensure_device(*t);  // octrees stay resident while the server lives
        if (r) {
            free_server(s);
            return r;
        }
        maxDepth = std::max(maxDepth, t->depth);
        s->wide = s->wide || t->words.size() >= (uint64_t(1) << 32);
    }
    s->reg = reg_entries(maxDepth);
    s->stackSlots = lds_slots(maxDepth, s->reg);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDevice(&s->device);
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, s->device);
    if (e == hipSuccess) {
        s->cus = prop.multiProcessorCount;
        const unsigned pinned = hipHostMallocCoherent | hipHostMallocMapped;
        e = hipHostMalloc(reinterpret_cast<void **>(&s->ctrl), sizeof(ServeCtrl), pinned);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&s->hslots), sizeof(ServeSlot) * ring, pinned);
        if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&s->complete), sizeof(uint64_t) * ring, pinned);
        if (e == hipSuccess) e = hipMalloc(&s->st, sizeof(ServeState));
        if (e == hipSuccess) e = hipMalloc(&s->dslots, sizeof(ViewDesc) * ring);
        if (e == hipSuccess) e = hipMalloc(&s->dbase, sizeof(uint64_t) * (size_t(s->maxReqs) + 1));
        const size_t doneBytes = sizeof(uint32_t) * kLineWords * (kCursors + 1) * size_t(ring);
        if (e == hipSuccess) e = hipMalloc(&s->done, doneBytes);
        if (e == hipSuccess) e = hipMemset(s->done, 0, doneBytes);
        if (e == hipSuccess) {
            if (s->hostOutputs) e = hipHostMalloc(reinterpret_cast<void **>(&s->arena), s->budget, pinned);
            else e = hipMalloc(&s->arena, s->budget);
        }
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copyStream, hipStreamNonBlocking);
        for (hipStream_t &ds : s->directStreams)
            if (e == hipSuccess) e = hipStreamCreateWithFlags(&ds, hipStreamNonBlocking);
        if (s->batchedLaunches)
            for (svo_renderer *&rd : s->renderers)
                if (e == hipSuccess && renderer_create(svo_server::kMaxGroup, &rd) != SVO_OK) e = hipErrorOutOfMemory;
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ended, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? SVO_ERR_NOMEM : SVO_ERR_DEVICE;
        free_server(s);
        return fail(code, std::string("server: ") + hipGetErrorString(e));
    }
    std::memset(s->ctrl, 0, sizeof(ServeCtrl));
    std::memset(s->hslots, 0, sizeof(ServeSlot) * ring);
    std::memset(s->complete, 0, sizeof(uint64_t) * ring);
    s->slotTicket.assign(size_t(ring), 0);
    s->alloc.reset(s->budget);
    s->t0 = Clock::now();
    s->lastBusy = s->t0;
    try {
        s->loop = std::thread(loop_main, s);
    } catch (...) {
        free_server(s);
        return fail(SVO_ERR_NOMEM, "server: cannot start the request loop thread");
    }
    *out = s;
    return SVO_OK;
}

void server_destroy(svo_server *s) { free_server(s); }

int server_submit(svo_server *s, const svo_request *r, uint64_t *ticket) {
    if (!s || !r || !ticket) return fail(SVO_ERR_INVALID, "server: null argument");
    if (r->tree < 0 || size_t(r->tree) >= s->trees.size()) return fail(SVO_ERR_INVALID, "server: bad octree index");
    if (r->width <= 0 || r->height <= 0 || r->width > 65536 || r->height > 65536)
        return fail(SVO_ERR_INVALID, "server: bad request size");
    if (!camera_ok(r->camera)) return fail(SVO_ERR_INVALID, "camera entries must be finite and within +-1e6");
    Ticket tk;
    tk.req = *r;
    tk.bytes = request_bytes(*r);
    if (tk.bytes > s->budget) return fail(SVO_ERR_LIMIT, "server: a request's outputs exceed the output budget");
    tk.tiles = uint32_t(((r->width + 7) / 8) * ((r->height + 7) / 8));
    std::lock_guard<std::mutex> lock(s->mu);
    if (s->error) return fail(s->error, s->errorMsg);
    tk.submitted = now_s(*s);
    const uint64_t id = s->nextTicket++;
    s->tickets.emplace(id, tk);
    s->waiting.push_back(id);
    admit(*s);
    if (s->error) return fail(s->error, s->errorMsg);
    *ticket = id;
    return SVO_OK;
}

int server_poll(svo_server *s, uint64_t ticket, int *state, double *latency) {
    if (!s || !state) return fail(SVO_ERR_INVALID, "server: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SVO_ERR_INVALID, "server: unknown ticket");
    if (s->error && it->second.state != 2) return fail(s->error, s->errorMsg);
    *state = it->second.state;
    if (latency) *latency = it->second.state == 2 ? it->second.completed - it->second.submitted : 0.0;
    return SVO_OK;
}

int server_wait(svo_server *s, uint64_t ticket, double timeout, double *latency) {
    if (!s) return fail(SVO_ERR_INVALID, "server: null argument");
    std::unique_lock<std::mutex> lock(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SVO_ERR_INVALID, "server: unknown ticket");
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double>(timeout > 0 ? timeout : 1e9));
    while (it->second.state != 2 && !s->error) {
        if (s->cv.wait_until(lock, deadline) == std::cv_status::timeout && it->second.state != 2)
            return fail(SVO_ERR_LIMIT, "server: wait timed out");
    }
    if (it->second.state != 2) return fail(s->error, s->errorMsg);
    if (latency) *latency = it->second.completed - it->second.submitted;
    return SVO_OK;
}

int server_outputs(svo_server *s, uint64_t ticket, uint32_t **normal, float **t, uint64_t **leaf) {
    if (!s) return fail(SVO_ERR_INVALID, "server: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SVO_ERR_INVALID, "server: unknown ticket");
    if (it->second.state != 2) return fail(SVO_ERR_INVALID, "server: request not complete");
    const Ticket &tk = it->second;
    const uint64_t n = uint64_t(tk.req.width) * tk.req.height;
    if (normal) *normal = reinterpret_cast<uint32_t *>(s->arena + tk.off);
    if (t) *t = reinterpret_cast<float *>(s->arena + tk.off + round_up(n * 4));
    if (leaf) *leaf = reinterpret_cast<uint64_t *>(s->arena + tk.off + 2 * round_up(n * 4));
    return SVO_OK;
}

int server_copy_outputs(svo_server *s, uint64_t ticket, uint32_t *normal, float *t, uint64_t *leaf) {
    uint32_t *dn = nullptr;
    float *dt = nullptr;
    uint64_t *dl = nullptr;
    int r = server_outputs(s, ticket, &dn, &dt, &dl);
    if (r) return r;
    uint64_t n = 0;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        const Ticket &tk = s->tickets[ticket];
        n = uint64_t(tk.req.width) * tk.req.height;
    }
    if (s->hostOutputs) {
        if (normal) std::memcpy(normal, dn, n * 4);
        if (t) std::memcpy(t, dt, n * 4);
        if (leaf) std::memcpy(leaf, dl, n * 8);
        return SVO_OK;
    }
    // the copy stream is non-blocking, so the copies do not wait for the
    // running launch (the request's outputs were released by it)
    hipError_t e = hipSuccess;
    if (normal) e = hipMemcpyAsync(normal, dn, n * 4, hipMemcpyDeviceToHost, s->copyStream);
    if (e == hipSuccess && t) e = hipMemcpyAsync(t, dt, n * 4, hipMemcpyDeviceToHost, s->copyStream);
    if (e == hipSuccess && leaf) e = hipMemcpyAsync(leaf, dl, n * 8, hipMemcpyDeviceToHost, s->copyStream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->copyStream);
    if (e != hipSuccess) return fail(SVO_ERR_DEVICE, std::string("server copy: ") + hipGetErrorString(e));
    return SVO_OK;
}

int server_release(svo_server *s, uint64_t ticket) {
    if (!s) return fail(SVO_ERR_INVALID, "server: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return fail(SVO_ERR_INVALID, "server: unknown ticket");
    if (it->second.state == 1) return fail(SVO_ERR_INVALID, "server: request still in flight");
    alloc_release(*s, it);
    return SVO_OK;
}

int server_stats(svo_server *s, svo_server_stats *out) {
    if (!s || !out) return fail(SVO_ERR_INVALID, "server: null argument");
    std::lock_guard<std::mutex> lock(s->mu);
    std::memset(out, 0, sizeof(*out));
    out->launches = s->launches;
    out->published = s->published;
    uint64_t grouped = 0;
    for (auto &g : s->groups) grouped += g.first.size();
    out->in_flight = s->inflight.size() + s->direct.size() + grouped;
    out->waiting = s->waiting.size();
    out->live = s->live ? 1 : 0;
    out->direct = s->directLaunches;
    out->batched = s->grouped;
    return SVO_OK;
}

}  // namespace svo

namespace svo {

int server_run_schedule(svo_server *s, const svo_request *reqs, const double *arrival, int n, uint64_t *tickets,
                        double *latency, double *makespan) {
    if (!s || !reqs || !arrival || !latency || n <= 0) return fail(SVO_ERR_INVALID, "schedule: null or empty argument");
    for (int i = 1; i < n; ++i)
        if (!(arrival[i] >= arrival[i - 1])) return fail(SVO_ERR_INVALID, "schedule: arrivals must be ascending");
    std::vector<uint64_t> ids(size_t(n), 0);
    std::vector<char> finished(size_t(n), 0);
    double base;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        base = now_s(*s);
    }
    const auto t0 = Clock::now();
    int next = 0, completed = 0, first = 0;
    double lastDone = 0.0;
    auto stalled = Clock::now();
    while (completed < n) {
        const double t = std::chrono::duration<double>(Clock::now() - t0).count();
        while (next < n && arrival[next] <= t) {
            int r = server_submit(s, &reqs[next], &ids[size_t(next)]);
            if (r) return r;
            next++;
        }
        bool progress = false;
        {
            std::unique_lock<std::mutex> lock(s->mu);
            if (s->error) return fail(s->error, s->errorMsg);
            for (int i = first; i < next; ++i) {
                if (finished[size_t(i)]) continue;
                auto it = s->tickets.find(ids[size_t(i)]);
                if (it == s->tickets.end()) return fail(SVO_ERR_INVALID, "schedule: a ticket was released by another caller");
                if (it->second.state != 2) continue;
                latency[i] = it->second.completed - (base + arrival[i]);
                lastDone = std::max(lastDone, it->second.completed - base);
                finished[size_t(i)] = 1;
                completed++;
                progress = true;
                if (!tickets) {
                    alloc_release(*s, it);
                }
            }
            while (first < next && finished[size_t(first)]) first++;
            if (progress) {
                stalled = Clock::now();
            } else if (next == n && s->inflight.empty() && !s->waiting.empty() &&
                       std::chrono::duration<double>(Clock::now() - stalled).count() > 0.5) {
                return fail(SVO_ERR_LIMIT, "schedule: the kept outputs fill the arena; raise output_budget");
            } else if (completed < n) {
                // sleep on the request loop's completion signal (which releases
                // the lock to it), at most 50 us; within 200 us of the next
                // arrival poll instead (a timed sleep can wake tens of
                // microseconds late, and latency counts from the arrival)
                const auto nowc = Clock::now();
                auto until = nowc + std::chrono::microseconds(50);
                bool poll = false;
                if (next < n) {
                    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
                                              std::chrono::duration<double>(arrival[next]));
                    poll = due - nowc < std::chrono::microseconds(200);
                    if (due - std::chrono::microseconds(100) < until) until = due - std::chrono::microseconds(100);
                }
                if (poll) {
                    lock.unlock();
                    for (int k = 0; k < 64; ++k) __builtin_ia32_pause();
                    lock.lock();
                } else {
                    s->cv.wait_until(lock, until);
                }
            }
        }
    }
    if (tickets)
        for (int i = 0; i < n; ++i) tickets[i] = ids[size_t(i)];
    if (makespan) *makespan = lastDone - arrival[0];
    return SVO_OK;
}

}  // namespace svo

namespace svo {

namespace {
__global__ void code_probe_kernel() {}
}  // namespace

// Makes the HIP runtime load this source file's device code on the current
// device now (it loads a code object at the first launch of one of its kernels).
hipError_t load_code_server() {
    hipFuncAttributes a;
    return hipFuncGetAttributes(&a, reinterpret_cast<const void *>(&code_probe_kernel));
}

}  // namespace svo
