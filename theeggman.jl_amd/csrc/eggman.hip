// eggman HIP library: device caches, level drivers and the C ABI of include/eggman_hip.h.
//
// The file is compiled once per part, in parallel: EGG_PART = 0 holds the shared state and the
// C ABI; EGG_PART = EGGMAN_F32 .. EGGMAN_I128 each instantiate the level drivers (and so the
// kernels) of one element type.  Without EGG_PART the whole library is one translation unit.
#ifndef EGG_PART
#define EGG_PART (-1)
#endif
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "kernels_fwd.hpp"
#include "kernels_plan.hpp"
#include "kernels_rev.hpp"
#include "plan_host.hpp"

#define EGGMAN_VERSION_STRING "eggman-hip 0.1 (gfx950)"

namespace eggman {

// ------------------------------------------------------------------------------ errors

#if EGG_PART <= 0
static thread_local std::string t_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_err = buf;
    return code;
}
#endif

// ------------------------------------------------------------------------------ state

// A plan resident on one device.  Only the transition words live on the device; the level
// boundaries are host-side launch parameters (as in ext/TheEggmanKernelAbstractionsExt.jl:28-30).
struct DevPlan {
    uint64_t id = 0;
    int device = 0;
    PlanLayout lay;
    void *d_trans = nullptr;
    size_t bytes = 0;
    // reverse-sweep index structures (kernels_rev.hpp), built on the first gradient call
    bool rev_ready = false;
    void *d_cent = nullptr, *d_coff = nullptr, *d_pent = nullptr, *d_segs = nullptr;
    void *d_qoff = nullptr, *d_qseg = nullptr;
    std::vector<int64_t> lvl_seg;
    int64_t nseg = 0;
    // parent-major dP tasks (used when the plan uses the Int32 layout)
    void *d_prank = nullptr, *d_tasks = nullptr, *d_roff = nullptr, *d_rtask = nullptr;
    void *d_qrow = nullptr, *d_qcol = nullptr;
    void *d_red = nullptr, *d_srow0 = nullptr, *d_sruns = nullptr;
    std::vector<int64_t> lvl_task;
    int64_t part_rows = 0, nred = 0;
    // fused child-major reverse step (plan_host.hpp: FusePlanHost), Int32 layout
    bool fuse = false;
    void *d_fent = nullptr, *d_fco = nullptr, *d_cmask = nullptr, *d_fprow0 = nullptr;
    void *d_fterms = nullptr, *d_ftoff = nullptr, *d_ftasks = nullptr, *d_frtask = nullptr, *d_fred = nullptr;
    void *d_fsrow0 = nullptr, *d_fsruns = nullptr;
    std::vector<FuseLevel> flv;
    int64_t fpart_rows = 0, fnred = 0;
    size_t rev_bytes = 0;
    GradLevels gl_full, gl_ckpt;   // where the gradient keeps forward levels (plan_host.hpp)
    // parent-pair items of each level (k_level_fwd_lds_vec2), built on first use: level L's
    // items at d_items + item_off[L-1], nitems[L-1] of them
    mutable bool items_ready = false;
    mutable uint32_t *d_items = nullptr;
    mutable std::vector<int64_t> item_off, nitems;
    mutable size_t items_bytes = 0;
};

// Grow-only device buffer (ext/TheEggmanKernelAbstractionsExt.jl:68-88).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// What the launch heuristics need to know about a device.
struct DevCaps {
    int cus = 256;                   // compute units
    int xcds = 8;                    // chiplets, each with its own L2
    size_t max_lds_block = 65536;    // LDS one workgroup may use
    size_t lds_per_cu = 65536;       // LDS per compute unit
};

// A captured forward walk and the arguments it was captured for.
struct GraphEntry {
    uint64_t plan = 0;
    int dtype = 0;
    const void *P = nullptr;
    int64_t ldp = 0, B = 0;
    void *out = nullptr;
    int bt = 0;
    int64_t chunk = 0;
    int flags = 0;                   // tune_flags of the launch choices
    int occ = 0;
    uint64_t gen = 0;                // work-buffer generation the graph's addresses belong to
    hipGraphExec_t exec = nullptr;
};

// Bumped whenever a work buffer is reallocated, which invalidates captured graphs.
#if EGG_PART <= 0
uint64_t g_buf_gen = 1;
#else
extern uint64_t g_buf_gen;
#endif

// Grow-only pinned host buffer (hipHostMalloc), the staging memory of the host entry points.
struct PinBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct DevState {
    DevCaps caps;
    hipStream_t cap = nullptr;       // capture stream for graphs
    std::vector<GraphEntry> graphs;
    hipStream_t stream = nullptr;    // the library's own stream on this device
    hipStream_t last = nullptr;      // stream of the most recent enqueued call
    hipEvent_t order = nullptr;      // orders calls on different streams (shared buffers)
    DevBuf H;                        // state values (ping-pong levels)
    DevBuf Hw;                       // whole-batch rows of the small levels
    DevBuf Hg;                       // every level of H, for the reverse sweep
    DevBuf G;                        // adjoint values (ping-pong levels)
    DevBuf part;                     // per-segment partial gradients
    DevBuf A;                        // staged matrix (repeated-pattern batches)
    DevBuf bnd;                      // per-instance range bounds of exact device-pointer calls
    DevBuf Hm;                       // the walks of a mixed call's large groups, one slice each
    // side streams of mixed calls (default priority), forked from and joined to the call's stream
    static constexpr int kSide = 4;
    hipStream_t side[kSide] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSide] = {};
    // second streams of the walks that run on side streams (their passes two at a time), and the
    // event that links one stream to another
    hipStream_t sub[kSide] = {};
    hipEvent_t ev_link = nullptr;
    void *scratch_dev = nullptr;     // this device's address of the no-allocation scratch array
    // host entry points: a two-slot pipeline of pinned host buffers and device buffers, with an
    // upload and a download stream beside the compute stream (run_pipe)
    hipStream_t up = nullptr, down = nullptr;
    hipEvent_t ev_up[2] = {}, ev_comp[2] = {}, ev_down[2] = {};
    PinBuf pin_in[2], pin_out[2];
    DevBuf din[2], dout[2];
    size_t bytes() const {
        return H.bytes + Hw.bytes + Hg.bytes + G.bytes + part.bytes + A.bytes + bnd.bytes + Hm.bytes +
               din[0].bytes + din[1].bytes + dout[0].bytes + dout[1].bytes;
    }
};

#if EGG_PART <= 0
std::recursive_mutex g_mu;
std::map<uint64_t, std::unique_ptr<DevPlan>> g_plans;
std::map<std::tuple<int, int, int>, uint64_t> g_plan_key;   // (device, K, index_bytes)
uint64_t g_next_id = 1;
std::map<int, DevState> g_dev;
#else
extern std::recursive_mutex g_mu;
extern std::map<uint64_t, std::unique_ptr<DevPlan>> g_plans;
extern std::map<std::tuple<int, int, int>, uint64_t> g_plan_key;
extern uint64_t g_next_id;
extern std::map<int, DevState> g_dev;
#endif

static int current_device(int *dev) {
    EGG_HIP(hipGetDevice(dev));
    return EGGMAN_OK;
}

static int dev_state(int device, DevState **ds) {
    DevState &d = g_dev[device];
    if (!d.stream) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
            d.caps.cus = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
            d.caps.max_lds_block = (size_t)v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess && v > 0)
            d.caps.lds_per_cu = (size_t)v;
        else
            d.caps.lds_per_cu = d.caps.max_lds_block;
        EGG_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
        EGG_HIP(hipEventCreateWithFlags(&d.order, hipEventDisableTiming));
        d.last = d.stream;
    }
    *ds = &d;
    return EGGMAN_OK;
}

// Make `s` wait for everything enqueued by the previous call, which may have used the same
// work buffers on another stream.
static int chain_stream(DevState &d, hipStream_t s) {
    if (s != d.last) {
        EGG_HIP(hipEventRecord(d.order, d.last));
        EGG_HIP(hipStreamWaitEvent(s, d.order, 0));
        d.last = s;
    }
    return EGGMAN_OK;
}

static int ensure(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return EGGMAN_OK;
    g_buf_gen++;
    if (b.p) {
        EGG_HIP(hipDeviceSynchronize());
        EGG_HIP(hipFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return fail(EGGMAN_ENOMEM, "device allocation of %.3f GB failed: %s", bytes / 1e9,
                    hipGetErrorString(e));
    }
    b.bytes = bytes;
    return EGGMAN_OK;
}

// The side streams (default priority) and their fork/join events, created on first use.
static int ensure_side(DevState &ds) {
    if (ds.ev_fork) return EGGMAN_OK;
    for (int i = 0; i < DevState::kSide; i++) {
        EGG_HIP(hipStreamCreateWithFlags(&ds.side[i], hipStreamNonBlocking));
        EGG_HIP(hipStreamCreateWithFlags(&ds.sub[i], hipStreamNonBlocking));
        EGG_HIP(hipEventCreateWithFlags(&ds.ev_join[i], hipEventDisableTiming));
    }
    EGG_HIP(hipEventCreateWithFlags(&ds.ev_link, hipEventDisableTiming));
    EGG_HIP(hipEventCreateWithFlags(&ds.ev_fork, hipEventDisableTiming));
    return EGGMAN_OK;
}

// `waiter` waits for what `signaler` holds so far.
static int link_streams(DevState &ds, hipStream_t waiter, hipStream_t signaler) {
    EGG_TRY(ensure_side(ds));
    EGG_HIP(hipEventRecord(ds.ev_link, signaler));
    EGG_HIP(hipStreamWaitEvent(waiter, ds.ev_link, 0));
    return EGGMAN_OK;
}

// Work of one call split over `st` and ns - 1 side streams: fork makes the side streams wait for
// what st holds so far, join makes st wait for everything the side streams were given.
static int fork_side(DevState &ds, hipStream_t st, int ns) {
    if (ns <= 1) return EGGMAN_OK;
    EGG_TRY(ensure_side(ds));
    EGG_HIP(hipEventRecord(ds.ev_fork, st));
    for (int i = 0; i < ns - 1; i++) EGG_HIP(hipStreamWaitEvent(ds.side[i], ds.ev_fork, 0));
    return EGGMAN_OK;
}
static int join_side(DevState &ds, hipStream_t st, int ns) {
    for (int i = 0; i < ns - 1; i++) {
        EGG_HIP(hipEventRecord(ds.ev_join[i], ds.side[i]));
        EGG_HIP(hipStreamWaitEvent(st, ds.ev_join[i], 0));
    }
    return EGGMAN_OK;
}
// Stream i of such a split: st itself, then the side streams.
static hipStream_t split_stream(const DevState &ds, hipStream_t st, int i) {
    return i == 0 ? st : ds.side[i - 1];
}

// Passes (or a gradient pass's sub-passes) run two at a time, alternating between the call's
// stream and a side stream, when there are at least four of them and the state values are at most
// 8 bytes: each launch's ramp-up and drain, and the short latency-bound launches of a pass's
// small levels (a few µs each on a few thousand states), then overlap the other pass's work.  The
// two passes run in step, level by level, reading the same plan words; passes write disjoint
// columns or buffers, so the values are the same bits.  Measured (forward, ms, one stream
// against two): FP64 N=28 B=1024 3.22 -> 3.07, B=4096 13.2 -> 12.3, B=256 0.879 -> 0.840; FP32
// N=28 1.73 -> 1.63; complex64 N=28 3.38 -> 3.06; FP64 N=32 B=256 7.43 -> 7.20; N=24 B=4096 1.79
// -> 1.78.  Two passes only (N=24 B=1024: 0.472 -> 0.483) and 16-byte values (complex128 N=28:
// 6.32 -> 6.86, their two working sets contend in L2) stay on one stream; three or four streams
// were slower than two (N=28: 3.31 and 3.44 ms).
constexpr int kPassStreams = 2;
// Gradient passes in flight at once (forward_grad_levels): N=28 FP64, forward + gradient per
// call, against one pass at a time with its forward sub-passes split over two streams:
// B=1024 11.79 -> 11.65 ms (11.82 -> 11.49 in another run), B=2048 23.5 -> 22.9,
// N=26 B=2048 9.99 -> 9.80; splitting each pass's sub-passes too was no better (11.79, 23.3,
// 10.07).
constexpr int kGradStreams = 2;
static int pass_streams(int dtype, int64_t npass) {
    return npass >= 4 && elt_hbytes(dtype) <= 8 ? kPassStreams : 1;
}

static int find_plan(uint64_t handle, DevPlan **pl) {
    auto it = g_plans.find(handle);
    if (it == g_plans.end())
        return fail(EGGMAN_EINVAL, "unknown plan handle %llu (released by eggman_cache_clear?)",
                    (unsigned long long)handle);
    *pl = it->second.get();
    return EGGMAN_OK;
}

static int ensure_pinned(PinBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return EGGMAN_OK;
    if (b.p) {
        EGG_HIP(hipHostFree(b.p));
        b.p = nullptr;
        b.bytes = 0;
    }
    hipError_t e = hipHostMalloc(&b.p, std::max<size_t>(bytes, 64), hipHostMallocDefault);
    if (e != hipSuccess) {
        b.p = nullptr;
        (void)hipGetLastError();
        return fail(EGGMAN_ENOMEM, "pinned host allocation of %.3f GB failed: %s", bytes / 1e9,
                    hipGetErrorString(e));
    }
    b.bytes = bytes;
    return EGGMAN_OK;
}

// Host copies of the staging.  One thread moves about 5-10 GB/s, below the host link's rate, so
// copies above a few MB are split over up to 8 threads (by rows, or by bytes for one block).
template <class F>
static void split_rows(int64_t rows, size_t bytes, F &&fn) {
    constexpr size_t kPerThread = size_t(4) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int64_t nt = std::min<int64_t>(std::min<int64_t>(std::min<int64_t>(8, hw), rows),
                                         (int64_t)(bytes / kPerThread));
    if (nt <= 1) {
        fn((int64_t)0, rows);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (rows + nt - 1) / nt;
    for (int64_t r0 = per; r0 < rows; r0 += per)
        th.emplace_back([&fn, r0, per, rows] { fn(r0, std::min(rows, r0 + per)); });
    fn((int64_t)0, std::min(rows, per));
    for (auto &t : th) t.join();
}

static void copy_bytes(void *dst, const void *src, size_t n) {
    constexpr size_t kBlock = size_t(1) << 20;
    const int64_t blocks = (int64_t)((n + kBlock - 1) / kBlock);
    split_rows(blocks, n, [&](int64_t b0, int64_t b1) {
        const size_t o = (size_t)b0 * kBlock, e = std::min(n, (size_t)b1 * kBlock);
        std::memcpy((char *)dst + o, (const char *)src + o, e - o);
    });
}

// Columns [b0, b0 + nb) of a rows x ld matrix of esz-byte elements (row-major, rows ld apart)
// into a dense rows x nb block, and back.
static void copy_columns(void *dst, const void *src, int64_t rows, int64_t ld, int64_t b0,
                         int64_t nb, size_t esz) {
    if (nb == ld) {
        copy_bytes(dst, src, (size_t)rows * nb * esz);
        return;
    }
    split_rows(rows, (size_t)rows * nb * esz, [&](int64_t q0, int64_t q1) {
        for (int64_t q = q0; q < q1; q++)
            std::memcpy((char *)dst + (size_t)q * nb * esz,
                        (const char *)src + ((size_t)q * ld + b0) * esz, (size_t)nb * esz);
    });
}

static void scatter_columns(void *dst, const void *src, int64_t rows, int64_t ld, int64_t b0,
                            int64_t nb, size_t esz) {
    if (nb == ld) {
        copy_bytes(dst, src, (size_t)rows * nb * esz);
        return;
    }
    split_rows(rows, (size_t)rows * nb * esz, [&](int64_t q0, int64_t q1) {
        for (int64_t q = q0; q < q1; q++)
            std::memcpy((char *)dst + ((size_t)q * ld + b0) * esz,
                        (const char *)src + (size_t)q * nb * esz, (size_t)nb * esz);
    });
}

// One host-entry call as a pipeline over chunks c = 0 .. nchunks-1 through two slots.  For chunk c
// in slot s = c % 2: the host fills the pinned input (fill(c, pinned, &in_bytes, &out_bytes)), the
// upload stream copies it to the device, the compute stream runs the chunk (run(c, din, dout,
// stream)), the download stream copies out_bytes of results back to pinned memory, and the host
// hands them to the caller (drain(c, pinned)).  Chunk c+1's fill and upload overlap chunk c's
// compute, and chunk c's download overlaps chunk c+1's compute.  Before slot s is refilled the
// host waits for its download, which follows its compute, which follows its upload, so no slot is
// reused while a copy or kernel still reads it.  Every copy has ended when the call returns.
template <class Fill, class Run, class Drain>
static int run_pipe(DevState &ds, int64_t nchunks, size_t in_max, size_t out_max, Fill &&fill,
                    Run &&run, Drain &&drain) {
    if (!ds.up) {
        EGG_HIP(hipStreamCreateWithFlags(&ds.up, hipStreamNonBlocking));
        EGG_HIP(hipStreamCreateWithFlags(&ds.down, hipStreamNonBlocking));
        for (int s = 0; s < 2; s++) {
            EGG_HIP(hipEventCreateWithFlags(&ds.ev_up[s], hipEventDisableTiming));
            EGG_HIP(hipEventCreateWithFlags(&ds.ev_comp[s], hipEventDisableTiming));
            EGG_HIP(hipEventCreateWithFlags(&ds.ev_down[s], hipEventDisableTiming));
        }
    }
    for (int s = 0; s < 2 && s < nchunks; s++) {
        EGG_TRY(ensure_pinned(ds.pin_in[s], in_max));
        EGG_TRY(ensure_pinned(ds.pin_out[s], out_max));
        EGG_TRY(ensure(ds.din[s], in_max));
        EGG_TRY(ensure(ds.dout[s], out_max));
    }
    hipStream_t st = ds.stream;
    EGG_TRY(chain_stream(ds, st));
    auto drained = [&](int rc) {
        // on an error, let every enqueued copy finish before the caller's memory is released
        (void)hipStreamSynchronize(ds.up);
        (void)hipStreamSynchronize(st);
        (void)hipStreamSynchronize(ds.down);
        return rc;
    };
    if (nchunks == 1) {
        // One chunk: both copies and the compute go in order on the call's stream.  Nothing
        // overlaps, so the cross-stream events of the pipeline would only add their latency.
        size_t ib = 0, ob = 0;
        int rc = fill(0, ds.pin_in[0].p, &ib, &ob);
        if (rc != EGGMAN_OK) return rc;
        hipError_t e = hipSuccess;
        if (ib) e = hipMemcpyAsync(ds.din[0].p, ds.pin_in[0].p, ib, hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return drained(fail(EGGMAN_EHIP, "upload: %s", hipGetErrorString(e)));
        if ((rc = run(0, ds.din[0].p, ds.dout[0].p, st)) != EGGMAN_OK) return drained(rc);
        if (ob) e = hipMemcpyAsync(ds.pin_out[0].p, ds.dout[0].p, ob, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            return drained(fail(EGGMAN_EHIP, "download: %s", hipGetErrorString(e)));
        return drain(0, ds.pin_out[0].p);
    }
    for (int64_t c = 0; c < nchunks; c++) {
        const int s = (int)(c & 1);
        int rc = EGGMAN_OK;
        if (c >= 2) {
            if (hipEventSynchronize(ds.ev_down[s]) != hipSuccess)
                return drained(fail(EGGMAN_EHIP, "download of chunk %lld failed", (long long)(c - 2)));
            if ((rc = drain(c - 2, ds.pin_out[s].p)) != EGGMAN_OK) return drained(rc);
        }
        size_t ib = 0, ob = 0;
        if ((rc = fill(c, ds.pin_in[s].p, &ib, &ob)) != EGGMAN_OK) return drained(rc);
        hipError_t e = hipSuccess;
        if (ib) e = hipMemcpyAsync(ds.din[s].p, ds.pin_in[s].p, ib, hipMemcpyHostToDevice, ds.up);
        if (e == hipSuccess) e = hipEventRecord(ds.ev_up[s], ds.up);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ds.ev_up[s], 0);
        if (e != hipSuccess) return drained(fail(EGGMAN_EHIP, "upload: %s", hipGetErrorString(e)));
        if ((rc = run(c, ds.din[s].p, ds.dout[s].p, st)) != EGGMAN_OK) return drained(rc);
        e = hipEventRecord(ds.ev_comp[s], st);
        if (e == hipSuccess) e = hipStreamWaitEvent(ds.down, ds.ev_comp[s], 0);
        if (e == hipSuccess && ob)
            e = hipMemcpyAsync(ds.pin_out[s].p, ds.dout[s].p, ob, hipMemcpyDeviceToHost, ds.down);
        if (e == hipSuccess) e = hipEventRecord(ds.ev_down[s], ds.down);
        if (e != hipSuccess) return drained(fail(EGGMAN_EHIP, "download: %s", hipGetErrorString(e)));
    }
    for (int64_t c = std::max<int64_t>(0, nchunks - 2); c < nchunks; c++) {
        const int s = (int)(c & 1);
        if (hipEventSynchronize(ds.ev_down[s]) != hipSuccess)
            return drained(fail(EGGMAN_EHIP, "download of chunk %lld failed", (long long)c));
        int rc = drain(c, ds.pin_out[s].p);
        if (rc != EGGMAN_OK) return drained(rc);
    }
    return EGGMAN_OK;
}

// Pinned staging per slot and direction, at most (a chunk of a few hundred ms of work at N=28).
constexpr size_t kPipeSlotBytes = size_t(256) << 20;

// Chunks a host-entry call splits into when the caller gives no max_batch: two when the device
// work per instance is large against its transfer (at least 1000 transitions per pair value,
// N >= 24), so each chunk's fixed costs (its small levels, launches, the first upload) stay few;
// four when transfers dominate.  N=28, B=1024: 3.63-3.70 ms in two chunks, 3.72 ms in four;
// N=24, B=4096: 2.07-2.13 against 2.14 ms.
static int pipe_parts(const PlanLayout &lay) {
    return lay.ntrans >= 1000 * lay.npairs ? 2 : 4;
}

// Staged bytes above which a call is split into parts even when its tuned launch takes the whole
// batch in one pass: N=20, B=16384 (25 MB) takes 1.60 ms in four chunks against 1.80 ms in one,
// N=16, B=65536 (63 MB) 2.0 against 2.5 ms.
constexpr size_t kPipeSplitBytes = size_t(8) << 20;

// Instances per chunk of a host-entry call: max_batch when the caller gives one; otherwise the
// whole batch when it is under two passes of the tuned launch (and its staging under
// kPipeSplitBytes), else 1/parts of it rounded up to whole passes, so that host staging and
// transfers overlap compute -- and never more than two slots of per_inst device bytes fit in
// half of the free device memory, nor a slot's pinned staging (per_inst bytes an instance) above
// kPipeSlotBytes.
static int host_chunk(int64_t B, int64_t max_batch, int64_t pass, size_t per_inst, int parts,
                      int64_t *chunk) {
    if (max_batch > 0) {
        *chunk = std::min(max_batch, B);
        return EGGMAN_OK;
    }
    size_t free_b = 0, total_b = 0;
    EGG_HIP(hipMemGetInfo(&free_b, &total_b));
    per_inst = std::max<size_t>(per_inst, 1);
    const int64_t fit = (int64_t)(free_b / 2 / (2 * per_inst));
    pass = std::max<int64_t>(pass, 1);
    if (B < 2 * pass && (size_t)B * per_inst >= kPipeSplitBytes) pass = (B + parts - 1) / parts;
    int64_t want = B;
    if (B >= 2 * pass) want = std::max(pass, ((B + parts - 1) / parts + pass - 1) / pass * pass);
    const int64_t pin = std::max<int64_t>(pass, (int64_t)(kPipeSlotBytes / per_inst) / pass * pass);
    *chunk = std::max<int64_t>(1, std::min(std::min(std::min(want, B), fit), pin));
    return EGGMAN_OK;
}

// Scoped switch to a plan's device.
struct DeviceGuard {
    int prev = -1;
    int rc = EGGMAN_OK;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev && hipSetDevice(dev) != hipSuccess)
            rc = fail(EGGMAN_EHIP, "hipSetDevice(%d) failed", dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// ------------------------------------------------------------------------------ exact range

// The 128-bit walk computes the hafnian modulo 2^128 (it only adds and multiplies), so its
// two's-complement result is the hafnian whenever |haf| < 2^127.  An instance is accepted only
// when an upper bound proves that: log2 of min((K-1)!! * m^(K/2), sqrt(prod_i max(r_i, 1))),
// m = max(1, max |p|), r_i = sum_{j != i} |p_ij| (for entrywise absolute values
// haf^2 <= perm <= the product of the row sums).  The same bound covers every (K-2)-minor, so
// it also covers the gradient.  The margin absorbs the rounding of the logarithms.
constexpr double kExactLog2Limit = 127.0 - 1e-6;

template <class PV>
static double exact_log2_bound(int K, PV &&pv) {
    double r[64] = {};
    double m = 1.0;
    int64_t q = 0;
    for (int i = 0; i < K; i++)
        for (int j = i + 1; j < K; j++, q++) {
            const double a = std::fabs((double)pv(q));
            r[i] += a;
            r[j] += a;
            m = std::max(m, a);
        }
    double lf = 0.5 * K * std::log2(m), lr = 0.0;
    for (int k = K - 1; k > 1; k -= 2) lf += std::log2((double)k);
    for (int i = 0; i < K; i++) lr += std::log2(std::max(r[i], 1.0));
    return std::min(lf, 0.5 * lr);
}

static int exact_overflow(int64_t b, double lb) {
    return fail(EGGMAN_EOVERFLOW,
                "exact hafnian of instance %lld may need %.1f bits: the 128-bit walk is exact only "
                "below 2^127 in magnitude",
                (long long)(b + 1), lb + 1.0);
}

// Instances [0, B) of P ([pair][ld], int64 host memory); b_first numbers them in messages.
static int check_exact_host(int K, const int64_t *P, int64_t ld, int64_t B, int64_t b_first = 0) {
    for (int64_t b = 0; b < B; b++) {
        const double lb = exact_log2_bound(K, [&](int64_t q) { return P[q * ld + b]; });
        if (lb >= kExactLog2Limit) return exact_overflow(b_first + b, lb);
    }
    return EGGMAN_OK;
}

// The boson-sampling shape: instance b reads A[idx_b[i], idx_b[j]] (A column-major, N x N).
static int check_exact_repeated(int K, const int64_t *A, int64_t N, const int32_t *idx, int64_t B) {
    std::vector<int64_t> pv((size_t)K * (K - 1) / 2);
    for (int64_t b = 0; b < B; b++) {
        const int32_t *ix = idx + b * K;
        int64_t q = 0;
        for (int i = 0; i < K; i++)
            for (int j = i + 1; j < K; j++) pv[q++] = A[(int64_t)(ix[j] - 1) * N + (ix[i] - 1)];
        const double lb = exact_log2_bound(K, [&](int64_t q) { return pv[q]; });
        if (lb >= kExactLog2Limit) return exact_overflow(b, lb);
    }
    return EGGMAN_OK;
}

// Device-resident P: the bounds are computed on the device (stream st), copied back, and the
// stream is synchronised before the check.
static int check_exact_device(DevState &ds, int K, const void *P, int64_t ldp, int64_t B,
                              hipStream_t st, int64_t b_first = 0) {
    if (B <= 0) return EGGMAN_OK;
    EGG_TRY(ensure(ds.bnd, (size_t)B * sizeof(double)));
    k_exact_log2_bound<<<dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st>>>(
        (const long long *)P, ldp, B, K, (double *)ds.bnd.p);
    EGG_HIP(hipGetLastError());
    std::vector<double> lb((size_t)B);
    EGG_HIP(hipMemcpyAsync(lb.data(), ds.bnd.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, st));
    EGG_HIP(hipStreamSynchronize(st));
    for (int64_t b = 0; b < B; b++)
        if (lb[(size_t)b] >= kExactLog2Limit) return exact_overflow(b_first + b, lb[(size_t)b]);
    return EGGMAN_OK;
}

// ------------------------------------------------------------------------------ tuning

// Launch choices for one forward call.  Every choice gives bit-identical results.
struct Tune {
    int bt = 64;        // lanes per state (batch tile width), power of two in [1, 64]
    int64_t chunk = 0;  // instances per pass over the levels (0 = all)
    bool lds = true;    // stage the level's pair rows in LDS (bt == 64 only)
    int occ = 0;        // workgroups per CU for the LDS kernel (0 = as many as LDS allows, <= 4)
    bool graph = false; // replay the level sequence from a captured HIP graph
    bool fused = false; // one workgroup per instance walks every level with H in LDS (small K)
    bool wide = false;  // 16 waves per workgroup in the LDS kernel instead of 8
    bool ckpt = false;  // gradient: keep checkpoint levels only and recompute the others
    bool narrow = false;// LDS kernel with 16-instance tiles, 4 states per wave
    bool vec = false;   // LDS kernel with 16-byte loads: 8 lanes per state, 8 states per wave
    bool deep = false;  // with vec: child loads in batches of 6 transitions instead of 4
    int64_t sub = 0;    // gradient: instances per forward sub-pass inside a pass (0 = the pass)
    bool whole_small = true; // walk the small end levels once for the whole batch (extra buffers)
    bool pairs = false; // with vec: slots take parent pairs sharing their children (k_level_fwd_lds_vec2)
};

// Tuning flags as they cross the C ABI (eggman_tune_set / eggman_tune_get).
enum {
    TUNE_LDS = 1, TUNE_GRAPH = 2, TUNE_FUSED = 4, TUNE_WIDE = 8, TUNE_CKPT = 16, TUNE_NARROW = 32,
    TUNE_VEC = 64, TUNE_DEEP = 128, TUNE_PAIRS = 256,
    TUNE_ALL = 511
};
static int tune_flags(const Tune &t) {
    return (t.lds ? TUNE_LDS : 0) | (t.graph ? TUNE_GRAPH : 0) | (t.fused ? TUNE_FUSED : 0) |
           (t.wide ? TUNE_WIDE : 0) | (t.ckpt ? TUNE_CKPT : 0) | (t.narrow ? TUNE_NARROW : 0) |
           (t.vec ? TUNE_VEC : 0) | (t.deep ? TUNE_DEEP : 0) | (t.pairs ? TUNE_PAIRS : 0);
}
static Tune tune_from_flags(int bt, int64_t chunk, int flags, int occ) {
    Tune t;
    t.bt = bt;
    t.chunk = chunk;
    t.occ = occ;
    t.lds = (flags & TUNE_LDS) != 0;
    t.graph = (flags & TUNE_GRAPH) != 0;
    t.fused = (flags & TUNE_FUSED) != 0;
    t.wide = (flags & TUNE_WIDE) != 0;
    t.ckpt = (flags & TUNE_CKPT) != 0;
    t.narrow = (flags & TUNE_NARROW) != 0;
    t.vec = (flags & TUNE_VEC) != 0;
    t.deep = (flags & TUNE_DEEP) != 0;
    t.pairs = (flags & TUNE_PAIRS) != 0;
    return t;
}

static int pow2_at_least(int64_t n) {
    int b = 1;
    while (b < 64 && b < n) b <<= 1;
    return b;
}

// Launch choices for a (degree, element type) the table does not cover, following what the
// autotuner picks for the tuned FP64 degrees: the fused LDS-resident walk whenever the plan's
// states fit in LDS (it declines otherwise); from 64 instances the 16-byte-load LDS kernel with 2
// workgroups per CU, in passes of about 2^26 state values (512, 128 and 64 instances at N=24,
// 28 and 32, the tuned passes); below 64 instances, as many lanes per state as instances.
static Tune default_tune(const PlanLayout &lay, int, int64_t B) {
    Tune t;
    t.bt = pow2_at_least(B);
    t.chunk = B;
    t.fused = true;
    t.lds = B >= 64;
    if (B >= 16 && B < 64 && lay.K >= 26) {
        // 16 to 63 instances of a degree whose levels are large: the 16-byte-load kernel on
        // 16-instance tiles (one to four of them) rather than a lane per instance.  N=28: 16
        // instances 0.152 -> 0.121 ms, 32 0.250 -> 0.157; N=32: 16 0.90 -> 0.47, 24 1.69 -> 0.90
        // (at N=24 both take about the same time, and small degrees take the fused kernel)
        t.bt = 64;
        t.lds = true;
        t.vec = true;
        t.deep = true;
        t.occ = 2;
        return t;
    }
    if (B >= 64) {
        t.lds = true;
        t.vec = true;
        t.occ = 2;
        int64_t c = 64;
        while (c * 2 * lay.nstates <= (int64_t(1) << 26)) c *= 2;
        t.chunk = std::min(c, B);
    }
    return t;
}

// Tuned choices per (device, degree, index width, element type, batch rounded up to a power of
// two), written by eggman_autotune or eggman_tune_set.
using TuneKey = std::tuple<int, int, int, int, int64_t>;
#if EGG_PART <= 0
std::map<TuneKey, Tune> g_tune;
#else
extern std::map<TuneKey, Tune> g_tune;
#endif

static int64_t batch_bucket(int64_t B) {
    int64_t b = 1;
    while (b < B) b <<= 1;
    return b;
}

// The entry for B's bucket; else the nearest tuned bucket of the same degree and element type
// (the largest one below B, or failing that the smallest above); else the defaults.
static Tune pick_tune(const DevPlan &pl, int dtype, int64_t B) {
    const TuneKey key(pl.device, pl.lay.K, pl.lay.index_bytes, dtype, batch_bucket(B));
    auto it = g_tune.find(key);
    if (it == g_tune.end()) {
        auto hi = g_tune.lower_bound(key);
        auto same = [&](std::map<TuneKey, Tune>::iterator i) {
            return i != g_tune.end() && std::get<0>(i->first) == pl.device &&
                   std::get<1>(i->first) == pl.lay.K && std::get<2>(i->first) == pl.lay.index_bytes &&
                   std::get<3>(i->first) == dtype;
        };
        // a one-instance entry (one lane per state) says nothing about a batch, nor a batch
        // entry about one instance: below 64 instances without an entry of their own, the
        // defaults (as many lanes per state as instances) serve it
        if (B < 64) return default_tune(pl.lay, dtype, B);
        if (hi != g_tune.begin() && same(std::prev(hi))) it = std::prev(hi);
        else if (same(hi)) it = hi;
        else return default_tune(pl.lay, dtype, B);
    }
    Tune t = it->second;
    if (t.bt > pow2_at_least(B) && t.bt != 64) t.bt = pow2_at_least(B);
    if (t.chunk <= 0 || t.chunk > B) t.chunk = B;
    return t;
}

// ------------------------------------------------------------------------------ forward

// The parent-pair items of every level (k_level_fwd_lds_vec2), built once per plan from its own
// device transitions: per level the flags of k_item_flags, compacted on the host into item words
// (state number in the level | second parent follows << 31).
static int ensure_items(const DevPlan &pl) {
    if (pl.items_ready) return EGGMAN_OK;
    const PlanLayout &lay = pl.lay;
    if (lay.nstates >= (int64_t(1) << 31))
        return fail(EGGMAN_EUNSUPPORTED, "parent-pair items need fewer than 2^31 states");
    const int top = lay.K / 2 + 1;
    std::vector<int64_t> off(top + 1, 0), cnt(top + 1, 0);
    std::vector<uint32_t> items;
    std::vector<uint8_t> flags((size_t)std::max<int64_t>(lay.max_level, 1));
    uint8_t *d_flags = nullptr;
    EGG_HIP(hipMalloc(&d_flags, flags.size()));
    int rc = EGGMAN_OK;
    for (int L = 2; L <= top && rc == EGGMAN_OK; L++) {
        const int64_t nst = lay.nst[L - 1];
        const int k = lay.size[L - 1] - 1;
        const unsigned blocks = (unsigned)((nst + 255) / 256);
        if (lay.index_bytes == 4)
            k_item_flags<int32_t><<<dim3(blocks), dim3(256)>>>((const int32_t *)pl.d_trans + lay.tstart[L - 1],
                                                              nst, k, lay.K, lay.lo[L - 1], d_flags);
        else
            k_item_flags<int64_t><<<dim3(blocks), dim3(256)>>>((const int64_t *)pl.d_trans + lay.tstart[L - 1],
                                                              nst, k, lay.K, lay.lo[L - 1], d_flags);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(flags.data(), d_flags, (size_t)nst, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            rc = fail(EGGMAN_EHIP, "item flags: %s", hipGetErrorString(e));
            break;
        }
        off[L - 1] = (int64_t)items.size();
        for (int64_t r = 0; r < nst; r++)
            if (flags[(size_t)r] & 1) items.push_back((uint32_t)r | ((flags[(size_t)r] & 2) ? 0x80000000u : 0u));
        cnt[L - 1] = (int64_t)items.size() - off[L - 1];
    }
    (void)hipFree(d_flags);
    if (rc != EGGMAN_OK) return rc;
    uint32_t *d = nullptr;
    const size_t bytes = std::max<size_t>(items.size(), 1) * sizeof(uint32_t);
    hipError_t e = hipMalloc(&d, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(EGGMAN_ENOMEM, "parent-pair items: %s", hipGetErrorString(e));
    }
    e = hipMemcpy(d, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(EGGMAN_EHIP, "parent-pair items: %s", hipGetErrorString(e));
    }
    pl.d_items = d;
    pl.items_bytes = bytes;
    pl.item_off = off;
    pl.nitems = cnt;
    pl.items_ready = true;
    return EGGMAN_OK;
}

template <int DT, typename W, int BT>
static void launch_level(const DevPlan &pl, int L, const void *P, int64_t ldp, const void *Hc,
                         int64_t ldh, void *Ho, int64_t ldo, int64_t nb, hipStream_t st) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    const int64_t nst = lay.nst[L - 1];
    const int k = lay.size[L - 1] - 1;
    const int64_t ntiles = (nb + BT - 1) / BT;
    const int64_t threads = nst * ntiles * BT;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (threads + kBlock - 1) / kBlock));
    const W *tl = (const W *)pl.d_trans + lay.tstart[L - 1];
    k_level_fwd<DT, W, BT><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(
        tl, nst, k, (const typename E::P *)P, ldp, (const typename E::H *)Hc, lay.levels[L - 2],
        ldh, (typename E::H *)Ho, ldo, nb, ntiles);
}

constexpr int kLdsWaves = 8;   // waves per workgroup of the LDS kernels (16 with Tune::wide)

// Workgroup-to-(tile, states) mapping of the LDS kernels for `nst` states of `nb` instances:
// enough workgroups per XCD to fill its CUs at the occupancy the LDS use allows (and at most 32
// waves per CU), and no more than there are states.
// `tile` instances per tile (64, or 16 for the narrow kernel) and `spw` states per wave step
// (1, or 4 for the narrow kernel).
static LdsMap make_lds_map(const DevCaps &caps, int occ_cap, int64_t nst, int64_t nb, size_t lds,
                           int wpb = kLdsWaves, int tile = 64, int spw = 1) {
    int occ = (int)std::max<size_t>(1, std::min<size_t>(32 / wpb, caps.lds_per_cu / std::max<size_t>(lds, 1)));
    if (occ_cap > 0) occ = std::min(occ, occ_cap);
    const int cus_xcd = std::max(1, caps.cus / caps.xcds);
    const int64_t target = (int64_t)cus_xcd * occ;
    const int64_t per_wg = (int64_t)wpb * spw;   // states one workgroup step covers
    LdsMap map;
    map.ntiles = (nb + tile - 1) / tile;
    if (map.ntiles >= 8) {
        map.tiles_xcd = (map.ntiles + 7) / 8;
        int64_t bpt = std::max<int64_t>(1, target / map.tiles_xcd);
        bpt = std::min<int64_t>(bpt, std::max<int64_t>(1, (nst + per_wg - 1) / per_wg));
        map.per_xcd = map.tiles_xcd * bpt;
    } else {
        map.tiles_xcd = 0;
        const int64_t groups = (8 + map.ntiles - 1) / map.ntiles;
        const int64_t per_group = (nst + groups - 1) / groups;
        map.per_xcd = std::max<int64_t>(1, std::min<int64_t>(target, (per_group + per_wg - 1) / per_wg));
    }
    return map;
}

// The contiguous range of pair indices read by level L: rows p_lo .. p_hi of the upper
// triangle, where p = min(S) of the level's states (plan_rank.hpp).  qlo is 0-based.
static void level_pairs(const PlanLayout &lay, int L, int *qlo, int *nq) {
    const int K = lay.K, m = lay.size[L - 1];
    const int plo = lay.lo[L - 1];
    const int phi = std::min(K - m + 1, K - 1);
    *qlo = (int)pair_index(plo, plo + 1, K) - 1;
    *nq = (int)pair_index(phi, K, K) - 1 - *qlo + 1;
}

// The LDS level kernels: 64-instance rows, one state per wave (KIND_WIDE); 16-instance tiles,
// 4 states per wave (KIND_NARROW); 16-byte loads, 8 states per wave (KIND_VEC).
enum { KIND_WIDE = 0, KIND_NARROW = 1, KIND_VEC = 2, KIND_VEC2 = 3 };

template <int DT> constexpr int kind_tile(int kind) {
    return kind >= KIND_VEC ? 128 / (int)sizeof(typename Elt<DT>::H) : kind == KIND_NARROW ? 16 : 64;
}

// The 16-byte-load kernel reads whole 128-byte tile rows: it needs rows a multiple of its tile
// apart and 16-byte aligned, with readable padding past the batch.
template <int DT>
static bool vec_rows_ok(const void *Hc, int64_t ldh) {
    return ldh % kind_tile<DT>(KIND_VEC) == 0 && ((uintptr_t)Hc & 15) == 0;
}

// Level rows of internal state buffers: padded to whole 128-byte tile rows of the 16-byte-load
// kernel (32 instances covers every element type), so any kernel may walk them.
inline int64_t pad_rows(int64_t nb) { return (nb + 31) / 32 * 32; }

// Launch one level with an LDS kernel; returns false (launching nothing) when the level's pair
// rows do not fit in one workgroup's LDS.  TAG only separates the instantiation used by the
// gradient's forward pass from the plain forward in profiles; the code is the same.
template <int DT, typename W, int WPB, int TAG, int KIND = KIND_WIDE, int UB = 4>
static bool launch_level_lds_w(const DevCaps &caps, const Tune &tune, const DevPlan &pl, int L,
                               const void *P, int64_t ldp, const void *Hc, int64_t ldh, void *Ho,
                               int64_t ldo, int64_t nb, hipStream_t st) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    constexpr int TW = kind_tile<DT>(KIND);
    int qlo, nq;
    level_pairs(lay, L, &qlo, &nq);
    const size_t lds = (size_t)nq * TW * sizeof(typename E::P);
    if (lds > caps.max_lds_block) return false;
    auto kern = KIND == KIND_VEC2     ? k_level_fwd_lds_vec2<DT, W, WPB, TAG, UB>
                : KIND == KIND_VEC    ? k_level_fwd_lds_vec<DT, W, WPB, TAG, UB>
                : KIND == KIND_NARROW ? k_level_fwd_lds_narrow<DT, W, WPB, TAG>
                                      : k_level_fwd_lds<DT, W, WPB, TAG>;
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)caps.max_lds_block) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        attr_set = true;
    }
    const int64_t nst = lay.nst[L - 1];
    const int64_t nwork = KIND == KIND_VEC2 ? pl.nitems[L - 1] : nst;   // items or states
    const LdsMap map = make_lds_map(caps, tune.occ, nwork, nb, lds, WPB, TW,
                                    KIND >= KIND_VEC ? 8 : KIND == KIND_NARROW ? 4 : 1);
    LevelArgs<DT, W> a;
    a.trans_lvl = (const W *)pl.d_trans + lay.tstart[L - 1];
    a.nst = nst;
    a.k = lay.size[L - 1] - 1;
    a.qlo = qlo;
    a.nq = nq;
    a.P = (const typename E::P *)P;
    a.ldp = ldp;
    a.Hc = (const typename E::H *)Hc;
    a.cfirst = lay.levels[L - 2];
    a.ldh = ldh;
    a.Ho = (typename E::H *)Ho;
    a.ldo = ldo;
    a.nb = nb;
    a.map = map;
    a.addr32 = child_addr32(lay.nstates, ldh * (int64_t)sizeof(typename E::H));
    a.items = KIND == KIND_VEC2 ? pl.d_items + pl.item_off[L - 1] : nullptr;
    a.nitems = KIND == KIND_VEC2 ? pl.nitems[L - 1] : 0;
    a.K = lay.K;
    a.plo = lay.lo[L - 1];
    a.vstore = ((uintptr_t)Ho & 15) == 0 && ldo % (16 / (int64_t)sizeof(typename E::H)) == 0;
    kern<<<dim3((unsigned)(8 * map.per_xcd)), dim3(WPB * 64), lds, st>>>(a);
    return true;
}

template <int DT, typename W, int TAG = 0>
static bool launch_level_lds(const DevCaps &caps, const Tune &tune, const DevPlan &pl, int L,
                             const void *P, int64_t ldp, const void *Hc, int64_t ldh, void *Ho,
                             int64_t ldo, int64_t nb, hipStream_t st) {
    if (tune.vec && tune.pairs && pl.items_ready && vec_rows_ok<DT>(Hc, ldh)) {
        if (tune.deep) {
            if (tune.wide)
                return launch_level_lds_w<DT, W, 16, TAG, KIND_VEC2, 6>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
            return launch_level_lds_w<DT, W, kLdsWaves, TAG, KIND_VEC2, 6>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        }
        if (tune.wide)
            return launch_level_lds_w<DT, W, 16, TAG, KIND_VEC2>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        return launch_level_lds_w<DT, W, kLdsWaves, TAG, KIND_VEC2>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
    }
    if (tune.vec && vec_rows_ok<DT>(Hc, ldh)) {
        if (tune.deep) {
            if (tune.wide)
                return launch_level_lds_w<DT, W, 16, TAG, KIND_VEC, 6>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
            return launch_level_lds_w<DT, W, kLdsWaves, TAG, KIND_VEC, 6>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        }
        if (tune.wide)
            return launch_level_lds_w<DT, W, 16, TAG, KIND_VEC>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        return launch_level_lds_w<DT, W, kLdsWaves, TAG, KIND_VEC>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
    }
    if (tune.narrow || tune.vec) {
        if (tune.wide)
            return launch_level_lds_w<DT, W, 16, TAG, KIND_NARROW>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        return launch_level_lds_w<DT, W, kLdsWaves, TAG, KIND_NARROW>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
    }
    if (tune.wide)
        return launch_level_lds_w<DT, W, 16, TAG>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
    return launch_level_lds_w<DT, W, kLdsWaves, TAG>(caps, tune, pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
}

// One or two instances (BT lanes per state): k_level_fwd_plain with the pair values in LDS when
// they fit in 64 KB; false (nothing launched) otherwise.  One N=28 FP64 matrix: 64.0 -> 61.5 us;
// one complex N=40 matrix: 12.2 -> 11.1 ms (with batches of 8).  From 4 instances on, the
// generic kernel's coalesced pair reads beat LDS (N=28, B=8: 103 against 121 us).
template <int DT, typename W, int BT>
static bool launch_level_plain(const DevPlan &pl, int L, const void *P, int64_t ldp, const void *Hc,
                               int64_t ldh, void *Ho, int64_t ldo, int64_t nb, hipStream_t st) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    const int64_t nst = lay.nst[L - 1];
    const int k = lay.size[L - 1] - 1;
    const size_t lds = (size_t)lay.npairs * BT * sizeof(typename E::P);
    if (lds > 65536 || k < 1 || nb > BT) return false;
    const int64_t per_block = kPlainBlock / BT;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (nst + per_block - 1) / per_block));
    const W *tl = (const W *)pl.d_trans + lay.tstart[L - 1];
    k_level_fwd_plain<DT, W, BT><<<dim3((unsigned)blocks), dim3(kPlainBlock), lds, st>>>(
        tl, nst, k, (const typename E::P *)P, ldp, (int)lay.npairs, (const typename E::H *)Hc,
        lay.levels[L - 2], ldh, (typename E::H *)Ho, ldo, nb);
    return true;
}

template <int DT, typename W>
static void launch_level_bt(int bt, const DevPlan &pl, int L, const void *P, int64_t ldp,
                            const void *Hc, int64_t ldh, void *Ho, int64_t ldo, int64_t nb,
                            hipStream_t st) {
    switch (bt) {
    case 1:
        if (nb == 1 && launch_level_plain<DT, W, 1>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st)) break;
        launch_level<DT, W, 1>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        break;
    case 2:
        if (!launch_level_plain<DT, W, 2>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st))
            launch_level<DT, W, 2>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st);
        break;
    case 4: launch_level<DT, W, 4>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 8: launch_level<DT, W, 8>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 16: launch_level<DT, W, 16>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 32: launch_level<DT, W, 32>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    default: launch_level<DT, W, 64>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    }
}

// LDS bytes of the fused kernel: every state value plus the instance's pair values.
static size_t fused_lds_bytes(const PlanLayout &lay, int dtype) {
    return (((size_t)lay.nstates * elt_hbytes(dtype) + 15) & ~(size_t)15) +
           (size_t)lay.npairs * elt_pbytes(dtype);
}

// Launch the fused kernel; false (nothing launched) if the state array does not fit in LDS.
template <int DT, typename W>
static bool launch_fused(const DevCaps &caps, const DevPlan &pl, const void *P, int64_t ldp,
                         int64_t B, void *out, hipStream_t st) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    const size_t lds = fused_lds_bytes(lay, DT);
    if (lds > caps.max_lds_block || lay.nlevels - 1 > kFusedMaxLevels || lay.nstates > INT32_MAX)
        return false;
    auto kern = k_fused_lds<DT, W>;
    static bool attr_set = false;
    if (!attr_set) {
        if (hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)caps.max_lds_block) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        attr_set = true;
    }
    FusedLevels lv;
    lv.nlev = lay.K / 2 + 1;
    for (int L = 1; L <= lv.nlev; L++) {
        lv.first[L - 1] = (int32_t)(lay.levels[L - 1] - 1);
        lv.nst[L - 1] = (int32_t)lay.nst[L - 1];
        lv.tstart[L - 1] = lay.tstart[L - 1];
        lv.k[L - 1] = std::max(0, lay.size[L - 1] - 1);
    }
    const int64_t occ = std::max<int64_t>(1, (int64_t)(caps.lds_per_cu / lds));
    const int64_t blocks = std::min<int64_t>(B, (int64_t)caps.cus * occ);
    kern<<<dim3((unsigned)blocks), dim3(1024), lds, st>>>(
        (const W *)pl.d_trans, lv, (int)lay.nstates, (int)lay.npairs, (const typename E::P *)P, ldp,
        (typename E::H *)out, B);
    return true;
}

// Runs of small levels in a one-pass walk of few instances (the plain path): the levels at the
// bottom and at the top of the plan with at most kFuseLevel states each walk in one launch of
// k_fused_range per run, their values in at most 64 KB of LDS, instead of one launch per level.
// One N=28 FP64 matrix: 67.2 -> 64.2 us; N=24: 46.3 -> 40.8 us (512 or 4096 states per level
// fused less or made the single workgroup the bottleneck).  2048 rather than 1024 also takes N=28's
// second bottom level (1820 states): 52.6 -> 50.4 us (N=24 35.4 -> 34.7, N=32 unchanged; 3000:
// 55.7 us at N=28).
constexpr int64_t kFuseLevel = 2048;
constexpr size_t kFuseLds = 65536;

// LDS bytes of a fused run La .. Lb: levels La .. Lb-1 (and the empty set when La = 2), then P.
static size_t range_lds_bytes(const PlanLayout &lay, int La, int Lb, int dtype) {
    int64_t keep = 0;
    for (int L = La == 2 ? 1 : La; L < Lb; L++) keep += lay.nst[L - 1];
    return (((size_t)keep * elt_hbytes(dtype) + 15) & ~(size_t)15) + (size_t)lay.npairs * elt_pbytes(dtype);
}

template <int DT, typename W>
static void launch_fused_range(const DevCaps &caps, const DevPlan &pl, int La, int Lb, const void *P,
                               int64_t ldp, const void *Hin, int64_t ldh, void *Hout, int64_t ldo,
                               int64_t B, hipStream_t st) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    FusedLevels lv;
    lv.nlev = lay.K / 2 + 1;
    for (int L = 1; L <= lv.nlev; L++) {
        lv.first[L - 1] = (int32_t)(lay.levels[L - 1] - 1);
        lv.nst[L - 1] = (int32_t)lay.nst[L - 1];
        lv.tstart[L - 1] = lay.tstart[L - 1];
        lv.k[L - 1] = std::max(0, lay.size[L - 1] - 1);
    }
    int64_t keep = 0;
    for (int L = La == 2 ? 1 : La; L < Lb; L++) keep += lay.nst[L - 1];
    const int base = La == 2 ? 0 : lv.first[La - 1];
    const size_t lds = range_lds_bytes(lay, La, Lb, DT);
    const int64_t occ = std::max<int64_t>(1, (int64_t)(caps.lds_per_cu / std::max<size_t>(lds, 1)));
    const int64_t blocks = std::min<int64_t>(B, (int64_t)caps.cus * std::min<int64_t>(occ, 2));
    k_fused_range<DT, W><<<dim3((unsigned)blocks), dim3(1024), lds, st>>>(
        (const W *)pl.d_trans, lv, La, Lb, base, (int)keep, (int)lay.npairs,
        (const typename E::P *)P, ldp, (const typename E::H *)Hin, ldh, (typename E::H *)Hout, ldo, B);
}

// Levels with at most kSmallLevel states are walked once for the whole batch rather than once per
// pass, while their whole-batch rows fit in kSmallLevelBytes.
constexpr int64_t kSmallLevel = 4096;
constexpr size_t kSmallLevelBytes = size_t(1) << 30;

// How a forward call of B instances walks its levels, and the state memory it needs.
struct WalkPlan {
    bool fused = false;        // one fused launch (states in LDS), no state memory
    int64_t chunk = 0;         // instances per pass
    bool pad = false;          // level rows padded to whole 128-byte tile rows (16-byte loads)
    int64_t rows = 0, ldmax = 0, ldw = 0;
    int lb = 1, lt = 0;        // levels 2..lb and lt..top walk once for the whole batch
    int fb = 1, ft = 0;        // one-pass walks: levels 2..fb and ft..top in one fused launch each
    int64_t wrows_b = 0, wrows_t = 0;
    size_t h_bytes = 0;        // two levels of `ldmax`-wide rows (ping-pong)
    size_t w_bytes = 0;        // whole-batch rows of the small bottom and top levels
};

static bool fused_fits(const DevCaps &caps, const PlanLayout &lay, int dtype) {
    return fused_lds_bytes(lay, dtype) <= caps.max_lds_block && lay.nlevels - 1 <= kFusedMaxLevels &&
           lay.nstates <= INT32_MAX;
}

static void walk_plan(const DevCaps &caps, const PlanLayout &lay, const Tune &tune, int64_t B, int dtype,
                      WalkPlan *w) {
    *w = WalkPlan();
    const int top = lay.K / 2 + 1;
    w->lt = w->ft = top + 1;
    if (tune.fused && fused_fits(caps, lay, dtype)) {
        w->fused = true;
        return;
    }
    const size_t hb = (size_t)elt_hbytes(dtype);
    w->chunk = tune.chunk > 0 ? std::min(tune.chunk, B) : B;
    w->rows = std::max<int64_t>(lay.max_level, 1);
    // level rows: nb apart, or for the 16-byte-load kernel padded to whole 128-byte tile rows
    w->pad = tune.bt == 64 && tune.lds && tune.vec;
    w->ldmax = w->pad ? pad_rows(w->chunk) : w->chunk;
    w->ldw = w->pad ? pad_rows(B) : B;
    // Small levels at the bottom (2 .. lb) and the top (lt .. top) are walked once for the whole
    // batch; the passes walk lb+1 .. lt-1, reading level lb's whole-batch rows at their columns
    // and writing level lt-1's.  Small levels are bound by latency (few states, long sums), so
    // one launch over every instance costs about what one pass's launch did.
    if (w->chunk < B && tune.bt == 64 && tune.whole_small) {
        int lb = 1, lt = top + 1;
        int64_t wb = 0, wt = 0;
        while (lb + 1 < top && lay.nst[lb] <= kSmallLevel) lb++;
        while (lt - 1 > lb + 1 && lay.nst[lt - 2] <= kSmallLevel) lt--;
        for (int L = 1; L <= lb && lb > 1; L++) wb = std::max(wb, lay.nst[L - 1]);
        for (int L = lt - 1; L < top && lt <= top; L++) wt = std::max(wt, lay.nst[L - 1]);
        if ((size_t)2 * (wb + wt) * w->ldw * hb <= kSmallLevelBytes) {
            w->lb = lb;
            w->lt = lt;
            w->wrows_b = wb;
            w->wrows_t = wt;
        }
    }
    // one pass of few instances: the run of small levels 2 .. fb and the run ft .. top each walk
    // in one fused launch (runs of one level gain nothing and are not fused)
    if (w->chunk == B && tune.bt < 64 && lay.nstates < INT32_MAX && top <= kFusedMaxLevels) {
        int fb = 1, ft = top + 1;
        while (fb + 1 <= top && lay.nst[fb] <= kFuseLevel && range_lds_bytes(lay, 2, fb + 1, dtype) <= kFuseLds)
            fb++;
        if (fb == 2) fb = 1;
        if (fb < top) {
            while (ft - 1 > fb + 1 && lay.nst[ft - 2] <= kFuseLevel &&
                   range_lds_bytes(lay, ft - 1, top, dtype) <= kFuseLds)
                ft--;
            if (ft == top) ft = top + 1;
        }
        w->fb = fb;
        w->ft = ft;
    }
    w->h_bytes = ((size_t)2 * w->rows * w->ldmax * hb + 255) & ~(size_t)255;
    w->w_bytes = (size_t)2 * (w->wrows_b + w->wrows_t) * w->ldw * hb;
}

// Caller-provided state memory (the no-allocation entry point, the concurrent walks of a mixed
// call): forward_levels then carves its level buffers from it instead of the library's grow-only
// buffers, and allocates nothing.  `aux`: a second stream the walk may run every other pass on,
// when the memory holds a second pair of level buffers (null: one stream, nothing created).
struct WorkSpan {
    char *p = nullptr;
    size_t bytes = 0;
    hipStream_t aux = nullptr;
};

// Walk all levels for instances [0, B) of P (rows ldp apart) into out, `chunk` instances per
// pass, keeping only two levels of state values (the level being written and the level below,
// which holds every child it reads).
template <int DT, typename W>
static int forward_levels(DevState &ds, const DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                          void *out, hipStream_t st, const Tune &tune, const WorkSpan *ws = nullptr) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    WalkPlan w;
    walk_plan(ds.caps, lay, tune, B, DT, &w);
    if (w.fused && launch_fused<DT, W>(ds.caps, pl, P, ldp, B, out, st)) {
        EGG_HIP(hipGetLastError());
        return EGGMAN_OK;
    }
    if (w.fused) {   // the kernel declined after all: walk level by level as the tuning says
        Tune t = tune;
        t.fused = false;
        walk_plan(ds.caps, lay, t, B, DT, &w);
    }
    const int64_t chunk = w.chunk;
    const size_t hb = sizeof(typename E::H), pb = sizeof(typename E::P);
    const bool pad = w.pad;
    const int64_t ldw = w.ldw;
    const int top = lay.K / 2 + 1;
    const int lb = w.lb, lt = w.lt, fb = w.fb, ft = w.ft;
    auto launch = [&](int L, const char *Pc, const void *Hc, int64_t ldh, void *Ho, int64_t ldo,
                      int64_t nb) {
        if (tune.bt == 64 && tune.lds &&
            launch_level_lds<DT, W>(ds.caps, tune, pl, L, Pc, ldp, Hc, ldh, Ho, ldo, nb, st))
            return;
        launch_level_bt<DT, W>(tune.bt, pl, L, Pc, ldp, Hc, ldh, Ho, ldo, nb, st);
    };
    if (tune.pairs) EGG_TRY(ensure_items(pl));
    // passes two at a time (pass_streams), each stream with its own pair of level buffers; not
    // under graph capture, and in caller-provided memory only with a second stream and room for
    // the second pair
    int ns = tune.graph ? 1 : pass_streams(DT, (B + chunk - 1) / chunk);
    if (ws && (!ws->aux || (size_t)ns * w.h_bytes + w.w_bytes > ws->bytes)) ns = 1;
    const hipStream_t aux = ns > 1 && ws ? ws->aux : nullptr;
    char *hmem, *wmem = nullptr;
    if (ws) {
        if (w.h_bytes + w.w_bytes > ws->bytes)
            return fail(EGGMAN_ENOMEM, "the walk needs %.1f MiB of state memory, %.1f MiB are provided",
                        (w.h_bytes + w.w_bytes) / 1048576.0, ws->bytes / 1048576.0);
        hmem = ws->p;
        wmem = ws->p + (size_t)ns * w.h_bytes;
    } else {
        EGG_TRY(ensure(ds.H, w.h_bytes * ns));
        hmem = (char *)ds.H.p;
        if (w.w_bytes > 0) {
            EGG_TRY(ensure(ds.Hw, w.w_bytes));
            wmem = (char *)ds.Hw.p;
        }
    }
    char *buf[2] = {hmem, hmem + (size_t)w.rows * w.ldmax * hb};
    char *wb[2] = {nullptr, nullptr}, *wt[2] = {nullptr, nullptr};
    if (w.w_bytes > 0) {
        wb[0] = wmem;
        wb[1] = wmem + (size_t)w.wrows_b * ldw * hb;
        wt[0] = wmem + (size_t)2 * w.wrows_b * ldw * hb;
        wt[1] = wt[0] + (size_t)w.wrows_t * ldw * hb;
    }
    if (lb > 1) {
        k_fill_one<DT><<<dim3((unsigned)((B + 255) / 256)), dim3(256), 0, st>>>(
            (typename E::H *)wb[1], B);
        for (int L = 2; L <= lb; L++)
            launch(L, (const char *)P, wb[(L - 1) & 1], ldw, wb[L & 1], ldw, B);
    }
    const hipStream_t st0 = st;
    if (aux) EGG_TRY(link_streams(ds, aux, st0));
    else EGG_TRY(fork_side(ds, st0, ns));
    for (int64_t b0 = 0, ip = 0; b0 < B; b0 += chunk, ip++) {
        const int64_t nb = std::min(chunk, B - b0);
        const int64_t ld = pad ? pad_rows(nb) : nb;
        const char *Pc = (const char *)P + (size_t)b0 * pb;
        char *outc = (char *)out + (size_t)b0 * hb;
        st = aux && ip % ns ? aux : split_stream(ds, st0, (int)(ip % ns));
        buf[0] = hmem + (size_t)(ip % ns) * w.h_bytes;
        buf[1] = buf[0] + (size_t)w.rows * w.ldmax * hb;
        if (fb > 1)
            launch_fused_range<DT, W>(ds.caps, pl, 2, fb, Pc, ldp, nullptr, 0,
                                      fb == top ? (void *)outc : (void *)buf[fb & 1], ld, nb, st);
        else if (lb == 1)
            k_fill_one<DT><<<dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st>>>(
                (typename E::H *)buf[1], nb);
        for (int L = std::max(lb, fb) + 1; L <= std::min(lt, ft) - 1; L++) {
            const bool from_w = L - 1 == lb && lb > 1;
            const void *Hc = from_w ? (const void *)(wb[lb & 1] + (size_t)b0 * hb) : buf[(L - 1) & 1];
            const int64_t ldh = from_w ? ldw : ld;
            void *Ho = buf[L & 1];
            int64_t ldo = ld;
            if (L == top) {
                Ho = outc;
            } else if (L == lt - 1 && lt <= top) {
                Ho = wt[L & 1] + (size_t)b0 * hb;
                ldo = ldw;
            }
            launch(L, Pc, Hc, ldh, Ho, ldo, nb);
        }
        if (ft <= top)
            launch_fused_range<DT, W>(ds.caps, pl, ft, top, Pc, ldp, buf[(ft - 1) & 1], ld, outc, 1, nb, st);
    }
    st = st0;
    if (aux) EGG_TRY(link_streams(ds, st0, aux));
    else EGG_TRY(join_side(ds, st, ns));
    for (int L = lt; L <= top; L++)
        launch(L, (const char *)P, wt[(L - 1) & 1], ldw, L == top ? out : (void *)wt[L & 1], ldw, B);
    EGG_HIP(hipGetLastError());
    return EGGMAN_OK;
}

template <int DT>
int forward_dt(DevState &ds, const DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                      void *out, hipStream_t st, const Tune &tune, const WorkSpan *ws) {
    if (pl.lay.index_bytes == 4) return forward_levels<DT, int32_t>(ds, pl, P, ldp, B, out, st, tune, ws);
    return forward_levels<DT, int64_t>(ds, pl, P, ldp, B, out, st, tune, ws);
}

// One part of the library instantiates this for its element type; the others use that.
#if EGG_PART >= 0
extern template int forward_dt<EGGMAN_F32>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
extern template int forward_dt<EGGMAN_F64>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
extern template int forward_dt<EGGMAN_C32>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
extern template int forward_dt<EGGMAN_C64>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
extern template int forward_dt<EGGMAN_I128>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
#endif
#if EGG_PART > 0
template int forward_dt<EGG_PART>(DevState &, const DevPlan &, const void *, int64_t, int64_t, void *, hipStream_t, const Tune &, const WorkSpan *);
#endif

static int forward_eager(int dtype, DevState &ds, const DevPlan &pl, const void *P, int64_t ldp,
                         int64_t B, void *out, hipStream_t st, const Tune &tune,
                         const WorkSpan *ws = nullptr) {
    switch (dtype) {
    case EGGMAN_F64: return forward_dt<EGGMAN_F64>(ds, pl, P, ldp, B, out, st, tune, ws);
    case EGGMAN_F32: return forward_dt<EGGMAN_F32>(ds, pl, P, ldp, B, out, st, tune, ws);
    case EGGMAN_C64: return forward_dt<EGGMAN_C64>(ds, pl, P, ldp, B, out, st, tune, ws);
    case EGGMAN_C32: return forward_dt<EGGMAN_C32>(ds, pl, P, ldp, B, out, st, tune, ws);
    case EGGMAN_I128: return forward_dt<EGGMAN_I128>(ds, pl, P, ldp, B, out, st, tune, ws);
    }
    return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
}

// The level sequence of one forward call captured once into a HIP graph and replayed: a call
// repeated with the same buffers (a serving loop, the plain path) then costs one graph launch
// instead of K/2 + 1 kernel launches from the host.  Graphs hold the library's work-buffer
// addresses, so a graph is dropped whenever a work buffer is reallocated.
static int forward_graph(int dtype, DevState &ds, const DevPlan &pl, const void *P, int64_t ldp,
                         int64_t B, void *out, hipStream_t st, const Tune &tune) {
    for (auto it = ds.graphs.begin(); it != ds.graphs.end(); ++it) {
        if (it->gen != g_buf_gen) continue;
        if (it->plan == pl.id && it->dtype == dtype && it->P == P && it->ldp == ldp && it->B == B &&
            it->out == out && it->bt == tune.bt && it->chunk == tune.chunk &&
            it->flags == tune_flags(tune) && it->occ == tune.occ) {
            EGG_HIP(hipGraphLaunch(it->exec, st));
            return EGGMAN_OK;
        }
    }
    // size the work buffers outside the capture (an eager pass also validates the launch)
    EGG_TRY(forward_eager(dtype, ds, pl, P, ldp, B, out, st, tune));
    if (!ds.cap) EGG_HIP(hipStreamCreateWithFlags(&ds.cap, hipStreamNonBlocking));
    EGG_HIP(hipStreamBeginCapture(ds.cap, hipStreamCaptureModeThreadLocal));
    const int rc = forward_eager(dtype, ds, pl, P, ldp, B, out, ds.cap, tune);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(ds.cap, &graph);
    if (rc != EGGMAN_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return rc != EGGMAN_OK ? rc : fail(EGGMAN_EHIP, "graph capture failed: %s", hipGetErrorString(e));
    }
    GraphEntry ge;
    ge.plan = pl.id;
    ge.dtype = dtype;
    ge.P = P;
    ge.ldp = ldp;
    ge.B = B;
    ge.out = out;
    ge.bt = tune.bt;
    ge.chunk = tune.chunk;
    ge.flags = tune_flags(tune);
    ge.occ = tune.occ;
    ge.gen = g_buf_gen;
    const hipError_t ei = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(EGGMAN_EHIP, "graph instantiation failed: %s", hipGetErrorString(ei));
    if (ds.graphs.size() >= 16) {
        (void)hipGraphExecDestroy(ds.graphs.front().exec);
        ds.graphs.erase(ds.graphs.begin());
    }
    ds.graphs.push_back(ge);
    return EGGMAN_OK;   // this call's results came from the eager pass
}

#if EGG_PART <= 0
// ------------------------------------------------------------------------------ no-allocation walk

// State memory of eggman_dp_forward_device_noalloc: one array per device, held from the library's
// first use of that device until the process ends.  On the host the array's name is not a device
// address: host code reaches the memory only through the address hipGetSymbolAddress returns for
// the current device (DevState::scratch_dev), looked up once per device.
constexpr size_t kScratchBytes = size_t(128) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static int scratch_address(DevState &ds, void **p) {
    if (!ds.scratch_dev) {
        void *a = nullptr;
        EGG_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
        ds.scratch_dev = a;
    }
    *p = ds.scratch_dev;
    return EGGMAN_OK;
}

// Launch choices of the no-allocation walk: the tuned choices without a captured graph, fitted to
// the scratch array -- first without the whole-batch rows of the small end levels, then with the
// pass halved until its two levels of rows fit (below 64 instances per pass, lanes per state
// follow the pass).  Every choice gives the same bits.  Fails if one instance per pass does not
// fit.
static int noalloc_tune(const DevState &ds, const DevPlan &pl, int dtype, int64_t B, Tune *out) {
    Tune t = pick_tune(pl, dtype, B);
    t.graph = false;
    if (!pl.items_ready) t.pairs = false;   // building the items would allocate (same bits)
    WalkPlan w;
    for (;;) {
        walk_plan(ds.caps, pl.lay, t, B, dtype, &w);
        if (w.fused || w.h_bytes + w.w_bytes <= kScratchBytes) break;
        if (w.w_bytes > 0 && t.whole_small) {
            t.whole_small = false;
            continue;
        }
        if (w.chunk <= 1)
            return fail(EGGMAN_ENOMEM,
                        "degree %d needs %.1f MiB of state memory for one instance; the "
                        "no-allocation entry point holds %zu MiB (use eggman_dp_forward_device)",
                        pl.lay.K, w.h_bytes / 1048576.0, kScratchBytes >> 20);
        t.chunk = (w.chunk + 1) / 2;
        if (t.chunk < 64) t.bt = pow2_at_least(t.chunk);
    }
    *out = t;
    return EGGMAN_OK;
}

#endif

static int forward_any(int dtype, DevState &ds, const DevPlan &pl, const void *P, int64_t ldp,
                       int64_t B, void *out, hipStream_t st, const Tune &tune) {
    if (tune.graph) return forward_graph(dtype, ds, pl, P, ldp, B, out, st, tune);
    return forward_eager(dtype, ds, pl, P, ldp, B, out, st, tune);
}

// ------------------------------------------------------------------------------ mixed degrees

// Groups of instances with different degrees (and so different plans) walked together: launch t
// advances every group by one level (level t + 2 of each group that still has levels left), so
// one launch carries the work of all degrees and small levels of one degree ride along with large
// levels of another.  Each group uses its own tuned pass size (instances per pass, as the
// single-degree walk does) and keeps two-level state buffers of that size; the groups' passes run
// in rounds, round r walking pass r of every group that has one.  P is the concatenation of the
// groups' [pair][count] blocks, out of their results.
template <int DT, typename W, int WPB, int KIND>
static int forward_mixed_levels(DevState &ds, const std::vector<DevPlan *> &pls,
                                const std::vector<int64_t> &counts, const std::vector<const void *> &Pv,
                                const std::vector<void *> &outv, hipStream_t st) {
    using E = Elt<DT>;
    using HT = typename E::H;
    using PT = typename E::P;
    const size_t ng = pls.size();
    std::vector<const PT *> Pg(ng);
    std::vector<HT *> outg(ng), buf0(ng), buf1(ng);
    std::vector<Tune> tg(ng);
    std::vector<int64_t> chunk(ng);
    size_t hoff = 0;
    int64_t rounds = 0;
    for (size_t g = 0; g < ng; g++) {
        Pg[g] = (const PT *)Pv[g];
        outg[g] = (HT *)outv[g];
        tg[g] = pick_tune(*pls[g], DT, std::max<int64_t>(counts[g], 1));
        chunk[g] = std::max<int64_t>(1, std::min<int64_t>(tg[g].chunk > 0 ? tg[g].chunk : counts[g], counts[g]));
        hoff += (size_t)2 * pls[g]->lay.max_level * pad_rows(chunk[g]);
        if (counts[g] > 0) rounds = std::max<int64_t>(rounds, (counts[g] + chunk[g] - 1) / chunk[g]);
    }
    EGG_TRY(ensure(ds.H, hoff * sizeof(HT)));
    hoff = 0;
    for (size_t g = 0; g < ng; g++) {
        const size_t rows = (size_t)pls[g]->lay.max_level * pad_rows(chunk[g]);
        buf0[g] = (HT *)ds.H.p + hoff;
        buf1[g] = buf0[g] + rows;
        hoff += 2 * rows;
    }
    auto kern = k_level_fwd_lds_multi<DT, W, WPB, KIND>;
    constexpr int TW = kind_tile<DT>(KIND);
    static bool attr_set = false;
    if (!attr_set) {
        attr_set = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)ds.caps.max_lds_block) == hipSuccess;
        (void)hipGetLastError();
    }
    for (int64_t r = 0; r < rounds; r++) {
        std::vector<int64_t> b0(ng), nbg(ng, 0);
        int steps = 0;
        for (size_t g = 0; g < ng; g++) {
            b0[g] = r * chunk[g];
            if (b0[g] >= counts[g]) continue;
            nbg[g] = std::min(chunk[g], counts[g] - b0[g]);
            steps = std::max(steps, pls[g]->lay.K / 2);
            k_fill_one<DT><<<dim3((unsigned)((nbg[g] + 255) / 256)), dim3(256), 0, st>>>(buf1[g], nbg[g]);
        }
        for (int t = 0; t < steps; t++) {
            const int L = t + 2;
            MultiLevelArgs<DT, W> m;
            m.n = 0;
            m.first[0] = 0;
            size_t lds_max = 0;
            auto flush = [&]() {
                if (m.n == 0) return;
                kern<<<dim3((unsigned)m.first[m.n]), dim3(WPB * 64), lds_max, st>>>(m);
                m.n = 0;
                lds_max = 0;
            };
            for (size_t g = 0; g < ng; g++) {
                const DevPlan &pl = *pls[g];
                const PlanLayout &lay = pl.lay;
                const int64_t nb = nbg[g];
                if (nb == 0 || L > lay.K / 2 + 1) continue;
                const int64_t ld = pad_rows(nb);
                const PT *Pc = Pg[g] + b0[g];
                HT *Hc = (L - 1) & 1 ? buf1[g] : buf0[g];
                HT *Ho = L == lay.K / 2 + 1 ? outg[g] + b0[g] : (L & 1 ? buf1[g] : buf0[g]);
                int qlo, nq;
                level_pairs(lay, L, &qlo, &nq);
                const size_t lds = (size_t)nq * TW * sizeof(PT);
                if (!attr_set || !tg[g].lds || lds > ds.caps.max_lds_block) {
                    launch_level_bt<DT, W>(std::min(tg[g].bt, pow2_at_least(nb)), pl, L, Pc, counts[g],
                                           Hc, ld, Ho, ld, nb, st);
                    continue;
                }
                LevelArgs<DT, W> &a = m.g[m.n];
                a.trans_lvl = (const W *)pl.d_trans + lay.tstart[L - 1];
                a.nst = lay.nst[L - 1];
                a.k = lay.size[L - 1] - 1;
                a.qlo = qlo;
                a.nq = nq;
                a.P = Pc;
                a.ldp = counts[g];
                a.Hc = Hc;
                a.cfirst = lay.levels[L - 2];
                a.ldh = ld;
                a.Ho = Ho;
                a.ldo = ld;
                a.nb = nb;
                a.map = make_lds_map(ds.caps, tg[g].occ, a.nst, nb, lds, WPB, TW,
                                     KIND == KIND_VEC ? 8 : KIND == KIND_NARROW ? 4 : 1);
                a.addr32 = child_addr32(lay.nstates, ld * (int64_t)sizeof(HT));
                a.items = nullptr;
                a.nitems = 0;
                a.K = lay.K;
                a.plo = lay.lo[L - 1];
                a.vstore = ((uintptr_t)Ho & 15) == 0 && ld % (16 / (int64_t)sizeof(HT)) == 0;
                m.first[m.n + 1] = m.first[m.n] + 8 * a.map.per_xcd;
                m.n++;
                lds_max = std::max(lds_max, lds);
                if (m.n == kMaxGroups) flush();
            }
            flush();
        }
    }
    EGG_HIP(hipGetLastError());
    return EGGMAN_OK;
}

// The workgroup width and the kernel follow the tuned choice of the group with the most
// transitions (one kernel serves every group of a launch).  Group g reads its P block at Pv[g]
// ([pair][counts[g]]) and writes its results at outv[g].
template <int DT, typename W>
static int forward_mixed_w(DevState &ds, const std::vector<DevPlan *> &pls,
                           const std::vector<int64_t> &counts, const std::vector<const void *> &Pv,
                           const std::vector<void *> &outv, hipStream_t st, bool wide, int kind) {
    switch (kind) {
    case KIND_VEC:
        return wide ? forward_mixed_levels<DT, W, 16, KIND_VEC>(ds, pls, counts, Pv, outv, st)
                    : forward_mixed_levels<DT, W, kLdsWaves, KIND_VEC>(ds, pls, counts, Pv, outv, st);
    case KIND_NARROW:
        return wide ? forward_mixed_levels<DT, W, 16, KIND_NARROW>(ds, pls, counts, Pv, outv, st)
                    : forward_mixed_levels<DT, W, kLdsWaves, KIND_NARROW>(ds, pls, counts, Pv, outv, st);
    }
    return wide ? forward_mixed_levels<DT, W, 16, KIND_WIDE>(ds, pls, counts, Pv, outv, st)
                : forward_mixed_levels<DT, W, kLdsWaves, KIND_WIDE>(ds, pls, counts, Pv, outv, st);
}

// A group of at least kMixedSolo instances fills the device by itself, so it walks as a
// single-degree forward with its own tuned launch, one group after another (largest plan first,
// so the work buffers grow once).  The smaller groups share one walk, where each launch carries
// a level of every one of them.  On 2048 matrices of degrees 16..32, about 400 of each, the
// groups walked one after another take 13.6 ms and the shared walk 16.5 ms.
constexpr int64_t kMixedSolo = 256;

template <int DT>
int forward_mixed_dt(DevState &ds, const std::vector<DevPlan *> &pls,
                            const std::vector<int64_t> &counts, const void *P, void *out,
                            hipStream_t st) {
    using HT = typename Elt<DT>::H;
    using PT = typename Elt<DT>::P;
    const size_t ng = pls.size();
    std::vector<const void *> Pv(ng);
    std::vector<void *> outv(ng);
    size_t poff = 0, ooff = 0;
    for (size_t g = 0; g < ng; g++) {
        Pv[g] = (const PT *)P + poff;
        outv[g] = (HT *)out + ooff;
        poff += (size_t)pls[g]->lay.npairs * counts[g];
        ooff += (size_t)counts[g];
    }
    std::vector<size_t> solo;
    std::vector<DevPlan *> spl;
    std::vector<int64_t> scnt;
    std::vector<const void *> sP;
    std::vector<void *> sout;
    for (size_t g = 0; g < ng; g++) {
        if (counts[g] >= kMixedSolo) {
            solo.push_back(g);
        } else if (counts[g] > 0) {
            spl.push_back(pls[g]);
            scnt.push_back(counts[g]);
            sP.push_back(Pv[g]);
            sout.push_back(outv[g]);
        }
    }
    // the large groups walk concurrently, each on a side stream with its own slice of ds.Hm (the
    // largest plans first, on separate streams), beside the shared walk of the small groups on st
    std::stable_sort(solo.begin(), solo.end(), [&](size_t a, size_t b) {
        return (double)pls[a]->lay.ntrans * counts[a] > (double)pls[b]->lay.ntrans * counts[b];
    });
    std::vector<Tune> tunes;
    std::vector<size_t> span;
    size_t total = 0;
    for (size_t g : solo) {
        Tune t = pick_tune(*pls[g], DT, counts[g]);
        t.graph = false;
        if (t.pairs) EGG_TRY(ensure_items(*pls[g]));
        WalkPlan w;
        walk_plan(ds.caps, pls[g]->lay, t, counts[g], DT, &w);
        // room for a second pair of level buffers when the walk runs its passes two at a time
        const int ns = w.fused ? 1 : pass_streams(DT, (counts[g] + w.chunk - 1) / w.chunk);
        const size_t bytes = ((size_t)ns * w.h_bytes + w.w_bytes + 255) & ~(size_t)255;
        tunes.push_back(t);
        span.push_back(bytes);
        total += bytes;
    }
    if (!solo.empty()) {
        EGG_TRY(ensure(ds.Hm, std::max<size_t>(total, 256)));
        const int nside = (int)std::min<size_t>(solo.size(), DevState::kSide);
        EGG_TRY(fork_side(ds, st, nside + 1));
        size_t off = 0;
        int rc = EGGMAN_OK;
        for (size_t i = 0; i < solo.size() && rc == EGGMAN_OK; i++) {
            const size_t g = solo[i];
            const WorkSpan ws{(char *)ds.Hm.p + off, span[i], ds.sub[i % DevState::kSide]};
            rc = forward_eager(DT, ds, *pls[g], Pv[g], counts[g], counts[g], outv[g],
                               ds.side[i % DevState::kSide], tunes[i], &ws);
            off += span[i];
        }
        EGG_TRY(join_side(ds, st, nside + 1));
        EGG_TRY(rc);
    }
    if (spl.empty()) return EGGMAN_OK;
    size_t big = 0;
    for (size_t g = 1; g < spl.size(); g++)
        if ((double)spl[g]->lay.ntrans * scnt[g] > (double)spl[big]->lay.ntrans * scnt[big]) big = g;
    const Tune t = pick_tune(*spl[big], DT, std::max<int64_t>(scnt[big], 1));
    const int kind = t.vec ? KIND_VEC : t.narrow ? KIND_NARROW : KIND_WIDE;
    if (spl[0]->lay.index_bytes == 4)
        return forward_mixed_w<DT, int32_t>(ds, spl, scnt, sP, sout, st, t.wide, kind);
    return forward_mixed_w<DT, int64_t>(ds, spl, scnt, sP, sout, st, t.wide, kind);
}

// One part of the library instantiates this for its element type; the others use that.
#if EGG_PART >= 0
extern template int forward_mixed_dt<EGGMAN_F32>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
extern template int forward_mixed_dt<EGGMAN_F64>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
extern template int forward_mixed_dt<EGGMAN_C32>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
extern template int forward_mixed_dt<EGGMAN_C64>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
extern template int forward_mixed_dt<EGGMAN_I128>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
#endif
#if EGG_PART > 0
template int forward_mixed_dt<EGG_PART>(DevState &, const std::vector<DevPlan *> &, const std::vector<int64_t> &, const void *, void *, hipStream_t);
#endif

static int forward_mixed_any(int dtype, DevState &ds, const std::vector<DevPlan *> &pls,
                             const std::vector<int64_t> &counts, const void *P, void *out,
                             hipStream_t st) {
    switch (dtype) {
    case EGGMAN_F64: return forward_mixed_dt<EGGMAN_F64>(ds, pls, counts, P, out, st);
    case EGGMAN_F32: return forward_mixed_dt<EGGMAN_F32>(ds, pls, counts, P, out, st);
    case EGGMAN_C64: return forward_mixed_dt<EGGMAN_C64>(ds, pls, counts, P, out, st);
    case EGGMAN_C32: return forward_mixed_dt<EGGMAN_C32>(ds, pls, counts, P, out, st);
    case EGGMAN_I128: return forward_mixed_dt<EGGMAN_I128>(ds, pls, counts, P, out, st);
    }
    return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
}

// Validate a mixed request: known handles on one device with one index width.
static int mixed_plans(int ngroups, const uint64_t *plans, const int64_t *counts,
                       std::vector<DevPlan *> *pls, std::vector<int64_t> *cnt, int64_t *npairs_total,
                       int64_t *total) {
    if (ngroups < 1 || !plans || !counts) return fail(EGGMAN_EINVAL, "need at least one group");
    *npairs_total = 0;
    *total = 0;
    for (int g = 0; g < ngroups; g++) {
        DevPlan *pl;
        EGG_TRY(find_plan(plans[g], &pl));
        if (counts[g] < 0) return fail(EGGMAN_EINVAL, "negative count in group %d", g);
        if (g > 0 && (pl->device != (*pls)[0]->device || pl->lay.index_bytes != (*pls)[0]->lay.index_bytes))
            return fail(EGGMAN_EINVAL, "all groups of a mixed call must share a device and an index width");
        pls->push_back(pl);
        cnt->push_back(counts[g]);
        *npairs_total += pl->lay.npairs * counts[g];
        *total += counts[g];
    }
    return EGGMAN_OK;
}

// ------------------------------------------------------------------------------ reverse sweep

// Transitions per dP segment: short enough that every level yields many (segment, tile) waves
// to hide the latency of the two gathers per transition.
constexpr int kRevSegLen = 256;
// Parent-major dP: waves per workgroup, parents per wave of a task (so a task holds up to
// kParWaves * kParSub parents), and tasks per run of the first reduction stage.
constexpr int kParWaves = 8;
constexpr int kParSub = 16;
constexpr int kParTaskLen = kParWaves * kParSub;
constexpr int kRedFanIn = 64;
constexpr int kFuseMinK = 26;   // smallest degree of the fused reverse step (use_fused)

static int upload_vec(void **dst, const void *src, size_t bytes, size_t *acc) {
    hipError_t e = hipMalloc(dst, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
        *dst = nullptr;
        (void)hipGetLastError();
        return fail(EGGMAN_ENOMEM, "device allocation of %.3f GB for the reverse sweep failed",
                    bytes / 1e9);
    }
    *acc += bytes;
    if (bytes) EGG_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return EGGMAN_OK;
}

// Build the reverse-sweep structures from the plan's own device transitions.
static int ensure_rev(DevPlan &pl) {
    if (pl.rev_ready) return EGGMAN_OK;
    const PlanLayout &lay = pl.lay;
    std::vector<uint8_t> trans((size_t)lay.ntrans * lay.index_bytes);
    EGG_HIP(hipDeviceSynchronize());
    EGG_HIP(hipMemcpy(trans.data(), pl.d_trans, trans.size(), hipMemcpyDeviceToHost));
    RevPlanHost rp;
    EGG_TRY(build_rev_plan(lay, trans.data(), kRevSegLen, kParTaskLen, kRedFanIn, &rp));
    std::vector<RevSegDev> segs(rp.segs.size());
    for (size_t i = 0; i < segs.size(); i++)
        segs[i] = RevSegDev{rp.segs[i].start, rp.segs[i].len, rp.segs[i].level};
    size_t acc = 0;
    EGG_TRY(upload_vec(&pl.d_cent, rp.cent.data(), rp.cent.size(), &acc));
    EGG_TRY(upload_vec(&pl.d_coff, rp.coff.data(), rp.coff.size() * sizeof(int64_t), &acc));
    EGG_TRY(upload_vec(&pl.d_pent, rp.pent.data(), rp.pent.size() * sizeof(uint64_t), &acc));
    EGG_TRY(upload_vec(&pl.d_segs, segs.data(), segs.size() * sizeof(RevSegDev), &acc));
    EGG_TRY(upload_vec(&pl.d_qoff, rp.qoff.data(), rp.qoff.size() * sizeof(int64_t), &acc));
    EGG_TRY(upload_vec(&pl.d_qseg, rp.qseg.data(), rp.qseg.size() * sizeof(int32_t), &acc));
    std::vector<ParTaskDev> tasks(rp.tasks.size());
    for (size_t i = 0; i < tasks.size(); i++) {
        const ParTask &t = rp.tasks[i];
        tasks[i] = ParTaskDev{t.start, t.row0, t.len, t.p, t.level, 0};
    }
    EGG_TRY(upload_vec(&pl.d_prank, rp.prank.data(), rp.prank.size() * sizeof(int32_t), &acc));
    EGG_TRY(upload_vec(&pl.d_tasks, tasks.data(), tasks.size() * sizeof(ParTaskDev), &acc));
    EGG_TRY(upload_vec(&pl.d_roff, rp.roff.data(), rp.roff.size() * sizeof(int64_t), &acc));
    EGG_TRY(upload_vec(&pl.d_rtask, rp.rtask.data(), rp.rtask.size() * sizeof(int32_t), &acc));
    EGG_TRY(upload_vec(&pl.d_qrow, rp.qrow.data(), rp.qrow.size() * sizeof(int32_t), &acc));
    EGG_TRY(upload_vec(&pl.d_qcol, rp.qcol.data(), rp.qcol.size() * sizeof(int32_t), &acc));
    std::vector<RedRowDev> red(rp.red.size());
    for (size_t i = 0; i < red.size(); i++)
        red[i] = RedRowDev{rp.red[i].tb, rp.red[i].te, rp.red[i].u, 0};
    EGG_TRY(upload_vec(&pl.d_red, red.data(), red.size() * sizeof(RedRowDev), &acc));
    EGG_TRY(upload_vec(&pl.d_srow0, rp.srow0.data(), rp.srow0.size() * sizeof(int32_t), &acc));
    EGG_TRY(upload_vec(&pl.d_sruns, rp.sruns.data(), rp.sruns.size() * sizeof(int32_t), &acc));
    pl.lvl_task = rp.lvl_task;
    pl.part_rows = rp.part_rows;
    pl.nred = (int64_t)rp.red.size();
    pl.lvl_seg = rp.lvl_seg;
    pl.nseg = (int64_t)rp.segs.size();
    if (lay.index_bytes == 4 && lay.K >= kFuseMinK) {
        FusePlanHost fp;
        EGG_TRY(build_fuse_plan(lay, rp, kRedFanIn, &fp));
        std::vector<ParTaskDev> ft(fp.tasks.size());
        for (size_t i = 0; i < ft.size(); i++) {
            const ParTask &t = fp.tasks[i];
            ft[i] = ParTaskDev{t.start, t.row0, t.len, t.p, t.level, 0};
        }
        std::vector<RedRowDev> fr(fp.red.size());
        for (size_t i = 0; i < fr.size(); i++)
            fr[i] = RedRowDev{fp.red[i].tb, fp.red[i].te, fp.red[i].u, 0};
        EGG_TRY(upload_vec(&pl.d_fent, fp.fent.data(), fp.fent.size() * sizeof(int32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_fco, fp.fco.data(), fp.fco.size() * sizeof(int32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_cmask, fp.cmask.data(), fp.cmask.size() * sizeof(uint32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_fprow0, fp.prow0.data(), fp.prow0.size() * sizeof(int64_t), &acc));
        EGG_TRY(upload_vec(&pl.d_fterms, fp.terms.data(), fp.terms.size() * sizeof(int32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_ftoff, fp.toff.data(), fp.toff.size() * sizeof(int64_t), &acc));
        EGG_TRY(upload_vec(&pl.d_ftasks, ft.data(), ft.size() * sizeof(ParTaskDev), &acc));
        EGG_TRY(upload_vec(&pl.d_frtask, fp.rtask.data(), fp.rtask.size() * sizeof(int32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_fred, fr.data(), fr.size() * sizeof(RedRowDev), &acc));
        EGG_TRY(upload_vec(&pl.d_fsrow0, fp.srow0.data(), fp.srow0.size() * sizeof(int32_t), &acc));
        EGG_TRY(upload_vec(&pl.d_fsruns, fp.sruns.data(), fp.sruns.size() * sizeof(int32_t), &acc));
        pl.flv = fp.lv;
        pl.fpart_rows = fp.part_rows;
        pl.fnred = (int64_t)fp.red.size();
        pl.fuse = true;
    }
    grad_levels_full(lay, &pl.gl_full);
    grad_levels_checkpointed(lay, &pl.gl_ckpt);
    pl.rev_bytes = acc;
    pl.rev_ready = true;
    return EGGMAN_OK;
}

// Device bytes one instance needs for forward + reverse: the kept levels of H and the segment
// region (every level, without checkpoints), two levels of G, the dP partials (segment sums, or
// task rows plus stage-one rows), and its P and dP columns.
// The fused child-major reverse step serves degrees from kFuseMinK up on the Int32 layout, for
// elements of at most 8 bytes (its LDS rows and registers are sized for them); the others keep
// the parent-major or segment dP.  Below kFuseMinK the levels are small, most of them would take
// the stripe-list kernels, and the parent-major dP is faster (N=22, B=4096: 3.0 against 4.5 ms;
// N=24 even; N=26, B=2048: 11.5 against 10.1 ms).  The choice depends on the degree and element
// type only, so each degree's gradient has one summation order for every pass size.
static bool use_fused(const DevPlan &pl, int dtype) {
    return pl.fuse && pl.lay.K >= kFuseMinK && elt_hbytes(dtype) <= 8;
}

static size_t grad_bytes_per_instance(const DevPlan &pl, int dtype, bool ckpt) {
    const size_t hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    const PlanLayout &lay = pl.lay;
    const GradLevels &gl = ckpt ? pl.gl_ckpt : pl.gl_full;
    const int64_t part = use_fused(pl, dtype) ? pl.fpart_rows + pl.fnred
                                              : std::max<int64_t>(pl.nseg, pl.part_rows + pl.nred);
    return (size_t)(gl.kept_rows + gl.seg_rows + 2 * lay.max_level) * hb +
           (size_t)std::max<int64_t>(part, 1) * hb + (size_t)lay.npairs * (pb + hb);
}

// Waves per workgroup of the parent-major dP kernel at level L: the most (of 8, 4, 2) whose
// LDS rows fit -- each wave holds rows p+1..K of dP for 64 instances, p >= p_lo -- or 0 if none
// does.  The choice depends on the degree, the element type and the device only, so the
// summation order (each wave takes a fixed slice of a task's parents) does too.  A wave's slice
// is at most 64 parents (one rank per lane).
static int par_waves(const DevCaps &caps, const PlanLayout &lay, int L, int dtype, int *rows,
                     size_t *lds) {
    *rows = lay.K - lay.lo[L - 1];
    for (int w = kParWaves; w >= 2; w /= 2) {
        *lds = (size_t)w * (*rows) * 64 * elt_hbytes(dtype);
        if (*lds <= caps.max_lds_block && (kParTaskLen + w - 1) / w <= 64) return w;
    }
    return 0;
}

// The parent-major dP path needs the Int32 layout and every level's rows in LDS.
static bool use_par_dp(const DevCaps &caps, const DevPlan &pl, int dtype) {
    if (pl.lay.index_bytes != 4 || pl.lvl_task.empty()) return false;
    for (int L = 2; L <= pl.lay.K / 2 + 1; L++) {
        int rows;
        size_t lds;
        if (!par_waves(caps, pl.lay, L, dtype, &rows, &lds)) return false;
    }
    return true;
}

template <int DT, int WPB, int U>
static void launch_dp_par_wu(const DevCaps &caps, const DevPlan &pl, int L, int rows, size_t lds,
                             const void *Hc, int64_t ldh, const void *Gs, int64_t ldg, void *part,
                             int64_t nb, hipStream_t st) {
    using HT = typename Elt<DT>::H;
    const PlanLayout &lay = pl.lay;
    auto kern = k_rev_dp_par<DT, WPB, U>;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)caps.max_lds_block);
        (void)hipGetLastError();
        attr_set = true;
    }
    const int64_t t0 = pl.lvl_task[L - 1], nt = pl.lvl_task[L] - t0;
    const int64_t ntiles = (nb + 63) / 64;
    if (nt == 0) return;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, nt * ntiles));
    kern<<<dim3((unsigned)blocks), dim3(WPB * 64), lds, st>>>(
        (const int32_t *)pl.d_trans + lay.tstart[L - 1], lay.size[L - 1] - 1, lay.K,
        (const int32_t *)pl.d_prank, (const ParTaskDev *)pl.d_tasks, t0, nt, (const HT *)Hc,
        lay.levels[L - 2], ldh, (const HT *)Gs, ldg, (HT *)part, nb, ntiles, rows,
        (kParTaskLen + WPB - 1) / WPB);
}

// Transitions per batch: for out-degree k <= 16 the smallest of 8 / 12 / 16 that holds a whole
// parent; above that, batches of 12 up to k = 24 and of 16 beyond, which keeps the repeated
// tail of the last batch short.  16-byte elements always use 8 (their registers are twice as
// wide).
template <int DT, int WPB>
static void launch_dp_par_w(const DevCaps &caps, const DevPlan &pl, int L, int rows, size_t lds,
                            const void *Hc, int64_t ldh, const void *Gs, int64_t ldg, void *part,
                            int64_t nb, hipStream_t st) {
    const int k = pl.lay.size[L - 1] - 1;
    if constexpr (sizeof(typename Elt<DT>::H) > 8) {
        launch_dp_par_wu<DT, WPB, 8>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st);
    } else {
        if (k <= 8) launch_dp_par_wu<DT, WPB, 8>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st);
        else if (k <= 12 || (k > 16 && k <= 24))
            launch_dp_par_wu<DT, WPB, 12>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st);
        else launch_dp_par_wu<DT, WPB, 16>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st);
    }
}

// One instance: k_rev_dp_par_one with the same slicing (WPB slices of the task) as the 64-wide
// kernel, so the same bits.
template <int DT, int WPB>
static void launch_dp_par_one(const DevPlan &pl, int L, int rows, const void *Hc, int64_t ldh,
                              const void *Gs, int64_t ldg, void *part, hipStream_t st) {
    using HT = typename Elt<DT>::H;
    const PlanLayout &lay = pl.lay;
    const int64_t t0 = pl.lvl_task[L - 1], nt = pl.lvl_task[L] - t0;
    if (nt == 0) return;
    const size_t lds = (size_t)WPB * rows * sizeof(HT);
    k_rev_dp_par_one<DT, WPB><<<dim3((unsigned)std::min<int64_t>(kMaxBlocks, nt)), dim3(WPB * 32), lds, st>>>(
        (const int32_t *)pl.d_trans + lay.tstart[L - 1], lay.size[L - 1] - 1, lay.K,
        (const int32_t *)pl.d_prank, (const ParTaskDev *)pl.d_tasks, t0, nt, (const HT *)Hc,
        lay.levels[L - 2], ldh, (const HT *)Gs, ldg, (HT *)part, 1, rows,
        (kParTaskLen + WPB - 1) / WPB);
}

template <int DT>
static void launch_dp_par(const DevCaps &caps, const DevPlan &pl, int L, const void *Hc,
                          int64_t ldh, const void *Gs, int64_t ldg, void *part, int64_t nb,
                          hipStream_t st) {
    int rows;
    size_t lds;
    const int wpb = par_waves(caps, pl.lay, L, DT, &rows, &lds);
    if (nb == 1 && pl.lay.size[L - 1] - 1 <= 32) {
        switch (wpb) {
        case 8: launch_dp_par_one<DT, 8>(pl, L, rows, Hc, ldh, Gs, ldg, part, st); return;
        case 4: launch_dp_par_one<DT, 4>(pl, L, rows, Hc, ldh, Gs, ldg, part, st); return;
        case 2: launch_dp_par_one<DT, 2>(pl, L, rows, Hc, ldh, Gs, ldg, part, st); return;
        }
    }
    switch (wpb) {
    case 8: launch_dp_par_w<DT, 8>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st); break;
    case 4: launch_dp_par_w<DT, 4>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st); break;
    case 2: launch_dp_par_w<DT, 2>(caps, pl, L, rows, lds, Hc, ldh, Gs, ldg, part, nb, st); break;
    }
}

// `lists`: the child-major parent entries the sums follow (pl.d_cent, parents ascending, or the
// fused step's pl.d_fent, pairs ascending); both share the offsets pl.d_coff.
template <int DT, typename W>
static void launch_rev_g(const DevCaps &caps, const Tune &tune, const DevPlan &pl, int L,
                         const void *P, int64_t ldp, const void *Gs, void *Gc, int64_t ldg,
                         int64_t nb, hipStream_t st, const void *lists) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    int qlo, nq;
    level_pairs(lay, L, &qlo, &nq);
    const int64_t nch = lay.nst[L - 2];
    const size_t lds = (size_t)nq * 64 * sizeof(typename E::P);
    const bool use_lds = tune.lds && lds <= caps.max_lds_block;
    const W *cent = (const W *)lists;
    const int64_t *coff = (const int64_t *)pl.d_coff + (lay.levels[L - 2] - 1);
    if (pow2_at_least(nb) < 64) {
        // batches of at most 32: as many lanes per child as instances (next power of two)
        const int bt = pow2_at_least(nb);
        const int64_t ntiles = (nb + bt - 1) / bt;
        const int64_t blocks = std::max<int64_t>(
            1, std::min<int64_t>(kMaxBlocks, (nch * ntiles * bt + kBlock - 1) / kBlock));
        const auto *Pp = (const typename E::P *)P;
        const auto *Gsp = (const typename E::H *)Gs;
        auto *Gcp = (typename E::H *)Gc;
        switch (bt) {
        case 1: k_rev_g_bt<DT, W, 1><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        case 2: k_rev_g_bt<DT, W, 2><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        case 4: k_rev_g_bt<DT, W, 4><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        case 8: k_rev_g_bt<DT, W, 8><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        case 16: k_rev_g_bt<DT, W, 16><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        case 32: k_rev_g_bt<DT, W, 32><<<dim3((unsigned)blocks), dim3(kBlock), 0, st>>>(cent, coff, nch, Pp, ldp, Gsp, ldg, Gcp, nb, ntiles); break;
        }
        return;
    }
    if (use_lds) {
        auto kern = k_rev_g<DT, W, kLdsWaves, true>;
        static bool attr_set = false;
        if (!attr_set) {
            attr_set = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)caps.max_lds_block) == hipSuccess;
            (void)hipGetLastError();
        }
        if (attr_set) {
            const LdsMap map = make_lds_map(caps, tune.occ, nch, nb, lds);
            kern<<<dim3((unsigned)(8 * map.per_xcd)), dim3(kLdsWaves * 64), lds, st>>>(
                cent, coff, nch, qlo, nq, (const typename E::P *)P, ldp,
                (const typename E::H *)Gs, ldg, (typename E::H *)Gc, nb, map);
            return;
        }
    }
    const LdsMap map = make_lds_map(caps, 2, nch, nb, 0);
    k_rev_g<DT, W, kLdsWaves, false><<<dim3((unsigned)(8 * map.per_xcd)), dim3(kLdsWaves * 64), 0, st>>>(
        cent, coff, nch, qlo, nq, (const typename E::P *)P, ldp, (const typename E::H *)Gs, ldg,
        (typename E::H *)Gc, nb, map);
}

// The reverse step of level L by the fused order: G of level L-1 into Gc and the level's dP
// partial rows.  Passes of fused_min_tiles(K) tiles of 64 instances or more take k_rev_fused,
// except on levels whose wave-stripes hold at most four blocks: there, and for smaller passes, the
// phase loop would be bound by latency (a tile's stripes are S workgroups, too few to fill the
// device with one or two tiles), so the adjoint gather over the sorted lists and the stripe sums
// run every phase at once (the same sums in the same order).
//
// Forward + gradient, ms, fused kernel against the lists: one tile, N=28 B=64 2.79 / 1.57, N=32
// B=64 18.7 / 12.3; two tiles, N=26 B=128 1.31 / 1.13, N=28 B=128 3.08 / 2.78, but N=32 B=128
// 20.8 / 23.9; three tiles, N=26 B=192 1.51 / 1.82, N=28 B=192 3.41 / 5.26, N=32 B=192 23.3 /
// 41.6; four and more, N=28 B=256 3.62 / 5.55, N=32 B=256 25.2 / 47.5.
static int64_t fused_min_tiles(int K) { return K >= 30 ? 2 : 3; }
// Children per level up to which the adjoint gather of batches of at most four takes one wave per
// child (rev_g_lanes_body): at N=28 the levels of 1, 105, 1820 and 12376 children, whose lists
// average 105, 52, 34 and 25 entries; the larger levels' lists average 19 or fewer.  (One N=28
// gradient: 0.277 -> 0.273 ms; at 32 instances it was 4% slower than a lane per child.)
constexpr int64_t kGLanesMax = 16384;

template <int DT, typename W>
static void launch_rev_fused(DevState &ds, const DevPlan &pl, const Tune &tune, int L, const void *P,
                             int64_t ldp, const void *Hc, int64_t ldh, const void *Gs, void *Gc,
                             int64_t ldg, void *part, int64_t nb, hipStream_t st) {
    using E = Elt<DT>;
    using HT = typename E::H;
    using PT = typename E::P;
    const FuseLevel &f = pl.flv[L];
    const int nph = f.phi - f.plo + 1;
    const int64_t V = (int64_t)f.S * kFuseWaves;
    if (nb >= 64 && (nb + 63) / 64 >= fused_min_tiles(pl.lay.K) && f.nblk > 4 * V) {
        const int64_t ntiles = (nb + 63) / 64;
        const size_t lds = (size_t)f.nslot * 64 * (sizeof(PT) + kFuseWaves * sizeof(HT));
        const int64_t grid = ntiles >= 8 ? 8 * (int64_t)f.S * ((ntiles + 7) / 8)
                                         : 8 * (((int64_t)f.S + 8 / ntiles - 1) / (8 / ntiles));
        // 32-bit row offsets when every level row this step touches lies within 4 GB of its base
        const int64_t rows_max = std::max<int64_t>(pl.lay.nst[L - 1], f.nch);
        const bool a32 = rows_max * std::max(ldh, ldg) * (int64_t)sizeof(HT) < (int64_t(1) << 32);
        auto kern = a32 ? k_rev_fused<DT, true> : k_rev_fused<DT, false>;
        static bool attr_set[2] = {false, false};
        if (!attr_set[a32]) {
            (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)ds.caps.max_lds_block);
            (void)hipGetLastError();
            attr_set[a32] = true;
        }
        kern<<<dim3((unsigned)grid), dim3(kFuseWaves * 64), lds, st>>>(
            (const int32_t *)pl.d_fent + f.ent0, (const int32_t *)pl.d_fco + f.cfirst,
            (const uint32_t *)pl.d_cmask + f.cfirst, f.nch, pl.lay.K, f.csize, f.plo, f.phi, f.S,
            (const int64_t *)pl.d_fprow0 + f.prow, (const PT *)P, ldp, (const HT *)Hc, ldh,
            (const HT *)Gs, (HT *)Gc, ldg, (HT *)part, nb, ntiles);
        return;
    }
    const int bt = pow2_at_least(nb);
    if (bt <= 32) {
        // batches of at most 32: the adjoint gather and the stripe sums in one launch
        const int64_t nch = pl.lay.nst[L - 2];
        const int64_t ntiles = (nb + bt - 1) / bt;
        // levels of at most kGLanesMax children (their parent lists are the longest) gather with
        // one wave per child when the batch is at most four instances
        const int glanes = nch <= kGLanesMax && bt <= 4 ? 1 : 0;
        const int64_t nga = glanes ? std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (nch + kBlock / 64 - 1) / (kBlock / 64)))
                                   : std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (nch * ntiles * bt + kBlock - 1) / kBlock));
        const int64_t ngb = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (int64_t)nph * f.S * f.nslot));
        const int32_t *cent = (const int32_t *)pl.d_fent;
        const int64_t *coff = (const int64_t *)pl.d_coff + (pl.lay.levels[L - 2] - 1);
#define EGG_FSMALL(BTV)                                                                          \
    k_rev_fsmall<DT, BTV><<<dim3((unsigned)(nga + ngb)), dim3(kBlock), fstripe_lds_bytes<DT, BTV>(), st>>>( \
        nga, cent, coff, nch, (const PT *)P, ldp, (HT *)Gc, ntiles, (const int2 *)pl.d_fterms,    \
        (const int64_t *)pl.d_ftoff + f.toff, (const int64_t *)pl.d_fprow0 + f.prow, f.S, f.nslot, \
        pl.lay.K, f.plo, nph, (const HT *)Hc, ldh, (const HT *)Gs, ldg, (HT *)part, nb, glanes)
        switch (bt) {
        case 1: EGG_FSMALL(1); break;
        case 2: EGG_FSMALL(2); break;
        case 4: EGG_FSMALL(4); break;
        case 8: EGG_FSMALL(8); break;
        case 16: EGG_FSMALL(16); break;
        case 32: EGG_FSMALL(32); break;
        }
#undef EGG_FSMALL
        return;
    }
    launch_rev_g<DT, W>(ds.caps, tune, pl, L, P, ldp, Gs, Gc, ldg, nb, st, pl.d_fent);
    if (nb >= 64) {
        const int64_t ntiles = (nb + 63) / 64;
        const int64_t nw = (int64_t)nph * f.S * f.nslot * ntiles;
        k_rev_fstripe_w<DT><<<dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, nw))),
                              dim3(kFuseWaves * 64), 0, st>>>(
            (const int2 *)pl.d_fterms, (const int64_t *)pl.d_ftoff + f.toff,
            (const int64_t *)pl.d_fprow0 + f.prow, f.S, f.nslot, pl.lay.K, f.plo, nph, (const HT *)Hc,
            ldh, (const HT *)Gs, ldg, (HT *)part, nb, ntiles);
        return;
    }
    const int64_t n = (int64_t)nph * f.S * f.nslot * nb * kFuseWaves;
    k_rev_fstripe<DT><<<dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (n + kBlock - 1) / kBlock))),
                        dim3(kBlock), 0, st>>>(
        (const int2 *)pl.d_fterms, (const int64_t *)pl.d_ftoff + f.toff,
        (const int64_t *)pl.d_fprow0 + f.prow, f.S, f.nslot, pl.lay.K, f.plo, nph, (const HT *)Hc, ldh,
        (const HT *)Gs, ldg, (HT *)part, nb);
}

// Forward walk, then the reverse sweep: out[b] = haf, dP[q][b] = d haf/d P.  The forward keeps
// every level, or (tune.ckpt) only the checkpoint levels of pl.gl_ckpt; a level that was not
// kept is recomputed with the rest of its segment from the kept level below when the reverse
// sweep first needs it.  Recomputed values are the same bits (same kernel, same inputs).
template <int DT, typename W>
static int forward_grad_levels(DevState &ds, DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                               void *out, void *dP, int64_t lddp, hipStream_t st,
                               const Tune &tune) {
    using E = Elt<DT>;
    using HT = typename E::H;
    const PlanLayout &lay = pl.lay;
    EGG_TRY(ensure_rev(pl));
    if (tune.pairs) EGG_TRY(ensure_items(pl));
    const size_t hb = sizeof(HT), pb = sizeof(typename E::P);
    const int64_t chunk = tune.chunk > 0 ? std::min(tune.chunk, B) : B;
    const bool fused = sizeof(W) == 4 && use_fused(pl, DT);
    const bool par = !fused && sizeof(W) == 4 && use_par_dp(ds.caps, pl, DT);
    const GradLevels &gl = tune.ckpt ? pl.gl_ckpt : pl.gl_full;
    // forward level rows: padded to whole tile rows for the 16-byte-load kernel
    const bool pad = tune.lds && tune.vec;
    const int64_t ldmax = pad ? pad_rows(chunk) : chunk;
    // one pass's buffers: every kept level (and a segment), two adjoint levels, the dP partials
    const size_t hg_bytes = (size_t)std::max<int64_t>(gl.kept_rows + gl.seg_rows, 1) * ldmax * hb;
    const size_t g_bytes = (size_t)2 * lay.max_level * ldmax * hb;
    const size_t p_bytes = (size_t)std::max<int64_t>(fused ? pl.fpart_rows + pl.fnred
                                                           : par ? pl.part_rows + pl.nred : pl.nseg, 1) *
                           chunk * hb;
    // two passes in flight (the call's stream and a side stream, each with its own buffers) when
    // there are two or more, the values are at most 8 bytes and a second set of buffers fits in
    // 80% of the free memory; the sub-passes of a pass then run one at a time
    int np = 1;
    if ((B + chunk - 1) / chunk >= 2 && elt_hbytes(DT) <= 8) {
        size_t free_b = 0, total_b = 0;
        EGG_HIP(hipMemGetInfo(&free_b, &total_b));
        const size_t have = ds.Hg.bytes + ds.G.bytes + ds.part.bytes;
        if (2 * (hg_bytes + g_bytes + p_bytes) <= have + (size_t)(0.8 * (double)free_b)) np = kGradStreams;
    }
    EGG_TRY(ensure(ds.Hg, hg_bytes * np));
    EGG_TRY(ensure(ds.G, g_bytes * np));
    EGG_TRY(ensure(ds.part, p_bytes * np));
    const int top = lay.K / 2 + 1;
    const hipStream_t stc = st;
    EGG_TRY(fork_side(ds, stc, np));
    for (int64_t b0 = 0, ip = 0; b0 < B; b0 += chunk, ip++) {
        const int si = (int)(ip % np);
        st = split_stream(ds, stc, si);
        HT *H = (HT *)((char *)ds.Hg.p + si * hg_bytes);
        HT *G[2] = {(HT *)((char *)ds.G.p + si * g_bytes),
                    (HT *)((char *)ds.G.p + si * g_bytes) + (size_t)lay.max_level * ldmax};
        HT *part = (HT *)((char *)ds.part.p + si * p_bytes);
        const int64_t nb = std::min(chunk, B - b0);
        const char *Pc = (const char *)P + (size_t)b0 * pb;
        HT *outc = (HT *)out + b0;
        HT *dPc = (HT *)dP + b0;
        const int64_t ldh = pad ? pad_rows(nb) : nb;
        // rows of level L (1 .. top-1) in the kept region or the segment region
        auto level_rows = [&](int L) {
            return H + (size_t)((gl.kept[L] ? 0 : gl.kept_rows) + gl.off[L]) * ldh;
        };
        // levels L0 .. L1 of the pass, walked in sub-passes of the forward's own tuned size (its
        // child rows stay in L2 as in the forward alone), each writing its columns of the rows;
        // the small levels at either end of the range (at most kSmallLevel states, bound by
        // latency) are walked once for the whole pass
        const int64_t sub = tune.sub > 0 ? std::min(tune.sub, nb) : nb;
        auto launch_cols = [&](int L, int64_t s0, int64_t ns) {
            const char *Ps = Pc + (size_t)s0 * pb;
            const void *Hc = level_rows(L - 1) + s0;
            void *Ho = L == top ? (void *)(outc + s0) : (void *)(level_rows(L) + s0);
            // the LDS kernels work on tiles of 64 instances (16 for the 16-byte-load kernel):
            // fewer take the small-batch ones (N=28, 16 to 48 instances, forward + gradient with
            // the 16-byte-load kernel from 16: 4-8% faster)
            if (!(tune.lds && (ns >= 64 || (tune.vec && ns >= 16)) &&
                  launch_level_lds<DT, W, 1>(ds.caps, tune, pl, L, Ps, ldp, Hc, ldh, Ho, ldh, ns, st)))
                launch_level_bt<DT, W>(pow2_at_least(ns), pl, L, Ps, ldp, Hc, ldh, Ho, ldh, ns, st);
        };
        // (sub-passes two at a time as the forward's passes, pass_streams: they write disjoint
        // columns)
        auto forward_range = [&](int L0, int L1) -> int {
            int lb = L0, lt = L1;
            while (lb <= lt && lay.nst[lb - 1] <= kSmallLevel) launch_cols(lb++, 0, nb);
            while (lt >= lb && lay.nst[lt - 1] <= kSmallLevel) lt--;
            const int ns = lb <= lt && np == 1 ? pass_streams(DT, (nb + sub - 1) / sub) : 1;
            const hipStream_t st0 = st;
            EGG_TRY(fork_side(ds, st0, ns));
            for (int64_t s0 = 0, i = 0, n; s0 < nb && lb <= lt; s0 += n, i++) {
                // a remainder of fewer than 64 instances joins the sub-pass before it (alone it
                // would take the small-batch kernels, one lane per instance)
                n = std::min(sub, nb - s0);
                if (nb - s0 - n > 0 && nb - s0 - n < 64) n = nb - s0;
                st = split_stream(ds, st0, (int)(i % ns));
                for (int L = lb; L <= lt; L++) launch_cols(L, s0, n);
            }
            st = st0;
            EGG_TRY(join_side(ds, st, ns));
            for (int L = lt + 1; L <= L1; L++) launch_cols(L, 0, nb);
            return EGGMAN_OK;
        };
        k_fill_one<DT><<<dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st>>>(level_rows(1), nb);
        EGG_TRY(forward_range(2, top));
        // the segment the forward left in the segment region: the run of levels below top
        int seg_lo = top, seg_hi = top - 1;
        if (!gl.kept[top - 1]) {
            seg_lo = top - 1;
            while (!gl.kept[seg_lo - 1]) seg_lo--;
        }
        // reverse: G of the full set is one
        k_fill_one<DT><<<dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st>>>(G[top & 1], nb);
        const int64_t ntiles = (nb + 63) / 64;
        for (int L = top; L >= 2; L--) {
            if (!gl.kept[L - 1] && !(seg_lo <= L - 1 && L - 1 <= seg_hi)) {
                // level L-1 heads a segment that is not resident: recompute the segment
                seg_hi = L - 1;
                seg_lo = L - 1;
                while (!gl.kept[seg_lo - 1]) seg_lo--;
                EGG_TRY(forward_range(seg_lo, seg_hi));
            }
            if (fused) {
                launch_rev_fused<DT, W>(ds, pl, tune, L, Pc, ldp, level_rows(L - 1), ldh, G[L & 1],
                                        G[(L - 1) & 1], ldh, part, nb, st);
                continue;
            }
            if (par) {
                launch_dp_par<DT>(ds.caps, pl, L, level_rows(L - 1), ldh, G[L & 1], ldh, part, nb, st);
            } else {
                const int64_t s0 = pl.lvl_seg[L - 1], ns = pl.lvl_seg[L] - s0;
                if (ns > 0) {
                    const int64_t tasks = ns * ntiles;
                    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (tasks + 3) / 4));
                    k_rev_dp_seg<DT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
                        (const uint64_t *)pl.d_pent, (const RevSegDev *)pl.d_segs, s0, ns,
                        level_rows(L - 1), ldh, G[L & 1], ldh, part, nb, ntiles);
                }
            }
            if (L > 2)
                launch_rev_g<DT, W>(ds.caps, tune, pl, L, Pc, ldp, G[L & 1], G[(L - 1) & 1], ldh, nb, st,
                                    pl.d_cent);
        }
        const int64_t total = lay.npairs * nb;
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (total + 255) / 256));
        if (fused) {
            HT *runs = part + (size_t)pl.fpart_rows * nb;
            const int64_t rtotal = pl.fnred * nb;
            const int64_t rblocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (rtotal + 255) / 256));
            k_rev_dp_reduce_runs<DT><<<dim3((unsigned)rblocks), dim3(256), 0, st>>>(
                (const RedRowDev *)pl.d_fred, pl.fnred, (const int32_t *)pl.d_frtask,
                (const ParTaskDev *)pl.d_ftasks, part, runs, nb);
            k_rev_dp_reduce_pairs<DT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
                (const int32_t *)pl.d_qrow, (const int32_t *)pl.d_qcol, (const int32_t *)pl.d_fsrow0,
                (const int32_t *)pl.d_fsruns, lay.K, runs, dPc, lddp, lay.npairs, nb);
        } else if (par) {
            HT *runs = part + (size_t)pl.part_rows * nb;
            const int64_t rtotal = pl.nred * nb;
            const int64_t rblocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (rtotal + 255) / 256));
            k_rev_dp_reduce_runs<DT><<<dim3((unsigned)rblocks), dim3(256), 0, st>>>(
                (const RedRowDev *)pl.d_red, pl.nred, (const int32_t *)pl.d_rtask,
                (const ParTaskDev *)pl.d_tasks, part, runs, nb);
            k_rev_dp_reduce_pairs<DT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
                (const int32_t *)pl.d_qrow, (const int32_t *)pl.d_qcol, (const int32_t *)pl.d_srow0,
                (const int32_t *)pl.d_sruns, lay.K, runs, dPc, lddp, lay.npairs, nb);
        } else
            k_rev_dp_reduce<DT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>(
                (const int64_t *)pl.d_qoff, (const int32_t *)pl.d_qseg, part, dPc, lddp, lay.npairs, nb);
    }
    st = stc;
    EGG_TRY(join_side(ds, st, np));
    EGG_HIP(hipGetLastError());
    return EGGMAN_OK;
}

template <int DT>
int forward_grad_dt(DevState &ds, DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                           void *out, void *dP, int64_t lddp, hipStream_t st, const Tune &tune) {
    if (pl.lay.index_bytes == 4)
        return forward_grad_levels<DT, int32_t>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    return forward_grad_levels<DT, int64_t>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
}

// One part of the library instantiates this for its element type; the others use that.
#if EGG_PART >= 0
extern template int forward_grad_dt<EGGMAN_F32>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
extern template int forward_grad_dt<EGGMAN_F64>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
extern template int forward_grad_dt<EGGMAN_C32>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
extern template int forward_grad_dt<EGGMAN_C64>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
extern template int forward_grad_dt<EGGMAN_I128>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
#endif
#if EGG_PART > 0
template int forward_grad_dt<EGG_PART>(DevState &, DevPlan &, const void *, int64_t, int64_t, void *, void *, int64_t, hipStream_t, const Tune &);
#endif

static int forward_grad_any(int dtype, DevState &ds, DevPlan &pl, const void *P, int64_t ldp,
                            int64_t B, void *out, void *dP, int64_t lddp, hipStream_t st,
                            const Tune &tune) {
    switch (dtype) {
    case EGGMAN_F64: return forward_grad_dt<EGGMAN_F64>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    case EGGMAN_F32: return forward_grad_dt<EGGMAN_F32>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    case EGGMAN_C64: return forward_grad_dt<EGGMAN_C64>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    case EGGMAN_C32: return forward_grad_dt<EGGMAN_C32>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    case EGGMAN_I128: return forward_grad_dt<EGGMAN_I128>(ds, pl, P, ldp, B, out, dP, lddp, st, tune);
    }
    return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
}

// Instances per pass of the gradient: the forward's tuned pass raised to at least
// grad_min_pass, within the batch and the memory cap.  The reverse kernels spread a level over
// (task, 64-instance tile) pairs and want eight or more tiles to fill the device (512
// instances); small degrees want more, about 2^27 state values per pass, since each launch then
// does too little work: 2048 instances at N=24 take 8.4 ms per 4096 gradients where 512 take
// 9.3 ms, while N=28 is fastest at 512 (16.9 ms per 1024 against 17.3 ms at 1024).
constexpr int64_t kGradMinPass = 512;
static int64_t grad_min_pass(int64_t nstates) {
    int64_t p = kGradMinPass;
    while (p * nstates < (int64_t(1) << 27)) p *= 2;
    return p;
}
static int64_t grad_pass(const Tune &t, int64_t nstates, int64_t B, int64_t cap) {
    const int64_t c = std::max<int64_t>(t.chunk > 0 ? t.chunk : B, grad_min_pass(nstates));
    return std::max<int64_t>(1, std::min(std::min(c, B), cap));
}

// Instances whose gradient buffers fit in 80% of the free device memory (0 if not one does).
static int grad_fit(const DevPlan &pl, int dtype, bool ckpt, int64_t *fit) {
    size_t free_b = 0, total_b = 0;
    EGG_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per = grad_bytes_per_instance(pl, dtype, ckpt);
    int64_t f = (int64_t)((double)free_b * 0.8 / (double)per);
    if (f >= 64) f -= f % 64;
    *fit = std::max<int64_t>(f, 0);
    return EGGMAN_OK;
}

// The gradient's launch choices for up to `want` instances: the forward's tuned choices
// (waves per workgroup, LDS staging); checkpointed levels when the tuning entry asks for them,
// or when keeping every level would hold a pass below min(want, kGradMinPass) instances and
// checkpoints allow a larger one; `cap` = the instances one call may hold in memory, and the
// pass size within it.
static int grad_tune(const DevPlan &pl, int dtype, int64_t want, Tune *tune, int64_t *cap) {
    Tune t = pick_tune(pl, dtype, std::max<int64_t>(want, 1));
    const int64_t fwd_pass = t.chunk;
    int64_t fit;
    EGG_TRY(grad_fit(pl, dtype, t.ckpt, &fit));
    if (!t.ckpt && fit < std::min(want, kGradMinPass)) {
        int64_t fit_ck;
        EGG_TRY(grad_fit(pl, dtype, true, &fit_ck));
        if (fit_ck > fit) {
            t.ckpt = true;
            fit = fit_ck;
        }
    }
    if (fit < 1)
        return fail(EGGMAN_ENOMEM, "one instance of the degree-%d gradient needs %.2f GB (%.2f GB "
                    "with checkpointed levels); not enough free device memory", pl.lay.K,
                    grad_bytes_per_instance(pl, dtype, false) / 1e9,
                    grad_bytes_per_instance(pl, dtype, true) / 1e9);
    *cap = std::min(want, fit);
    t.chunk = grad_pass(t, pl.lay.nstates, want, *cap);
    t.sub = fwd_pass > 0 && fwd_pass < t.chunk ? fwd_pass : 0;
    // the forward part walks sub-passes of the forward's tuned size (t.sub); narrow tiles pay off
    // when a walk is at most 8 tiles of 16 (one per XCD), so a larger walk takes the 64-wide
    // kernel with 16 waves per workgroup, 2 per CU instead.  The 16-byte-load kernel keeps its
    // choice.
    if (t.narrow && !t.vec && t.chunk > 128 && (t.sub == 0 || t.sub > 128)) {
        t.narrow = false;
        t.wide = true;
        t.occ = 2;
    }
    *tune = t;
    return EGGMAN_OK;
}

// ------------------------------------------------------------------------------ plans

static int register_plan(std::unique_ptr<DevPlan> p, uint64_t *handle) {
    p->id = g_next_id++;
    const uint64_t id = p->id;
    g_plan_key[std::make_tuple(p->device, p->lay.K, p->lay.index_bytes)] = id;
    g_plans[id] = std::move(p);
    *handle = id;
    return EGGMAN_OK;
}

static bool cached_plan(int device, int K, int ib, uint64_t *handle) {
    auto it = g_plan_key.find(std::make_tuple(device, K, ib));
    if (it == g_plan_key.end()) return false;
    *handle = it->second;
    return true;
}

static int alloc_plan(DevPlan &p) {
    p.bytes = (size_t)p.lay.ntrans * p.lay.index_bytes;
    hipError_t e = hipMalloc(&p.d_trans, std::max<size_t>(p.bytes, 16));
    if (e != hipSuccess) {
        p.d_trans = nullptr;
        (void)hipGetLastError();
        return fail(EGGMAN_ENOMEM, "device allocation of the degree-%d plan (%.3f GB) failed: %s",
                    p.lay.K, p.bytes / 1e9, hipGetErrorString(e));
    }
    return EGGMAN_OK;
}

// ------------------------------------------------------------------------------ device gather

template <typename PT>
void launch_gather(const void *A, int64_t N, const int32_t *idx, int K, int64_t B, void *P,
                          int64_t ldp, hipStream_t st) {
    const int64_t total = (int64_t)K * (K - 1) / 2 * B;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (total + 255) / 256));
    k_gather_pairs<PT><<<dim3((unsigned)blocks), dim3(256), 0, st>>>((const PT *)A, N, idx, K, B,
                                                                      (PT *)P, ldp);
}

// One part of the library instantiates this for its element type; the others use that.
#if EGG_PART >= 0
extern template void launch_gather<float>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
extern template void launch_gather<double>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
extern template void launch_gather<float2>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
extern template void launch_gather<double2>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
extern template void launch_gather<long long>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
#endif
#if EGG_PART > 0
template void launch_gather<typename Elt<EGG_PART>::P>(const void *, int64_t, const int32_t *, int, int64_t, void *, int64_t, hipStream_t);
#endif

static void gather_any(int dtype, const void *A, int64_t N, const int32_t *idx, int K, int64_t B,
                       void *P, int64_t ldp, hipStream_t st) {
    switch (dtype) {
    case EGGMAN_F32: launch_gather<float>(A, N, idx, K, B, P, ldp, st); break;
    case EGGMAN_F64: launch_gather<double>(A, N, idx, K, B, P, ldp, st); break;
    case EGGMAN_C32: launch_gather<float2>(A, N, idx, K, B, P, ldp, st); break;
    case EGGMAN_C64: launch_gather<double2>(A, N, idx, K, B, P, ldp, st); break;
    case EGGMAN_I128: launch_gather<long long>(A, N, idx, K, B, P, ldp, st); break;
    }
}

}  // namespace eggman

using namespace eggman;

// ================================================================================ C ABI

#if EGG_PART <= 0
extern "C" {

const char *eggman_version(void) { return EGGMAN_VERSION_STRING; }

const char *eggman_last_error(void) { return t_err.c_str(); }

int eggman_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int eggman_dp_counts(int K, int64_t *nstates, int64_t *ntrans) {
    if (!nstates || !ntrans) return fail(EGGMAN_EINVAL, "null output pointer");
    PlanLayout lay;
    // counts exist for layouts that cannot be packed too, so only the degree is checked here
    int rc = plan_layout(K, 8, &lay);
    if (rc != EGGMAN_OK && rc != EGGMAN_EPLAN) return rc;
    if (rc == EGGMAN_EPLAN) t_err.clear();
    *nstates = lay.nstates;
    *ntrans = lay.ntrans;
    return EGGMAN_OK;
}

int eggman_plan_build(int K, int index_bytes, void *starts, void *trans, void *levels,
                      int nthreads) {
    if (!starts || !trans || !levels) return fail(EGGMAN_EINVAL, "null plan array");
    PlanLayout lay;
    EGG_TRY(plan_layout(K, index_bytes, &lay));
    return build_plan_host(lay, starts, trans, levels, nthreads);
}

static int check_plan_arrays(const PlanLayout &lay, const void *starts, int64_t nstarts,
                             const void *trans, int64_t ntrans, const void *levels,
                             int64_t nlevels) {
    if (!starts || !trans || !levels) return fail(EGGMAN_EINVAL, "null plan array");
    if (nstarts != lay.nstates + 1 || ntrans != lay.ntrans || nlevels != lay.nlevels)
        return fail(EGGMAN_EPLAN,
                    "plan arrays have %lld/%lld/%lld entries, degree %d needs %lld/%lld/%d",
                    (long long)nstarts, (long long)ntrans, (long long)nlevels, lay.K,
                    (long long)(lay.nstates + 1), (long long)lay.ntrans, lay.nlevels);
    return validate_plan(lay, starts, trans, levels, 0);
}

int eggman_plan_validate(int K, int index_bytes, const void *starts, int64_t nstarts,
                         const void *trans, int64_t ntrans, const void *levels, int64_t nlevels) {
    PlanLayout lay;
    EGG_TRY(plan_layout(K, index_bytes, &lay));
    return check_plan_arrays(lay, starts, nstarts, trans, ntrans, levels, nlevels);
}

int eggman_plan_upload(int K, int index_bytes, const void *starts, int64_t nstarts,
                       const void *trans, int64_t ntrans, const void *levels, int64_t nlevels,
                       uint64_t *handle) {
    if (!handle) return fail(EGGMAN_EINVAL, "null handle pointer");
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int dev;
    EGG_TRY(current_device(&dev));
    if (cached_plan(dev, K, index_bytes, handle)) return EGGMAN_OK;
    auto p = std::make_unique<DevPlan>();
    p->device = dev;
    EGG_TRY(plan_layout(K, index_bytes, &p->lay));
    EGG_TRY(check_plan_arrays(p->lay, starts, nstarts, trans, ntrans, levels, nlevels));
    EGG_TRY(alloc_plan(*p));
    hipError_t e = hipMemcpy(p->d_trans, trans, p->bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p->d_trans);
        return fail(EGGMAN_EHIP, "plan upload failed: %s", hipGetErrorString(e));
    }
    return register_plan(std::move(p), handle);
}

int eggman_plan_generate(int K, int index_bytes, uint64_t *handle) {
    if (!handle) return fail(EGGMAN_EINVAL, "null handle pointer");
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int dev;
    EGG_TRY(current_device(&dev));
    if (cached_plan(dev, K, index_bytes, handle)) return EGGMAN_OK;
    auto p = std::make_unique<DevPlan>();
    p->device = dev;
    EGG_TRY(plan_layout(K, index_bytes, &p->lay));
    DevState *ds;
    EGG_TRY(dev_state(dev, &ds));
    EGG_TRY(alloc_plan(*p));
    const PlanLayout &lay = p->lay;
    for (int L = 2; L <= K / 2 + 1; L++) {
        const int64_t nst = lay.nst[L - 1];
        const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(kMaxBlocks, (nst + 255) / 256));
        if (index_bytes == 4)
            k_plan_gen<int32_t><<<dim3((unsigned)blocks), dim3(256), 0, ds->stream>>>(
                (int32_t *)p->d_trans + lay.tstart[L - 1], nst, K, lay.size[L - 1], lay.lo[L - 1],
                lay.levels[L - 2], lay.bits);
        else
            k_plan_gen<int64_t><<<dim3((unsigned)blocks), dim3(256), 0, ds->stream>>>(
                (int64_t *)p->d_trans + lay.tstart[L - 1], nst, K, lay.size[L - 1], lay.lo[L - 1],
                lay.levels[L - 2], lay.bits);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ds->stream);
    if (e != hipSuccess) {
        (void)hipFree(p->d_trans);
        return fail(EGGMAN_EHIP, "device plan generation failed: %s", hipGetErrorString(e));
    }
    return register_plan(std::move(p), handle);
}

int eggman_plan_info(uint64_t handle, int *K, int *index_bytes, int64_t *nstates,
                     int64_t *ntrans) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(handle, &pl));
    if (K) *K = pl->lay.K;
    if (index_bytes) *index_bytes = pl->lay.index_bytes;
    if (nstates) *nstates = pl->lay.nstates;
    if (ntrans) *ntrans = pl->lay.ntrans;
    return EGGMAN_OK;
}

int eggman_plan_download(uint64_t handle, void *trans_host, int64_t ntrans) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(handle, &pl));
    if (!trans_host || ntrans != pl->lay.ntrans)
        return fail(EGGMAN_EINVAL, "download needs a buffer of %lld words",
                    (long long)pl->lay.ntrans);
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    EGG_HIP(hipDeviceSynchronize());
    EGG_HIP(hipMemcpy(trans_host, pl->d_trans, pl->bytes, hipMemcpyDeviceToHost));
    return EGGMAN_OK;
}

int eggman_dp_forward(uint64_t plan, int dtype, const void *P_host, int64_t B, int64_t npairs,
                      int64_t max_batch, void *out_host) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    const int hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    if (!hb) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 0) return fail(EGGMAN_EINVAL, "negative batch size");
    if (npairs != pl->lay.npairs)
        return fail(EGGMAN_EINVAL, "P has %lld pair rows, the degree-%d plan expects %lld",
                    (long long)npairs, pl->lay.K, (long long)pl->lay.npairs);
    if (B == 0) return EGGMAN_OK;
    if (!P_host || !out_host) return fail(EGGMAN_EINVAL, "null P or out pointer");
    if (dtype == EGGMAN_I128) EGG_TRY(check_exact_host(pl->lay.K, (const int64_t *)P_host, B, B));
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    const Tune t0 = pick_tune(*pl, dtype, B);
    int64_t chunk;
    EGG_TRY(host_chunk(B, max_batch, t0.chunk > 0 ? t0.chunk : B, (size_t)npairs * pb + hb,
                       pipe_parts(pl->lay), &chunk));
    const Tune tune = pick_tune(*pl, dtype, chunk);
    const int64_t nchunks = (B + chunk - 1) / chunk;
    auto span = [&](int64_t c, int64_t *b0, int64_t *nb) {
        *b0 = c * chunk;
        *nb = std::min(chunk, B - *b0);
    };
    return run_pipe(
        *ds, nchunks, (size_t)chunk * npairs * pb, (size_t)chunk * hb,
        [&](int64_t c, void *pin, size_t *ib, size_t *ob) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            // this chunk's pairs, relaid out to [pair][nb] (ext/TheEggmanKernelAbstractionsExt.jl:127-129)
            copy_columns(pin, P_host, npairs, B, b0, nb, pb);
            *ib = (size_t)nb * npairs * pb;
            *ob = (size_t)nb * hb;
            return EGGMAN_OK;
        },
        [&](int64_t c, void *din, void *dout, hipStream_t st) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            return forward_any(dtype, *ds, *pl, din, nb, nb, dout, st, tune);
        },
        [&](int64_t c, const void *pin) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            std::memcpy((char *)out_host + (size_t)b0 * hb, pin, (size_t)nb * hb);
            return EGGMAN_OK;
        });
}

int eggman_dp_forward_device(uint64_t plan, int dtype, const void *P_dev, int64_t ldp, int64_t B,
                             void *out_dev, void *stream) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    if (!elt_hbytes(dtype)) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 0 || ldp < B) return fail(EGGMAN_EINVAL, "need 0 <= B <= ldp (B=%lld, ldp=%lld)",
                                      (long long)B, (long long)ldp);
    if (B == 0) return EGGMAN_OK;
    if (!P_dev || !out_dev) return fail(EGGMAN_EINVAL, "null P or out pointer");
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    hipStream_t st = stream ? (hipStream_t)stream : ds->stream;
    EGG_TRY(chain_stream(*ds, st));
    if (dtype == EGGMAN_I128) EGG_TRY(check_exact_device(*ds, pl->lay.K, P_dev, ldp, B, st));
    const Tune tune = pick_tune(*pl, dtype, B);
    return forward_any(dtype, *ds, *pl, P_dev, ldp, B, out_dev, st, tune);
}

int eggman_dp_forward_device_noalloc(uint64_t plan, int dtype, const void *P_dev, int64_t ldp,
                                     int64_t B, void *out_dev, void *stream) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    if (!elt_hbytes(dtype)) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (dtype == EGGMAN_I128)
        return fail(EGGMAN_EUNSUPPORTED, "the exact walk checks its range through host memory and "
                                         "a synchronisation; the no-allocation entry point serves "
                                         "the floating-point types (use eggman_dp_forward_device)");
    if (B < 0 || ldp < B) return fail(EGGMAN_EINVAL, "need 0 <= B <= ldp (B=%lld, ldp=%lld)",
                                      (long long)B, (long long)ldp);
    if (B == 0) return EGGMAN_OK;
    if (!P_dev || !out_dev) return fail(EGGMAN_EINVAL, "null P or out pointer");
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    void *sp;
    EGG_TRY(scratch_address(*ds, &sp));
    Tune tune;
    EGG_TRY(noalloc_tune(*ds, *pl, dtype, B, &tune));
    hipStream_t st = stream ? (hipStream_t)stream : ds->stream;
    EGG_TRY(chain_stream(*ds, st));
    const WorkSpan ws{(char *)sp, kScratchBytes};
    return forward_eager(dtype, *ds, *pl, P_dev, ldp, B, out_dev, st, tune, &ws);
}

// Time every launch choice for (plan, dtype, B) on synthetic pair values and keep the fastest.
// All choices give bit-identical results, so tuning never changes an answer.
int eggman_autotune(uint64_t plan, int dtype, int64_t B, int reps, char *report,
                    int64_t report_len) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    const int hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    if (!hb) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 1) return fail(EGGMAN_EINVAL, "batch must be positive");
    if (reps < 1) reps = 3;
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    EGG_TRY(chain_stream(*ds, ds->stream));
    std::vector<Tune> cands;
    if (B >= 64) {
        // passes of B / 1 .. B / 16 instances, and of 64, 128 and 256 (the level kernels' own
        // sweet spots, whatever the batch)
        std::vector<int64_t> chunks;
        for (int64_t d : {1, 2, 4, 8, 16}) {
            int64_t c = (B / d + 63) / 64 * 64;
            if (c >= 64 && std::find(chunks.begin(), chunks.end(), c) == chunks.end()) chunks.push_back(c);
        }
        for (int64_t c : {64, 128, 256})
            if (c <= B && std::find(chunks.begin(), chunks.end(), c) == chunks.end()) chunks.push_back(c);
        for (int64_t c : chunks) {
            for (int occ : {1, 2}) {
                Tune t;
                t.chunk = c;
                t.occ = occ;
                cands.push_back(t);
                t.wide = true;
                cands.push_back(t);
            }
            for (int occ : {1, 2, 4}) {
                Tune t;
                t.chunk = c;
                t.occ = occ;
                t.narrow = true;
                cands.push_back(t);
                t.wide = true;
                cands.push_back(t);
            }
            for (int occ : {1, 2, 4})
                for (bool pairs : {false, true}) {
                    Tune t;
                    t.chunk = c;
                    t.occ = occ;
                    t.vec = true;
                    t.pairs = pairs;
                    cands.push_back(t);
                    t.wide = true;
                    cands.push_back(t);
                    t.wide = false;
                    t.deep = true;
                    cands.push_back(t);
                    t.wide = true;
                    cands.push_back(t);
                }
            Tune g;
            g.chunk = c;
            g.lds = false;
            cands.push_back(g);
        }
    } else {
        Tune t;
        t.bt = pow2_at_least(B);
        t.chunk = B;
        t.lds = false;
        cands.push_back(t);
        if (t.bt != 64) {
            t.bt = 64;
            cands.push_back(t);
        }
        t.lds = true;
        cands.push_back(t);
    }
    if (fused_lds_bytes(pl->lay, dtype) <= ds->caps.max_lds_block) {
        Tune f;
        f.fused = true;
        f.chunk = B;
        cands.push_back(f);
    }
    DevBuf Pd, Od;
    EGG_TRY(ensure(Pd, (size_t)B * pl->lay.npairs * pb));
    EGG_TRY(ensure(Od, (size_t)B * hb));
    // pair values: the bit pattern 0x3f in every byte, a moderate real or integer value
    EGG_HIP(hipMemsetAsync(Pd.p, 0x3f, Pd.bytes, ds->stream));
    hipEvent_t e0, e1;
    EGG_HIP(hipEventCreate(&e0));
    EGG_HIP(hipEventCreate(&e1));
    std::string rep;
    float best = 1e30f;
    Tune best_t = cands[0];
    int rc = EGGMAN_OK;
    for (size_t ci = 0; ci < cands.size(); ci++) {
        const Tune t = cands[ci];
        // warm-up (sizes the buffers; a graph candidate captures here and replays below)
        rc = forward_any(dtype, *ds, *pl, Pd.p, B, B, Od.p, ds->stream, t);
        if (rc == EGGMAN_OK && t.graph) rc = forward_any(dtype, *ds, *pl, Pd.p, B, B, Od.p, ds->stream, t);
        if (rc != EGGMAN_OK) break;
        if (hipEventRecord(e0, ds->stream) != hipSuccess) { rc = fail(EGGMAN_EHIP, "event record"); break; }
        for (int r = 0; r < reps && rc == EGGMAN_OK; r++)
            rc = forward_any(dtype, *ds, *pl, Pd.p, B, B, Od.p, ds->stream, t);
        if (rc != EGGMAN_OK) break;
        float ms = 0;
        if (hipEventRecord(e1, ds->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(EGGMAN_EHIP, "event timing failed");
            break;
        }
        ms /= reps;
        char line[160];
        snprintf(line, sizeof line, "bt=%d chunk=%lld flags=%d occ=%d ms=%.4f\n", t.bt,
                 (long long)t.chunk, tune_flags(t), t.occ, ms);
        rep += line;
        if (ms < best) {
            best = ms;
            best_t = t;
        }
        // after the eager candidates, try the best of them replayed from a graph
        if (ci + 1 == cands.size() && !best_t.graph && !t.graph) {
            Tune gt = best_t;
            gt.graph = true;
            cands.push_back(gt);
        }
    }
    (void)hipStreamSynchronize(ds->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(Pd.p);
    (void)hipFree(Od.p);
    if (rc != EGGMAN_OK) return rc;
    g_tune[TuneKey(pl->device, pl->lay.K, pl->lay.index_bytes, dtype, batch_bucket(B))] = best_t;
    char line[160];
    snprintf(line, sizeof line, "best bt=%d chunk=%lld flags=%d occ=%d ms=%.4f\n", best_t.bt,
             (long long)best_t.chunk, tune_flags(best_t), best_t.occ, best);
    rep += line;
    if (report && report_len > 0) {
        const size_t n = std::min<size_t>(rep.size(), (size_t)report_len - 1);
        std::memcpy(report, rep.data(), n);
        report[n] = 0;
    }
    return EGGMAN_OK;
}

int eggman_tune_set(int K, int index_bytes, int dtype, int64_t B, int bt, int64_t chunk,
                    int flags, int occ) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (!elt_hbytes(dtype) || B < 1 || (bt & (bt - 1)) || bt < 1 || bt > 64 || occ < 0 ||
        (flags & ~TUNE_ALL))
        return fail(EGGMAN_EINVAL, "bad tuning entry");
    int dev;
    EGG_TRY(current_device(&dev));
    g_tune[TuneKey(dev, K, index_bytes, dtype, batch_bucket(B))] = tune_from_flags(bt, chunk, flags, occ);
    return EGGMAN_OK;
}

int eggman_tune_get(int K, int index_bytes, int dtype, int64_t B, int *bt, int64_t *chunk,
                    int *flags, int *occ) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int dev;
    EGG_TRY(current_device(&dev));
    auto it = g_tune.find(TuneKey(dev, K, index_bytes, dtype, batch_bucket(B)));
    if (it == g_tune.end()) return fail(EGGMAN_EINVAL, "no tuning entry");
    if (bt) *bt = it->second.bt;
    if (chunk) *chunk = it->second.chunk;
    if (flags) *flags = tune_flags(it->second);
    if (occ) *occ = it->second.occ;
    return EGGMAN_OK;
}

int eggman_dp_forward_repeated(uint64_t plan, int dtype, const void *A_host, int64_t N,
                               const int32_t *idx_host, int64_t B, int64_t max_batch,
                               void *out_host) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    const int hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    if (!hb) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 0 || N < 1) return fail(EGGMAN_EINVAL, "need N >= 1 and B >= 0");
    if (B == 0) return EGGMAN_OK;
    if (!A_host || !idx_host || !out_host) return fail(EGGMAN_EINVAL, "null A, idx or out pointer");
    const int K = pl->lay.K;
    for (int64_t t = 0; t < B * K; t++)
        if (idx_host[t] < 1 || idx_host[t] > N)
            return fail(EGGMAN_EINVAL, "index list %lld holds %d, outside 1..%lld",
                        (long long)(t / K + 1), idx_host[t], (long long)N);
    if (dtype == EGGMAN_I128) EGG_TRY(check_exact_repeated(K, (const int64_t *)A_host, N, idx_host, B));
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    EGG_TRY(chain_stream(*ds, ds->stream));
    const int64_t npairs = pl->lay.npairs;
    const Tune t0 = pick_tune(*pl, dtype, B);
    int64_t chunk;
    EGG_TRY(host_chunk(B, max_batch, t0.chunk > 0 ? t0.chunk : B,
                       (size_t)K * sizeof(int32_t) + hb + (size_t)npairs * pb / 2,
                       pipe_parts(pl->lay), &chunk));
    const Tune tune = pick_tune(*pl, dtype, chunk);
    const int64_t nchunks = (B + chunk - 1) / chunk;
    // the matrix once, then the pair values of each chunk gathered on the device into one buffer
    // (chunks run one after another on the compute stream)
    const size_t aoff = ((size_t)N * N * pb + 255) & ~(size_t)255;
    EGG_TRY(ensure(ds->A, aoff + (size_t)chunk * npairs * pb));
    char *Pg = (char *)ds->A.p + aoff;
    EGG_HIP(hipMemcpyAsync(ds->A.p, A_host, (size_t)N * N * pb, hipMemcpyHostToDevice, ds->stream));
    auto span = [&](int64_t c, int64_t *b0, int64_t *nb) {
        *b0 = c * chunk;
        *nb = std::min(chunk, B - *b0);
    };
    return run_pipe(
        *ds, nchunks, (size_t)chunk * K * sizeof(int32_t), (size_t)chunk * hb,
        [&](int64_t c, void *pin, size_t *ib, size_t *ob) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            std::memcpy(pin, idx_host + b0 * K, (size_t)nb * K * sizeof(int32_t));
            *ib = (size_t)nb * K * sizeof(int32_t);
            *ob = (size_t)nb * hb;
            return EGGMAN_OK;
        },
        [&](int64_t c, void *din, void *dout, hipStream_t st) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            gather_any(dtype, ds->A.p, N, (const int32_t *)din, K, nb, Pg, nb, st);
            return forward_any(dtype, *ds, *pl, Pg, nb, nb, dout, st, tune);
        },
        [&](int64_t c, const void *pin) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            std::memcpy((char *)out_host + (size_t)b0 * hb, pin, (size_t)nb * hb);
            return EGGMAN_OK;
        });
}

int eggman_dp_forward_mixed(int ngroups, const uint64_t *plans, const int64_t *counts, int dtype,
                            const void *P_host, int64_t max_batch, void *out_host) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    std::vector<DevPlan *> pls;
    std::vector<int64_t> cnt;
    int64_t np_total, total;
    EGG_TRY(mixed_plans(ngroups, plans, counts, &pls, &cnt, &np_total, &total));
    const int hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    if (!hb) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (total == 0) return EGGMAN_OK;
    if (!P_host || !out_host) return fail(EGGMAN_EINVAL, "null P or out pointer");
    const size_t ng = pls.size();
    std::vector<int64_t> poff(ng), ooff(ng);   // each group's block in P_host and out_host
    for (size_t g = 0, po = 0, oo = 0; g < ng; g++) {
        poff[g] = (int64_t)po;
        ooff[g] = (int64_t)oo;
        po += (size_t)pls[g]->lay.npairs * cnt[g];
        oo += (size_t)cnt[g];
    }
    if (dtype == EGGMAN_I128)
        for (size_t g = 0; g < ng; g++)
            EGG_TRY(check_exact_host(pls[g]->lay.K, (const int64_t *)P_host + poff[g], cnt[g], cnt[g],
                                     ooff[g]));
    DeviceGuard g(pls[0]->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pls[0]->device, &ds));
    // Chunks take the same share of every group (chunk c: instances [c*n/nc, (c+1)*n/nc) of a
    // group of n), so each keeps all degrees together: enough chunks to keep two slots of staging
    // within half the free device memory and one slot's pinned staging within kPipeSlotBytes, and
    // at most max_batch instances of any group in a chunk.  Every instance's value is the same
    // bits however it is chunked.
    const size_t staged = (size_t)np_total * pb + (size_t)total * hb;
    size_t free_b = 0, total_b = 0;
    EGG_HIP(hipMemGetInfo(&free_b, &total_b));
    int64_t maxcnt = 0;
    for (int64_t c : cnt) maxcnt = std::max(maxcnt, c);
    int64_t nchunks = 1;
    nchunks = std::max<int64_t>(nchunks, (int64_t)((4 * staged + free_b - 1) / std::max<size_t>(free_b, 1)));
    nchunks = std::max<int64_t>(nchunks, (int64_t)((staged + kPipeSlotBytes - 1) / kPipeSlotBytes));
    nchunks = std::min(nchunks, maxcnt);
    if (max_batch > 0) nchunks = std::max<int64_t>(nchunks, (maxcnt + max_batch - 1) / max_batch);
    auto share = [&](int64_t c, size_t gi, int64_t *lo, int64_t *n) {
        *lo = c * cnt[gi] / nchunks;
        *n = (c + 1) * cnt[gi] / nchunks - *lo;
    };
    size_t in_max = 0, out_max = 0;
    for (int64_t c = 0; c < nchunks; c++) {
        size_t ib = 0, ob = 0;
        for (size_t gi = 0; gi < ng; gi++) {
            int64_t lo, n;
            share(c, gi, &lo, &n);
            ib += (size_t)pls[gi]->lay.npairs * n * pb;
            ob += (size_t)n * hb;
        }
        in_max = std::max(in_max, ib);
        out_max = std::max(out_max, ob);
    }
    return run_pipe(
        *ds, nchunks, in_max, out_max,
        [&](int64_t c, void *pin, size_t *ib, size_t *ob) {
            size_t o = 0, r = 0;
            for (size_t gi = 0; gi < ng; gi++) {
                int64_t lo, n;
                share(c, gi, &lo, &n);
                copy_columns((char *)pin + o, (const char *)P_host + (size_t)poff[gi] * pb,
                             pls[gi]->lay.npairs, cnt[gi], lo, n, pb);
                o += (size_t)pls[gi]->lay.npairs * n * pb;
                r += (size_t)n * hb;
            }
            *ib = o;
            *ob = r;
            return EGGMAN_OK;
        },
        [&](int64_t c, void *din, void *dout, hipStream_t st) {
            std::vector<int64_t> nc(ng);
            for (size_t gi = 0; gi < ng; gi++) {
                int64_t lo;
                share(c, gi, &lo, &nc[gi]);
            }
            return forward_mixed_any(dtype, *ds, pls, nc, din, dout, st);
        },
        [&](int64_t c, const void *pin) {
            size_t r = 0;
            for (size_t gi = 0; gi < ng; gi++) {
                int64_t lo, n;
                share(c, gi, &lo, &n);
                std::memcpy((char *)out_host + (size_t)(ooff[gi] + lo) * hb, (const char *)pin + r,
                            (size_t)n * hb);
                r += (size_t)n * hb;
            }
            return EGGMAN_OK;
        });
}

int eggman_dp_forward_mixed_device(int ngroups, const uint64_t *plans, const int64_t *counts,
                                   int dtype, const void *P_dev, void *out_dev, void *stream) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    std::vector<DevPlan *> pls;
    std::vector<int64_t> cnt;
    int64_t np_total, total;
    EGG_TRY(mixed_plans(ngroups, plans, counts, &pls, &cnt, &np_total, &total));
    if (!elt_hbytes(dtype)) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (total == 0) return EGGMAN_OK;
    if (!P_dev || !out_dev) return fail(EGGMAN_EINVAL, "null P or out pointer");
    DeviceGuard g(pls[0]->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pls[0]->device, &ds));
    hipStream_t st = stream ? (hipStream_t)stream : ds->stream;
    EGG_TRY(chain_stream(*ds, st));
    if (dtype == EGGMAN_I128) {
        int64_t off = 0, first = 0;
        for (size_t gi = 0; gi < pls.size(); gi++) {
            EGG_TRY(check_exact_device(*ds, pls[gi]->lay.K, (const int64_t *)P_dev + off, cnt[gi],
                                       cnt[gi], st, first));
            off += pls[gi]->lay.npairs * cnt[gi];
            first += cnt[gi];
        }
    }
    return forward_mixed_any(dtype, *ds, pls, cnt, P_dev, out_dev, st);
}

int eggman_dp_forward_grad(uint64_t plan, int dtype, const void *P_host, int64_t B,
                           int64_t npairs, int64_t max_batch, void *out_host, void *dP_host) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    const int hb = elt_hbytes(dtype), pb = elt_pbytes(dtype);
    if (!hb) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 0) return fail(EGGMAN_EINVAL, "negative batch size");
    if (npairs != pl->lay.npairs)
        return fail(EGGMAN_EINVAL, "P has %lld pair rows, the degree-%d plan expects %lld",
                    (long long)npairs, pl->lay.K, (long long)pl->lay.npairs);
    if (B == 0) return EGGMAN_OK;
    if (!P_host || !out_host || !dP_host) return fail(EGGMAN_EINVAL, "null P, out or dP pointer");
    if (dtype == EGGMAN_I128) EGG_TRY(check_exact_host(pl->lay.K, (const int64_t *)P_host, B, B));
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    EGG_TRY(ensure_rev(*pl));
    Tune tune;
    int64_t cap, chunk;
    EGG_TRY(grad_tune(*pl, dtype, max_batch > 0 ? std::min(max_batch, B) : B, &tune, &cap));
    EGG_TRY(host_chunk(B, max_batch, tune.chunk, (size_t)npairs * (pb + hb) + hb, 2, &chunk));
    chunk = std::min(chunk, cap);
    const int64_t nchunks = (B + chunk - 1) / chunk;
    auto span = [&](int64_t c, int64_t *b0, int64_t *nb) {
        *b0 = c * chunk;
        *nb = std::min(chunk, B - *b0);
    };
    // a chunk's results: its nb values, then its gradient rows [pair][nb]
    return run_pipe(
        *ds, nchunks, (size_t)chunk * npairs * pb, (size_t)chunk * (npairs + 1) * hb,
        [&](int64_t c, void *pin, size_t *ib, size_t *ob) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            copy_columns(pin, P_host, npairs, B, b0, nb, pb);
            *ib = (size_t)nb * npairs * pb;
            *ob = (size_t)nb * (npairs + 1) * hb;
            return EGGMAN_OK;
        },
        [&](int64_t c, void *din, void *dout, hipStream_t st) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            return forward_grad_any(dtype, *ds, *pl, din, nb, nb, dout,
                                    (char *)dout + (size_t)nb * hb, nb, st, tune);
        },
        [&](int64_t c, const void *pin) {
            int64_t b0, nb;
            span(c, &b0, &nb);
            std::memcpy((char *)out_host + (size_t)b0 * hb, pin, (size_t)nb * hb);
            scatter_columns(dP_host, (const char *)pin + (size_t)nb * hb, npairs, B, b0, nb, hb);
            return EGGMAN_OK;
        });
}

int eggman_dp_forward_grad_device(uint64_t plan, int dtype, const void *P_dev, int64_t ldp,
                                  int64_t B, void *out_dev, void *dP_dev, int64_t lddp,
                                  void *stream) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DevPlan *pl;
    EGG_TRY(find_plan(plan, &pl));
    if (!elt_hbytes(dtype)) return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
    if (B < 0 || ldp < B || lddp < B)
        return fail(EGGMAN_EINVAL, "need 0 <= B <= ldp, lddp (B=%lld, ldp=%lld, lddp=%lld)",
                    (long long)B, (long long)ldp, (long long)lddp);
    if (B == 0) return EGGMAN_OK;
    if (!P_dev || !out_dev || !dP_dev) return fail(EGGMAN_EINVAL, "null P, out or dP pointer");
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    hipStream_t st = stream ? (hipStream_t)stream : ds->stream;
    EGG_TRY(chain_stream(*ds, st));
    if (dtype == EGGMAN_I128) EGG_TRY(check_exact_device(*ds, pl->lay.K, P_dev, ldp, B, st));
    EGG_TRY(ensure_rev(*pl));
    Tune tune;
    int64_t cap;
    EGG_TRY(grad_tune(*pl, dtype, B, &tune, &cap));
    return forward_grad_any(dtype, *ds, *pl, P_dev, ldp, B, out_dev, dP_dev, lddp, st, tune);
}

int eggman_exact_bound(int K, const int64_t *P, int64_t ldp, int64_t B, double *log2_bound) {
    if (K < 2 || K > 64 || (K & 1) || B < 0 || ldp < B)
        return fail(EGGMAN_EINVAL, "need an even degree 2..64 and 0 <= B <= ldp");
    if (B > 0 && !P) return fail(EGGMAN_EINVAL, "null P");
    int rc = EGGMAN_OK;
    for (int64_t b = 0; b < B; b++) {
        const double lb = exact_log2_bound(K, [&](int64_t q) { return P[q * ldp + b]; });
        if (log2_bound) log2_bound[b] = lb;
        if (lb >= kExactLog2Limit && rc == EGGMAN_OK) rc = exact_overflow(b, lb);
    }
    return rc;
}

int eggman_gather_rate(int64_t working_set_bytes, double *lines_per_s) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (working_set_bytes < 128 || !lines_per_s)
        return fail(EGGMAN_EINVAL, "need a working set of at least one 128-byte line");
    int dev;
    EGG_TRY(current_device(&dev));
    DevState *ds;
    EGG_TRY(dev_state(dev, &ds));
    EGG_TRY(chain_stream(*ds, ds->stream));
    const uint64_t nlines = (uint64_t)working_set_bytes / 128;
    DevBuf buf, sink;
    EGG_TRY(ensure(buf, (size_t)nlines * 128));
    EGG_TRY(ensure(sink, 64));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = EGGMAN_OK;
    float ms = 0;
    const int blocks = ds->caps.cus * 4, iters = 64, reps = 5;
    // zero the working set, then one warm-up launch
    if (hipMemsetAsync(buf.p, 0, buf.bytes, ds->stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(EGGMAN_EHIP, "gather probe setup failed");
    } else {
        k_gather_probe<<<dim3((unsigned)blocks), dim3(512), 0, ds->stream>>>(
            (const double2 *)buf.p, nlines, iters, 1ull, (double *)sink.p);
        (void)hipEventRecord(e0, ds->stream);
        for (int r = 0; r < reps; r++)
            k_gather_probe<<<dim3((unsigned)blocks), dim3(512), 0, ds->stream>>>(
                (const double2 *)buf.p, nlines, iters, (uint64_t)r + 2, (double *)sink.p);
        (void)hipEventRecord(e1, ds->stream);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            rc = fail(EGGMAN_EHIP, "gather probe timing failed");
    }
    (void)hipStreamSynchronize(ds->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf.p);
    (void)hipFree(sink.p);
    if (rc != EGGMAN_OK) return rc;
    *lines_per_s = (double)reps * blocks * 64.0 * 8.0 * iters / (ms / 1e3);
    return EGGMAN_OK;
}

int64_t eggman_cache_bytes(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int64_t total = 0;
    for (auto &kv : g_plans) total += (int64_t)kv.second->bytes;
    for (auto &kv : g_plans) total += (int64_t)kv.second->rev_bytes;
    for (auto &kv : g_plans) total += (int64_t)kv.second->items_bytes;
    for (auto &kv : g_dev) total += (int64_t)kv.second.bytes();
    return total;
}

int eggman_cache_clear(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int rc = EGGMAN_OK;
    for (auto &kv : g_dev) {
        DeviceGuard g(kv.first);
        if (hipDeviceSynchronize() != hipSuccess) rc = fail(EGGMAN_EHIP, "synchronize failed");
        for (GraphEntry &ge : kv.second.graphs) (void)hipGraphExecDestroy(ge.exec);
        kv.second.graphs.clear();
        DevState &d = kv.second;
        for (DevBuf *b : {&d.H, &d.Hw, &d.Hg, &d.G, &d.part, &d.A, &d.bnd, &d.Hm,
                          &d.din[0], &d.din[1], &d.dout[0], &d.dout[1]}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->bytes = 0;
        }
        for (PinBuf *b : {&d.pin_in[0], &d.pin_in[1], &d.pin_out[0], &d.pin_out[1]}) {
            if (b->p) (void)hipHostFree(b->p);
            b->p = nullptr;
            b->bytes = 0;
        }
    }
    for (auto &kv : g_plans) {
        DeviceGuard g(kv.second->device);
        DevPlan &p = *kv.second;
        for (void *ptr : {p.d_trans, p.d_cent, p.d_coff, p.d_pent, p.d_segs, p.d_qoff, p.d_qseg,
                          p.d_prank, p.d_tasks, p.d_roff, p.d_rtask, p.d_qrow, p.d_qcol,
                          p.d_red, p.d_srow0, p.d_sruns, p.d_fent, p.d_fco, p.d_cmask, p.d_fprow0,
                          p.d_fterms, p.d_ftoff, p.d_ftasks, p.d_frtask, p.d_fred, p.d_fsrow0, p.d_fsruns,
                          (void *)p.d_items})
            if (ptr) (void)hipFree(ptr);
    }
    g_plans.clear();
    g_plan_key.clear();
    return rc;
}

int64_t eggman_batch_bytes(int K, int dtype, int64_t B) {
    const int hb = elt_hbytes(dtype);
    if (!hb || B < 0) return fail(EGGMAN_EINVAL, "bad dtype or batch size");
    PlanLayout lay;
    int rc = plan_layout(K, 8, &lay);
    if (rc != EGGMAN_OK && rc != EGGMAN_EPLAN) return rc;
    return B * (lay.nstates + lay.npairs) * (int64_t)hb;
}

int eggman_grad_levels(int K, int checkpointed, int64_t *kept_rows, int64_t *seg_rows,
                       uint8_t *kept, int64_t nkept) {
    PlanLayout lay;
    int rc = plan_layout(K, 8, &lay);
    if (rc != EGGMAN_OK && rc != EGGMAN_EPLAN) return rc;
    GradLevels gl;
    if (checkpointed) grad_levels_checkpointed(lay, &gl);
    else grad_levels_full(lay, &gl);
    if (kept_rows) *kept_rows = gl.kept_rows;
    if (seg_rows) *seg_rows = gl.seg_rows;
    if (kept) {
        if (nkept < (int64_t)gl.kept.size())
            return fail(EGGMAN_EINVAL, "kept needs %lld entries", (long long)gl.kept.size());
        for (size_t L = 0; L < gl.kept.size(); L++) kept[L] = L == 0 ? 0 : (uint8_t)gl.kept[L];
    }
    return EGGMAN_OK;
}

int eggman_fuse_validate(int K, int64_t *entries_checked, int64_t *terms_checked) {
    if (K < 4 || K > 32) return fail(EGGMAN_EINVAL, "the fused reverse step's tables need 4 <= K <= 32, got %d", K);
    PlanLayout lay;
    EGG_TRY(plan_layout(K, 4, &lay));
    std::vector<int32_t> starts((size_t)lay.nstates + 1), trans((size_t)lay.ntrans), levels(lay.nlevels);
    EGG_TRY(build_plan_host(lay, starts.data(), trans.data(), levels.data(), 0));
    RevPlanHost rp;
    EGG_TRY(build_rev_plan(lay, trans.data(), kRevSegLen, kParTaskLen, kRedFanIn, &rp));
    FusePlanHost fp;
    EGG_TRY(build_fuse_plan(lay, rp, kRedFanIn, &fp));
    return validate_fuse_plan(lay, trans.data(), rp, fp, entries_checked, terms_checked);
}

}  // extern "C"
#endif

