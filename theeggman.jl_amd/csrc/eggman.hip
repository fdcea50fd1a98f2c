// eggman HIP library: device caches, level drivers and the C ABI of include/eggman_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "kernels_fwd.hpp"
#include "kernels_plan.hpp"
#include "plan_host.hpp"

#define EGGMAN_VERSION_STRING "eggman-hip 0.1 (gfx950)"

namespace eggman {

// ------------------------------------------------------------------------------ errors

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

// ------------------------------------------------------------------------------ state

// A plan resident on one device.  Only the transition words live on the device; the level
// boundaries are host-side launch parameters (as in ext/TheEggmanKernelAbstractionsExt.jl:28-30).
struct DevPlan {
    uint64_t id = 0;
    int device = 0;
    PlanLayout lay;
    void *d_trans = nullptr;
    size_t bytes = 0;
};

// Grow-only device buffer (ext/TheEggmanKernelAbstractionsExt.jl:68-88).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct DevState {
    hipStream_t stream = nullptr;    // the library's own stream on this device
    hipStream_t last = nullptr;      // stream of the most recent enqueued call
    hipEvent_t order = nullptr;      // orders calls on different streams (shared buffers)
    DevBuf H;                        // state values (ping-pong levels)
    DevBuf P;                        // staged pair values (host entry points)
    DevBuf out;                      // staged results (host entry points)
};

static std::recursive_mutex g_mu;
static std::map<uint64_t, std::unique_ptr<DevPlan>> g_plans;
static std::map<std::tuple<int, int, int>, uint64_t> g_plan_key;   // (device, K, index_bytes)
static uint64_t g_next_id = 1;
static std::map<int, DevState> g_dev;

static int current_device(int *dev) {
    EGG_HIP(hipGetDevice(dev));
    return EGGMAN_OK;
}

static int dev_state(int device, DevState **ds) {
    DevState &d = g_dev[device];
    if (!d.stream) {
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

static int find_plan(uint64_t handle, DevPlan **pl) {
    auto it = g_plans.find(handle);
    if (it == g_plans.end())
        return fail(EGGMAN_EINVAL, "unknown plan handle %llu (released by eggman_cache_clear?)",
                    (unsigned long long)handle);
    *pl = it->second.get();
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

// ------------------------------------------------------------------------------ tuning

// Launch choices for one forward call.  Every choice gives bit-identical results.
struct Tune {
    int bt = 64;        // lanes per state (batch tile width), power of two in [1, 64]
    int64_t chunk = 0;  // instances per pass over the levels (0 = all)
};

static int pow2_at_least(int64_t n) {
    int b = 1;
    while (b < 64 && b < n) b <<= 1;
    return b;
}

static Tune default_tune(const PlanLayout &, int, int64_t B) {
    Tune t;
    t.bt = pow2_at_least(B);
    t.chunk = B;
    return t;
}

// ------------------------------------------------------------------------------ forward

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

template <int DT, typename W>
static void launch_level_bt(int bt, const DevPlan &pl, int L, const void *P, int64_t ldp,
                            const void *Hc, int64_t ldh, void *Ho, int64_t ldo, int64_t nb,
                            hipStream_t st) {
    switch (bt) {
    case 1: launch_level<DT, W, 1>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 2: launch_level<DT, W, 2>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 4: launch_level<DT, W, 4>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 8: launch_level<DT, W, 8>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 16: launch_level<DT, W, 16>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    case 32: launch_level<DT, W, 32>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    default: launch_level<DT, W, 64>(pl, L, P, ldp, Hc, ldh, Ho, ldo, nb, st); break;
    }
}

// Walk all levels for instances [0, B) of P (rows ldp apart) into out, `chunk` instances per
// pass, keeping only two levels of state values (the level being written and the level below,
// which holds every child it reads).
template <int DT, typename W>
static int forward_levels(DevState &ds, const DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                          void *out, hipStream_t st, const Tune &tune) {
    using E = Elt<DT>;
    const PlanLayout &lay = pl.lay;
    const int64_t chunk = tune.chunk > 0 ? std::min(tune.chunk, B) : B;
    const size_t hb = sizeof(typename E::H), pb = sizeof(typename E::P);
    const int64_t rows = std::max<int64_t>(lay.max_level, 1);
    EGG_TRY(ensure(ds.H, (size_t)2 * rows * chunk * hb));
    char *buf[2] = {(char *)ds.H.p, (char *)ds.H.p + (size_t)rows * chunk * hb};
    const int top = lay.K / 2 + 1;
    for (int64_t b0 = 0; b0 < B; b0 += chunk) {
        const int64_t nb = std::min(chunk, B - b0);
        const char *Pc = (const char *)P + (size_t)b0 * pb;
        char *outc = (char *)out + (size_t)b0 * hb;
        k_fill_one<DT><<<dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, st>>>(
            (typename E::H *)buf[1], nb);
        for (int L = 2; L <= top; L++) {
            const void *Hc = buf[(L - 1) & 1];
            void *Ho = L == top ? (void *)outc : (void *)buf[L & 1];
            launch_level_bt<DT, W>(tune.bt, pl, L, Pc, ldp, Hc, nb, Ho, nb, nb, st);
        }
    }
    EGG_HIP(hipGetLastError());
    return EGGMAN_OK;
}

template <int DT>
static int forward_dt(DevState &ds, const DevPlan &pl, const void *P, int64_t ldp, int64_t B,
                      void *out, hipStream_t st, const Tune &tune) {
    if (pl.lay.index_bytes == 4) return forward_levels<DT, int32_t>(ds, pl, P, ldp, B, out, st, tune);
    return forward_levels<DT, int64_t>(ds, pl, P, ldp, B, out, st, tune);
}

static int forward_any(int dtype, DevState &ds, const DevPlan &pl, const void *P, int64_t ldp,
                       int64_t B, void *out, hipStream_t st, const Tune &tune) {
    switch (dtype) {
    case EGGMAN_F64: return forward_dt<EGGMAN_F64>(ds, pl, P, ldp, B, out, st, tune);
    case EGGMAN_F32: return forward_dt<EGGMAN_F32>(ds, pl, P, ldp, B, out, st, tune);
    case EGGMAN_C64: return forward_dt<EGGMAN_C64>(ds, pl, P, ldp, B, out, st, tune);
    case EGGMAN_C32: return forward_dt<EGGMAN_C32>(ds, pl, P, ldp, B, out, st, tune);
    case EGGMAN_I128: return forward_dt<EGGMAN_I128>(ds, pl, P, ldp, B, out, st, tune);
    }
    return fail(EGGMAN_EINVAL, "unknown dtype %d", dtype);
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

}  // namespace eggman

using namespace eggman;

// ================================================================================ C ABI

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
    DeviceGuard g(pl->device);
    EGG_TRY(g.rc);
    DevState *ds;
    EGG_TRY(dev_state(pl->device, &ds));
    EGG_TRY(chain_stream(*ds, ds->stream));
    const int64_t chunk = max_batch > 0 ? std::min(max_batch, B) : B;
    EGG_TRY(ensure(ds->P, (size_t)chunk * npairs * pb));
    EGG_TRY(ensure(ds->out, (size_t)chunk * hb));
    Tune tune = default_tune(pl->lay, dtype, chunk);
    tune.chunk = chunk;
    std::vector<char> stage;
    for (int64_t b0 = 0; b0 < B; b0 += chunk) {
        const int64_t nb = std::min(chunk, B - b0);
        const void *src = P_host;
        if (nb != B) {
            // this chunk's pairs, relaid out to [pair][nb] (ext/TheEggmanKernelAbstractionsExt.jl:127-129)
            stage.resize((size_t)nb * npairs * pb);
            for (int64_t q = 0; q < npairs; q++)
                std::memcpy(stage.data() + (size_t)q * nb * pb,
                            (const char *)P_host + ((size_t)q * B + b0) * pb, (size_t)nb * pb);
            src = stage.data();
        }
        EGG_HIP(hipMemcpyAsync(ds->P.p, src, (size_t)nb * npairs * pb, hipMemcpyHostToDevice,
                               ds->stream));
        EGG_TRY(forward_any(dtype, *ds, *pl, ds->P.p, nb, nb, ds->out.p, ds->stream, tune));
        EGG_HIP(hipMemcpyAsync((char *)out_host + (size_t)b0 * hb, ds->out.p, (size_t)nb * hb,
                               hipMemcpyDeviceToHost, ds->stream));
        EGG_HIP(hipStreamSynchronize(ds->stream));
    }
    return EGGMAN_OK;
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
    Tune tune = default_tune(pl->lay, dtype, B);
    return forward_any(dtype, *ds, *pl, P_dev, ldp, B, out_dev, st, tune);
}

int64_t eggman_cache_bytes(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int64_t total = 0;
    for (auto &kv : g_plans) total += (int64_t)kv.second->bytes;
    for (auto &kv : g_dev) total += (int64_t)(kv.second.H.bytes + kv.second.P.bytes + kv.second.out.bytes);
    return total;
}

int eggman_cache_clear(void) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    int rc = EGGMAN_OK;
    for (auto &kv : g_dev) {
        DeviceGuard g(kv.first);
        if (hipDeviceSynchronize() != hipSuccess) rc = fail(EGGMAN_EHIP, "synchronize failed");
        for (DevBuf *b : {&kv.second.H, &kv.second.P, &kv.second.out}) {
            if (b->p) (void)hipFree(b->p);
            b->p = nullptr;
            b->bytes = 0;
        }
    }
    for (auto &kv : g_plans) {
        DeviceGuard g(kv.second->device);
        if (kv.second->d_trans) (void)hipFree(kv.second->d_trans);
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

}  // extern "C"
