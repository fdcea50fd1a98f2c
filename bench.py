"""Benchmark of the MI355X subset-DP hafnian path.

Workload (one "step"): the forward walk of a batch of B = 1024 real Float64 symmetric N = 28
matrices (entries N(0, 1/N), seeded) sharing one plan, with P and the results resident in HBM
when the timed region starts.  Reported metric: DP transitions per second (one transition = one
(matrix, transition) multiply-add, 6 052 062 per N = 28 matrix), with hafnians/s, the plain
one-matrix-per-call path at the same N (Float64, and the ComplexF64 single-hafnian case), the
HBM roofline of the level kernel and the CPU restatement of the reference timed on the host.

    python bench.py [--steps K] [--warmup W] [--batch B] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "theeggman.jl_amd"))

HBM_PEAK_GBS = 8000.0
N_DEFAULT = 28


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def alg_bytes_fwd(K, B, tbytes, ibytes, ns, nt):
    """(ntrans + nstates + 1) * sizeof(I) + B * (ntrans + nstates + npairs) * sizeof(T): the
    plan streamed once per batch; per matrix one child read per transition, one write per
    state and one read per pair."""
    npairs = K * (K - 1) // 2
    return (nt + ns + 1) * ibytes + B * (nt + ns + npairs) * tbytes


def walk_split(nst, B, chunk, small=4096):
    """How the library walks B instances in passes of `chunk`: (passes, lb, lt) where levels
    2..lb and lt..top (at most `small` states each) run once for the whole batch when there is
    more than one pass, and levels lb+1..lt-1 once per pass.  nst[L-1] = states of level L."""
    top = len(nst)
    passes = -(-B // chunk)
    lb, lt = 1, top + 1
    if passes > 1:
        while lb + 1 < top and nst[lb] <= small:
            lb += 1
        while lt - 1 > lb + 1 and nst[lt - 2] <= small:
            lt -= 1
    return passes, lb, lt


def compulsory_bytes(nst, K, B, chunk, tbytes, ibytes, small=4096):
    """HBM bytes one forward call cannot avoid: every state value written once and read once as
    a child by the level above (the top level's one value per instance is the output), the plan
    words of a level once per walk of that level (once per pass, or once for the whole batch for
    the small end levels), and the pair values once."""
    reads, writes = compulsory_split(nst, K, B, chunk, tbytes, ibytes, small)
    return reads + writes


def grad_compulsory_bytes(nst, K, B, passes, tbytes, ibytes):
    """HBM bytes one forward + gradient call cannot avoid, in `passes` passes: the forward part
    writes every state value once and reads each once as a child; the reverse sweep reads every
    parent adjoint once and every child value once, and writes every child adjoint once; the
    plan's transition words are read once per pass by each part; P is read by each part and dP
    written once."""
    states = sum(int(n) for n in nst)
    top = len(nst)
    ntrans = sum(int(nst[L - 1]) * (2 * (L - 1) - 1) for L in range(2, top + 1))
    npairs = K * (K - 1) // 2
    below_top = states - int(nst[-1])
    reads = B * below_top * tbytes + B * (states - 1) * tbytes + B * below_top * tbytes
    reads += 2 * passes * ntrans * ibytes + 2 * B * npairs * tbytes
    writes = B * (states - 1) * tbytes + B * below_top * tbytes + B * npairs * tbytes
    return reads, writes


def compulsory_split(nst, K, B, chunk, tbytes, ibytes, small=4096):
    """(reads, writes) of compulsory_bytes."""
    passes, lb, lt = walk_split(nst, B, chunk, small)
    top = len(nst)
    states = sum(int(n) for n in nst)
    writes = B * (states - 1) * tbytes                     # levels 2..top
    reads = B * (states - int(nst[-1])) * tbytes           # levels 1..top-1, each once as a child
    for L in range(2, top + 1):
        k = 2 * (L - 1) - 1                       # out-degree of level L (subsets of size 2(L-1))
        walks = passes if lb < L < lt else 1
        reads += walks * int(nst[L - 1]) * k * ibytes
    return reads + B * (K * (K - 1) // 2) * tbytes, writes


def level_launches(nst, B, chunk, small=4096):
    """Level launches of one forward call of B instances in passes of `chunk`, as the library
    makes them: every level once per pass, except that when there is more than one pass the
    small levels at the bottom and the top (at most `small` states each) run once for the whole
    batch.  nst[L-1] is the number of states of level L."""
    top = len(nst)
    passes = -(-B // chunk)
    if passes == 1:
        return top - 1
    lb, lt = 1, top + 1
    while lb + 1 < top and nst[lb] <= small:
        lb += 1
    while lt - 1 > lb + 1 and nst[lt - 2] <= small:
        lt -= 1
    return (lb - 1) + passes * ((lt - 1) - lb) + (top - lt + 1)


def time_device(hip, stream, fn, steps, warmup):
    """Run fn warmup times, then time exactly `steps` calls between synchronisations, with HIP
    events recorded on `stream`, the stream the kernels are launched on."""
    for _ in range(warmup):
        fn()
    hip.device_synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        fn()
    ev1.record(stream)
    ev1.synchronize()
    hip.device_synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_ms(ev1) / steps, wall * 1e3 / steps


TRAFFIC_PROFILE = os.path.join("profiles", "hbm_traffic_n28_b1024.json")


def traffic_from_profiles(kernel, chunk):
    """HBM bytes per launch of the level kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs of this workload, scaled by the copy-kernel calibration), with
    its provenance -- or None when the summary is missing or was taken for another kernel or pass
    size than the launch timed here."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    if prof.get("kernel", "").split("<")[0].split("::")[-1] != kernel or prof.get("chunk", chunk) != chunk:
        return None
    return prof


# The reference's published single-call CPU figure for this path: `hafnian(A; check_symmetric =
# false)` on one ComplexF64 N=28 matrix, min over samples, 32 threads of a Ryzen 9 9950X, Julia
# 1.12.7 (assets/data/jl-eggman-hafnian-bench.json, key haf.eggman.rpt1.N=28).
PUBLISHED_CPU_N28_CF64_MS = 3.28


def host_cpu():
    """CPU model and SMT state of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = None
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    return {"model": model, "smt_active": smt, "logical_cpus": os.cpu_count()}


def native_oracle():
    """Build the oracle's C restatement with -march=native for this host (the CPU baseline's
    flags: -O3 -march=native -ffp-contract=off) into a temporary directory and load it; fall back
    to the portable build when no compiler is available."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    out = os.path.join(tempfile.mkdtemp(prefix="eggman_oracle_"), "liboracle_native.so")
    cmd = ["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
           "-o", out, os.path.join(ROOT, "oracle", "dp_oracle.c"), "-lquadmath"]
    try:
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
    except (OSError, subprocess.SubprocessError):
        return oracle, "portable build (-march=x86-64-v3): native build failed"
    oracle.use_library(out)
    return oracle, "gcc -O3 -march=native -ffp-contract=off -fopenmp"


def cpu_baseline(K, budget_s=12.0, plain_budget_s=6.0):
    """The oracle's C restatement of the reference CPU walk on this host's cores, min over reps
    after one warmup, plan build excluded (as benchmark/gpu/run_gpu_validation.jl:73-83 times):

    * batched (the reported value): _batch_cpu! over _haf_dp (src/hafnian.jl:1226-1242), B split
      into nthreads contiguous chunks, each instance serial;
    * plain: one N=K matrix by _haf_dp with its level-chunk policy (2000 transitions per chunk,
      parallel from 60000 transitions; src/dp.jl:370-442) on all threads, in Float64 and in
      ComplexF64 -- the latter beside the reference's published 3.28 ms."""
    import numpy as np
    from theeggman import synth
    oracle, build = native_oracle()
    nthreads = min(os.cpu_count() or 1, oracle.max_threads())
    Bs = max(nthreads * 32, 256)
    As = np.stack([synth.sym_normal(K, 3, b, 1 / np.sqrt(K)) for b in range(Bs)])
    P = synth.pair_batch(As)
    oracle.build_plan(K)

    def best_of(fn, budget, cap):
        fn()                                               # warmup
        best, spent, reps = float("inf"), 0.0, 0
        while spent < budget and reps < cap:
            t0 = time.perf_counter()
            fn()
            dt = time.perf_counter() - t0
            best, spent, reps = min(best, dt), spent + dt, reps + 1
        return best, spent, reps

    best, spent, reps = best_of(lambda: oracle.haf_dp(P, K, oracle.F64, nthreads), budget_s, 200)
    _, nt = oracle.dp_counts(K)
    plain = {}
    for name, T, dt, A in (("f64", np.float64, oracle.F64, synth.sym_normal(K, 3, 0, 1 / np.sqrt(K))),
                           ("c64", np.complex128, oracle.C64, synth.sym_complex_normal(K, 2, 0))):
        P1 = synth.pair_batch(A[None]).astype(T)
        b1, sp1, r1 = best_of(lambda: oracle.haf_dp(P1, K, dt, nthreads), plain_budget_s / 2, 2000)
        plain[name] = {"ms_per_hafnian": b1 * 1e3, "transitions_per_s": nt / b1,
                       "reps": r1, "threads": nthreads}
    plain["c64"]["published_reference_ms"] = PUBLISHED_CPU_N28_CF64_MS if K == 28 else None
    if K == 28:
        plain["c64"]["ratio_to_published"] = plain["c64"]["ms_per_hafnian"] / PUBLISHED_CPU_N28_CF64_MS
    return {"value": Bs * nt / best, "unit": "DP transitions/s", "cores": nthreads,
            "kind": "port", "hafnians_per_s": Bs / best,
            "sample": f"{Bs} real Float64 N={K} matrices per rep, _batch_cpu! policy over "
                      f"{nthreads} threads, min of {reps} reps ({spent:.1f} s)",
            "build": build, "host": host_cpu(), "plain": plain}


def oracle_checks(P, res, K, nsample=16):
    """The timed batch's results against the oracle: bit-exact equality with its Float64 walk
    (the reference's `_dp_states!` order, src/dp.jl:344-363) for three instances, and the relative
    error of `nsample` instances against its __float128 walk (the BigFloat stand-in)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    B = P.shape[1]
    idx = [0, B // 2, B - 1]
    want = oracle.haf_dp(P[:, idx].copy(), K, oracle.F64, 8)
    exact = bool(res[idx].tobytes() == want.tobytes())
    sample = np.linspace(0, B - 1, min(nsample, B)).astype(int)
    q = oracle.haf_dp(P[:, sample].copy(), K, oracle.Q, 8)        # (hi, lo) pairs
    ref = q[:, 0] + q[:, 1]
    rel = np.abs(res[sample] - ref) / np.abs(ref)
    return exact, {"vs": "__float128 oracle walk", "instances": int(len(sample)),
                   "max_rel_err": float(rel.max()), "median_rel_err": float(np.median(rel)),
                   "tolerance": 1e-10, "within_tolerance": bool(rel.max() <= 1e-10)}


def run_configs(path):
    """Measure the five workload configurations of the path (device times with inputs resident
    in HBM; the CPU restatement beside the small ones) and write them to `path` as JSON."""
    import numpy as np
    import theeggman as te
    from theeggman import _lib, hip, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    st = hip.Stream()
    res = {}

    def dev_forward(As_or_P, K, T, steps, backend=None, P=None):
        be = backend or te.HIPBackend()
        plan = te.PlanRef(K, 8) if be.plan_source == "generate" else te._dp_plan(K)
        if P is None:
            P = synth.pair_batch(np.asarray(As_or_P)).astype(T)
        B = P.shape[1]
        Pd = hip.DeviceBuffer.from_array(P)
        out = hip.DeviceBuffer(B * 16)
        fn = lambda: te._haf_dp_device(be, plan, T, Pd.addr, B, B, out.addr, st.handle)
        ms, _ = time_device(hip, st, fn, steps, 2)
        return ms, B

    # 1: plain N=20 real Float64, Uniform(-1, 1); the reference runs it on the CPU as-is
    A = synth.sym_uniform(20, 1, 0)
    ms, _ = dev_forward(A[None], 20, np.float64, 200)
    P = synth.pair_batch(A[None])
    oracle.build_plan(20)
    t0 = time.perf_counter()
    for _ in range(20):
        oracle.haf_dp(P, 20, oracle.F64, 1)
    cpu_ms = (time.perf_counter() - t0) / 20 * 1e3
    res["1_plain_n20_f64"] = {"gpu_ms": ms, "cpu_restatement_1thread_ms": cpu_ms,
                              "tuned": te.tune_get(20, np.float64, 1)}
    # 2: plain N=28 ComplexF64, (g1 + i g2)/sqrt(2N)
    ms, _ = dev_forward(synth.sym_complex_normal(28, 2, 0)[None], 28, np.complex128, 100)
    res["2_plain_n28_c64"] = {"gpu_ms": ms, "hafnians_per_s": 1e3 / ms}
    # 3a: 4096 real N=24 matrices, N(0, 1/N)
    As = np.stack([synth.sym_normal(24, 3, b, 1 / np.sqrt(24)) for b in range(4096)])
    ms, B = dev_forward(As, 24, np.float64, 10)
    _, nt = te._dp_counts(24)
    res["3a_batched_n24_b4096_f64"] = {"gpu_ms": ms, "hafnians_per_s": B / ms * 1e3,
                                       "transitions_per_s": B * nt / ms * 1e3}
    # 3b: 2048 matrices, degrees uniform on {16, 20, 24, 28, 32}, one pass
    rng = np.random.default_rng(3)
    degs = rng.choice([16, 20, 24, 28, 32], size=2048)
    groups = {K: np.stack([synth.sym_normal(int(K), 5, i, 1 / np.sqrt(K))
                           for i in np.flatnonzero(degs == K)]) for K in (16, 20, 24, 28, 32)}
    Ps = [synth.pair_batch(groups[K]) for K in sorted(groups)]
    plans = [te._dp_plan(K) for K in sorted(groups)]
    hs = [te._device_plan(te.HIPBackend(), p) for p in plans]
    flat = np.concatenate([P.ravel() for P in Ps])
    Pd = hip.DeviceBuffer.from_array(flat)
    out = hip.DeviceBuffer(2048 * 8)
    import ctypes
    harr = (ctypes.c_uint64 * 5)(*hs)
    carr = (ctypes.c_int64 * 5)(*[P.shape[1] for P in Ps])
    fn = lambda: _lib.check(_lib.lib().eggman_dp_forward_mixed_device(
        5, harr, carr, _lib.F64, Pd.addr, out.addr, st.handle), "mixed")
    ms, _ = time_device(hip, st, fn, 5, 1)
    ntr = sum(P.shape[1] * te._dp_counts(K)[1] for P, K in zip(Ps, sorted(groups)))
    res["3b_mixed_2048_k16_32_f64"] = {"gpu_ms": ms, "hafnians_per_s": 2048 / ms * 1e3,
                                       "transitions_per_s": ntr / ms * 1e3,
                                       "counts": {int(K): int(groups[K].shape[0]) for K in groups}}
    # 4: 256 real N=24 matrices, Uniform(-1, 1), forward + full gradient
    As = np.stack([synth.sym_uniform(24, 4, b) for b in range(256)])
    P = synth.pair_batch(As)
    Pd = hip.DeviceBuffer.from_array(P)
    out = hip.DeviceBuffer(256 * 8)
    dP = hip.DeviceBuffer(256 * 8 * P.shape[0])
    h = te._device_plan(te.HIPBackend(), te._dp_plan(24))
    fn = lambda: _lib.check(_lib.lib().eggman_dp_forward_grad_device(
        h, _lib.F64, Pd.addr, 256, 256, out.addr, dP.addr, 256, st.handle), "grad")
    ms, _ = time_device(hip, st, fn, 10, 2)
    res["4_grad_n24_b256_f64"] = {"gpu_ms": ms, "hafnians_with_gradient_per_s": 256 / ms * 1e3}
    # 5a: one ComplexF64 matrix at N=36 and N=40, Int64 plan generated in HBM
    gen = te.HIPBackend(plan_source="generate")
    for K in (36, 40):
        ms, _ = dev_forward(synth.sym_complex_normal(K, 2, 0)[None], K, np.complex128, 5, backend=gen)
        res[f"5a_plain_n{K}_c64"] = {"gpu_ms": ms, "transitions_per_s": te._dp_counts(K)[1] / ms * 1e3}
    # 5b: 64 G(40, 1/2) graphs, exact perfect-matching counts on the 128-bit path
    As = np.stack([synth.gnp_adjacency(40, 52, b) for b in range(64)])
    ms, B = dev_forward(As, 40, np.int64, 2, backend=gen)
    res["5b_exact_n40_b64"] = {"gpu_ms": ms, "hafnians_per_s": B / ms * 1e3,
                               "transitions_per_s": B * te._dp_counts(40)[1] / ms * 1e3}
    res["command"] = "python bench.py --configs PATH"
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)


def run_plans(path):
    """Plan costs and device memory (the reference's GPU validation stages C4 and D,
    benchmark/gpu/run_gpu_validation.jl:315-366, and its published plan build times,
    src/dp.jl:54-55, 73-78): host build time of the library's plan builder, upload and
    device-generation times, the calls it takes to amortise an upload, and the device bytes
    a batch needs against the reference's dp_batch_bytes (src/dp.jl:584-587)."""
    import numpy as np
    import theeggman as te
    from theeggman import _lib, hip, synth
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    L = _lib.lib()
    st = hip.Stream()
    res = {"plan": {}, "amortise": {}, "memory": {}}
    for K in (20, 24, 28, 32, 36, 40):
        ib = 4 if K <= 32 else 8
        row = {"index_bytes": ib}
        ns, nt = te._dp_counts(K)
        row["bytes"] = (ns + nt + 1) * ib
        if K <= 36:
            t0 = time.perf_counter()
            plan = te._build_dp_plan(K, np.int32 if ib == 4 else np.int64)
            row["host_build_s"] = time.perf_counter() - t0
        if K <= 32:
            t0 = time.perf_counter()
            oracle.build_plan(K)
            row["oracle_build_s"] = time.perf_counter() - t0
            oracle._PLANS.clear()
            te.empty_gpu_cache()
            t0 = time.perf_counter()
            te._device_plan(te.HIPBackend(), plan)
            hip.device_synchronize()
            row["upload_s"] = time.perf_counter() - t0
        te.empty_gpu_cache()
        t0 = time.perf_counter()
        te._device_plan(te.HIPBackend(plan_source="generate"), te.PlanRef(K, ib))
        hip.device_synchronize()
        row["generate_s"] = time.perf_counter() - t0
        if K <= 36:
            del plan
        res["plan"][K] = row
        log(f"plan K={K}: {row}")
        te.empty_gpu_cache()
    # upload amortisation: the first plain call of a degree (plan upload included) against the
    # steady-state call, as stage C4 measures it
    for K in (24, 28, 32):
        plan = te._dp_plan(K)
        A = synth.sym_complex_normal(K, 2, 0)
        te.empty_gpu_cache()
        t0 = time.perf_counter()
        te.hafnian(A, backend=te.HIPBackend())
        first = time.perf_counter() - t0
        t0 = time.perf_counter()
        for _ in range(20):
            te.hafnian(A, backend=te.HIPBackend())
        steady = (time.perf_counter() - t0) / 20
        res["amortise"][K] = {"first_call_s": first, "steady_call_s": steady,
                              "calls_to_amortise": (first - steady) / steady}
        log(f"amortise K={K}: {res['amortise'][K]}")
    # device memory one call needs beside the plan: forward (two levels per pass) and forward +
    # gradient (the kept levels of a pass), measured as the library's cache bytes after a call
    for K in (24, 28, 32):
        B = 1024
        plan = te._dp_plan(K)
        te.empty_gpu_cache()
        h = te._device_plan(te.HIPBackend(), plan)
        base = L.eggman_cache_bytes()
        As = np.stack([synth.sym_normal(K, 3, b, 1 / np.sqrt(K)) for b in range(B)])
        P = synth.pair_batch(As)
        Pd = hip.DeviceBuffer.from_array(P)
        out = hip.DeviceBuffer(B * 8)
        te._haf_dp_device(te.HIPBackend(), plan, np.float64, Pd.addr, B, B, out.addr, st.handle)
        st.synchronize()
        fwd = L.eggman_cache_bytes() - base
        dP = hip.DeviceBuffer(B * 8 * P.shape[0])
        _lib.check(L.eggman_dp_forward_grad_device(h, _lib.F64, Pd.addr, B, B, out.addr, dP.addr, B,
                                                   st.handle))
        st.synchronize()
        grad = L.eggman_cache_bytes() - base
        ns, _ = te._dp_counts(K)
        npairs = K * (K - 1) // 2
        res["memory"][K] = {
            "batch": B, "reference_dp_batch_bytes": int(te.dp_batch_bytes(K, np.float64, B)),
            "forward_work_bytes": int(fwd), "forward_plus_gradient_work_bytes": int(grad),
            "caller_bytes_per_instance_forward": (npairs + 1) * 8,
            "caller_bytes_per_instance_gradient": (2 * npairs + 1) * 8,
            "note": "forward work memory is two levels of one pass (independent of B); the "
                    "reference holds every state of every instance (dp_batch_bytes)"}
        log(f"memory K={K}: {res['memory'][K]}")
        del Pd, out, dP
    res["command"] = "python bench.py --plans PATH"
    with open(path, "w") as f:
        json.dump(res, f, indent=1, default=str)
    print(json.dumps(res, default=str), flush=True)


def run_sweep(path, K=28):
    """Forward and forward + full gradient of real FP64 degree-K batches over a range of batch
    sizes (device times, inputs resident in HBM, tuned launches), written to `path` as JSON."""
    import numpy as np
    import theeggman as te
    from theeggman import _lib, hip, synth
    st = hip.Stream()
    be = te.HIPBackend()
    plan = te._dp_plan(K)
    h = te._device_plan(be, plan)
    _, nt = te._dp_counts(K)
    npairs = K * (K - 1) // 2
    rows = []
    for B in (1, 2, 4, 8, 16, 32, 64, 128, 192, 256, 384, 512, 1024, 2048, 4096):
        As = np.stack([synth.sym_normal(K, 9, b, 1 / np.sqrt(K)) for b in range(B)])
        P = synth.pair_batch(As).astype(np.float64)
        Pd = hip.DeviceBuffer.from_array(P)
        out = hip.DeviceBuffer(B * 8)
        dP = hip.DeviceBuffer(B * 8 * npairs)
        steps = 50 if B <= 64 else 10 if B <= 1024 else 4
        fms, _ = time_device(hip, st, lambda: te._haf_dp_device(be, plan, np.float64, Pd.addr, B, B,
                                                                out.addr, st.handle), steps, 2)
        gms, _ = time_device(hip, st, lambda: _lib.check(_lib.lib().eggman_dp_forward_grad_device(
            h, _lib.F64, Pd.addr, B, B, out.addr, dP.addr, B, st.handle), "grad"), max(2, steps // 2), 1)
        row = {"B": B, "forward_ms": fms, "hafnians_per_s": B / (fms / 1e3),
               "transitions_per_s": B * nt / (fms / 1e3), "forward_grad_ms": gms,
               "hafnians_with_gradient_per_s": B / (gms / 1e3)}
        rows.append(row)
        log(f"sweep K={K} B={B}: forward {fms:.3f} ms, forward+gradient {gms:.3f} ms")
        del Pd, out, dP
    res = {"workload": f"real FP64 N={K} batches, N(0,1/N), tuned launches, inputs resident in HBM",
           "command": "python bench.py --sweep PATH", "rows": rows}
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--configs", metavar="PATH",
                    help="instead of the benchmark line, measure the five workload configurations "
                         "and write them to PATH")
    ap.add_argument("--sweep", metavar="PATH",
                    help="instead of the benchmark line, time the forward and the forward + "
                         "gradient of N=28 FP64 batches of 1 to 4096 and write them to PATH")
    ap.add_argument("--plans", metavar="PATH",
                    help="instead of the benchmark line, measure plan build/upload/generation "
                         "times, upload amortisation and device memory, and write them to PATH")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--N", type=int, default=N_DEFAULT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path lines")
    ap.add_argument("--no-grad", action="store_true", help="skip the forward+gradient line")
    ap.add_argument("--autotune", action="store_true",
                    help="re-run the autotuner before timing instead of using the shipped table")
    ap.add_argument("--no-tune", action="store_true",
                    help="ignore every tuning table: built-in default launch choices")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="after the timed steps, run one device-to-device copy of this many MiB "
                         "(a kernel of known traffic, to scale the HBM counters of a profile)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested): the batched path "
              "shards by instance with no communication, so its whole-job rate is N times the "
              "one-GPU value", file=sys.stderr)
        sys.exit(2)
    if args.configs:
        run_configs(args.configs)
        return
    if args.sweep:
        run_sweep(args.sweep)
        return
    if args.plans:
        run_plans(args.plans)
        return

    import numpy as np
    import theeggman as te
    from theeggman import _lib, hip, synth

    K, B = args.N, args.batch
    be = te.HIPBackend()
    plan = te._dp_plan(K)
    ns, nt = te._dp_counts(K)
    ib = plan.index_bytes
    stream = hip.Stream()

    log(f"building inputs N={K} B={B}")
    As = np.stack([synth.sym_normal(K, 3, b, 1 / np.sqrt(K)) for b in range(B)])
    P = synth.pair_batch(As)                                    # [pair][batch]
    Pd = hip.DeviceBuffer.from_array(P)
    out = hip.DeviceBuffer(B * 8)

    def step():
        te._haf_dp_device(be, plan, np.float64, Pd.addr, B, B, out.addr, stream.handle)

    if args.no_tune:
        te.tune_set(K, np.float64, B, bt=64 if B >= 64 else 1 << max(0, (B - 1).bit_length()),
                    chunk=B, occ=0, lds=True)
    elif args.autotune:
        log("autotune batched")
        te.autotune(K, np.float64, B, reps=2)
    tuned = te.tune_get(K, np.float64, B)
    log("batched forward")
    ms, wall_ms = time_device(hip, stream, step, args.steps, args.warmup)
    res = out.download(np.float64, B)

    value = B * nt / (ms / 1e3)
    chunk = min((tuned or {}).get("chunk") or B, B)
    nst = np.diff(np.asarray(plan.levels, dtype=np.int64))
    launches = level_launches(nst, B, chunk)
    # the library walks passes two at a time on two streams when there are four or more (and the
    # values are at most 8 bytes), so two level launches are in flight at once
    in_flight = 2 if -(-B // chunk) >= 4 else 1
    tl = tuned or {}
    head = ("k_level_fwd_lds_vec" if tl.get("vec") else
            "k_level_fwd_lds_narrow" if tl.get("narrow") else "k_level_fwd_lds") \
        if tl.get("lds", 1) and not tl.get("fused") else "k_level_fwd"
    # the roofline's algorithmic bytes are the compulsory traffic of the step; the
    # per-transition model (one 8-byte child read per transition) is reported beside it, labelled
    nbytes = compulsory_bytes(nst, K, B, chunk, 8, ib)
    model_bytes = alg_bytes_fwd(K, B, 8, ib, ns, nt)
    achieved = nbytes / (ms / 1e3) / 1e9
    result = {
        "metric": "DP transitions/s (batched N=28 FP64 forward)",
        "value": value,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 normals, N(0,1/N), symmetric, zero diagonal)",
        "config": {"workload": f"batched forward, {B} real Float64 N={K} matrices sharing one "
                               f"Int32 plan, inputs resident in HBM", "N": K, "batch": B,
                   "plan_bytes": (nt + ns + 1) * ib},
        "hafnians_per_s": B / (ms / 1e3),
        "wall_ms_per_step": wall_ms,
        "oracle_spot_check": None,
        "tuned_launch": tuned,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": head, "launches_per_step": launches,
                     "alg_bytes_per_launch": nbytes / launches,
                     "alg_bytes": "compulsory: each state value written once and read once, "
                                  "plan words once per walk of a level, pair values once",
                     "avg_launch_ms": ms / launches,
                     "launches_in_flight": in_flight,
                     "launch_time_note": "avg_launch_ms is the step time per launch; with two "
                                         "passes in flight each launch lasts about twice that "
                                         "(the kernel trace's average duration), and the device "
                                         "rate is the same",
                     "per_transition_model": {
                         "bytes_per_step": model_bytes,
                         "gbs": model_bytes / (ms / 1e3) / 1e9,
                         "note": "one 8-byte child read per transition; the colex sweep serves "
                                 "most of them from L2/MALL, so this rate can exceed the HBM peak "
                                 "and is not a roofline fraction"}},
    }

    rl = result["roofline"]
    prof = traffic_from_profiles(head, chunk)
    if prof:
        rl["traffic"] = prof["bytes_per_launch"]
        rl["traffic_source"] = (f"{TRAFFIC_PROFILE}: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE "
                                f"runs of this workload ({prof.get('measured_at', 'see file')}), "
                                f"FETCH scaled by {prof.get('fetch_scale')}")
        creads, cwrites = compulsory_split(nst, K, B, chunk, 8, ib)
        rl["read_amplification"] = prof["fetch_bytes_per_launch"] * launches / creads
        rl["write_amplification"] = prof["write_bytes_per_launch"] * launches / cwrites
    if args.copy_calibration <= 0:
        # measured device-copy bandwidth on this GPU (reads + writes of a 2 GiB copy), the
        # practical ceiling beside the nominal 8 TB/s (skipped in counter runs, whose one copy
        # must be the calibration copy)
        cbytes = 2 << 30
        csrc, cdst = hip.DeviceBuffer(cbytes), hip.DeviceBuffer(cbytes)
        csrc.zero()
        cms, _ = time_device(hip, stream, lambda: hip.copy_d2d_async(cdst, csrc, cbytes, stream), 5, 1)
        copy_gbs = 2 * cbytes / (cms / 1e3) / 1e9
        del csrc, cdst
        rl["measured_copy_gbs"] = copy_gbs
        rl["frac_of_copy"] = rl["achieved"] / copy_gbs
        # the practical ceiling of the level kernel's gathers: random reads of whole 128-byte
        # lines (8 lanes x 16 bytes, as the kernel reads a child row) from a working set the size
        # of the level buffers of the passes in flight, measured on this GPU
        import ctypes
        probe = {}
        for name, ws in (("l2_resident", 2 << 20),
                         ("pass_buffers", in_flight * 2 * int(nst.max()) * chunk * 8)):
            lps = ctypes.c_double()
            _lib.check(_lib.lib().eggman_gather_rate(ws, ctypes.byref(lps)), "gather_rate")
            probe[name] = {"working_set_bytes": ws, "line_gbs": lps.value * 128 / 1e9}
        rl["gather_probe"] = probe
        if prof and prof.get("l1_to_l2_read_requests_per_launch"):
            # the kernel's L1-to-L2 line requests per second (counter run) against the probe's
            # rates blended by the kernel's L2 hit rate: its hits at the L2-resident rate, its
            # misses at the rate of a working set the size of the level buffers of the passes in
            # flight
            kl = prof["l1_to_l2_read_requests_per_launch"] * 128 / (rl["avg_launch_ms"] / 1e3) / 1e9
            h = prof["l2"]["hit_rate"]
            ceil = 1.0 / (h / probe["l2_resident"]["line_gbs"] + (1 - h) / probe["pass_buffers"]["line_gbs"])
            rl["l1_to_l2_line_gbs"] = kl
            rl["gather_ceiling_gbs"] = ceil
            rl["frac_of_gather_ceiling"] = kl / ceil
            # the blend is an estimate (the probe's misses are random over the whole working set,
            # the kernel's follow the colex sweep, and two launches overlap); the probe's
            # all-L2-hit rate bounds any line-request stream from above
            rl["frac_of_l2_resident_line_rate"] = kl / probe["l2_resident"]["line_gbs"]
    if rl["traffic"]:
        rl["traffic_gbs"] = rl["traffic"] / (rl["avg_launch_ms"] / 1e3) / 1e9
        rl["traffic_frac"] = rl["traffic_gbs"] / HBM_PEAK_GBS

    if args.copy_calibration > 0:
        nbytes_copy = args.copy_calibration << 20
        src, dst = hip.DeviceBuffer(nbytes_copy), hip.DeviceBuffer(nbytes_copy)
        src.zero()
        hip.copy_d2d(dst, src, nbytes_copy)
        hip.device_synchronize()
        result["copy_calibration_bytes"] = nbytes_copy

    if not args.no_plain:
        # the same batch through the host-pointer entry point (P and the results in host memory,
        # staged through pinned buffers; uploads and downloads overlap compute across chunks):
        # the rate a caller handing over host arrays sees, PCIe transfers included
        Pt = np.ascontiguousarray(P)                              # [pair][batch], host
        host_out = np.empty(B, np.float64)
        h = te._device_plan(be, plan)

        def host_call():
            _lib.check(_lib.lib().eggman_dp_forward(h, _lib.F64, Pt.ctypes.data, B, Pt.shape[0], 0,
                                                    host_out.ctypes.data), "forward")
        log("host-pointer entry")
        host_call()
        reps = max(5, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(reps):
            host_call()
        hms = (time.perf_counter() - t0) / reps * 1e3
        if host_out.tobytes() != res.tobytes():
            raise SystemExit("host-pointer entry differs from the device entry: refusing to report")
        result["host_entry"] = {"ms_per_call": hms, "hafnians_per_s": B / (hms / 1e3),
                                "transitions_per_s": B * nt / (hms / 1e3), "calls": reps,
                                "bytes_in_per_call": int(P.nbytes), "bytes_out_per_call": B * 8}

    if not args.no_grad:
        # forward + full gradient d haf / d P for the same batch, inputs resident in HBM
        dPd = hip.DeviceBuffer(B * 8 * (K * (K - 1) // 2))
        h = te._device_plan(be, plan)

        def grad_step():
            _lib.check(_lib.lib().eggman_dp_forward_grad_device(
                h, _lib.F64, Pd.addr, B, B, out.addr, dPd.addr, B, stream.handle), "grad")
        log("batched forward + gradient")
        gsteps = max(2, args.steps // 4)
        msg, _ = time_device(hip, stream, grad_step, gsteps, 1)
        result["fwd_grad"] = {"ms_per_step": msg, "hafnians_with_gradient_per_s": B / (msg / 1e3),
                              "forward_transitions_per_s": B * nt / (msg / 1e3), "steps": gsteps}
        # compulsory traffic of the call (passes of 512 instances at N=28; the library raises the
        # forward's tuned pass to at least 512 for the gradient)
        gpass = max(512, B if B < 512 else 512)
        grd, gwr = grad_compulsory_bytes(nst, K, B, -(-B // gpass), 8, plan.index_bytes)
        result["fwd_grad"]["roofline"] = {
            "bound": "hbm", "alg_bytes": "compulsory: forward part writes each state once and reads "
            "each once as a child; reverse reads each parent adjoint and child value once and writes "
            "each child adjoint once; plan words once per pass per part; P, dP",
            "compulsory_bytes": grd + gwr, "achieved": (grd + gwr) / (msg / 1e3) / 1e9,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (grd + gwr) / (msg / 1e3) / 1e9 / HBM_PEAK_GBS,
            "counters": "profiles/grad_trace_counters_n28_b1024.md"}
        del dPd

    if not args.no_plain:
        for name, T, gen in (("plain_f64", np.float64, lambda: synth.sym_normal(K, 3, 0, 1 / np.sqrt(K))),
                             ("plain_c64", np.complex128, lambda: synth.sym_complex_normal(K, 2, 0))):
            A = gen()
            P1 = synth.pair_batch(A[None]).astype(T)
            P1d = hip.DeviceBuffer.from_array(P1)
            o1 = hip.DeviceBuffer(16)

            def one():
                te._haf_dp_device(be, plan, T, P1d.addr, 1, 1, o1.addr, stream.handle)
            if args.no_tune:
                te.tune_set(K, T, 1, bt=1, chunk=1, occ=0, lds=False)
            elif args.autotune:
                log(f"autotune {name}")
                te.autotune(K, T, 1, reps=5)
            choice = te.tune_get(K, T, 1)
            log(f"{name}")
            steps1 = max(args.steps * 5, 50)
            ms1, wall1 = time_device(hip, stream, one, steps1, args.warmup)
            result[name] = {"ms_per_hafnian": ms1, "wall_ms_per_hafnian": wall1,
                            "hafnians_per_s": 1e3 / ms1, "transitions_per_s": nt / (ms1 / 1e3),
                            "tuned_launch": choice}
            if name == "plain_f64" and not args.no_grad:
                # one matrix per call, value and full gradient
                dP1 = hip.DeviceBuffer(8 * (K * (K - 1) // 2))
                h1 = te._device_plan(be, plan)

                def one_grad():
                    _lib.check(_lib.lib().eggman_dp_forward_grad_device(
                        h1, _lib.F64, P1d.addr, 1, 1, o1.addr, dP1.addr, 1, stream.handle), "grad")
                log("plain_f64 with gradient")
                msg1, _ = time_device(hip, stream, one_grad, max(args.steps, 20), args.warmup)
                result["plain_grad_f64"] = {"ms_per_hafnian_with_gradient": msg1,
                                            "hafnians_with_gradient_per_s": 1e3 / msg1}
        result["batched_vs_plain_f64_hafnians"] = result["hafnians_per_s"] / result["plain_f64"]["hafnians_per_s"]
        if "fwd_grad" in result:
            result["batched_grad_vs_plain_f64_hafnians"] = (
                result["fwd_grad"]["hafnians_with_gradient_per_s"] / result["plain_f64"]["hafnians_per_s"])
        if "fwd_grad" in result and "plain_grad_f64" in result:
            result["batched_grad_vs_plain_grad_f64_hafnians"] = (
                result["fwd_grad"]["hafnians_with_gradient_per_s"]
                / result["plain_grad_f64"]["hafnians_with_gradient_per_s"])

    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(K)
        # the CPU oracle as the checker of the timed results: a bit-exact spot check against its
        # Float64 walk, and the error of a sample against its __float128 walk
        log("oracle checks")
        result["oracle_spot_check"], result["accuracy"] = oracle_checks(P, res, K)
        if result["oracle_spot_check"] is False:
            raise SystemExit("device result differs from the oracle: refusing to report")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
