"""Parity of the HIP path with the oracle, through the C ABI, on an MI355X.

Bar: bit-exact for every element type.  The device accumulates each state in plan order with no
fused multiply-add, exactly as the oracle (and the reference's _dp_states!, src/dp.jl:344-363)
does, so equality is asserted with array_equal / tobytes, never a tolerance -- except against
the __float128 oracle, where the stated bound is the reference's 1e-13 for the DP
(test/runtests.jl:318-332).
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

pytestmark = pytest.mark.gpu

BE = te.HIPBackend()
BE_GEN = te.HIPBackend(plan_source="generate")
DT = {np.float64: o.F64, np.float32: o.F32, np.complex128: o.C64, np.complex64: o.C32}


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def stack(N, B, config, kind="uniform"):
    if kind == "uniform":
        return np.stack([synth.sym_uniform(N, config, b) for b in range(B)])
    if kind == "complex":
        return np.stack([synth.sym_complex_normal(N, config, b) for b in range(B)])
    return np.stack([synth.sym_normal(N, config, b, 1 / np.sqrt(N)) for b in range(B)])


def oracle_vals(As, T, nthreads=8):
    As = As.astype(T)
    return o.haf_dp(synth.pair_batch(As), As.shape[-1], DT[T], nthreads)


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


@pytest.mark.parametrize("N", [4, 8, 16, 20, 24])
@pytest.mark.parametrize("B", [1, 3, 64, 100])
def test_f64_batched_bit_exact(N, B):
    note(f"f64 N={N} B={B}")
    As = stack(N, B, 11)
    got = te.hafnian(As, backend=BE)
    assert got.tobytes() == oracle_vals(As, np.float64).tobytes()


@pytest.mark.parametrize("T", [np.float32, np.complex64, np.complex128])
@pytest.mark.parametrize("N", [12, 20])
def test_other_types_bit_exact(T, N):
    note(f"{np.dtype(T)} N={N}")
    As = stack(N, 37, 12, "complex" if np.dtype(T).kind == "c" else "uniform").astype(T)
    got = te.hafnian(As, backend=BE)
    assert got.dtype == np.dtype(T)
    assert got.tobytes() == oracle_vals(As, T).tobytes()


@pytest.mark.parametrize("T,N,B", [(np.float64, 22, 1), (np.float64, 26, 1), (np.float64, 30, 1),
                                   (np.complex128, 26, 1), (np.float32, 28, 2), (np.float64, 26, 5)])
def test_fused_level_runs_bit_exact(T, N, B):
    # a one-pass walk of few instances runs its small bottom and top levels as one fused launch
    # each; the values are the oracle's bits
    note(f"fused runs {np.dtype(T)} N={N} B={B}")
    As = stack(N, B, 13, "complex" if np.dtype(T).kind == "c" else "normal").astype(T)
    got = te.hafnian(As, backend=BE)
    assert got.tobytes() == oracle_vals(As, T).tobytes()


def test_fused_level_runs_exact_integers():
    G = np.stack([synth.gnp_adjacency(26, 14, b) for b in range(3)]).astype(np.int64)
    got = te.hafnian_exact(G)
    want = o.haf_dp(o.gather_pairs_batch(G), 26, o.I128, 8)
    assert [int(v) for v in got] == [int(v) for v in want]


def test_no_fma_on_device():
    x = 1.0 + 2.0**-30
    A = np.zeros((4, 4))
    for (i, j), v in {(0, 1): 1.0, (2, 3): -1.0, (0, 2): x, (1, 3): x}.items():
        A[i, j] = A[j, i] = v
    assert te.hafnian(A, backend=BE) == 2.0**-29       # fused would give 2^-29 + 2^-60


def test_docstring_examples_on_device():
    assert te.hafnian(np.array([[0, 1, 2, 3], [1, 0, 4, 5], [2, 4, 0, 6], [3, 5, 6, 0]])) == 28.0
    assert te.hafnian_repeated(np.array([[1.0, 2.0], [2.0, 3.0]]), [2, 2], method="dp") == 11.0
    assert te.hafnian(np.array([[0.0, 1.0], [1.0, 0.0]])) == 1.0


def test_max_batch_chunking_bit_identical():
    # test/gpu.jl:26-40
    As = stack(16, 7, 13, "complex")
    scalar = np.array([te.hafnian(A, backend=BE) for A in As])
    assert te.hafnian(As, backend=BE).tobytes() == scalar.tobytes()
    for mb in (1, 2, 3, 100):
        assert te.hafnian(As, backend=BE, max_batch=mb).tobytes() == scalar.tobytes()
    assert te.hafnian(As[:1], backend=BE).tobytes() == scalar[:1].tobytes()
    assert te.hafnian([], backend=BE).shape == (0,)


def test_repeated_batch_matches_scalar():
    # test/gpu.jl:42-56
    A = synth.sym_complex_normal(16, 14, 0)
    np.fill_diagonal(A, synth.normal(synth.stream(14, 1), 16))
    rpts = [[2, 2, 2, 1, 1, 0, 0, 0, 1, 1, 2, 2, 1, 1, 0, 0], [1] * 16,
            [4, 4, 4, 4] + [0] * 12]
    scalar = np.array([te.hafnian_repeated(A, r, backend=BE, method="dp") for r in rpts])
    assert te.hafnian_repeated(A, rpts, backend=BE).tobytes() == scalar.tobytes()
    assert te.hafnian_repeated(A, rpts, backend=BE, max_batch=2).tobytes() == scalar.tobytes()
    for r, v in zip(rpts, scalar):
        idx = [i + 1 for i, c in enumerate(r) for _ in range(c)]
        P = o.gather_pairs(A, idx).reshape(-1, 1)
        assert o.haf_dp(P, len(idx), o.C64)[0].tobytes() == v.tobytes()


def test_determinism():
    As = stack(18, 4, 15, "complex")
    assert te.hafnian(As[0], backend=BE).tobytes() == te.hafnian(As[0], backend=BE).tobytes()
    assert te.hafnian(As, backend=BE).tobytes() == te.hafnian(As, backend=BE).tobytes()


@pytest.mark.parametrize("K", [8, 12, 16, 20])
def test_int32_and_int64_layouts_bit_identical(K):
    As = stack(K, 5, 16, "complex")
    P = te._gather_pairs_batch(np.complex128, As, None, K)
    v32 = te._haf_dp_backend(BE, P, te._dp_plan(K, np.int32))
    v64 = te._haf_dp_backend(BE, P, te._dp_plan(K, np.int64))
    assert v32.tobytes() == v64.tobytes()


@pytest.mark.parametrize("K,I", [(4, np.int32), (16, np.int32), (28, np.int32), (16, np.int64),
                                  (28, np.int64)])
def test_device_generated_plan_equals_host_plan(K, I):
    note(f"plan_generate K={K} {np.dtype(I)}")
    hp = te._dp_plan(K, I)
    h = te._device_plan(BE_GEN, hp)
    got = np.empty_like(hp.trans)
    _lib.check(_lib.lib().eggman_plan_download(h, got.ctypes.data, got.size))
    assert np.array_equal(got, hp.trans)
    A = stack(K, 3, 17)
    P = te._gather_pairs_batch(np.float64, A, None, K)
    assert te._haf_dp_backend(BE_GEN, P, hp).tobytes() == te._haf_dp_backend(BE, P, hp).tobytes()


def test_cache_accounting():
    # test/gpu.jl:96-108
    te.empty_gpu_cache()
    assert te.gpu_cache_bytes() == 0
    A = synth.sym_complex_normal(16, 18, 0)
    te.hafnian(A, backend=BE)
    used = te.gpu_cache_bytes()
    assert used > 0
    te.hafnian(A, backend=BE)
    assert te.gpu_cache_bytes() == used
    te.empty_gpu_cache()
    assert te.gpu_cache_bytes() == 0
    assert te.hafnian(A, backend=BE) == te.hafnian(A, backend=BE)


def test_exact_integer_path():
    note("exact N=16,20")
    for N in (16, 20):
        As = [synth.gnp_adjacency(N, 19, b) for b in range(6)]
        got = te.hafnian_exact(As)
        want = [o.haf_dp(o.gather_pairs(A).reshape(-1, 1), N, o.I128)[0] for A in As]
        assert got == want
    assert te.hafnian_exact(np.ones((20, 20), np.int64)) == 654729075          # 19!!
    n = 9
    B = np.random.default_rng(3).integers(-2, 3, (n, n))
    A = np.zeros((2 * n, 2 * n), np.int64)
    A[:n, n:] = B
    A[n:, :n] = B.T
    assert te.hafnian_exact(A) == o.perm_ryser(B)


def test_config2_n28_complex_against_quad_oracle():
    # plain single hafnian, N=28 ComplexF64, (g1 + i g2)/sqrt(2N)
    note("N=28 complex single")
    A = synth.sym_complex_normal(28, 2, 0)
    got = te.hafnian(A, backend=BE)
    P = o.gather_pairs(A).reshape(-1, 1)
    assert got.tobytes() == o.haf_dp(P, 28, o.C64, 8)[0].tobytes()
    hi, lo = o.haf_dp(P, 28, o.CQ, 8)[0]
    err = abs((got - hi) - lo) / abs(hi)
    note(f"N=28 complex rel err vs __float128 oracle: {err:.3e}")
    assert err < 1e-13


def test_full_size_batch_n28_properties():
    # N=28, B=1024 real FP64: bit-exact on a sample of instances, and order-equivariant
    note("N=28 B=1024")
    B = 1024
    As = stack(28, B, 3, "normal")
    got = te.hafnian(As, backend=BE, check_symmetric=False)
    sample = [0, 1, 63, 64, 511, 1000, 1023]
    want = o.haf_dp(synth.pair_batch(As[sample]), 28, o.F64, 8)
    assert got[sample].tobytes() == want.tobytes()
    perm = np.random.default_rng(0).permutation(B)
    assert te.hafnian(As[perm], backend=BE, check_symmetric=False).tobytes() == got[perm].tobytes()


def test_device_resident_entry_point():
    from theeggman import hip
    K, B, ldp = 20, 200, 256
    As = stack(K, B, 20)
    P = synth.pair_batch(As)
    Ppad = np.zeros((P.shape[0], ldp))
    Ppad[:, :B] = P
    Pd = hip.DeviceBuffer.from_array(Ppad)
    out = hip.DeviceBuffer(B * 8)
    st = hip.Stream()
    te._haf_dp_device(BE, te._dp_plan(K), np.float64, Pd.addr, ldp, B, out.addr, st.handle)
    st.synchronize()
    assert out.download(np.float64, B).tobytes() == o.haf_dp(P, K, o.F64, 8).tobytes()
    # the same buffers through the library's own stream (stream 0)
    out.zero()
    te._haf_dp_device(BE, te._dp_plan(K), np.float64, Pd.addr, ldp, B, out.addr, 0)
    hip.device_synchronize()
    assert out.download(np.float64, B).tobytes() == o.haf_dp(P, K, o.F64, 8).tobytes()


@pytest.mark.parametrize("T,N,B", [(np.float64, 26, 100), (np.float64, 18, 70), (np.float32, 30, 64),
                                   (np.complex64, 26, 80), (np.float64, 18, 1), (np.float64, 28, 8),
                                   (np.complex128, 24, 33), (np.float64, 28, 24), (np.float32, 26, 17),
                                   (np.complex64, 30, 40)])
def test_untuned_defaults_bit_exact(T, N, B):
    # degrees and element types with no tuning entry use the default launch choices (the fused
    # walk, the 16-byte-load kernel on 16-instance tiles for 16 to 63 instances of degree 26 and
    # up, or that kernel in passes); the values are the oracle's bits
    note(f"defaults {np.dtype(T)} N={N} B={B}")
    assert te.tune_get(N, T, B) is None
    As = stack(N, B, 15, "complex" if np.dtype(T).kind == "c" else "normal").astype(T)
    got = te.hafnian(As, backend=BE, check_symmetric=False)
    assert got.tobytes() == oracle_vals(As, T).tobytes()


@pytest.mark.parametrize("T", [np.float64, np.float32, np.complex64, np.complex128])
def test_passes_in_flight_bit_identical(T):
    # four or more passes of at most 8-byte values run two at a time on two streams, each with
    # its own level buffers (a ragged last pass included); one pass, and 16-byte values (which stay
    # on one stream), are the references
    N, B = 20, 1000
    note(f"passes in flight N={N} B={B} {np.dtype(T)}")
    kind = "complex" if np.iscomplexobj(np.zeros(1, T)) else "uniform"
    As = stack(N, B, 71, kind).astype(T)
    P = te._gather_pairs_batch(T, As, None, N)
    plan = te._dp_plan(N)
    code = te.tuning.dtype_code(T)
    prev = te.tuning.tune_get(N, T, B)
    flags = te.tuning.LDS | te.tuning.VEC
    try:
        _lib.check(_lib.lib().eggman_tune_set(N, plan.index_bytes, code, B, 64, 64, flags, 2), "tune_set")
        split = te._haf_dp_backend(BE, P, plan)        # 16 passes
        _lib.check(_lib.lib().eggman_tune_set(N, plan.index_bytes, code, B, 64, 0, flags, 2), "tune_set")
        whole = te._haf_dp_backend(BE, P, plan)        # one pass
    finally:
        if prev is not None:
            _lib.check(_lib.lib().eggman_tune_set(N, plan.index_bytes, code, B, prev["bt"], prev["chunk"],
                                                  prev["flags"], prev["occ"]), "tune_set")
    assert split.tobytes() == whole.tobytes()
    if T in DT:
        assert split.tobytes() == oracle_vals(As, T).tobytes()


def test_calls_from_several_threads():
    # host threads calling at once, host-pointer calls and device-pointer calls on streams of
    # their own (passes two at a time inside, the library's buffers shared): every result is the
    # bits of the same call made alone
    import threading
    from theeggman import hip
    K = 28
    plan = te._dp_plan(K)
    note(f"threads K={K}")
    batches = [stack(K, 300 + 37 * i, 40 + i) for i in range(6)]
    want = [te.hafnian(As, backend=BE, check_symmetric=False) for As in batches]
    got, errors = [None] * 12, []

    def host_call(i):
        try:
            got[i] = te.hafnian(batches[i % 6], backend=BE, check_symmetric=False)
        except Exception as e:          # reported below
            errors.append(repr(e))

    def device_call(i):
        try:
            As = batches[i % 6]
            P = synth.pair_batch(As)
            Pd = hip.DeviceBuffer.from_array(P)
            out = hip.DeviceBuffer(As.shape[0] * 8)
            st = hip.Stream()
            for _ in range(3):
                te._haf_dp_device(BE, plan, np.float64, Pd.addr, As.shape[0], As.shape[0], out.addr,
                                  st.handle)
            st.synchronize()
            got[i] = out.download(np.float64, As.shape[0])
        except Exception as e:
            errors.append(repr(e))

    threads = [threading.Thread(target=host_call if i < 6 else device_call, args=(i,)) for i in range(12)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i in range(12):
        assert got[i] is not None and got[i].tobytes() == want[i % 6].tobytes(), i
