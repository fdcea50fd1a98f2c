"""The workload configurations at their full sizes on an MI355X, and the range guard of the exact
path.

* The target: N=28, B=1024 real Float64, forward and full gradient.  The values are the forward's
  bits; the gradient is the same bits when the batch is cut into passes of 64 (the gradient then
  runs other pass and sub-pass sizes); sampled entries equal the deletion identity (oracle at
  N=26); homogeneity holds for all 1024 gradients; the values are within 1e-10 of the
  __float128 oracle.
* 4096 real N=24 matrices, N(0, 1/N): a sample equals the oracle's bits, through the host entry
  point and through the device-resident entry point with the batch's own tuned launch, and the
  results follow any reordering of the batch bit for bit.
* One ComplexF64 matrix at N=36 on the Int64 plan generated in HBM: a relabeled block-diagonal
  matrix equals the product of its two N=18 blocks by the oracle, and a dense matrix is invariant
  under relabeling to 1e-10.
* The exact path refuses, before any walk, instances whose hafnian may not fit in 127 bits, at
  every entry point, and returns the exact value just inside the range.
"""
import math

import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, hip, synth

pytestmark = pytest.mark.gpu
BE = te.HIPBackend()
GEN = te.HIPBackend(plan_source="generate")


def note(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


def deletion(A, i, j, dtype):
    keep = [k for k in range(A.shape[0]) if k not in (i, j)]
    Ar = A[np.ix_(keep, keep)]
    return o.haf_dp(o.gather_pairs(Ar).reshape(-1, 1), len(keep), dtype, 8)[0]


def relabel(A, seed):
    pi = np.random.default_rng(seed).permutation(A.shape[0])
    return A[np.ix_(pi, pi)]


def test_target_n28_b1024_forward_and_gradient():
    note("target: N=28 B=1024 F64 forward + gradient")
    N, B = 28, 1024
    As = np.stack([synth.sym_normal(N, 3, b, 1 / np.sqrt(N)) for b in range(B)])
    P = te._gather_pairs_batch(np.float64, As, None, N)
    plan = te._dp_plan(N)
    vals, dP = te._haf_dp_backend_grad(BE, P, plan)
    fwd = te._haf_dp_backend(BE, P, plan)
    assert vals.tobytes() == fwd.tobytes()
    note("  passes of 64")
    v64, dP64 = te._haf_dp_backend_grad(BE, P, plan, max_batch=64)
    assert v64.tobytes() == vals.tobytes()
    assert np.asarray(dP64).tobytes() == np.asarray(dP).tobytes()
    # sampled entries against the deletion identity
    iu, ju = np.triu_indices(N, 1)
    for b in (0, 511, 1023):
        scale = np.abs(dP[b]).max()
        for (i, j) in [(0, 1), (0, 27), (9, 20), (26, 27)]:
            d = deletion(As[b], i, j, o.F64)
            q = te._pair_index(i + 1, j + 1, N) - 1
            assert abs(dP[b, q] - d) <= 1e-12 * scale, (b, i, j)
    # homogeneity for every instance: sum_{i<j} a_ij d haf / d a_ij = (N/2) haf
    lhs = np.einsum("bq,bq->b", P, dP)
    tol = 1e-11 * np.maximum(1.0, np.abs(P).sum(axis=1) * np.abs(dP).max(axis=1))
    assert (np.abs(lhs - N / 2 * vals) <= tol).all()
    # the values within 1e-10 of the __float128 oracle (the BigFloat stand-in)
    sample = [0, 100, 511, 777, 1023]
    q = o.haf_dp(synth.pair_batch(As[sample]), N, o.Q, 8)
    rel = np.abs(vals[sample] - (q[:, 0] + q[:, 1])) / np.abs(q[:, 0])
    note(f"  max rel err vs __float128 oracle {rel.max():.2e}")
    assert rel.max() <= 1e-10


def test_config3a_n24_b4096():
    note("config 3a: N=24 B=4096 F64")
    N, B = 24, 4096
    As = np.stack([synth.sym_normal(N, 3, b, 1 / np.sqrt(N)) for b in range(B)])
    got = te.hafnian(As, backend=BE, check_symmetric=False)
    sample = np.linspace(0, B - 1, 48).astype(int)
    want = o.haf_dp(synth.pair_batch(As[sample]), N, o.F64, 8)
    assert got[sample].tobytes() == want.tobytes()
    # the device-resident entry point runs the whole batch as one call with its tuned launch
    P = synth.pair_batch(As)
    Pd = hip.DeviceBuffer.from_array(P)
    out = hip.DeviceBuffer(B * 8)
    st = hip.Stream()
    te._haf_dp_device(BE, te._dp_plan(N), np.float64, Pd.addr, B, B, out.addr, st.handle)
    st.synchronize()
    assert out.download(np.float64, B).tobytes() == got.tobytes()
    perm = np.random.default_rng(3).permutation(B)
    assert te.hafnian(As[perm], backend=BE, check_symmetric=False).tobytes() == got[perm].tobytes()


def test_config5a_n36_complex_single():
    note("config 5a: one ComplexF64 N=36 matrix")
    A1, A2 = synth.sym_complex_normal(18, 55, 0), synth.sym_complex_normal(18, 55, 1)
    A = np.zeros((36, 36), complex)
    A[:18, :18] = A1
    A[18:, 18:] = A2
    h1 = o.haf_dp(o.gather_pairs(A1).reshape(-1, 1), 18, o.C64)[0]
    h2 = o.haf_dp(o.gather_pairs(A2).reshape(-1, 1), 18, o.C64)[0]
    got = te.hafnian(relabel(A, 8), method="dp", backend=GEN)
    assert abs(got - h1 * h2) <= 1e-12 * abs(h1 * h2)
    D = synth.sym_complex_normal(36, 2, 0)
    v = te.hafnian(D, method="dp", backend=GEN)
    w = te.hafnian(relabel(D, 9), method="dp", backend=GEN)
    note(f"  relabel rel diff {abs(v - w) / abs(v):.2e}")
    assert abs(v - w) <= 1e-10 * abs(v)


def test_exact_path_refuses_out_of_range_instances():
    note("exact range guard")
    big = np.full((6, 6), 1 << 62, np.int64)          # haf = 15 * 2^186
    with pytest.raises(OverflowError):
        te.hafnian_exact(big)
    with pytest.raises(OverflowError):
        te.hafnian_exact([np.ones((6, 6), np.int64), big])
    P = te._gather_pairs_batch(np.int64, big, None, 6)
    with pytest.raises(OverflowError):
        te._haf_dp_backend_grad(BE, P, te._dp_plan(6))
    with pytest.raises(OverflowError):
        te.backend._haf_dp_backend_repeated(BE, big, [[1, 2, 3, 4, 5, 6]], te._dp_plan(6))
    with pytest.raises(OverflowError):
        te._haf_dp_backend_mixed(BE, [te._gather_pairs_batch(np.int64, np.ones((4, 4), np.int64), None, 4), P],
                                 [te._dp_plan(4), te._dp_plan(6)])
    # the device-resident entry point checks on the device
    Pd = hip.DeviceBuffer.from_array(np.ascontiguousarray(P.T))
    out = hip.DeviceBuffer(16)
    with pytest.raises(OverflowError):
        te._haf_dp_device(BE, te._dp_plan(6), np.int64, Pd.addr, 1, 1, out.addr, 0)
    # just inside the range: K_40 with every edge 5, haf = 39!! * 5^20 ~ 2^124.5, exact
    A5 = np.full((40, 40), 5, np.int64)
    assert te.hafnian_exact(A5, backend=GEN, check_symmetric=False) == \
        math.prod(range(39, 0, -2)) * 5 ** 20
    # and the same on the device entry point, both limbs
    P5 = np.ascontiguousarray(synth.pair_batch(A5[None]).astype(np.int64))
    Pd5 = hip.DeviceBuffer.from_array(P5)
    te._haf_dp_device(GEN, te.PlanRef(40, 8), np.int64, Pd5.addr, 1, 1, out.addr, 0)
    hip.device_synchronize()
    lo, hi = out.download(np.uint64, 2)
    assert te.backend._limbs_to_int(int(lo), int(hi)) == math.prod(range(39, 0, -2)) * 5 ** 20


def test_gather_probe_reports_line_rates():
    # the measurement helper behind the benchmark's gather ceiling: a line rate for an L2-sized
    # and a MALL-sized working set, the smaller one at least as fast
    import ctypes
    rates = []
    for ws in (2 << 20, 128 << 20):
        r = ctypes.c_double()
        _lib.check(_lib.lib().eggman_gather_rate(ws, ctypes.byref(r)), "gather_rate")
        rates.append(r.value * 128)
        note(f"  gather probe {ws >> 20} MiB: {r.value * 128 / 1e12:.2f} TB/s")
    assert 1e11 < rates[1] <= rates[0] * 1.05
    assert _lib.lib().eggman_gather_rate(64, ctypes.byref(ctypes.c_double())) == _lib.EINVAL
