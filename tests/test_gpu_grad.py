"""The reverse sweep (d haf / d A) on an MI355X, checked against what the mathematics fixes.

* Deletion identity: d haf / d a_ij = haf(A with rows and columns i, j removed), evaluated by the
  CPU oracle at degree N-2 -- exactly, on the 128-bit integer path.
* Central differences with h = 1e-6.  The hafnian is affine in each symmetric pair variable, so
  the central difference has no truncation error: it equals the derivative up to the rounding of
  the two forward values, whose size is measured against the __float128 oracle and sets the
  tolerance.
* Homogeneity (sum_{i<j} a_ij d haf/d a_ij = (N/2) haf) and the fold of a repeated-row gradient.
"""
import sys

import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

pytestmark = pytest.mark.gpu
BE = te.HIPBackend()


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


@pytest.mark.parametrize("N", [8, 12, 16])
def test_exact_gradient_is_the_deletion_hafnian(N):
    note(f"exact grad N={N}")
    A = np.random.default_rng(N).integers(-3, 4, (N, N)).astype(np.int64)
    A = np.triu(A, 1)
    A = A + A.T
    P = te._gather_pairs_batch(np.int64, A, None, N)
    vals, dP = te._haf_dp_backend_grad(BE, P, te._dp_plan(N))
    assert vals[0] == o.haf_dp(o.gather_pairs(A).reshape(-1, 1), N, o.I128)[0]
    q = 0
    for i in range(N):
        for j in range(i + 1, N):
            assert dP[0][q] == deletion(A, i, j, o.I128), (i, j)
            q += 1
    # homogeneity, exactly
    assert sum(int(A[i, j]) * dP[0][te._pair_index(i + 1, j + 1, N) - 1]
               for i in range(N) for j in range(i + 1, N)) * 2 == N * vals[0]


def sparse_adjacency(N, config, b, sparsity):
    # 0/1 adjacency of a G(N, 2^-sparsity) graph: the AND of independent G(N, 1/2) graphs
    A = synth.gnp_adjacency(N, config, b)
    for x in range(1, sparsity):
        A = A & synth.gnp_adjacency(N, config + 100 * x, b)
    return A


@pytest.mark.parametrize("T,N,B,sparsity", [(np.float64, 26, 70, 1), (np.float64, 28, 1024, 1),
                                            (np.float32, 26, 70, 2), (np.complex64, 26, 70, 4),
                                            (np.float32, 26, 300, 2), (np.complex64, 26, 300, 4)])
def test_fused_gradient_is_exact_on_integer_matrices(T, N, B, sparsity):
    # 0/1 adjacency matrices (Gaussian integers a + bi, a, b in {0, 1}, for complex): every value
    # and partial sum of the walk and of the reverse sweep is a non-negative integer (a Gaussian
    # integer) of at most the hafnian of a subgraph, far below 2^53 for G(N, 1/2), and below 2^24
    # for the sparser graphs of the single-precision cases (G(26, 1/4) real, G(26, 1/16) per
    # component complex: about 10^5 and 10^3), so the
    # floating-point gradient of the fused reverse step (degrees 26 and up) must equal the exact
    # one bit for bit whatever its summation order -- through the fused kernel (passes of four
    # tiles of 64 and more), through the stripe lists (passes of 64 to 255) and through the
    # small-batch kernels (passes of 7, and of 1 and 3, whose small levels gather with one wave per
    # child)
    note(f"fused grad exact {np.dtype(T)} N={N} B={B}")
    if np.dtype(T).kind == "c":
        As = np.stack([sparse_adjacency(N, 71, b, sparsity) + 1j * sparse_adjacency(N, 72, b, sparsity)
                       for b in range(B)])
    else:
        As = np.stack([sparse_adjacency(N, 71, b, sparsity) for b in range(B)]).astype(np.float64)
    P = te._gather_pairs_batch(T, As, None, N)
    plan = te._dp_plan(N)
    runs = [te._haf_dp_backend_grad(BE, P, plan)]
    for mb in ((64, 7, 3, 1) if B <= 100 else (256, 64) if B < 1024 else (512, 64)):
        runs.append(te._haf_dp_backend_grad(BE, P, plan, max_batch=mb))
    if np.dtype(T).kind == "c":
        # the deletion hafnians of a few instances from the oracle's complex128 walk (exact here)
        want = np.zeros((3, N * (N - 1) // 2), np.complex128)
        for b in range(3):
            q = 0
            for i in range(N):
                for j in range(i + 1, N):
                    want[b, q] = deletion(As[b], i, j, o.C64)
                    q += 1
        for v, g in runs:
            assert np.array_equal(np.asarray(g)[:3].astype(np.complex128), want)
        return
    Pi = te._gather_pairs_batch(np.int64, As.astype(np.int64), None, N)
    vi, gi = te._haf_dp_backend_grad(BE, Pi, plan)
    want_v = np.array([float(x) for x in vi], T)
    want_g = np.array([[float(x) for x in row] for row in gi], T)
    for v, g in runs:
        assert np.asarray(v).tobytes() == want_v.tobytes()
        assert np.asarray(g).tobytes() == want_g.tobytes()


@pytest.mark.parametrize("T,N", [(np.float64, 16), (np.float64, 20), (np.complex128, 16)])
def test_float_gradient_matches_deletion_identity(T, N):
    note(f"grad {np.dtype(T)} N={N}")
    A = (synth.sym_complex_normal(N, 30, 0) if T == np.complex128 else synth.sym_uniform(N, 30, 0))
    val, G = te.hafnian_grad(A.astype(T))
    assert val == te.hafnian(A.astype(T))                    # same forward, bit for bit
    D = np.zeros((N, N), T)
    for i in range(N):
        for j in range(i + 1, N):
            D[i, j] = D[j, i] = deletion(A, i, j, o.C64 if T == np.complex128 else o.F64)
    scale = np.abs(D).max()
    err = np.abs(G - D).max() / scale
    note(f"  max |G - deletion| / max|deletion| = {err:.2e}")
    assert err < 1e-12
    assert (np.diag(G) == 0).all()
    assert np.allclose(G, G.T, rtol=0, atol=0)
    lhs = np.sum(np.triu(A, 1) * np.triu(G, 1))
    assert abs(lhs - N / 2 * val) <= 1e-11 * max(1.0, abs(N / 2 * val), np.abs(A).sum() * scale)


def test_config4_batched_gradient_n24():
    # 256 real Float64 N=24 matrices, Uniform(-1, 1): full gradient, checked on a sample against
    # the deletion identity and central differences (h = 1e-6)
    note("config 4: N=24 B=256 forward+gradient")
    N, B, h = 24, 256, 1e-6
    As = np.stack([synth.sym_uniform(N, 4, b) for b in range(B)])
    vals, Gs = te.hafnian_grad(As, check_symmetric=False)
    assert vals.tobytes() == te.hafnian(As, check_symmetric=False).tobytes()
    rng = np.random.default_rng(4)
    pairs = [(0, 1), (0, 23), (5, 17), (11, 12), (22, 23)]
    for b in (0, 100, 255):
        A = As[b]
        for (i, j) in pairs:
            d = deletion(A, i, j, o.F64)
            assert abs(Gs[b, i, j] - d) <= 1e-12 * max(abs(d), np.abs(Gs[b]).max() * 1e-3)
        # central differences, batched: A +- h E_ij for each sampled pair
        Ap, Am = [], []
        for (i, j) in pairs:
            E = np.zeros((N, N))
            E[i, j] = E[j, i] = h
            Ap.append(A + E)
            Am.append(A - E)
        fp = te.hafnian(np.stack(Ap), check_symmetric=False)
        fm = te.hafnian(np.stack(Am), check_symmetric=False)
        # rounding of the forward values, measured against the __float128 oracle
        qp = o.haf_dp(synth.pair_batch(np.stack(Ap)), N, o.Q, 8)
        qm = o.haf_dp(synth.pair_batch(np.stack(Am)), N, o.Q, 8)
        for k, (i, j) in enumerate(pairs):
            fd = (fp[k] - fm[k]) / (2 * h)
            exact_fd = ((qp[k, 0] - qm[k, 0]) + (qp[k, 1] - qm[k, 1])) / (2 * h)
            rounding = (abs(fp[k] - qp[k, 0] - qp[k, 1]) + abs(fm[k] - qm[k, 0] - qm[k, 1])) / (2 * h)
            assert abs(exact_fd - Gs[b, i, j]) <= 1e-9 * max(1.0, abs(Gs[b, i, j]))
            assert abs(fd - Gs[b, i, j]) <= 4 * rounding + 1e-9 * max(1.0, abs(Gs[b, i, j]))
    _ = rng


@pytest.mark.parametrize("N", [16, 24, 28])
def test_gradient_is_deterministic_and_chunk_independent(N):
    B = 70
    note(f"grad chunks N={N}")
    As = np.stack([synth.sym_uniform(N, 31, b) for b in range(B)])
    P = te._gather_pairs_batch(np.float64, As, None, N)
    v1, g1 = te._haf_dp_backend_grad(BE, P, te._dp_plan(N))
    v2, g2 = te._haf_dp_backend_grad(BE, P, te._dp_plan(N))
    v3, g3 = te._haf_dp_backend_grad(BE, P, te._dp_plan(N), max_batch=7)
    assert v1.tobytes() == v2.tobytes() == v3.tobytes()
    assert np.asarray(g1).tobytes() == np.asarray(g2).tobytes() == np.asarray(g3).tobytes()
    # passes of 1, 32 and 33 instances: one lane per child, 32 lanes, and the 64-wide kernels
    for mb in (1, 32, 33):
        v, g = te._haf_dp_backend_grad(BE, P, te._dp_plan(N), max_batch=mb)
        assert v.tobytes() == v1.tobytes(), mb
        assert np.asarray(g).tobytes() == np.asarray(g1).tobytes(), mb


@pytest.mark.parametrize("N,T", [(26, np.float64), (28, np.float64), (28, np.float32)])
def test_fused_kernel_and_stripe_lists_agree(N, T):
    # a pass of five tiles takes the fused kernel, passes of 64 the stripe lists, passes of 256
    # (four tiles) and of 44 one each, passes of 7 the small-batch kernels: the same sums in the
    # same order, so the same bits on random matrices
    B = 300
    note(f"fused vs lists N={N} {np.dtype(T)}")
    As = np.stack([synth.sym_uniform(N, 33, b) for b in range(B)]).astype(T)
    P = te._gather_pairs_batch(T, As, None, N)
    plan = te._dp_plan(N)
    v0, g0 = te._haf_dp_backend_grad(BE, P, plan)
    for mb in (256, 64, 7):
        v, g = te._haf_dp_backend_grad(BE, P, plan, max_batch=mb)
        assert np.asarray(v).tobytes() == np.asarray(v0).tobytes(), mb
        assert np.asarray(g).tobytes() == np.asarray(g0).tobytes(), mb


def test_repeated_gradient_folds_the_expanded_gradient():
    A = synth.sym_uniform(6, 32, 0)
    np.fill_diagonal(A, synth.uniform(synth.stream(32, 1), 6))
    rpt = [2, 1, 3, 0, 1, 1]
    val, G = te.hafnian_repeated_grad(A, rpt)
    R = te.reduction(A, rpt)
    vR, GR = te.hafnian_grad(R)
    assert val == vR
    idx = np.repeat(np.arange(6), rpt)
    F = np.zeros((6, 6))
    n = len(idx)
    for a in range(n):
        for b in range(a + 1, n):
            u, v = idx[a], idx[b]
            F[u, v] += GR[a, b]
            if u != v:
                F[v, u] += GR[a, b]
    assert np.allclose(F, G, rtol=1e-13, atol=1e-13)


def test_rrule_pullback():
    A = synth.sym_uniform(12, 33, 0)
    y, pb = te.rrule(te.hafnian, A)
    _, G = te.hafnian_grad(A)
    assert y == te.hafnian(A)
    none, dA = pb(2.5)
    # the cotangent of a function of a symmetric matrix: each half of a pair carries G_ij / 2
    assert none is None and np.array_equal(dA, 2.5 * G / 2)
    # a symmetric tangent: <dA_bar, E> = sum_{i<j} G_ij E_ij, the directional derivative
    E = synth.sym_uniform(12, 33, 1)
    assert np.isclose(np.sum(dA / 2.5 * E), np.sum(np.triu(G, 1) * np.triu(E, 1)), rtol=1e-13)


@pytest.mark.parametrize("T", [np.float64, np.complex128])
def test_rrule_through_symmetrising_chain(T):
    # f(X) = hafnian(X + X^T) for a non-symmetric X: the pullback of X -> X + X^T is
    # dX = dA + dA^T, and dX[i, j] must be the derivative of f in X[i, j] -- central differences
    # (exact up to rounding: f is affine in each entry), conjugated for complex X (ChainRules'
    # dy * conj(f') convention)
    N, h = 10, 1e-4
    r = np.random.default_rng(34)
    X = r.uniform(-1, 1, (N, N)).astype(T)
    if np.dtype(T).kind == "c":
        X = X + 1j * r.uniform(-1, 1, (N, N))
    _, pb = te.rrule(te.hafnian, X + X.T)
    _, dA = pb(1.0)
    dX = dA + dA.T
    for (i, j) in [(0, 1), (3, 7), (8, 2), (9, 0)]:
        E = np.zeros((N, N), T)
        E[i, j] = h
        fp = te.hafnian((X + E) + (X + E).T)
        fm = te.hafnian((X - E) + (X - E).T)
        fd = (fp - fm) / (2 * h)
        assert abs(np.conj(dX[i, j]) - fd) <= 1e-8 * max(1.0, abs(fd)), (i, j)


@pytest.mark.parametrize("N,B", [(28, 4), (32, 1)])
def test_exact_gradient_large_graphs_sampled_deletion(N, B):
    # G(N, 1/2) adjacency matrices on the 128-bit path: sampled entries equal the deletion
    # hafnian exactly, and the full gradient satisfies homogeneity exactly
    note(f"exact grad graphs N={N} B={B}")
    As = np.stack([synth.gnp_adjacency(N, 61, b) for b in range(B)]).astype(np.int64)
    P = te._gather_pairs_batch(np.int64, As, None, N)
    vals, dP = te._haf_dp_backend_grad(BE, P, te._dp_plan(N))
    want = o.haf_dp(synth.pair_batch(As).astype(np.int64), N, o.I128, 8)
    pairs = [(0, 1), (0, N - 1), (3, N // 2), (N // 2 - 1, N // 2), (N - 2, N - 1)]
    for b in range(B):
        assert vals[b] == want[b]
        for (i, j) in pairs[: 5 if N <= 28 else 2]:
            assert dP[b][te._pair_index(i + 1, j + 1, N) - 1] == deletion(As[b], i, j, o.I128), (b, i, j)
        hom = sum(int(As[b][i, j]) * dP[b][te._pair_index(i + 1, j + 1, N) - 1]
                  for i in range(N) for j in range(i + 1, N))
        assert hom * 2 == N * vals[b]


def test_float_gradient_n28_sampled_deletion():
    note("grad f64 N=28 B=64")
    N, B = 28, 64
    As = np.stack([synth.sym_normal(N, 62, b, 1 / np.sqrt(N)) for b in range(B)])
    vals, Gs = te.hafnian_grad(As, check_symmetric=False)
    assert vals.tobytes() == te.hafnian(As, check_symmetric=False).tobytes()
    for b in (0, 37, 63):
        scale = np.abs(Gs[b]).max()
        for (i, j) in [(0, 1), (0, 27), (9, 20), (13, 14), (26, 27)]:
            d = deletion(As[b], i, j, o.F64)
            assert abs(Gs[b, i, j] - d) <= 1e-12 * scale, (b, i, j)
        lhs = np.sum(np.triu(As[b], 1) * np.triu(Gs[b], 1))
        assert abs(lhs - N / 2 * vals[b]) <= 1e-11 * max(1.0, np.abs(As[b]).sum() * scale)


@pytest.mark.parametrize("lo,hi", [(np.float32, np.float64), (np.complex64, np.complex128)])
def test_single_precision_gradient_follows_double(lo, hi):
    # the 4- and 8-byte element paths of the reverse sweep: the single-precision gradient of a
    # matrix exactly representable in single precision agrees with the double-precision one to
    # single-precision rounding
    note(f"grad {np.dtype(lo)} vs {np.dtype(hi)} N=20")
    N = 20
    A = synth.sym_complex_normal(N, 63, 0) if hi == np.complex128 else synth.sym_uniform(N, 63, 0)
    A = A.astype(lo).astype(hi)
    v_lo, G_lo = te.hafnian_grad(A.astype(lo))
    v_hi, G_hi = te.hafnian_grad(A)
    assert G_lo.dtype == lo
    scale = np.abs(G_hi).max()
    assert np.abs(G_lo.astype(hi) - G_hi).max() <= 1e-4 * scale
    assert abs(v_lo - v_hi) <= 1e-4 * max(abs(v_hi), np.abs(A).sum() * scale / N)


@pytest.mark.parametrize("N", [16, 22])
def test_exact_gradient_same_on_both_layouts(N):
    # the Int32 layout takes the parent-major dP pass, the Int64 layout the pair-major segments;
    # on the 128-bit integer path both must give the same integers
    note(f"exact grad Int32 vs Int64 N={N}")
    As = np.stack([synth.gnp_adjacency(N, 64, b) for b in range(5)]).astype(np.int64)
    As[:, 0, 1] = As[:, 1, 0] = 3                      # a non-0/1 entry
    P = te._gather_pairs_batch(np.int64, As, None, N)
    v32, g32 = te._haf_dp_backend_grad(BE, P, te._dp_plan(N, np.int32))
    v64, g64 = te._haf_dp_backend_grad(BE, P, te._dp_plan(N, np.int64))
    assert list(v32) == list(v64)
    assert [list(r) for r in g32] == [list(r) for r in g64]
    for b in (0, 4):
        assert g64[b][te._pair_index(2, N, N) - 1] == deletion(As[b], 1, N - 1, o.I128)


@pytest.mark.parametrize("N,T,I", [(20, np.float64, np.int32), (28, np.float64, np.int32),
                                   (18, np.complex128, np.int32), (16, np.int64, np.int32),
                                   (20, np.float64, np.int64)])
def test_checkpointed_gradient_is_bit_identical(N, T, I):
    # keeping only checkpoint levels and recomputing the others must not change a bit
    note(f"checkpointed grad N={N} {np.dtype(T)} {np.dtype(I)}")
    B = 70
    if T == np.int64:
        As = np.stack([synth.gnp_adjacency(N, 65, b) for b in range(B)]).astype(np.int64)
    elif T == np.complex128:
        As = np.stack([synth.sym_complex_normal(N, 65, b) for b in range(B)])
    else:
        As = np.stack([synth.sym_uniform(N, 65, b) for b in range(B)])
    P = te._gather_pairs_batch(T, As, None, N)
    plan = te._dp_plan(N, I)
    kept_rows, seg_rows, kept = te.grad_levels(N)
    assert not all(kept[1:])                           # some level is recomputed
    code = te.tuning.dtype_code(T)
    ib = np.dtype(I).itemsize
    v0, g0 = te._haf_dp_backend_grad(BE, P, plan)
    v1, g1 = te._haf_dp_backend_grad(BE, P, plan, max_batch=16)

    def entries(flags):   # the tuning entries the two call shapes look up (B=70 and B=16)
        for bb in (B, 16):
            _lib.check(_lib.lib().eggman_tune_set(N, ib, code, bb, 64, 0, flags, 0), "tune_set")
    entries(te.tuning.LDS | te.tuning.CKPT)
    try:
        v2, g2 = te._haf_dp_backend_grad(BE, P, plan)
        v3, g3 = te._haf_dp_backend_grad(BE, P, plan, max_batch=16)
    finally:
        entries(te.tuning.LDS)     # launch choices only change speed, never results
    if T == np.int64:
        assert list(v0) == list(v1) == list(v2) == list(v3)
        assert [list(r) for r in g0] == [list(r) for r in g2] == [list(r) for r in g3]
    else:
        assert v0.tobytes() == v1.tobytes() == v2.tobytes() == v3.tobytes()
        assert np.asarray(g0).tobytes() == np.asarray(g2).tobytes() == np.asarray(g3).tobytes()


def test_fused_gradient_with_64bit_level_offsets():
    # one pass of 4352 instances puts the largest N=28 level's rows (125970 states) beyond 4 GB of
    # their base, so the fused reverse step addresses them with 64-bit offsets; its gradient must
    # be the same bits as in passes of 512 instances (32-bit offsets)
    N, B = 28, 4352
    plan = te._dp_plan(N)
    assert int(np.diff(plan.levels).max()) * B * 8 >= 1 << 32
    note(f"fused grad 64-bit rows N={N} B={B}")
    As = np.stack([synth.sym_uniform(N, 73, b) for b in range(B)])
    P = te._gather_pairs_batch(np.float64, As, None, N)
    code, ib = te.tuning.dtype_code(np.float64), plan.index_bytes
    # the entry B's bucket resolves to now (its own, else the nearest tuned bucket below), so that
    # only the pass size changes and the entry can be put back as it was
    prev = next((e for bb in (B, 4096, 2048, 1024, 512, 256, 128, 64)
                 if (e := te.tuning.tune_get(N, np.float64, bb)) is not None), None)
    assert prev is not None, "no tuned entry for N=28 Float64"

    def entry(chunk):
        _lib.check(_lib.lib().eggman_tune_set(N, ib, code, B, prev["bt"], chunk, prev["flags"],
                                              prev["occ"]), "tune_set")
    entry(B)
    try:
        v1, g1 = te._haf_dp_backend_grad(BE, P, plan)
    finally:
        entry(prev["chunk"])
    v2, g2 = te._haf_dp_backend_grad(BE, P, plan, max_batch=512)
    assert np.asarray(v1).tobytes() == np.asarray(v2).tobytes()
    assert np.asarray(g1).tobytes() == np.asarray(g2).tobytes()


def test_grad_batch_forms_agree():
    # a (B, N, N) stack, a list of the same matrices and one matrix at a time give the same bits;
    # G is symmetric with a zero diagonal; a stack with one asymmetric matrix is refused
    note("hafnian_grad batch forms: N=12 B=5")
    As = np.stack([synth.sym_normal(12, 9, b, 0.3) for b in range(5)])
    v1, G1 = te.hafnian_grad(As)
    v2, G2 = te.hafnian_grad(list(As))
    assert v1.tobytes() == v2.tobytes() and G1.tobytes() == G2.tobytes()
    for b in range(5):
        v, G = te.hafnian_grad(As[b])
        assert v == v1[b] and G.tobytes() == G1[b].tobytes()
    assert np.array_equal(G1, G1.transpose(0, 2, 1))
    assert not np.diagonal(G1, axis1=1, axis2=2).any()
    bad = As.copy()
    bad[3, 0, 1] += 1.0
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(1, 2\)"):
        te.hafnian_grad(bad)
