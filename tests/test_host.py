"""Host-side checks of the product that need no GPU: the C ABI exports, the library's host plan
builder and validator against the oracle, and the router's validation and refusals (which
raise before any device call)."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "golden.json")))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "eggman_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(eggman_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 10
    for name in syms:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(syms)


def test_version_and_device_count():
    L = _lib.lib()
    assert b"gfx950" in L.eggman_version()
    assert L.eggman_device_count() >= 0


@pytest.mark.parametrize("K", list(range(2, 65, 2)))
def test_counts_equal_oracle(K):
    assert te._dp_counts(K) == o.py_dp_counts(K)


@pytest.mark.parametrize("K", list(range(2, 29, 2)))
@pytest.mark.parametrize("I", [np.int32, np.int64])
def test_host_builder_equals_oracle(K, I):
    p = te._build_dp_plan(K, I)
    q = o.build_plan(K, np.dtype(I).itemsize)
    assert np.array_equal(p.starts, q.starts)
    assert np.array_equal(p.trans, q.trans)
    assert np.array_equal(p.levels, q.levels)
    if K <= 24:
        assert q.sha256() == GOLD["plan_sha256"][f"{K}_{np.dtype(I).itemsize}"]


def test_host_builder_k32():
    p = te._build_dp_plan(32, np.int32)
    q = o.build_plan(32, 4)
    assert np.array_equal(p.trans, q.trans) and np.array_equal(p.starts, q.starts)


def test_plan_cache_and_layouts():
    # test/runtests.jl:263-264
    assert te._dp_plan(8) is te._dp_plan(8)
    assert te._dp_plan(8, np.int32) is not te._dp_plan(8, np.int64)
    assert te._dp_index_type(32) == np.int32 and te._dp_index_type(34) == np.int64


def test_build_refusals():
    # test/runtests.jl:293-297
    with pytest.raises(te.ArgumentError):
        te._build_dp_plan(66, np.int64)
    with pytest.raises(te.ArgumentError):
        te._build_dp_plan(50, np.int64)        # 3.3 TB: more than this machine has
    with pytest.raises(te.ArgumentError):
        te._build_dp_plan(34, np.int32)        # will not pack


def _plan_args(p):
    return (p.K, p.index_bytes, p.starts.ctypes.data, p.starts.size, p.trans.ctypes.data,
            p.trans.size, p.levels.ctypes.data, p.levels.size)


def test_validator_accepts_plans_and_rejects_corruption():
    L = _lib.lib()
    for K in (4, 12, 20):
        for I in (np.int32, np.int64):
            p = te._build_dp_plan(K, I)
            assert L.eggman_plan_validate(*_plan_args(p)) == 0
    p = te._build_dp_plan(12, np.int32)
    bad = te.HafnianDPPlan(p.K, p.nstates, p.starts.copy(), p.trans.copy(), p.levels.copy())
    bad.trans[100] = (bad.trans[100] & 511) | ((p.nstates + 5) << 9)   # child out of range
    assert L.eggman_plan_validate(*_plan_args(bad)) == _lib.EPLAN
    bad = te.HafnianDPPlan(p.K, p.nstates, p.starts.copy(), p.trans.copy(), p.levels.copy())
    bad.trans[7] = (bad.trans[7] >> 9) << 9                            # pair 0
    assert L.eggman_plan_validate(*_plan_args(bad)) == _lib.EPLAN
    bad = te.HafnianDPPlan(p.K, p.nstates, p.starts.copy(), p.trans.copy(), p.levels.copy())
    bad.starts[50] += 1                                                # ragged out-degree
    assert L.eggman_plan_validate(*_plan_args(bad)) == _lib.EPLAN
    bad = te.HafnianDPPlan(p.K, p.nstates, p.starts.copy(), p.trans.copy(), p.levels.copy())
    bad.levels[3] += 1
    assert L.eggman_plan_validate(*_plan_args(bad)) == _lib.EPLAN
    assert "level" in _lib.last_error()


def test_dp_batch_bytes():
    # test/gpu.jl:149-154
    ns, _ = te._dp_counts(20)
    assert te.dp_batch_bytes(20, np.complex128, 1) == (ns + 190) * 16
    assert te.dp_batch_bytes(20, np.complex64, 8) == 8 * (ns + 190) * 8
    assert _lib.lib().eggman_batch_bytes(20, _lib.C64, 1) == (ns + 190) * 16


def test_gather_matches_oracle():
    A = synth.sym_uniform(9, 7, 0)
    P = te._gather_pairs_batch(np.float64, A, None, 9)
    assert P.shape == (1, 36)
    assert np.array_equal(P[0], o.gather_pairs(A))
    idx = [1, 1, 2, 3, 3, 3]
    P = te._gather_pairs_batch(np.float64, A, [idx], 6)
    assert np.array_equal(P[0], o.gather_pairs(A, idx))
    As = np.stack([synth.sym_uniform(6, 7, b) for b in range(3)])
    P = te._gather_pairs_batch(np.float64, As, None, 6)
    assert P.flags.f_contiguous and np.array_equal(P.T, synth.pair_batch(As))


def test_router_validation_before_any_device_call():
    # src/hafnian.jl:1058-1066, 1111-1147, 1263-1301, 1317-1359; test/gpu.jl:110-119
    A = synth.sym_uniform(10, 1, 0)
    assert te.hafnian(np.zeros((0, 0))) == 1.0
    assert te.hafnian(synth.sym_uniform(5, 1, 1)) == 0.0
    with pytest.raises(te.DimensionMismatch):
        te.hafnian(np.zeros((3, 4)))
    with pytest.raises(te.ArgumentError):
        te.hafnian(A, method="fast")
    with pytest.raises(te.ArgumentError):
        te.hafnian(np.ones((66, 66)), method="dp")
    with pytest.raises(te.ArgumentError):
        te.hafnian(np.ones((14, 14)), method="unrolled")
    with pytest.raises(NotImplementedError):
        te.hafnian(A, method="sieve")
    with pytest.raises(NotImplementedError):
        te.hafnian(np.ones((34, 34)))           # :auto above DP_MAX is the CPU sieve
    bad = A.copy()
    bad[0, 1] += 1
    with pytest.raises(te.ArgumentError, match="not symmetric at"):
        te.hafnian(bad)
    with pytest.raises(te.ArgumentError):
        te.hafnian_repeated(A, [np.ones(10, int), np.r_[1, 1, 1, 1, 1, 1, 1, 1, 1, 0]])
    with pytest.raises(te.ArgumentError):
        te.hafnian_repeated(A, [np.ones(10, int), np.r_[-1, np.ones(9, int)]])
    with pytest.raises(te.DimensionMismatch):
        te.hafnian_repeated(A, [np.ones(10, int), np.ones(9, int)])
    with pytest.raises(te.DimensionMismatch):
        te.hafnian([A, synth.sym_uniform(12, 1, 2)])
    assert te.hafnian_repeated(A, np.zeros(10, int)) == 1.0
    assert te.hafnian_repeated(A, np.r_[1, np.zeros(9, int)]) == 0.0
    assert te.hafnian([]).shape == (0,)
    assert te.hafnian_exact(np.ones((3, 3), int)) == 0
    assert te.hafnian_exact(np.zeros((0, 0), int)) == 1
    with pytest.raises(te.ArgumentError):
        te.hafnian_exact(np.ones((4, 4)))       # not an integer matrix


def test_reduction():
    A = np.array([[1.0, 2.0], [2.0, 3.0]])
    R = te.reduction(A, [2, 2])
    assert R.shape == (4, 4) and R[0, 1] == 1.0 and R[0, 2] == 2.0 and R[2, 3] == 3.0


def test_synthetic_generator_is_stable():
    # fixed outputs of the documented generator (SplitMix64, 2^-53 uniform, 2^-30 normals)
    x = synth.splitmix64(0, 2)
    assert int(x[0]) == 0xE220A8397B1DCDAF and int(x[1]) == 0x6E789E6AA1B965F4
    u = synth.uniform(synth.stream(3, 5), 1000)
    assert (u >= -1).all() and (u < 1).all()
    g = synth.normal(synth.stream(2, 0), 20000)
    assert abs(g.mean()) < 0.03 and abs(g.std() - 1) < 0.03
    assert (np.round(g * 2**30) == g * 2**30).all()


def _level_rows(K):
    """State rows of levels 1..K/2 from the closed form: level L holds the 2(L-1)-subsets of
    {lo..K}, lo = (K - m + 1) // 2 + 1 (the counts the oracle's plan builder reproduces)."""
    from math import comb
    rows = [0]
    for L in range(1, K // 2 + 1):
        m = 2 * (L - 1)
        lo = (K - m + 1) // 2 + 1
        rows.append(comb(K - lo + 1, m))
    return rows


@pytest.mark.parametrize("K", [4, 8, 12, 16, 20, 24, 28])
def test_gradient_checkpoints_are_optimal(K):
    # brute force over every set of kept levels: level 1 kept, cost = kept rows + the largest
    # run of levels that are not kept; the library's choice must reach the minimum
    rows = _level_rows(K)
    n = K // 2
    assert sum(rows) == o.dp_counts(K)[0] - 1          # every level but the top one
    best = None
    for bits in range(1 << max(n - 1, 0)):
        kept = [False, True] + [bool(bits >> (L - 2) & 1) for L in range(2, n + 1)]
        seg = cur = 0
        for L in range(2, n + 1):
            cur = 0 if kept[L] else cur + rows[L]
            seg = max(seg, cur)
        cost = sum(rows[L] for L in range(1, n + 1) if kept[L]) + seg
        best = cost if best is None else min(best, cost)
    kr, sr, kept = te.grad_levels(K, checkpointed=True)
    assert kept[1]
    assert kr == sum(rows[L] for L in range(1, n + 1) if kept[L])
    seg = cur = 0
    for L in range(2, n + 1):
        cur = 0 if kept[L] else cur + rows[L]
        seg = max(seg, cur)
    assert sr == seg
    assert kr + sr == best
    kf, sf, kept_f = te.grad_levels(K, checkpointed=False)
    assert (kf, sf) == (sum(rows), 0) and all(kept_f[1:])


def test_gradient_checkpoints_save_memory_at_large_degree():
    for K in (28, 32, 36, 40):
        kr, sr, _ = te.grad_levels(K, checkpointed=True)
        full, _, _ = te.grad_levels(K, checkpointed=False)
        assert kr + sr < 0.75 * full, K


# ------------------------------------------------------------- range check of the exact walk

def _exact_bound(As):
    As = np.asarray(As, dtype=np.int64)
    K = As.shape[-1]
    P = np.ascontiguousarray(synth.pair_batch(As).astype(np.int64))
    B = P.shape[1]
    lb = np.empty(B)
    rc = _lib.lib().eggman_exact_bound(K, P.ctypes.data, B, B, lb.ctypes.data)
    return rc, lb


def test_exact_bound_is_an_upper_bound():
    # log2 |haf| <= the bound, with haf from the oracle's 128-bit walk (exact in this range),
    # for random signed matrices, sparse ones and the all-equal matrices that meet (N-1)!! m^(N/2)
    rng = np.random.default_rng(7)
    for N in (4, 8, 12, 16):
        for m in (1, 3, 100, 1 << 20):
            A = np.triu(rng.integers(-m, m + 1, (N, N)), 1)
            A = A + A.T
            if m == 3:
                A[rng.random((N, N)) < 0.5] = 0
                A = np.triu(A, 1) + np.triu(A, 1).T
            rc, lb = _exact_bound([A, np.full((N, N), m)])
            assert (rc == _lib.OK) == (lb.max() < 127 - 1e-6)
            if rc != _lib.OK:
                continue       # beyond the 128-bit walk's range, so no exact value to compare
            for M, b in ((A, lb[0]), (np.full((N, N), m), lb[1])):
                h = o.haf_dp(o.gather_pairs(M).astype(np.int64).reshape(-1, 1), N, o.I128)[0]
                assert h == 0 or abs(h).bit_length() - 1 <= b + 1e-9, (N, m)
                assert h == 0 or math.log2(abs(h)) <= b + 1e-9, (N, m)
    # the all-ones matrix meets the first bound: log2 (N-1)!!
    _, lb = _exact_bound([np.ones((16, 16), np.int64)])
    assert abs(lb[0] - np.log2(float(np.prod(np.arange(15, 0, -2))))) < 1e-9


def test_exact_bound_refuses_what_128_bits_cannot_hold():
    # every off-diagonal entry 2^62 at N=6: haf = 15 * 2^186, which the 128-bit walk wraps to 0
    A = np.full((6, 6), 1 << 62, np.int64)
    assert o.haf_dp(o.gather_pairs(A).reshape(-1, 1), 6, o.I128)[0] == 0
    rc, lb = _exact_bound([A])
    assert rc == _lib.EOVERFLOW and lb[0] > 186
    assert "127" in _lib.last_error()
    # K_40 with every edge 5: 39!! * 5^20 ~ 2^124.5 fits; with 6 (~2^129.8) it may not
    rc5, lb5 = _exact_bound([np.full((40, 40), 5, np.int64)])
    rc6, _ = _exact_bound([np.full((40, 40), 6, np.int64)])
    assert rc5 == _lib.OK and 124 < lb5[0] < 125 and rc6 == _lib.EOVERFLOW
    # G(40, 1/2) adjacency matrices are far inside the range (row-sum bound ~ 2^88)
    rc, lb = _exact_bound([synth.gnp_adjacency(40, 52, b) for b in range(4)])
    assert rc == _lib.OK and (lb < 90).all()
    # extreme entries do not break the bound's arithmetic
    rc, _ = _exact_bound([np.full((4, 4), np.iinfo(np.int64).min, np.int64)])
    assert rc == _lib.EOVERFLOW


@pytest.mark.parametrize("K", [4, 6, 12, 16, 20, 26, 28])
def test_fused_reverse_tables_follow_the_plan(K):
    # the gradient's fused reverse step reads each child's parents sorted by pair, phase by phase,
    # and sums dP over stripe term lists; the library rebuilds those tables from its host plan and
    # checks every entry and term against the plan's own transitions (a state's element set is its
    # child's plus the transition's pair), so every transition appears once in each
    ne, nt = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(_lib.lib().eggman_fuse_validate(K, ctypes.byref(ne), ctypes.byref(nt)), "fuse_validate")
    ntrans = te._dp_counts(K)[1]
    assert ne.value == ntrans and nt.value == ntrans


def test_symmetry_check_semantics():
    # isapprox's default tolerance on the strict triangle (src/hafnian.jl:765-783); the diagonal is
    # ignored; a batch reports the first matrix that fails, and an exactly symmetric stack passes
    from theeggman.hafnian import _check_symmetric
    A = synth.sym_uniform(8, 3, 0)
    _check_symmetric(A)
    B = A.copy()
    B[2, 5] *= 1 + 1e-17                      # within eps
    B[4, 4] = np.nan                          # diagonal: ignored
    _check_symmetric(B)
    C = A.copy()
    C[5, 2] += 1e-6
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(3, 6\)"):
        _check_symmetric(C)
    stack = np.stack([A, A, C])
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(3, 6\)"):
        te.hafnian(stack)
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(3, 6\)"):
        te.hafnian([A, C])
    Z = (A + 1j * A).astype(np.complex128)
    _check_symmetric(Z)
    Z2 = Z.copy()
    Z2[0, 7] += 1j * 1e-3
    with pytest.raises(te.ArgumentError):
        _check_symmetric(Z2)


def test_batch_refusals_before_the_device():
    # the batch forms name the first bad pattern or matrix before any device call
    A = np.ones((4, 4))
    with pytest.raises(te.DimensionMismatch, match="rpt 2 has length 3"):
        te.hafnian_repeated(A, [[1, 1, 0, 0], [1, 1, 0]])
    with pytest.raises(te.ArgumentError, match="rpt 2 contains a negative entry"):
        te.hafnian_repeated(A, [[1, 1, 0, 0], [3, -1, 0, 0]])
    with pytest.raises(te.ArgumentError, match="rpt 3 sums to 4, but rpt 1 sums to 2"):
        te.hafnian_repeated(A, [[1, 1, 0, 0], [0, 0, 1, 1], [1, 1, 1, 1]])
    S = synth.sym_normal(6, 1, 0, 0.3)
    bad = S.copy()
    bad[2, 4] += 0.5
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(3, 5\)"):
        te.hafnian([S, S, bad])
    with pytest.raises(te.DimensionMismatch, match="matrix 2 is 4×4, expected 6×6"):
        te.hafnian([S, np.ones((4, 4))])
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(3, 5\)"):
        te.hafnian_mixed([np.ones((4, 4)), S, bad])
    odd = np.zeros((5, 5))
    odd[0, 1] = 1.0
    with pytest.raises(te.ArgumentError, match=r"not symmetric at \(1, 2\)"):
        te.hafnian_mixed([S, odd])
