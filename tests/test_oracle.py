"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible slip in the restatement -- a wrong pair index, a
transposed operand, a dropped level, a wrong sort key, an FMA where Julia has none -- fails one
of them: the reference's own tables and docstring examples, the hand-derived K=4 plan, an
independent pure-Python restatement, brute-force matching enumeration, exact rational and
128-bit integer arithmetic, and identities of the hafnian (closed forms for graph families,
haf([0 B; B^T 0]) = perm(B), permutation invariance, expansion along a row).
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as o

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def rng_sym(seed, N, kind="real"):
    r = np.random.default_rng(seed)
    B = r.standard_normal((N, N))
    if kind == "complex":
        B = B + 1j * r.standard_normal((N, N))
    return B + B.T


def haf_f64(A, dtype=o.F64, nthreads=1, ib=None):
    K = A.shape[0]
    return o.haf_dp(o.gather_pairs(A).reshape(-1, 1), K, dtype, nthreads, ib)[0]


def haf_int(A):
    A = np.asarray(A, dtype=np.int64)
    return o.haf_dp(o.gather_pairs(A).reshape(-1, 1), A.shape[0], o.I128)[0]


# ----------------------------------------------------------------------------- the plan

@pytest.mark.parametrize("K", [12, 16, 20, 24, 28, 32])
def test_counts_match_reference_table(K):
    # src/dp.jl:34-37, and nstates = F(K+1) (src/dp.jl:26)
    ns, nt = o.dp_counts(K)
    assert nt == GOLD["reference_ntrans"][str(K)]
    fib = [0, 1]
    while len(fib) < K + 2:
        fib.append(fib[-1] + fib[-2])
    assert ns == fib[K + 1]
    assert (ns, nt) == o.py_dp_counts(K)


@pytest.mark.parametrize("K", [24, 26, 28, 30, 32])
def test_plan_sizes_match_reference_table(K):
    ns, nt = o.dp_counts(K)
    assert round((nt + ns + 1) * 4 / 2**20, 1) == GOLD["reference_plan_mib"][str(K)]


def test_k4_plan_hand_derived():
    g = GOLD["k4_plan"]
    for key, q in g["pairs"].items():
        i, j = map(int, key.split(","))
        assert o.pair_index(i, j, 4) == q
    p = o.build_plan(4, 4)
    assert p.starts.tolist() == g["starts"]
    assert p.trans.tolist() == g["trans"]
    assert p.levels.tolist() == g["levels"]


@pytest.mark.parametrize("K", range(2, 25, 2))
@pytest.mark.parametrize("ib", [4, 8])
def test_plan_hash_c_equals_python_restatement(K, ib):
    assert o.build_plan(K, ib).sha256() == GOLD["plan_sha256"][f"{K}_{ib}"]


@pytest.mark.parametrize("K", [4, 8, 12, 16])
@pytest.mark.parametrize("ib", [4, 8])
def test_plan_invariants(K, ib):
    # test/runtests.jl:241-267
    p = o.build_plan(K, ib)
    ns, nt = o.dp_counts(K)
    bits = o.pair_bits(ib)
    assert p.nstates == ns and len(p.trans) == nt
    assert p.starts[0] == 1 and p.starts[-1] == nt + 1
    assert (np.diff(p.starts) >= 0).all()
    for s in range(2, ns + 1):
        ch = p.trans[p.starts[s - 1] - 1:p.starts[s] - 1] >> bits
        assert (ch < s).all()
    assert p.starts[1] == 1                                  # state 1 is the empty set
    assert p.starts[ns] - p.starts[ns - 1] == K - 1          # the full set, last
    q = p.trans & ((1 << bits) - 1)
    assert ((q >= 1) & (q <= K * (K - 1) // 2)).all()


def test_packing_limits():
    # test/runtests.jl:222-239, 295-297: DP_MAX is exactly the Int32 packing limit
    ns, _ = o.dp_counts(32)
    assert 496 <= (1 << 9) - 1 and ns <= (2**31 - 1) >> 9
    ns2, _ = o.dp_counts(34)
    assert (34 * 33 // 2).bit_length() + ns2.bit_length() > 31
    with pytest.raises(ValueError):
        o.build_plan(34, 4)
    with pytest.raises(ValueError):
        o.build_plan(66, 8)


# ----------------------------------------------------------------------------- values

def test_docstring_examples():
    A, want = GOLD["docstring"]["hafnian_4x4"]
    assert haf_f64(np.array(A, float)) == want
    assert haf_int(A) == want
    A2, rpt, want2 = GOLD["docstring"]["hafnian_repeated"]
    A2 = np.array(A2, float)
    idx = [i + 1 for i, r in enumerate(rpt) for _ in range(r)]
    P = o.gather_pairs(A2, idx).reshape(-1, 1)
    assert o.haf_dp(P, len(idx), o.F64)[0] == want2
    assert haf_f64(np.array([[0.0, 1.0], [1.0, 0.0]])) == 1.0


def test_exact_golden_values():
    for case in GOLD["exact"]:
        assert haf_int(case["A"]) == case["haf"]


@pytest.mark.parametrize("N", [2, 4, 6, 8, 10])
def test_brute_force(N):
    # test/runtests.jl:75-83
    A = rng_sym(N, N)
    assert haf_f64(A) == pytest.approx(o.brute_hafnian(A.tolist()), rel=1e-12, abs=1e-12)
    Ac = rng_sym(N + 100, N, "complex")
    v = o.haf_dp(o.gather_pairs(Ac).reshape(-1, 1), N, o.C64)[0]
    w = o.brute_hafnian(Ac.tolist())
    assert abs(v - w) <= 1e-12 * max(1.0, abs(w))


def test_diagonal_ignored():
    A = rng_sym(7, 8)
    C = A.copy()
    np.fill_diagonal(C, 3 * np.diag(A) + 7)
    assert haf_f64(A) == haf_f64(C)


@pytest.mark.parametrize("N", [12, 14, 16])
def test_f64_against_exact_rationals(N):
    # dyadic inputs: Fractions evaluate the same walk exactly
    r = np.random.default_rng(N)
    A = np.round(r.uniform(-1, 1, (N, N)) * 2**20) / 2**20
    A = np.triu(A, 1)
    A = A + A.T
    exact = o.py_haf_dp([Fraction(float(x)) for x in o.gather_pairs(A)], o.py_build_plan(N))
    v = haf_f64(A)
    assert abs(Fraction(float(v)) - exact) <= Fraction(1, 10**13) * abs(exact)
    hi, lo = o.haf_dp(o.gather_pairs(A).reshape(-1, 1), N, o.Q)[0]
    assert abs(Fraction(float(hi)) + Fraction(float(lo)) - exact) <= Fraction(1, 10**28) * abs(exact)


def test_i128_matches_python_ints():
    for N in (8, 12, 16):
        A = np.random.default_rng(N).integers(-5, 6, (N, N))
        A = np.triu(A, 1)
        A = A + A.T
        want = o.py_haf_dp([int(x) for x in o.gather_pairs(A)], o.py_build_plan(N))
        assert haf_int(A) == want


def test_i128_wraps_exactly_past_64_bits():
    # all-ones times 2^15: haf = (N-1)!! * 2^(15 N/2) ~ 2^103 at N = 12, far beyond 64 bits
    N = 12
    A = np.full((N, N), 1 << 15, np.int64)
    assert haf_int(A) == math.prod(range(N - 1, 0, -2)) * (1 << (15 * N // 2))
    assert haf_int(-A) == math.prod(range(N - 1, 0, -2)) * (1 << (15 * N // 2))   # N/2 even
    B = A.copy()
    B[0, 1] = B[1, 0] = -(1 << 15)     # one negated edge: the signed result is still exact
    want = -(1 << (15 * N // 2)) * math.prod(range(N - 3, 0, -2)) + \
        (1 << (15 * N // 2)) * (math.prod(range(N - 1, 0, -2)) - math.prod(range(N - 3, 0, -2)))
    assert haf_int(B) == want


# ---------------------------------------------------- single precision and the complex quad walk

def small_int_sym(seed, N, lo, hi):
    A = np.random.default_rng(seed).integers(lo, hi + 1, (N, N))
    A = np.triu(A, 1)
    return A + A.T


@pytest.mark.parametrize("N,m", [(10, 3), (12, 2), (14, 1)])
def test_f32_walk_exact_on_small_integers(N, m):
    # entries in [-m, m]: every state value, product and partial sum is an integer of magnitude
    # at most (N-1)!! * m^(N/2) < 2^24, so float32 arithmetic is exact and the Float32 walk must
    # equal the 128-bit integer walk (a wrong pair, child or level changes the integer)
    assert math.prod(range(N - 1, 0, -2)) * m ** (N // 2) < 2 ** 24
    for seed in range(3):
        A = small_int_sym(100 * N + seed, N, -m, m)
        v = haf_f64(A.astype(np.float32), o.F32)
        assert v.dtype == np.float32
        assert int(v) == v and int(v) == haf_int(A)
        # the Int64 layout walks the same sums
        assert haf_f64(A.astype(np.float32), o.F32, ib=8) == v


@pytest.mark.parametrize("N", [10, 12, 14])
def test_c32_walk_exact_on_gaussian_integers(N):
    # entries a + bi with a, b in {-1, 0, 1}: |state| <= (N-1)!! * sqrt(2)^(N/2) < 2^24, so the
    # ComplexF32 walk is exact and must equal the same walk over Python complex numbers, whose
    # double arithmetic is exact for these integers (a slip in (ac - bd, ad + bc) fails it)
    assert math.prod(range(N - 1, 0, -2)) * 2 ** (N / 4) < 2 ** 24
    for seed in range(3):
        A = small_int_sym(200 * N + seed, N, -1, 1) + 1j * small_int_sym(300 * N + seed, N, -1, 1)
        v = haf_f64(A.astype(np.complex64), o.C32)
        assert v.dtype == np.complex64
        want = o.py_haf_dp([complex(x) for x in o.gather_pairs(A)], o.py_build_plan(N))
        assert complex(v) == want
        assert want == complex(haf_int(A.real.astype(np.int64)), 0) or A.imag.any()


class _CFrac:
    """Exact complex rationals, for the complex walk over Fractions."""

    def __init__(self, re, im=Fraction(0)):
        self.re, self.im = Fraction(re), Fraction(im)

    def __add__(self, o_):
        return _CFrac(self.re + o_.re, self.im + o_.im)

    def __mul__(self, o_):
        return _CFrac(self.re * o_.re - self.im * o_.im, self.re * o_.im + self.im * o_.re)


@pytest.mark.parametrize("N", [12, 16])
def test_cq_walk_against_exact_complex_rationals(N):
    # complex dyadic inputs: the complex __float128 walk is within its 113-bit rounding of the
    # exact value of the same walk over complex Fractions, component by component
    r = np.random.default_rng(400 + N)
    A = (np.round(r.uniform(-1, 1, (N, N)) * 2**12) + 1j * np.round(r.uniform(-1, 1, (N, N)) * 2**12)) / 2**12
    A = np.triu(A, 1)
    A = A + A.T
    P = o.gather_pairs(A)
    exact = o.py_haf_dp([_CFrac(x.real, x.imag) for x in P], o.py_build_plan(N),
                        one=_CFrac(1), zero=_CFrac(0))
    hi, lo = o.haf_dp(P.reshape(-1, 1), N, o.CQ)[0]
    scale = abs(exact.re) + abs(exact.im)
    got_re = Fraction(float(hi.real)) + Fraction(float(lo.real))
    got_im = Fraction(float(hi.imag)) + Fraction(float(lo.imag))
    assert abs(got_re - exact.re) <= Fraction(1, 10**28) * scale
    assert abs(got_im - exact.im) <= Fraction(1, 10**28) * scale
    # and the ComplexF64 walk within double rounding of it
    v = haf_f64(A, o.C64)
    assert abs(Fraction(float(v.real)) - exact.re) <= Fraction(1, 10**12) * scale
    assert abs(Fraction(float(v.imag)) - exact.im) <= Fraction(1, 10**12) * scale


def double_factorial(n):
    return math.prod(range(n, 0, -2))


def graph(N, edges):
    A = np.zeros((N, N), np.int64)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    return A


@pytest.mark.parametrize("N", [2, 8, 16, 20])
def test_closed_form_complete_graph(N):
    assert haf_int(np.ones((N, N), np.int64)) == double_factorial(N - 1)


@pytest.mark.parametrize("n", [2, 4, 7, 9])
def test_closed_form_families(n):
    # complete bipartite K_{n,n}: n!; cycle C_{2n}: 2; ladder P_2 x P_n: F(n+1)
    Kbip = np.zeros((2 * n, 2 * n), np.int64)
    Kbip[:n, n:] = 1
    Kbip[n:, :n] = 1
    assert haf_int(Kbip) == math.factorial(n)
    cyc = graph(2 * n, [(i, (i + 1) % (2 * n)) for i in range(2 * n)])
    assert haf_int(cyc) == 2
    ladder = graph(2 * n, [(i, i + 1) for i in range(n - 1)] +
                   [(n + i, n + i + 1) for i in range(n - 1)] + [(i, n + i) for i in range(n)])
    fib = [1, 1]
    while len(fib) < n + 2:
        fib.append(fib[-1] + fib[-2])
    assert haf_int(ladder) == fib[n]


def test_closed_form_petersen_and_cube():
    petersen = graph(10, [(i, (i + 1) % 5) for i in range(5)] +
                     [(5 + i, 5 + (i + 2) % 5) for i in range(5)] + [(i, i + 5) for i in range(5)])
    assert haf_int(petersen) == 6
    cube = graph(8, [(a, a ^ (1 << k)) for a in range(8) for k in range(3) if a < a ^ (1 << k)])
    assert haf_int(cube) == 9


def brute_perm(B):
    import itertools
    n = len(B)
    return sum(math.prod(B[i][p[i]] for i in range(n)) for p in itertools.permutations(range(n)))


@pytest.mark.parametrize("n", [3, 5, 6, 8])
def test_bipartite_block_is_permanent(n):
    r = np.random.default_rng(n)
    B = r.integers(-2, 3, (n, n))
    if n <= 6:
        assert o.perm_ryser(B) == brute_perm(B.tolist())
    A = np.zeros((2 * n, 2 * n), np.int64)
    A[:n, n:] = B
    A[n:, :n] = B.T
    assert haf_int(A) == o.perm_ryser(B)


def test_permutation_invariance_exact():
    N = 18
    r = np.random.default_rng(5)
    A = r.integers(0, 2, (N, N))
    A = np.triu(A, 1)
    A = A + A.T
    v = haf_int(A)
    for s in range(3):
        pi = np.random.default_rng(100 + s).permutation(N)
        assert haf_int(A[np.ix_(pi, pi)]) == v


def test_row_expansion_exact():
    # haf(A) = sum_{j != i} A[i, j] haf(A with rows/cols i, j removed)
    N = 12
    A = np.random.default_rng(9).integers(-3, 4, (N, N))
    A = np.triu(A, 1)
    A = A + A.T
    total = 0
    for j in range(1, N):
        keep = [k for k in range(N) if k not in (0, j)]
        total += int(A[0, j]) * haf_int(A[np.ix_(keep, keep)])
    assert total == haf_int(A)


# ----------------------------------------------------------------------------- policies

@pytest.mark.parametrize("N", [16, 20, 24])
def test_threading_is_bit_exact(N):
    # test/runtests.jl:300-316
    A = rng_sym(200 + N, N, "complex")
    P = o.gather_pairs(A).reshape(-1, 1)
    ref = o.haf_dp(P, N, o.C64, 1)[0]
    for nt in (2, 3, 7):
        assert o.haf_dp(P, N, o.C64, nt)[0].tobytes() == ref.tobytes()


def test_batch_policy_equals_serial():
    N, B = 14, 9
    As = [rng_sym(300 + b, N) for b in range(B)]
    P = o.gather_pairs_batch(As)
    serial = np.array([o.haf_dp(P[:, b:b + 1], N, o.F64)[0] for b in range(B)])
    for nt in (1, 4, 16):
        assert o.haf_dp(P, N, o.F64, nt).tobytes() == serial.tobytes()


@pytest.mark.parametrize("K", [8, 12, 16])
def test_int32_and_int64_layouts_bit_identical(K):
    # test/runtests.jl:269-281
    A = rng_sym(400 + K, K, "complex")
    P = o.gather_pairs(A).reshape(-1, 1)
    assert o.haf_dp(P, K, o.C64, 1, 4).tobytes() == o.haf_dp(P, K, o.C64, 1, 8).tobytes()


def no_fma_case():
    """A 4x4 matrix whose hafnian, summed in plan order, rounds differently with a fused
    multiply-add: the full set sums P12*A34 + P13*A24 + P14*A23 in ascending j, and with
    A12 = 1, A34 = -1, A13 = A24 = x = 1 + 2^-30, A14 = 0 the second step is -1 + x*x, which is
    2^-29 when x*x is rounded first (Julia) and 2^-29 + 2^-60 when fused."""
    x = 1.0 + 2.0**-30
    A = np.zeros((4, 4))
    for (i, j), v in {(0, 1): 1.0, (2, 3): -1.0, (0, 2): x, (1, 3): x}.items():
        A[i, j] = A[j, i] = v
    return A, 2.0**-29, 2.0**-29 + 2.0**-60


def test_no_fma_in_restatement():
    A, unfused, fused = no_fma_case()
    assert unfused != fused
    assert haf_f64(A) == unfused
