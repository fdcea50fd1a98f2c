"""Degrees 36 and 40 on the Int64 plan layout, generated in HBM (21 GiB at K=40).

The oracle cannot walk these degrees in test time, so parity rests on committed exact counts of
eight G(40, 1/2) graphs (the oracle's 128-bit walk, run once by tests/golden/make_n40_exact.py)
and on identities that fix the exact answer and exercise the whole plan: closed forms of graph families, the permanent of a
bipartite block, the product rule for a block-diagonal matrix whose blocks the oracle evaluates
at half the degree, all hidden behind a random relabeling of the vertices -- and invariance under
relabeling itself, bit-exact on the 128-bit integer path.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

pytestmark = pytest.mark.gpu
GEN = te.HIPBackend(plan_source="generate")
# exact counts of eight of the 64 graphs, from the C oracle's 128-bit walk over its own host-built
# K=40 plan (tests/golden/make_n40_exact.py)
GOLD40 = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                     "n40_exact.json")))


def note(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


def relabel(A, seed):
    pi = np.random.default_rng(seed).permutation(A.shape[0])
    return A[np.ix_(pi, pi)]


def exact(As):
    return te.hafnian_exact(As, backend=GEN, check_symmetric=False)


@pytest.mark.parametrize("N", [36, 40])
def test_closed_forms(N):
    note(f"closed forms N={N}")
    n = N // 2
    Kbip = np.zeros((N, N), np.int64)
    Kbip[:n, n:] = 1
    Kbip[n:, :n] = 1
    cyc = np.zeros((N, N), np.int64)
    for i in range(N):
        cyc[i, (i + 1) % N] = cyc[(i + 1) % N, i] = 1
    ladder = np.zeros((N, N), np.int64)
    for i in range(n - 1):
        ladder[i, i + 1] = ladder[i + 1, i] = ladder[n + i, n + i + 1] = ladder[n + i + 1, n + i] = 1
    for i in range(n):
        ladder[i, n + i] = ladder[n + i, i] = 1
    fib = [1, 1]
    while len(fib) < n + 2:
        fib.append(fib[-1] + fib[-2])
    mats = [np.ones((N, N), np.int64), relabel(Kbip, 1), relabel(cyc, 2), relabel(ladder, 3)]
    got = exact(mats)
    assert got == [math.prod(range(N - 1, 0, -2)), math.factorial(n), 2, fib[n]]


def test_n40_bipartite_block_is_permanent():
    n = 20
    B = (synth.splitmix64(synth.stream(50, 0), n * n) >> np.uint64(63)).astype(np.int64).reshape(n, n)
    A = np.zeros((2 * n, 2 * n), np.int64)
    A[:n, n:] = B
    A[n:, :n] = B.T
    assert exact([relabel(A, 4)])[0] == o.perm_ryser(B)


def test_n40_block_diagonal_product():
    G1, G2 = synth.gnp_adjacency(20, 51, 0), synth.gnp_adjacency(20, 51, 1)
    A = np.zeros((40, 40), np.int64)
    A[:20, :20] = G1
    A[20:, 20:] = G2
    h1 = o.haf_dp(o.gather_pairs(G1).reshape(-1, 1), 20, o.I128)[0]
    h2 = o.haf_dp(o.gather_pairs(G2).reshape(-1, 1), 20, o.I128)[0]
    assert exact([relabel(A, 5)])[0] == h1 * h2


def test_n40_64_random_graphs_exact_and_relabel_invariant():
    # 64 G(40, 1/2) graphs: perfect-matching counts on the 128-bit path; every count must be
    # unchanged, bit for bit, by a random relabeling of the graph's vertices
    note("64 x G(40,1/2) exact")
    As = [synth.gnp_adjacency(40, 52, b) for b in range(64)]
    counts = exact(As)
    assert all(isinstance(c, int) and 0 <= c < math.prod(range(39, 0, -2)) for c in counts)
    for b, c in GOLD40["counts"].items():
        assert counts[int(b)] == c, b
    again = exact([relabel(A, 100 + b) for b, A in enumerate(As)])
    assert again == counts
    note(f"  counts[0:3] = {counts[:3]}")


def test_n40_complex_single_block_diagonal():
    # one ComplexF64 matrix at N=40 (entries (g1 + i g2)/sqrt(2N)), as a relabeled block-diagonal
    # of two N=20 blocks: haf = haf(A1) * haf(A2), each block by the oracle
    note("N=40 complex single")
    A1, A2 = synth.sym_complex_normal(20, 53, 0), synth.sym_complex_normal(20, 53, 1)
    A = np.zeros((40, 40), complex)
    A[:20, :20] = A1
    A[20:, 20:] = A2
    h1 = o.haf_dp(o.gather_pairs(A1).reshape(-1, 1), 20, o.C64)[0]
    h2 = o.haf_dp(o.gather_pairs(A2).reshape(-1, 1), 20, o.C64)[0]
    got = te.hafnian(relabel(A, 6), method="dp", backend=GEN)
    assert abs(got - h1 * h2) <= 1e-12 * abs(h1 * h2)
    # and a dense one: relabeling changes every summation path, the value only by rounding
    D = synth.sym_complex_normal(40, 2, 0)
    v = te.hafnian(D, method="dp", backend=GEN)
    w = te.hafnian(relabel(D, 7), method="dp", backend=GEN)
    assert abs(v - w) <= 1e-10 * abs(v)


def test_n36_uploaded_plan_equals_generated():
    note("K=36 host plan vs device plan")
    hp = te._dp_plan(36, np.int64)
    h = te._device_plan(GEN, te.PlanRef(36, 8))
    got = np.empty_like(hp.trans)
    _lib.check(_lib.lib().eggman_plan_download(h, got.ctypes.data, got.size))
    assert np.array_equal(got, hp.trans)
    A = synth.sym_uniform(36, 54, 0)
    up = te.hafnian(A, method="dp")
    gen = te.hafnian(A, method="dp", backend=GEN)
    assert up == gen
