"""Mixed-degree batches on an MI355X: one call over matrices of different sizes, each
result bit-identical to the single-degree path and to the oracle."""
import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


def test_mixed_small_sizes_match_single_calls():
    sizes = [4, 6, 16, 0, 5, 20, 16, 24, 2, 12, 20]
    As = [synth.sym_uniform(n, 40, i) if n else np.zeros((0, 0)) for i, n in enumerate(sizes)]
    got = te.hafnian_mixed(As)
    want = np.array([te.hafnian(A) for A in As])
    assert got.tobytes() == want.tobytes()
    for A, v in zip(As, got):
        n = A.shape[0]
        if n >= 2 and n % 2 == 0:
            assert o.haf_dp(o.gather_pairs(A).reshape(-1, 1), n, o.F64)[0] == v


def test_mixed_exact_and_complex():
    As = [synth.gnp_adjacency(n, 41, i) for i, n in enumerate([8, 12, 16, 20, 12])]
    Ps = [te._gather_pairs_batch(np.int64, A, None, A.shape[0]) for A in As]
    res = te._haf_dp_backend_mixed(te.HIPBackend(), Ps, [te._dp_plan(A.shape[0]) for A in As])
    for A, r in zip(As, res):
        assert r == [o.haf_dp(o.gather_pairs(A).reshape(-1, 1), A.shape[0], o.I128)[0]]
    Cs = [synth.sym_complex_normal(n, 42, i) for i, n in enumerate([10, 14, 18])]
    assert te.hafnian_mixed(Cs).tobytes() == np.array([te.hafnian(C) for C in Cs]).tobytes()


def test_mixed_2048_degrees_16_to_32():
    # 2048 real matrices with degrees drawn uniformly from {16, 20, 24, 28, 32}, one pass
    note("mixed 2048 x {16..32}")
    rng = np.random.default_rng(3)
    degs = rng.choice([16, 20, 24, 28, 32], size=2048)
    As = [synth.sym_normal(int(n), 5, i, 1 / np.sqrt(n)) for i, n in enumerate(degs)]
    got = te.hafnian_mixed(As, check_symmetric=False)
    for n in (16, 20, 24, 28, 32):
        idx = np.flatnonzero(degs == n)
        same = te.hafnian(np.stack([As[i] for i in idx]), check_symmetric=False)
        assert got[idx].tobytes() == same.tobytes(), n
        sample = idx[:2]
        P = synth.pair_batch(np.stack([As[i] for i in sample]))
        assert got[sample].tobytes() == o.haf_dp(P, n, o.F64, 8).tobytes(), n


def test_mixed_large_and_small_groups():
    # groups of 256 or more instances walk on their own and the others share one walk; one call
    # holding both kinds, the matrices of each degree scattered through the list
    note("mixed: large and small groups")
    sizes = np.array([16] * 300 + [20] * 5 + [12] * 9 + [24] * 260 + [8] * 3)
    np.random.default_rng(44).shuffle(sizes)
    As = [synth.sym_uniform(int(n), 44, i) for i, n in enumerate(sizes)]
    got = te.hafnian_mixed(As, check_symmetric=False)
    for n in (8, 12, 16, 20, 24):
        idx = np.flatnonzero(sizes == n)
        same = te.hafnian(np.stack([As[i] for i in idx]), check_symmetric=False)
        assert got[idx].tobytes() == same.tobytes(), n
    P = synth.pair_batch(np.stack([As[i] for i in np.flatnonzero(sizes == 20)]))
    assert got[sizes == 20].tobytes() == o.haf_dp(P, 20, o.F64, 8).tobytes()


def test_mixed_chunking_is_bit_identical():
    # the host entry point cuts a call into chunks that take the same share of every group
    # (max_batch instances of a group per chunk, ragged shares included); concurrent large groups
    # and the shared walk of the small ones give the same bits for every chunking, and the exact
    # type is chunked the same way
    note("mixed chunking")
    sizes = np.array([16] * 270 + [20] * 300 + [12] * 7 + [24] * 33)
    As = [synth.sym_uniform(int(n), 45, i) for i, n in enumerate(sizes)]
    whole = te.hafnian_mixed(As, check_symmetric=False)
    for mb in (1, 7, 64, 299):
        got = te.hafnian_mixed(As, check_symmetric=False, max_batch=mb)
        assert got.tobytes() == whole.tobytes(), mb
    Gs = [synth.gnp_adjacency(n, 46, i) for i, n in enumerate([12] * 5 + [16] * 3)]
    Ps = [te._gather_pairs_batch(np.int64, np.stack(Gs[:5]), None, 12),
          te._gather_pairs_batch(np.int64, np.stack(Gs[5:]), None, 16)]
    plans = [te._dp_plan(12), te._dp_plan(16)]
    a = te._haf_dp_backend_mixed(te.HIPBackend(), Ps, plans)
    b = te._haf_dp_backend_mixed(te.HIPBackend(), Ps, plans, max_batch=2)
    assert a == b
    assert a[1] == o.haf_dp(synth.pair_batch(np.stack(Gs[5:])).astype(np.int64), 16, o.I128)
