"""The host entry points' staging pipeline on an MI355X: pinned two-slot buffers, uploads and
downloads on their own streams overlapping compute, and chunk sizes chosen from the tuned pass
size and free device memory when the caller gives no max_batch.  Chunking never changes a bit
(test/gpu.jl:26-56 asks the same of max_batch): every result here is compared with the oracle or
with the one-chunk call, bit for bit."""
import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, synth

pytestmark = pytest.mark.gpu
BE = te.HIPBackend()


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


def note(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("K,B,pass_", [(16, 3001, 256), (20, 1000, 128), (12, 70, 16)])
def test_pipelined_forward_matches_the_oracle(K, B, pass_):
    # a tuned pass well below B: the call splits itself into about four chunks (the last ragged)
    note(f"pipeline forward K={K} B={B} pass={pass_}")
    te.tune_set(K, np.float64, B, bt=64, chunk=pass_, occ=2, lds=True, vec=True)
    As = np.stack([synth.sym_uniform(K, 70, b) for b in range(B)])
    P = te._gather_pairs_batch(np.float64, As, None, K)
    got = te._haf_dp_backend(BE, P, te._dp_plan(K))
    want = o.haf_dp(synth.pair_batch(As), K, o.F64, 8)
    assert got.tobytes() == want.tobytes()
    # an explicit max_batch that does not divide B, and the whole batch at once
    for mb in (B // 3 + 1, B):
        assert te._haf_dp_backend(BE, P, te._dp_plan(K), max_batch=mb).tobytes() == want.tobytes()


def test_pipelined_forward_complex_and_exact():
    K, B = 14, 900
    te.tune_set(K, np.complex128, B, bt=64, chunk=64, occ=2, lds=True, narrow=True)
    As = np.stack([synth.sym_complex_normal(K, 71, b) for b in range(B)])
    P = te._gather_pairs_batch(np.complex128, As, None, K)
    got = te._haf_dp_backend(BE, P, te._dp_plan(K))
    assert got.tobytes() == o.haf_dp(synth.pair_batch(As), K, o.C64, 8).tobytes()
    te.tune_set(K, np.int64, B, bt=64, chunk=64, occ=2, lds=True)
    G = np.stack([synth.gnp_adjacency(K, 71, b) for b in range(B)]).astype(np.int64)
    Pi = te._gather_pairs_batch(np.int64, G, None, K)
    exact = te._haf_dp_backend(BE, Pi, te._dp_plan(K))
    want = o.haf_dp(o.gather_pairs_batch(G), K, o.I128, 8)
    assert [int(v) for v in exact] == [int(v) for v in want]


def test_pipelined_gradient_is_chunk_independent():
    K, B = 16, 1500
    note(f"pipeline gradient K={K} B={B}")
    te.tune_set(K, np.float64, B, bt=64, chunk=128, occ=2, lds=True, vec=True)
    As = np.stack([synth.sym_uniform(K, 72, b) for b in range(B)])
    P = te._gather_pairs_batch(np.float64, As, None, K)
    plan = te._dp_plan(K)
    v_all, g_all = te._haf_dp_backend_grad(BE, P, plan, max_batch=B)
    for mb in (None, 700, 512):
        v, g = te._haf_dp_backend_grad(BE, P, plan, max_batch=mb)
        assert v.tobytes() == v_all.tobytes(), mb
        assert np.ascontiguousarray(g).tobytes() == np.ascontiguousarray(g_all).tobytes(), mb
    # the values are the forward's, and a sampled gradient entry is the deletion hafnian
    assert v_all.tobytes() == o.haf_dp(synth.pair_batch(As), K, o.F64, 8).tobytes()
    b, i, j = 777, 3, 11
    keep = [k for k in range(K) if k not in (i, j)]
    d = o.haf_dp(o.gather_pairs(As[b][np.ix_(keep, keep)]).reshape(-1, 1), K - 2, o.F64)[0]
    q = te._pair_index(i + 1, j + 1, K) - 1
    assert abs(g_all[b, q] - d) <= 1e-12 * max(1.0, abs(d))


def test_pipelined_repeated_patterns():
    # the boson-sampling shape: one matrix, many index lists, split into several chunks
    N, K, B = 24, 16, 2500
    te.tune_set(K, np.float64, B, bt=64, chunk=256, occ=2, lds=True, vec=True)
    A = synth.sym_uniform(N, 73, 0)
    rng = np.random.default_rng(73)
    idxs = np.stack([np.sort(rng.choice(N, K, replace=True)) + 1 for _ in range(B)]).astype(np.int32)
    got = te.backend._haf_dp_backend_repeated(BE, A, idxs, te._dp_plan(K))
    whole = te.backend._haf_dp_backend_repeated(BE, A, idxs, te._dp_plan(K), max_batch=B)
    assert got.tobytes() == whole.tobytes()
    P = o.gather_pairs_batch(np.broadcast_to(A, (B, N, N)), idxs)
    assert got.tobytes() == o.haf_dp(P, K, o.F64, 8).tobytes()
