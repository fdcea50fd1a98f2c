"""The autotuner and the graph-replay mode on an MI355X: every launch choice must give the same
bits as the oracle, and a tuned or replayed call must equal an eager one."""
import numpy as np
import pytest

import oracle as o
import theeggman as te
from theeggman import _lib, hip, synth

pytestmark = pytest.mark.gpu
BE = te.HIPBackend()


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


@pytest.mark.parametrize("K,B", [(16, 200), (20, 1), (20, 24)])
def test_autotune_choice_is_bit_exact(K, B):
    rep = te.autotune(K, np.float64, B, reps=2)
    assert rep["candidates"] and "bt" in rep["choice"]
    assert te.tune_get(K, np.float64, B)["bt"] == rep["choice"]["bt"]
    As = np.stack([synth.sym_uniform(K, 60, b) for b in range(B)])
    got = te.hafnian(As)
    want = o.haf_dp(synth.pair_batch(As), K, o.F64, 8)
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("setting", [dict(bt=64, chunk=64, lds=1, occ=1),
                                     dict(bt=64, chunk=64, lds=1, occ=2, wide=1),
                                     dict(bt=64, chunk=128, lds=0, occ=0),
                                     dict(bt=32, chunk=0, lds=0, occ=0),
                                     dict(bt=64, chunk=0, lds=1, occ=2, graph=1),
                                     dict(bt=64, chunk=0, lds=0, occ=0, fused=1),
                                     dict(bt=1, chunk=0, lds=0, occ=0, graph=1, fused=1),
                                     dict(bt=64, chunk=64, lds=1, occ=2, narrow=1),
                                     dict(bt=64, chunk=0, lds=1, occ=4, wide=1, narrow=1),
                                     dict(bt=64, chunk=0, lds=1, occ=2, graph=1, narrow=1),
                                     dict(bt=64, chunk=64, lds=1, occ=2, vec=1),
                                     dict(bt=64, chunk=0, lds=1, occ=4, wide=1, vec=1),
                                     dict(bt=64, chunk=0, lds=1, occ=2, graph=1, vec=1),
                                     dict(bt=64, chunk=64, lds=1, occ=2, vec=1, deep=1),
                                     dict(bt=64, chunk=0, lds=1, occ=4, wide=1, vec=1, deep=1),
                                     dict(bt=64, chunk=64, lds=1, occ=2, vec=1, pairs=1),
                                     dict(bt=64, chunk=0, lds=1, occ=2, graph=1, vec=1, pairs=1),
                                     dict(bt=64, chunk=64, lds=1, occ=4, wide=1, vec=1, deep=1, pairs=1)])
def test_every_launch_choice_gives_the_same_bits(setting):
    K, B = 18, 150
    te.tune_set(K, np.float64, B, **setting)
    As = np.stack([synth.sym_uniform(K, 61, b) for b in range(B)])
    want = o.haf_dp(synth.pair_batch(As), K, o.F64, 8)
    for _ in range(3):                      # a graph captures on the first call, replays after
        assert te.hafnian(As).tobytes() == want.tobytes()


def test_graph_replay_on_device_buffers():
    K, B = 20, 1
    te.tune_set(K, np.complex128, B, bt=1, chunk=1, occ=0, lds=False, graph=True)
    A = synth.sym_complex_normal(K, 62, 0)
    P = synth.pair_batch(A[None])
    Pd = hip.DeviceBuffer.from_array(P)
    out = hip.DeviceBuffer(16)
    st = hip.Stream()
    want = o.haf_dp(P, K, o.C64)[0]
    for i in range(4):
        out.zero()
        te._haf_dp_device(BE, te._dp_plan(K), np.complex128, Pd.addr, 1, 1, out.addr, st.handle)
        st.synchronize()
        assert out.download(np.complex128, 1)[0].tobytes() == want.tobytes(), i
    # new input values in the same buffer: the replayed graph reads them
    A2 = synth.sym_complex_normal(K, 62, 1)
    P2 = synth.pair_batch(A2[None])
    Pd.upload(P2)
    te._haf_dp_device(BE, te._dp_plan(K), np.complex128, Pd.addr, 1, 1, out.addr, st.handle)
    st.synchronize()
    assert out.download(np.complex128, 1)[0].tobytes() == o.haf_dp(P2, K, o.C64)[0].tobytes()


def test_tuning_table_round_trip(tmp_path):
    te.tune_set(22, np.float64, 512, bt=64, chunk=256, occ=2, lds=True, wide=True)
    path = tmp_path / "tuning.json"
    te.save_tuning(str(path))
    assert te.load_tuning(str(path)) >= 1
    assert te.tune_get(22, np.float64, 512) == {"bt": 64, "chunk": 256, "flags": 1 | 8, "occ": 2,
                                                "lds": 1, "graph": 0, "fused": 0, "wide": 1,
                                                "ckpt": 0, "narrow": 0, "vec": 0, "deep": 0,
                                                "pairs": 0}


@pytest.mark.parametrize("K,B,T", [(28, 1024, np.float64), (24, 37, np.float64), (22, 130, np.complex128),
                                   (20, 100, np.float32), (16, 70, np.int64)])
def test_narrow_tiles_give_the_same_bits(K, B, T):
    # 16-instance tiles, 4 states per wave: ragged last tiles, every element type, and the
    # gradient's forward pass, against the 64-wide kernel and the oracle
    note = print
    note(f"narrow K={K} B={B} {np.dtype(T)}", flush=True)
    if T == np.int64:
        As = np.stack([synth.gnp_adjacency(K, 66, b) for b in range(B)]).astype(np.int64)
    elif T == np.complex128:
        As = np.stack([synth.sym_complex_normal(K, 66, b) for b in range(B)])
    else:
        As = np.stack([synth.sym_uniform(K, 66, b) for b in range(B)]).astype(T)
    P = te._gather_pairs_batch(T, As, None, K)
    plan = te._dp_plan(K)
    te.tune_set(K, T, B, bt=64, chunk=0, occ=2, lds=True)
    wide = te._haf_dp_backend(BE, P, plan)
    te.tune_set(K, T, B, bt=64, chunk=0, occ=2, lds=True, narrow=True)
    narrow = te._haf_dp_backend(BE, P, plan)
    if T == np.int64:
        assert list(narrow) == list(wide)
    else:
        assert np.asarray(narrow).tobytes() == np.asarray(wide).tobytes()
        if K <= 24:
            code = {np.float64: o.F64, np.float32: o.F32, np.complex128: o.C64}[T]
            want = o.haf_dp(synth.pair_batch(As), K, code, 8)
            assert np.asarray(narrow).tobytes() == want.tobytes()
    if T != np.int64 and B <= 130:
        _, g_narrow = te._haf_dp_backend_grad(BE, P, plan)
        te.tune_set(K, T, B, bt=64, chunk=0, occ=2, lds=True)
        _, g_wide = te._haf_dp_backend_grad(BE, P, plan)
        assert np.asarray(g_narrow).tobytes() == np.asarray(g_wide).tobytes()


@pytest.mark.parametrize("K,B,T", [(28, 1024, np.float64), (24, 37, np.float64), (22, 130, np.complex128),
                                   (20, 100, np.float32), (16, 70, np.int64), (12, 3, np.complex64),
                                   ("20/Int64", 300, np.float64)])
@pytest.mark.parametrize("chunk", [0, 48])
def test_vector_loads_give_the_same_bits(K, B, T, chunk):
    I = None
    if isinstance(K, str):                  # the Int64 plan layout at a small degree
        K, I = int(K.split("/")[0]), np.int64
    # 16-byte loads (two or more instances per lane): ragged last tiles and passes, padded level
    # rows, every element type and both plan layouts, against the 64-wide kernel and the oracle
    print(f"vec K={K} B={B} {np.dtype(T)} chunk={chunk}", flush=True)
    if T == np.int64:
        As = np.stack([synth.gnp_adjacency(K, 67, b) for b in range(B)]).astype(np.int64)
    elif T in (np.complex128, np.complex64):
        As = np.stack([synth.sym_complex_normal(K, 67, b) for b in range(B)]).astype(T)
    else:
        As = np.stack([synth.sym_uniform(K, 67, b) for b in range(B)]).astype(T)
    P = te._gather_pairs_batch(T, As, None, K)
    plan = te._dp_plan(K, I)
    ib = np.dtype(plan.trans.dtype).itemsize
    te.tune_set(K, T, B, bt=64, chunk=0, occ=2, lds=True, index_bytes=ib)
    wide = te._haf_dp_backend(BE, P, plan)
    te.tune_set(K, T, B, bt=64, chunk=chunk, occ=2, lds=True, vec=True, index_bytes=ib)
    got = te._haf_dp_backend(BE, P, plan)
    # batches of 6 transitions instead of 4 (every remainder of k mod 6 occurs over the levels)
    te.tune_set(K, T, B, bt=64, chunk=chunk, occ=2, lds=True, vec=True, deep=True, index_bytes=ib)
    deep = te._haf_dp_backend(BE, P, plan)
    # parent pairs sharing their children, in batches of 4 and of 6
    te.tune_set(K, T, B, bt=64, chunk=chunk, occ=2, lds=True, vec=True, pairs=True, index_bytes=ib)
    pairs = te._haf_dp_backend(BE, P, plan)
    te.tune_set(K, T, B, bt=64, chunk=chunk, occ=2, lds=True, vec=True, deep=True, pairs=True,
                index_bytes=ib)
    pairs6 = te._haf_dp_backend(BE, P, plan)
    if T == np.int64:
        assert list(got) == list(wide) == list(deep) == list(pairs) == list(pairs6)
    else:
        assert np.asarray(got).tobytes() == np.asarray(wide).tobytes()
        assert np.asarray(deep).tobytes() == np.asarray(wide).tobytes()
        assert np.asarray(pairs).tobytes() == np.asarray(wide).tobytes()
        assert np.asarray(pairs6).tobytes() == np.asarray(wide).tobytes()
        if K <= 24:
            code = {np.float64: o.F64, np.float32: o.F32, np.complex128: o.C64, np.complex64: o.C32}[T]
            want = o.haf_dp(synth.pair_batch(As), K, code, 8)
            assert np.asarray(got).tobytes() == want.tobytes()
    if T != np.int64 and B <= 300:
        # the gradient's forward pass on padded level rows: the same bits as without
        _, g_vec = te._haf_dp_backend_grad(BE, P, plan)     # (the deep pairs entry is still set)
        te.tune_set(K, T, B, bt=64, chunk=chunk, occ=2, lds=True, index_bytes=ib)
        _, g_wide = te._haf_dp_backend_grad(BE, P, plan)
        assert np.asarray(g_vec).tobytes() == np.asarray(g_wide).tobytes()
