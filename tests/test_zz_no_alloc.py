"""The no-allocation entry point against eggman_dp_forward_device, through the C ABI.

eggman_dp_forward_device_noalloc keeps its state levels in one static device array per GPU and
fits its passes to it; its values must be the bits of eggman_dp_forward_device on the same
inputs -- on consecutive calls, each later call with a smaller input than the one before (the
array keeps the previous call's contents), across the fused, one-pass and multi-pass walks and
the element types it serves, and it must refuse what the array cannot hold without launching.
"""
import ctypes

import numpy as np
import pytest

import theeggman as te
from theeggman import _lib, hip, synth

pytestmark = pytest.mark.gpu
BE = te.HIPBackend()


def note(msg):
    print(msg, flush=True)


@pytest.fixture(autouse=True)
def _gpu_present():
    if _lib.lib().eggman_device_count() < 1:
        pytest.skip("no GPU visible")


def _inputs(N, B, T, seed):
    if np.dtype(T).kind == "c":
        As = np.stack([synth.sym_complex_normal(N, seed, b) for b in range(B)])
    else:
        As = np.stack([synth.sym_normal(N, seed, b, 1 / np.sqrt(N)) for b in range(B)])
    return np.ascontiguousarray(synth.pair_batch(As).astype(T))


def _both(N, B, T, seed, st):
    """(values of the no-allocation entry, values of the allocating entry) on one input."""
    P = _inputs(N, B, T, seed)
    Pd = hip.DeviceBuffer.from_array(P)
    esz = np.dtype(T).itemsize
    o1, o2 = hip.DeviceBuffer(B * esz), hip.DeviceBuffer(B * esz)
    o1.zero()
    o2.zero()
    h = te._device_plan(BE, te._dp_plan(N))
    code = te.backend.dtype_code(T)
    _lib.check(_lib.lib().eggman_dp_forward_device_noalloc(h, code, ctypes.c_void_p(Pd.addr), B, B,
                                                           ctypes.c_void_p(o1.addr),
                                                           ctypes.c_void_p(st.handle)), "noalloc")
    st.synchronize()
    got = o1.download(T, B)
    _lib.check(_lib.lib().eggman_dp_forward_device(h, code, ctypes.c_void_p(Pd.addr), B, B,
                                                   ctypes.c_void_p(o2.addr),
                                                   ctypes.c_void_p(st.handle)), "device")
    st.synchronize()
    return got, o2.download(T, B)


# each sequence: calls in a row, every later call smaller than the one before
SEQUENCES = [
    [(28, 1024, np.float64), (28, 300, np.float64)],     # tuned passes; the array is reused
    [(24, 4096, np.float64), (20, 64, np.float64)],      # passes shrunk to fit, then a fused-range walk
    [(26, 200, np.complex128), (16, 8, np.complex128)],  # complex; then the fused one-launch walk
    [(30, 64, np.float32), (22, 3, np.float32)],
    [(32, 100, np.float64), (28, 1, np.float64)],        # one instance of N=32 needs 11 MiB per level pair
]


@pytest.mark.parametrize("seq", SEQUENCES, ids=lambda s: "-".join(f"{n}x{b}" for n, b, _ in s))
def test_noalloc_equals_the_allocating_entry(seq):
    st = hip.Stream()
    for k, (N, B, T) in enumerate(seq):
        note(f"noalloc N={N} B={B} {np.dtype(T)}")
        got, want = _both(N, B, T, 90 + k, st)
        assert got.tobytes() == want.tobytes(), (N, B)


def test_noalloc_refuses_what_does_not_fit():
    # one complex N=36 instance needs 2 x 5.3 M states x 16 B = 162 MiB: refused before any launch
    h = te._device_plan(te.HIPBackend(plan_source="generate"), te.PlanRef(36, 8))
    P = hip.DeviceBuffer(630 * 16)
    out = hip.DeviceBuffer(16)
    rc = _lib.lib().eggman_dp_forward_device_noalloc(h, _lib.C64, ctypes.c_void_p(P.addr), 1, 1,
                                                     ctypes.c_void_p(out.addr), None)
    assert rc == _lib.ENOMEM and "MiB" in _lib.last_error()
    # the exact path is not served
    h28 = te._device_plan(BE, te._dp_plan(28))
    rc = _lib.lib().eggman_dp_forward_device_noalloc(h28, _lib.I128, ctypes.c_void_p(P.addr), 1, 1,
                                                     ctypes.c_void_p(out.addr), None)
    assert rc == _lib.EUNSUPPORTED
    # and the entry point still works afterwards
    st = hip.Stream()
    got, want = _both(20, 10, np.float64, 99, st)
    assert got.tobytes() == want.tobytes()
