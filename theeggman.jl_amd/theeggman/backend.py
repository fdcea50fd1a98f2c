"""The device hook: Python mirror of ``_haf_dp_backend`` and the device caches
(src/dp.jl:514-587, ext/TheEggmanKernelAbstractionsExt.jl:28-146), implemented by the HIP
library through its C ABI.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from ._lib import (C32, C64, F32, F64, I128, ArgumentError, DimensionMismatch, check, lib)
from .plan import HafnianDPPlan, _dp_counts

_DTYPE_CODE = {np.dtype(np.float32): F32, np.dtype(np.float64): F64,
               np.dtype(np.complex64): C32, np.dtype(np.complex128): C64,
               np.dtype(np.int64): I128}


class HIPBackend:
    """The MI355X backend.  Any singleton works as a backend for the router, which only tests
    ``backend is not None`` (src/hafnian.jl:1000).

    ``plan_source`` chooses how a degree's plan reaches the device: ``"upload"`` copies the
    host-built HafnianDPPlan (the reference's _device_plan, ext:31-42); ``"generate"`` builds it
    directly in device memory, bit-identical, without a host copy.
    """

    def __init__(self, plan_source: str = "upload"):
        if plan_source not in ("upload", "generate"):
            raise ArgumentError(f"plan_source must be 'upload' or 'generate', got {plan_source!r}")
        self.plan_source = plan_source

    def __repr__(self):
        return f"HIPBackend(plan_source={self.plan_source!r})"

    def __eq__(self, other):
        return isinstance(other, HIPBackend) and other.plan_source == self.plan_source

    def __hash__(self):
        return hash(("HIPBackend", self.plan_source))


# Handles of device-resident plans keyed by (backend, K, layout), src/dp.jl:536.
_DEVICE_PLANS: dict = {}
_DEVICE_PLAN_LOCK = threading.RLock()


def dtype_code(dt) -> int:
    dt = np.dtype(dt)
    if dt not in _DTYPE_CODE:
        raise ArgumentError(f"element type {dt} has no device path")
    return _DTYPE_CODE[dt]


def _device_plan(backend: HIPBackend, plan: HafnianDPPlan) -> int:
    """Device copy of ``plan`` (ext/TheEggmanKernelAbstractionsExt.jl:31-42), cached."""
    key = (backend, plan.K, plan.index_bytes)
    with _DEVICE_PLAN_LOCK:
        h = _DEVICE_PLANS.get(key)
        if h is None:
            handle = ctypes.c_uint64()
            if backend.plan_source == "generate":
                check(lib().eggman_plan_generate(plan.K, plan.index_bytes, ctypes.byref(handle)),
                      "plan_generate")
            else:
                s = np.ascontiguousarray(plan.starts)
                t = np.ascontiguousarray(plan.trans)
                lv = np.ascontiguousarray(plan.levels)
                check(lib().eggman_plan_upload(plan.K, plan.index_bytes, s.ctypes.data, s.size,
                                               t.ctypes.data, t.size, lv.ctypes.data, lv.size,
                                               ctypes.byref(handle)), "plan_upload")
            h = handle.value
            _DEVICE_PLANS[key] = h
        return h


def _as_pair_major(P: np.ndarray) -> np.ndarray:
    """Julia's Matrix{T}(B, npairs) is column-major, i.e. [pair][batch] with the batch index
    fastest; a numpy (B, npairs) array in Fortran order has exactly that memory."""
    return np.asfortranarray(P)


def _haf_dp_backend(backend: HIPBackend, P: np.ndarray, plan: HafnianDPPlan, *,
                    max_batch: int | None = None):
    """Hafnians of the B instances gathered in ``P`` (shape (B, npairs)) on the device
    (src/dp.jl:525; ext/TheEggmanKernelAbstractionsExt.jl:90-146).

    Returns a length-B array of P's element type -- or, for int64 P (the exact path), a list of
    Python ints.  ``max_batch`` bounds the instances on the device at once; results are
    bit-identical for every value.
    """
    if not isinstance(backend, HIPBackend):
        raise ArgumentError(f"no GPU backend method for {type(backend).__name__}; pass a "
                            "theeggman.HIPBackend()")
    P = np.asarray(P)
    if P.ndim != 2:
        raise DimensionMismatch(f"P must be B x npairs, got shape {P.shape}")
    B, npairs = P.shape
    if npairs != plan.K * (plan.K - 1) // 2:
        raise DimensionMismatch(f"P has {npairs} pair columns, the degree-{plan.K} plan "
                                f"expects {plan.K * (plan.K - 1) // 2}")
    code = dtype_code(P.dtype)
    if B == 0:
        return [] if code == I128 else np.empty(0, P.dtype)
    h = _device_plan(backend, plan)
    Pf = _as_pair_major(P)
    out = np.empty((B, 2), np.uint64) if code == I128 else np.empty(B, P.dtype)
    mb = 0 if max_batch is None else max(1, min(B, int(max_batch)))
    check(lib().eggman_dp_forward(h, code, Pf.ctypes.data, B, npairs, mb, out.ctypes.data),
          "dp_forward")
    if code == I128:
        return [_limbs_to_int(int(lo), int(hi)) for lo, hi in out]
    return out


def _haf_dp_backend_grad(backend: HIPBackend, P: np.ndarray, plan: HafnianDPPlan, *,
                         max_batch: int | None = None):
    """Hafnians and d haf / d P of the B instances in ``P`` (shape (B, npairs)).

    Returns ``(values, dP)`` with dP of shape (B, npairs) and P's element type (for int64 P:
    lists of Python ints).  New entry point: the reference has no derivative.
    """
    if not isinstance(backend, HIPBackend):
        raise ArgumentError(f"no GPU backend method for {type(backend).__name__}")
    P = np.asarray(P)
    if P.ndim != 2:
        raise DimensionMismatch(f"P must be B x npairs, got shape {P.shape}")
    B, npairs = P.shape
    if npairs != plan.K * (plan.K - 1) // 2:
        raise DimensionMismatch(f"P has {npairs} pair columns, the degree-{plan.K} plan "
                                f"expects {plan.K * (plan.K - 1) // 2}")
    code = dtype_code(P.dtype)
    if B == 0:
        return (np.empty(0, P.dtype), np.empty((0, npairs), P.dtype)) if code != I128 else ([], [])
    h = _device_plan(backend, plan)
    Pf = _as_pair_major(P)
    if code == I128:
        out = np.empty((B, 2), np.uint64)
        dP = np.empty((npairs, B, 2), np.uint64)
    else:
        out = np.empty(B, P.dtype)
        dP = np.empty((npairs, B), P.dtype)
    mb = 0 if max_batch is None else max(1, min(B, int(max_batch)))
    check(lib().eggman_dp_forward_grad(h, code, Pf.ctypes.data, B, npairs, mb, out.ctypes.data,
                                       dP.ctypes.data), "dp_forward_grad")
    if code == I128:
        vals = [_limbs_to_int(int(lo), int(hi)) for lo, hi in out]
        grads = [[_limbs_to_int(int(dP[q, b, 0]), int(dP[q, b, 1])) for q in range(npairs)]
                 for b in range(B)]
        return vals, grads
    return out, dP.T


def _haf_dp_backend_repeated(backend: HIPBackend, A: np.ndarray, idxs, plan, *,
                             max_batch: int | None = None):
    """The boson-sampling batch: hafnians of A[idx_b, idx_b] for every 1-based index list idx_b
    (all of length plan.K), with the pair values gathered on the device from one copy of A.
    Returns a length-B array of A's element type (lists of ints for int64 A).  Bit-identical to
    gathering on the host and calling _haf_dp_backend."""
    if not isinstance(backend, HIPBackend):
        raise ArgumentError(f"no GPU backend method for {type(backend).__name__}")
    A = np.asfortranarray(A)                      # column-major, as Julia stores it
    N = A.shape[0]
    code = dtype_code(A.dtype)
    idx = np.ascontiguousarray(np.asarray(idxs, dtype=np.int32).reshape(-1, plan.K))
    B = idx.shape[0]
    if B == 0:
        return [] if code == I128 else np.empty(0, A.dtype)
    h = _device_plan(backend, plan)
    out = np.empty((B, 2), np.uint64) if code == I128 else np.empty(B, A.dtype)
    mb = 0 if max_batch is None else max(1, min(B, int(max_batch)))
    check(lib().eggman_dp_forward_repeated(h, code, A.ctypes.data, N, idx.ctypes.data, B, mb,
                                           out.ctypes.data), "dp_forward_repeated")
    if code == I128:
        return [_limbs_to_int(int(lo), int(hi)) for lo, hi in out]
    return out


def _haf_dp_backend_mixed(backend: HIPBackend, Ps, plans, *, max_batch: int | None = None):
    """Hafnians of several groups of instances, group g of degree plans[g].K with pair values
    Ps[g] (shape (B_g, npairs_g)), in one device pass: groups of 256 or more instances walk
    concurrently on streams of their own, the smaller ones share a walk that advances every group
    by one level per launch.  ``max_batch`` bounds the instances of any one group on the device
    at once (chunks take the same share of every group; the values are the same bits).  Returns one array (or
    list of ints, exact path) per group.  New entry point."""
    if not isinstance(backend, HIPBackend):
        raise ArgumentError(f"no GPU backend method for {type(backend).__name__}")
    if len(Ps) != len(plans) or not Ps:
        raise DimensionMismatch("need one P block per plan")
    dt = np.result_type(*[np.asarray(P).dtype for P in Ps])
    code = dtype_code(dt)
    blocks, counts, handles = [], [], []
    for P, plan in zip(Ps, plans):
        P = np.asarray(P, dtype=dt)
        B, npairs = P.shape
        if npairs != plan.K * (plan.K - 1) // 2:
            raise DimensionMismatch(f"P block has {npairs} pair columns, its degree-{plan.K} plan "
                                    f"expects {plan.K * (plan.K - 1) // 2}")
        blocks.append(np.ascontiguousarray(P.T).ravel())      # [pair][batch]
        counts.append(B)
        handles.append(_device_plan(backend, plan))
    total = int(sum(counts))
    if total == 0:
        return [np.empty(0, dt) for _ in Ps]
    flat = np.concatenate(blocks)
    out = np.empty((total, 2), np.uint64) if code == I128 else np.empty(total, dt)
    h = (ctypes.c_uint64 * len(handles))(*handles)
    c = (ctypes.c_int64 * len(counts))(*counts)
    mb = 0 if max_batch is None else max(1, int(max_batch))
    check(lib().eggman_dp_forward_mixed(len(handles), h, c, code, flat.ctypes.data, mb,
                                        out.ctypes.data), "dp_forward_mixed")
    res, o = [], 0
    for B in counts:
        if code == I128:
            res.append([_limbs_to_int(int(lo), int(hi)) for lo, hi in out[o:o + B]])
        else:
            res.append(out[o:o + B])
        o += B
    return res


def _limbs_to_int(lo: int, hi: int) -> int:
    v = (hi << 64) | lo
    return v - (1 << 128) if v >> 127 else v


def _haf_dp_device(backend: HIPBackend, plan: HafnianDPPlan, dtype, P_ptr: int, ldp: int,
                   B: int, out_ptr: int, stream: int = 0) -> None:
    """Device-resident forward: P (``[pair][ldp]``) and out already in device memory, enqueued on
    ``stream`` (a raw hipStream_t, 0 = the library's stream).  Returns without synchronising."""
    h = _device_plan(backend, plan)
    check(lib().eggman_dp_forward_device(h, dtype_code(dtype), ctypes.c_void_p(P_ptr), ldp, B,
                                         ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream)),
          "dp_forward_device")


def gpu_cache_bytes() -> int:
    """Device memory held by cached plans and work buffers (src/dp.jl:552-560)."""
    with _DEVICE_PLAN_LOCK:
        return int(lib().eggman_cache_bytes())


def empty_gpu_cache() -> None:
    """Drop every cached device plan and work buffer (src/dp.jl:567-573)."""
    with _DEVICE_PLAN_LOCK:
        _DEVICE_PLANS.clear()
        check(lib().eggman_cache_clear(), "cache_clear")


def dp_batch_bytes(N: int, T, B: int) -> int:
    """Device memory one batch needs besides the plan: B*(nstates+npairs)*sizeof(T)
    (src/dp.jl:584-587)."""
    nstates, _ = _dp_counts(N)
    return B * (nstates + N * (N - 1) // 2) * np.dtype(T).itemsize


def grad_levels(N: int, checkpointed: bool = True):
    """Which forward levels the gradient keeps resident for degree N: ``(kept_rows, seg_rows,
    kept)`` with ``kept[L]`` true for level L (1..N/2) kept, the others recomputed segment by
    segment in ``seg_rows`` state rows (rows per instance).  New: the reference has no
    derivative."""
    import ctypes
    kr, sr = ctypes.c_int64(), ctypes.c_int64()
    mask = (ctypes.c_uint8 * (N // 2 + 1))()
    check(lib().eggman_grad_levels(N, 1 if checkpointed else 0, ctypes.byref(kr), ctypes.byref(sr),
                                   ctypes.cast(mask, ctypes.c_void_p), len(mask)), "grad_levels")
    return kr.value, sr.value, [bool(mask[L]) for L in range(len(mask))]
