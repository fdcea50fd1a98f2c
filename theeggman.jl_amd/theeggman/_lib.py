"""ctypes binding of libeggman_hip.so (the C ABI declared in include/eggman_hip.h).

The library is the only engine of this package: there is no CPU fallback.  If it is missing,
importing this module raises immediately.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeggman_hip.so")

OK, EINVAL, ENOMEM, EPLAN, EHIP, EUNSUPPORTED, EOVERFLOW = 0, -1, -2, -3, -4, -5, -6
F32, F64, C32, C64, I128 = 1, 2, 3, 4, 5


class ArgumentError(ValueError):
    """Julia's ArgumentError: an argument is invalid (bad degree, method, plan...)."""


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch: array sizes do not agree."""


class EggmanError(RuntimeError):
    """A device-side failure reported by the library (HIP error, out of memory)."""


_i64 = ctypes.c_int64
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_int = ctypes.c_int

# name -> (restype, argtypes); every symbol include/eggman_hip.h declares
SIGNATURES = {
    "eggman_version": (ctypes.c_char_p, []),
    "eggman_last_error": (ctypes.c_char_p, []),
    "eggman_device_count": (_int, []),
    "eggman_dp_counts": (_int, [_int, _i64p, _i64p]),
    "eggman_plan_build": (_int, [_int, _int, _vp, _vp, _vp, _int]),
    "eggman_plan_upload": (_int, [_int, _int, _vp, _i64, _vp, _i64, _vp, _i64, _u64p]),
    "eggman_plan_generate": (_int, [_int, _int, _u64p]),
    "eggman_plan_validate": (_int, [_int, _int, _vp, _i64, _vp, _i64, _vp, _i64]),
    "eggman_plan_info": (_int, [ctypes.c_uint64, ctypes.POINTER(_int), ctypes.POINTER(_int),
                                _i64p, _i64p]),
    "eggman_plan_download": (_int, [ctypes.c_uint64, _vp, _i64]),
    "eggman_dp_forward": (_int, [ctypes.c_uint64, _int, _vp, _i64, _i64, _i64, _vp]),
    "eggman_dp_forward_device": (_int, [ctypes.c_uint64, _int, _vp, _i64, _i64, _vp, _vp]),
    "eggman_dp_forward_device_noalloc": (_int, [ctypes.c_uint64, _int, _vp, _i64, _i64, _vp, _vp]),
    "eggman_dp_forward_repeated": (_int, [ctypes.c_uint64, _int, _vp, _i64, _vp, _i64, _i64, _vp]),
    "eggman_dp_forward_mixed": (_int, [_int, _u64p, _i64p, _int, _vp, _i64, _vp]),
    "eggman_dp_forward_mixed_device": (_int, [_int, _u64p, _i64p, _int, _vp, _vp, _vp]),
    "eggman_dp_forward_grad": (_int, [ctypes.c_uint64, _int, _vp, _i64, _i64, _i64, _vp, _vp]),
    "eggman_dp_forward_grad_device": (_int, [ctypes.c_uint64, _int, _vp, _i64, _i64, _vp, _vp,
                                             _i64, _vp]),
    "eggman_autotune": (_int, [ctypes.c_uint64, _int, _i64, _int, ctypes.c_char_p, _i64]),
    "eggman_tune_set": (_int, [_int, _int, _int, _i64, _int, _i64, _int, _int]),
    "eggman_tune_get": (_int, [_int, _int, _int, _i64, ctypes.POINTER(_int), _i64p,
                               ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "eggman_exact_bound": (_int, [_int, _vp, _i64, _i64, _vp]),
    "eggman_gather_rate": (_int, [_i64, ctypes.POINTER(ctypes.c_double)]),
    "eggman_cache_bytes": (_i64, []),
    "eggman_cache_clear": (_int, []),
    "eggman_batch_bytes": (_i64, [_int, _int, _i64]),
    "eggman_grad_levels": (_int, [_int, _int, _i64p, _i64p, _vp, _i64]),
    "eggman_fuse_validate": (_int, [_int, _i64p, _i64p]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library (loaded once).  Raises if the shared object was never built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"{LIB_PATH} is missing: build it first (python -c 'import "
                        "__graft_entry__ as g; g.build()'); there is no CPU fallback")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def last_error() -> str:
    return lib().eggman_last_error().decode(errors="replace")


def check(rc: int, what: str = "") -> None:
    """Map a status code to the exception the reference would raise."""
    if rc == OK:
        return
    msg = last_error() or what
    if rc in (EINVAL, EPLAN):
        raise ArgumentError(msg)
    if rc == ENOMEM:
        raise MemoryError(msg)
    if rc == EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == EOVERFLOW:
        raise OverflowError(msg)
    raise EggmanError(f"{what}: {msg}" if what else msg)
