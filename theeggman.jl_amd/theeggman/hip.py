"""Minimal HIP runtime plumbing over ctypes: device buffers, copies, streams and events.

Uses the very HIP runtime libeggman_hip.so is linked against (already loaded by the time this
module is used), so buffers and streams made here are valid arguments to the library's
device-pointer entry points.  This is what the benchmark and the tests use to keep inputs
resident in HBM; any other HIP allocator (torch included) works as well, provided only one HIP
runtime is loaded in the process -- import torch before theeggman if both are used.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import EggmanError, lib

_hip = None


def runtime():
    global _hip
    if _hip is None:
        lib()                                   # loads libamdhip64.so.7 as the library's dependency
        H = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        H.hipMemcpyAsync.argtypes = [vp, vp, sz, ctypes.c_int, vp]
        H.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        H.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        H.hipEventDestroy.argtypes = [vp]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        H.hipGetErrorString.restype = ctypes.c_char_p
        H.hipGetErrorString.argtypes = [ctypes.c_int]
        _hip = H
    return _hip


def _chk(rc, what):
    if rc != 0:
        raise EggmanError(f"{what}: {runtime().hipGetErrorString(rc).decode()}")


H2D, D2H, D2D = 1, 2, 3


class DeviceBuffer:
    """A device allocation of ``nbytes`` bytes, freed when the object is collected."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        _chk(runtime().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @property
    def addr(self) -> int:
        return self.ptr.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _chk(runtime().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more than the buffer holds")
        _chk(runtime().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self) -> None:
        _chk(runtime().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def __del__(self):
        try:
            if self.ptr and self.ptr.value and _hip is not None:
                _hip.hipFree(self.ptr)
                self.ptr = ctypes.c_void_p()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.ptr = ctypes.c_void_p()
        _chk(runtime().hipStreamCreate(ctypes.byref(self.ptr)), "hipStreamCreate")

    @property
    def handle(self) -> int:
        return self.ptr.value or 0

    def synchronize(self) -> None:
        _chk(runtime().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")


class Event:
    def __init__(self):
        self.ptr = ctypes.c_void_p()
        _chk(runtime().hipEventCreate(ctypes.byref(self.ptr)), "hipEventCreate")

    def record(self, stream: Stream) -> None:
        _chk(runtime().hipEventRecord(self.ptr, stream.ptr), "hipEventRecord")

    def synchronize(self) -> None:
        _chk(runtime().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float()
        _chk(runtime().hipEventElapsedTime(ctypes.byref(ms), self.ptr, end.ptr),
             "hipEventElapsedTime")
        return ms.value


def copy_d2d(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than a buffer")
    _chk(runtime().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def copy_d2d_async(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int, stream: "Stream") -> None:
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy larger than a buffer")
    _chk(runtime().hipMemcpyAsync(dst.ptr, src.ptr, nbytes, D2D, stream.handle), "hipMemcpyAsync D2D")


def device_synchronize() -> None:
    _chk(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
