"""The per-degree DP plan: Python mirror of HafnianDPPlan and its cache (src/dp.jl:96-329).

Names, arguments and refusals follow the reference.  The arrays are built by the library's host
builder (``eggman_plan_build``), which writes exactly the reference's layout.
"""
from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass

import numpy as np

from ._lib import ArgumentError, check, lib

DP_MAX = 32          # src/dp.jl:96 -- largest degree the Int32 layout packs
DP_HARD_MAX = 64     # src/dp.jl:124 -- largest degree `method=:dp` accepts at all
UNROLL_MAX = 12      # src/unrolled.jl:52


def _pair_bits(index_bytes: int) -> int:
    """Bits of the pair index in a packed transition (src/dp.jl:170-171)."""
    return 9 if index_bytes == 4 else 12


def _pair_mask(index_bytes: int) -> int:
    """src/dp.jl:172."""
    return (1 << _pair_bits(index_bytes)) - 1


def _dp_index_type(K: int):
    """numpy dtype of the plan words: int32 through DP_MAX, int64 above (src/dp.jl:180)."""
    return np.int32 if K <= DP_MAX else np.int64


def _pair_index(i: int, j: int, K: int) -> int:
    """Flat 1-based upper-triangle index of the pair i < j (src/unrolled.jl:60)."""
    return (i - 1) * K - (i * (i + 1)) // 2 + j


def _dp_counts(K: int) -> tuple[int, int]:
    """(nstates, ntransitions) of the degree-K plan without building it (src/dp.jl:192-208)."""
    ns, nt = ctypes.c_int64(), ctypes.c_int64()
    check(lib().eggman_dp_counts(K, ctypes.byref(ns), ctypes.byref(nt)), "dp_counts")
    return ns.value, nt.value


@dataclass(eq=False)
class HafnianDPPlan:
    """Matrix-independent CSR of reachable subsets (src/dp.jl:150-156); 1-based values."""
    K: int
    nstates: int
    starts: np.ndarray
    trans: np.ndarray
    levels: np.ndarray

    @property
    def index_bytes(self) -> int:
        return self.trans.dtype.itemsize


@dataclass(frozen=True)
class PlanRef:
    """A degree and layout without host arrays: enough for a backend that generates the plan
    in device memory (``HIPBackend(plan_source="generate")``)."""
    K: int
    index_bytes: int

    @property
    def nstates(self) -> int:
        return _dp_counts(self.K)[0]


def _build_dp_plan(K: int, I=None) -> HafnianDPPlan:
    """Build the degree-K plan in layout ``I`` (np.int32 / np.int64), src/dp.jl:223-246.

    Refuses, before allocating, exactly what the reference refuses: K above DP_HARD_MAX, a pair
    or state index that does not fit the packing, and a plan over half of physical memory.
    """
    if I is None:
        I = _dp_index_type(K)
    I = np.dtype(I)
    if I not in (np.dtype(np.int32), np.dtype(np.int64)):
        raise ArgumentError(f"plan layout must be Int32 or Int64, got {I}")
    if K > DP_HARD_MAX:
        raise ArgumentError(f"degree {K} exceeds DP_HARD_MAX = {DP_HARD_MAX}; subsets are "
                            f"enumerated as bitmasks and a UInt64 holds only {DP_HARD_MAX} of them")
    if K < 2 or K % 2:
        raise ArgumentError(f"plans exist for even degrees >= 2, got {K}")
    nstates, ntrans = _dp_counts(K)
    bits = _pair_bits(I.itemsize)
    npairs = K * (K - 1) // 2
    imax = np.iinfo(I).max
    if npairs > _pair_mask(I.itemsize):
        raise ArgumentError(f"degree {K} needs more than {bits} bits per pair index in Int{I.itemsize * 8}")
    if nstates > (imax >> bits):
        raise ArgumentError(f"degree {K} has too many states to pack into an Int{I.itemsize * 8} transition")
    if ntrans > imax:
        raise ArgumentError(f"degree {K} has too many transitions to index with Int{I.itemsize * 8}")
    nbytes = (ntrans + nstates + 1) * I.itemsize
    if 2 * nbytes >= _total_memory():
        raise ArgumentError(f"a degree-{K} plan needs {nbytes / 2**30:.1f} GB, and about twice "
                            "that while building -- more than this machine has; use method = :sieve")
    starts = np.empty(nstates + 1, I)
    trans = np.empty(ntrans, I)
    levels = np.empty(K // 2 + 2, I)
    check(lib().eggman_plan_build(K, I.itemsize, starts.ctypes.data, trans.ctypes.data,
                                  levels.ctypes.data, 0), "plan_build")
    return HafnianDPPlan(K, nstates, starts, trans, levels)


def _total_memory() -> int:
    import os
    return os.sysconf("SC_PHYS_PAGES") * os.sysconf("SC_PAGE_SIZE")


_DP_PLANS32: dict = {}
_DP_PLANS64: dict = {}
_DP_PLAN_LOCK = threading.RLock()


def _dp_plan(K: int, I=None) -> HafnianDPPlan:
    """The degree-K plan in layout ``I``, built on first use and cached for the session
    (src/dp.jl:325-329); one cache per layout (src/dp.jl:308-313)."""
    if I is None:
        I = _dp_index_type(K)
    cache = _DP_PLANS32 if np.dtype(I) == np.dtype(np.int32) else _DP_PLANS64
    with _DP_PLAN_LOCK:
        plan = cache.get(K)
        if plan is None:
            plan = _build_dp_plan(K, I)
            cache[K] = plan
        return plan
