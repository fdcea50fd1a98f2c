"""CPU oracle for the subset-DP hafnian path of TheEggman.jl.

TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` import this module, and only as the checker (or as the timed CPU
baseline).  The product package never imports it.

Two independent restatements of the reference live here:

* ``liboracle.so`` (``dp_oracle.c``), the fast C restatement: plan counts and build
  (src/dp.jl:192-304), the level walk (src/dp.jl:344-442) and the CPU batch policy
  (src/hafnian.jl:1226-1242), for Float32/64, ComplexF32/64, a wrapping 128-bit integer walk and a
  __float128 walk; plus exact Ryser permanents.
* pure-Python restatements (``py_*``) of the same plan build and walk over Python ints and
  ``fractions.Fraction``, and the brute-force matching enumeration of test/runtests.jl:11-25.
  They are slow and are used for small degrees only, to pin the C restatement.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
from fractions import Fraction
from math import comb

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

F32, F64, C32, C64, I128, Q, CQ = 1, 2, 3, 4, 5, 6, 7

DP_MAX = 32          # src/dp.jl:96
DP_HARD_MAX = 64     # src/dp.jl:124

_lib = None


def use_library(path: str) -> None:
    """Load the oracle from another build of dp_oracle.c (the CPU baseline times a
    ``-march=native`` build made on the machine it runs on)."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None
    _PLANS.clear()


def lib():
    """Load liboracle.so (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing; run `make -C oracle` (or build())")
        L = ctypes.CDLL(LIB_PATH)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.oracle_dp_counts.argtypes = [ctypes.c_int, i64p, i64p]
        L.oracle_build_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_haf_dp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]
        L.oracle_perm_ryser.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------------------ plan

def dp_counts(K: int) -> tuple[int, int]:
    """(nstates, ntrans) from the closed form of src/dp.jl:192-208 (C restatement)."""
    ns, nt = ctypes.c_int64(), ctypes.c_int64()
    rc = lib().oracle_dp_counts(K, ctypes.byref(ns), ctypes.byref(nt))
    if rc:
        raise ValueError(f"oracle_dp_counts({K}) -> {rc}")
    return ns.value, nt.value


def pair_bits(index_bytes: int) -> int:
    """src/dp.jl:170-171."""
    return 9 if index_bytes == 4 else 12


def pair_index(i: int, j: int, K: int) -> int:
    """src/unrolled.jl:60, 1-based, i < j."""
    return (i - 1) * K - (i * (i + 1)) // 2 + j


class Plan:
    """The arrays of HafnianDPPlan{I} (src/dp.jl:150-156), 1-based values as in Julia."""

    def __init__(self, K, index_bytes, starts, trans, levels):
        self.K = K
        self.index_bytes = index_bytes
        self.starts = starts
        self.trans = trans
        self.levels = levels
        self.nstates = len(starts) - 1

    def sha256(self) -> str:
        h = hashlib.sha256()
        for a in (self.starts, self.trans, self.levels):
            h.update(np.ascontiguousarray(a).astype("<i8").tobytes())
        return h.hexdigest()


_PLANS: dict = {}


def build_plan(K: int, index_bytes: int | None = None) -> Plan:
    """C restatement of _build_dp_plan (src/dp.jl:223-304), cached like _dp_plan (:325-329)."""
    if index_bytes is None:
        index_bytes = 4 if K <= DP_MAX else 8
    key = (K, index_bytes)
    if key not in _PLANS:
        ns, nt = dp_counts(K)
        dt = np.int32 if index_bytes == 4 else np.int64
        starts = np.zeros(ns + 1, dt)
        trans = np.zeros(nt, dt)
        levels = np.zeros(K // 2 + 2, dt)
        rc = lib().oracle_build_plan(K, index_bytes, _ptr(starts), _ptr(trans), _ptr(levels))
        if rc:
            raise ValueError(f"oracle_build_plan({K}, {index_bytes}) -> {rc}")
        _PLANS[key] = Plan(K, index_bytes, starts, trans, levels)
    return _PLANS[key]


# ------------------------------------------------------------------------------ evaluation

_P_DT = {F32: np.float32, F64: np.float64, C32: np.complex64, C64: np.complex128,
         I128: np.int64, Q: np.float64, CQ: np.complex128}


def haf_dp(P: np.ndarray, K: int, dtype: int, nthreads: int = 1,
           index_bytes: int | None = None):
    """Hafnians of the instances in ``P`` ([pair][batch] = shape (npairs, B), batch fastest in
    memory, i.e. Julia's column-major ``Matrix{T}(B, npairs)``), by the C restatement of
    _haf_dp / _batch_cpu!.

    Returns a numpy array of length B: float32/float64/complex64/complex128 for the plain types,
    Python ints (two's complement from 128 bits) for I128, and (hi, lo) double-double pairs
    (shape (B, 2), complex: (B, 2) of complex) for Q / CQ.
    """
    plan = build_plan(K, index_bytes)
    npairs = K * (K - 1) // 2
    P = np.ascontiguousarray(np.asarray(P, dtype=_P_DT[dtype]).reshape(npairs, -1))
    B = P.shape[1]
    if dtype in (F32, F64, C32, C64):
        out = np.zeros(B, _P_DT[dtype])
    elif dtype == I128:
        out = np.zeros((B, 2), np.uint64)
    elif dtype == Q:
        out = np.zeros((B, 2), np.float64)
    else:
        out = np.zeros((B, 4), np.float64)
    rc = lib().oracle_haf_dp(K, plan.index_bytes, _ptr(plan.starts), _ptr(plan.trans),
                             _ptr(plan.levels), dtype, _ptr(P), B, nthreads, _ptr(out))
    if rc:
        raise ValueError(f"oracle_haf_dp -> {rc}")
    if dtype == I128:
        return [limbs_to_int(int(lo), int(hi)) for lo, hi in out]
    if dtype == CQ:
        return np.stack([out[:, 0] + 1j * out[:, 2], out[:, 1] + 1j * out[:, 3]], axis=1)
    return out


def limbs_to_int(lo: int, hi: int) -> int:
    v = (hi << 64) | lo
    return v - (1 << 128) if v >> 127 else v


def perm_ryser(Bm) -> int:
    """Exact permanent of a small integer matrix (C, Ryser with Gray code, 128-bit)."""
    Bm = np.ascontiguousarray(np.asarray(Bm, dtype=np.int64))
    n = Bm.shape[0]
    out = np.zeros(2, np.uint64)
    rc = lib().oracle_perm_ryser(_ptr(Bm), n, _ptr(out))
    if rc:
        raise ValueError(f"oracle_perm_ryser -> {rc}")
    return limbs_to_int(int(out[0]), int(out[1]))


def max_threads() -> int:
    return lib().oracle_max_threads()


# ------------------------------------------------------------------------------ gathers

def gather_pairs(A: np.ndarray, idx=None) -> np.ndarray:
    """_gather_pairs! (src/dp.jl:452-462): P[pair_index(i,j)] = A[idx[i], idx[j]] for i < j.
    ``idx`` holds 1-based indices (default 1:N).  Returns the pair vector (0-based storage)."""
    A = np.asarray(A)
    if idx is None:
        idx = list(range(1, A.shape[0] + 1))
    K = len(idx)
    P = np.zeros(K * (K - 1) // 2, A.dtype)
    for i in range(1, K):
        for j in range(i + 1, K + 1):
            P[pair_index(i, j, K) - 1] = A[idx[i - 1] - 1, idx[j - 1] - 1]
    return P


def gather_pairs_batch(As, idxs=None) -> np.ndarray:
    """_gather_pairs_batch (src/dp.jl:464-512) as a (npairs, B) array ([pair][batch])."""
    cols = [gather_pairs(A, None if idxs is None else idxs[b]) for b, A in enumerate(As)]
    return np.stack(cols, axis=1)


# ------------------------------------------------------------------------------ pure Python

def py_dp_counts(K: int) -> tuple[int, int]:
    """src/dp.jl:192-208, in Python integers."""
    nstates, ntrans = 1, 0
    for p in range(1, K + 1):
        lo = p - 1
        hi = min(2 * (p - 1), K - 1)
        r = lo if lo % 2 == 0 else lo + 1
        while r <= hi:
            k = K - r - 1
            c = comb(K - p, k) if 0 <= k <= K - p else 0
            nstates += c
            ntrans += c * k
            r += 2
    return nstates, ntrans


def py_build_plan(K: int, index_bytes: int = 4) -> Plan:
    """src/dp.jl:248-304 in plain Python: DFS from the full mask into a set, sort by
    (popcount, mask), rank with a dict, emit packed transitions in ascending j, mark levels."""
    bits = pair_bits(index_bytes)
    full = (1 << K) - 1
    reached = {full}
    stack = [full]
    while stack:
        m = stack.pop()
        if m == 0:
            continue
        i = (m & -m).bit_length()
        rest = m & ~(1 << (i - 1))
        r = rest
        while r:
            j = (r & -r).bit_length()
            r &= ~(1 << (j - 1))
            sub = rest & ~(1 << (j - 1))
            if sub not in reached:
                reached.add(sub)
                stack.append(sub)
    masks = sorted(reached, key=lambda m: (bin(m).count("1"), m))
    number = {m: s for s, m in enumerate(masks, start=1)}
    starts, trans = [1], []
    for m in masks:
        if m:
            i = (m & -m).bit_length()
            rest = m & ~(1 << (i - 1))
            r = rest
            while r:
                j = (r & -r).bit_length()
                r &= ~(1 << (j - 1))
                c = number[rest & ~(1 << (j - 1))]
                trans.append((c << bits) | pair_index(i, j, K))
        starts.append(len(trans) + 1)
    levels = [1]
    for s in range(2, len(masks) + 1):
        if bin(masks[s - 1]).count("1") != bin(masks[s - 2]).count("1"):
            levels.append(s)
    levels.append(len(masks) + 1)
    dt = np.int32 if index_bytes == 4 else np.int64
    return Plan(K, index_bytes, np.array(starts, dt), np.array(trans, dt), np.array(levels, dt))


def py_haf_dp(P, plan: Plan, one=1, zero=0):
    """_haf_dp (src/dp.jl:408-442) over arbitrary Python numbers (int, Fraction, complex...).
    ``P`` is the pair vector (0-based storage of the 1-based pair index)."""
    bits = pair_bits(plan.index_bytes)
    mask = (1 << bits) - 1
    H = [zero] * plan.nstates
    H[0] = one
    starts = [int(x) for x in plan.starts]
    trans = [int(x) for x in plan.trans]
    levels = [int(x) for x in plan.levels]
    for L in range(2, len(levels)):
        for s in range(levels[L - 1], levels[L]):
            acc = zero
            for k in range(starts[s - 1], starts[s]):
                t = trans[k - 1]
                acc = acc + P[(t & mask) - 1] * H[(t >> bits) - 1]
            H[s - 1] = acc
    return H[plan.nstates - 1]


def brute_hafnian(A):
    """Direct matching enumeration (test/runtests.jl:11-25), exact for Python numbers."""
    N = len(A)
    if N == 0:
        return 1
    if N % 2:
        return 0

    def rec(rem):
        if not rem:
            return 1
        i = rem[0]
        s = 0
        for t in range(1, len(rem)):
            s = s + A[i][rem[t]] * rec([rem[k] for k in range(1, len(rem)) if k != t])
        return s

    return rec(list(range(N)))


def as_fraction_matrix(A) -> list:
    """Exact rationals of a float matrix (every float is a dyadic rational)."""
    return [[Fraction(float(x)) for x in row] for row in np.asarray(A, dtype=np.float64)]
