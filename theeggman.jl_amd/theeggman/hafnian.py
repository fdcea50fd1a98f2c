"""Public API: Python mirror of TheEggman's router for the `:dp` path (src/hafnian.jl:740-1374).

Same names, keyword meanings and refusals as the reference.  The reference's CPU strategies
(``:unrolled``, ``:sieve`` and the cost model choosing between them) are not part of this
package: every hafnian here is evaluated by the subset DP on the device.  Requests that the
reference would answer with a CPU strategy raise NotImplementedError instead of silently
running something else.
"""
from __future__ import annotations

import numpy as np

from ._lib import ArgumentError, DimensionMismatch, I128
from .backend import (HIPBackend, _haf_dp_backend, _haf_dp_backend_grad,
                      _haf_dp_backend_mixed, _haf_dp_backend_repeated, dtype_code)
from .plan import DP_HARD_MAX, DP_MAX, UNROLL_MAX, PlanRef, _dp_plan

_DEFAULT_BACKEND = None


def default_backend() -> HIPBackend:
    global _DEFAULT_BACKEND
    if _DEFAULT_BACKEND is None:
        _DEFAULT_BACKEND = HIPBackend()
    return _DEFAULT_BACKEND


# ------------------------------------------------------------------------------ helpers

def _haf_eltype(dt) -> np.dtype:
    """float(eltype(A)) (src/hafnian.jl:744): integers and Bool promote to Float64."""
    dt = np.dtype(dt)
    if dt.kind in "biu":
        return np.dtype(np.float64)
    if dt.kind == "c" and dt not in (np.dtype(np.complex64), np.dtype(np.complex128)):
        return np.dtype(np.complex128)
    if dt.kind == "f" and dt not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise ArgumentError(f"element type {dt} has no device path")
    return dt


def _checksquare(A) -> int:
    """LinearAlgebra.checksquare."""
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise DimensionMismatch(f"matrix is not square: dimensions are {A.shape}")
    return A.shape[0]


def _check_symmetric(A: np.ndarray) -> None:
    """Symmetric to within isapprox's default tolerance (src/hafnian.jl:765-783): exact equality
    first, then |x-y|^2 <= eps * max(|x|^2, |y|^2)."""
    n = A.shape[0]
    if n < 2 or (A == A.T).all():
        return
    iu = _triu(n)
    x = A[iu]
    y = A.T[iu]
    eq = x == y
    if eq.all():
        return
    real = np.dtype(_haf_eltype(A.dtype)).type(0).real.dtype
    tol = np.finfo(real).eps
    d = np.abs((x - y).astype(np.complex128 if np.iscomplexobj(x) else np.float64)) ** 2
    m = np.maximum(np.abs(x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)) ** 2,
                   np.abs(y.astype(np.complex128 if np.iscomplexobj(y) else np.float64)) ** 2)
    ok = eq | (d <= tol * m)
    if not ok.all():
        k = int(np.argmin(ok))
        i, j = int(iu[0][k]) + 1, int(iu[1][k]) + 1
        raise ArgumentError(f"matrix is not symmetric at ({i}, {j})")


def _check_method(method: str, K: int) -> None:
    """src/hafnian.jl:1058-1066."""
    if method not in ("auto", "unrolled", "dp", "sieve"):
        raise ArgumentError(f"method must be :auto, :unrolled, :dp or :sieve, got :{method}")
    if method == "unrolled" and K > UNROLL_MAX:
        raise ArgumentError(f"method=:unrolled needs degree ≤ {UNROLL_MAX}, got {K}")
    if method == "dp" and K > DP_HARD_MAX:
        raise ArgumentError(f"method=:dp needs degree ≤ {DP_HARD_MAX}, got {K}")


def _route(method: str, K: int) -> None:
    """What the reference's _steer_backend (src/hafnian.jl:999-1011) would run, restricted to the
    strategy this package provides."""
    if method in ("unrolled", "sieve"):
        raise NotImplementedError(f"method=:{method} is a CPU strategy of the reference; this "
                                  "package provides the device :dp path only")
    if method == "auto" and K > DP_MAX:
        raise NotImplementedError(f"degree {K} is above DP_MAX = {DP_MAX}, where the reference's "
                                  "cost model picks the CPU sieve; pass method='dp' to run the "
                                  "Int64-layout DP on the device")


def _plan_for(K: int, backend=None):
    """_dp_plan(K, Int32) through DP_MAX, Int64 above (src/hafnian.jl:1044-1049).  A backend that
    generates plans on the device only needs the degree and layout, so no host plan is built."""
    I = np.int32 if K <= DP_MAX else np.int64
    if isinstance(backend, HIPBackend) and backend.plan_source == "generate":
        return PlanRef(K, np.dtype(I).itemsize)
    return _dp_plan(K, I)


_TRIU_CACHE: dict = {}


def _triu(K: int):
    """Row-major strict upper triangle: the order of _pair_index (src/unrolled.jl:60)."""
    t = _TRIU_CACHE.get(K)
    if t is None:
        t = np.triu_indices(K, 1)
        _TRIU_CACHE[K] = t
    return t


def _gather_pairs_batch(T, A_or_As, idxs, K: int) -> np.ndarray:
    """_gather_pairs_batch (src/dp.jl:479-512): returns P of shape (B, npairs) whose memory is
    [pair][batch] (Fortran order), element (b, pair_index(i,j)) = A_b[idx_b[i], idx_b[j]], i < j.
    ``A_or_As`` is one matrix shared by every instance, or a (B, N, N) stack; ``idxs`` holds
    1-based index vectors (None = 1:N for every instance)."""
    iu, ju = _triu(K)
    if isinstance(A_or_As, np.ndarray) and A_or_As.ndim == 2:
        A = A_or_As
        if idxs is None:
            # one row: a (1, npairs) array is already in [pair][batch] order
            return A[iu, ju].astype(T, copy=False).reshape(1, -1)
        cols = []
        for idx in idxs:
            r = np.asarray(idx, dtype=np.int64) - 1
            cols.append(A[r[iu], r[ju]])
        Pt = np.stack(cols, axis=0)
    else:
        As = A_or_As
        if idxs is None:
            Pt = As[:, iu, ju]
        else:
            Pt = np.stack([As[b][np.asarray(idx) - 1][:, np.asarray(idx) - 1][iu, ju]
                           for b, idx in enumerate(idxs)], axis=0)
    return np.asfortranarray(Pt.astype(T, copy=False))


def _expanded_indices(rpt, total: int) -> np.ndarray:
    """Row i repeated rpt[i] times, 1-based (src/hafnian.jl:1014-1021)."""
    return np.repeat(np.arange(1, len(rpt) + 1, dtype=np.int64), np.asarray(rpt, dtype=np.int64))


def _as_matrix(A) -> np.ndarray:
    A = np.asarray(A)
    if A.dtype == object:
        raise ArgumentError("object arrays have no device path")
    return A


def _is_batch(A) -> bool:
    if isinstance(A, (list, tuple)):
        return True
    return isinstance(A, np.ndarray) and A.ndim == 3


def _backend_or_default(backend):
    return default_backend() if backend is None else backend


# ------------------------------------------------------------------------------ API

def hafnian(A, *, method: str = "auto", check_symmetric: bool = True, backend=None,
            max_batch: int | None = None, nthreads: int | None = None, glynn=None):
    """Hafnian of the square symmetric matrix ``A`` (src/hafnian.jl:1111-1147), or of every
    matrix of a batch (a list or a (B, N, N) array; src/hafnian.jl:1263-1301).

    The diagonal is ignored.  Odd N gives 0 and N == 0 gives 1 without touching the device.
    ``check_symmetric=False`` skips only the O(N^2) symmetry check; the DP then reads the upper
    triangle.  ``max_batch`` (batch form) bounds the instances on the device at once.
    ``nthreads`` and ``glynn`` are accepted for signature compatibility and have no effect on the
    device path.
    """
    if _is_batch(A):
        return _hafnian_batch(A, method=method, check_symmetric=check_symmetric,
                              backend=backend, max_batch=max_batch)
    A = _as_matrix(A)
    N = _checksquare(A)
    T = _haf_eltype(A.dtype)
    if N == 0:
        return T.type(1)
    if N % 2:
        return T.type(0)
    _check_method(method, N)
    if check_symmetric:
        _check_symmetric(A)
    _route(method, N)
    P = _gather_pairs_batch(T, A, None, N)
    return _haf_dp_backend(_backend_or_default(backend), P, _plan_for(N, backend))[0]


def _hafnian_batch(As, *, method, check_symmetric, backend, max_batch):
    As = [_as_matrix(A) for A in As] if isinstance(As, (list, tuple)) else As
    B = len(As)
    if B == 0:
        return np.empty(0, np.float64)
    T = _haf_eltype(np.result_type(*[A.dtype for A in As]) if isinstance(As, list)
                    else As.dtype)
    N = _checksquare(As[0])
    if isinstance(As, list):
        for b, A in enumerate(As):
            if A.ndim != 2 or A.shape != (N, N):
                raise DimensionMismatch(f"matrix {b + 1} is {A.shape[0]}×{A.shape[-1]}, "
                                        f"expected {N}×{N}; a batch shares one strategy and one "
                                        "plan")
    stack = np.asarray(As) if isinstance(As, list) else As
    # a stack whose every matrix equals its transpose needs no per-matrix check
    if check_symmetric and not np.array_equal(stack, stack.transpose(0, 2, 1)):
        for A in As:
            _check_symmetric(A)
    if N == 0:
        return np.ones(B, T)
    if N % 2:
        return np.zeros(B, T)
    _check_method(method, N)
    _route(method, N)
    P = _gather_pairs_batch(T, stack, None, N)
    return _haf_dp_backend(_backend_or_default(backend), P, _plan_for(N, backend), max_batch=max_batch)


def hafnian_repeated(A, rpt, *, method: str = "auto", check_symmetric: bool = True,
                     backend=None, max_batch: int | None = None, nthreads: int | None = None,
                     glynn=None):
    """Hafnian of A with row/column i repeated rpt[i] times (src/hafnian.jl:1194-1222), or, when
    ``rpt`` is a list of patterns, of every pattern against the one A -- the Gaussian
    boson-sampling batch (src/hafnian.jl:1317-1359).  Repeated rows do read the diagonal."""
    A = _as_matrix(A)
    if len(rpt) > 0 and np.ndim(rpt[0]) == 1:
        return _hafnian_repeated_batch(A, rpt, method=method, check_symmetric=check_symmetric,
                                       backend=backend, max_batch=max_batch)
    N = _checksquare(A)
    if len(rpt) != N:
        raise DimensionMismatch(f"rpt has length {len(rpt)}, expected {N}")
    r = np.asarray(rpt, dtype=np.int64)
    if (r < 0).any():
        raise ArgumentError("rpt must contain non-negative integers")
    T = _haf_eltype(A.dtype)
    total = int(r.sum())
    if total == 0:
        return T.type(1)
    if total % 2:
        return T.type(0)
    _check_method(method, total)
    if check_symmetric:
        _check_symmetric(A)
    _route(method, total)
    idx = _expanded_indices(r, total)
    P = _gather_pairs_batch(T, A, [idx], total)
    return _haf_dp_backend(_backend_or_default(backend), P, _plan_for(total, backend))[0]


def _hafnian_repeated_batch(A, rpts, *, method, check_symmetric, backend, max_batch):
    T = _haf_eltype(A.dtype)
    B = len(rpts)
    if B == 0:
        return np.empty(0, T)
    N = _checksquare(A)
    total = int(np.sum(rpts[0]))
    try:
        R = np.asarray(rpts, dtype=np.int64)
    except (ValueError, TypeError):
        R = None
    if not (R is not None and R.ndim == 2 and R.shape[1] == N and not (R < 0).any()
            and (R.sum(axis=1) == total).all()):
        # find and name the first bad pattern
        for b, rpt in enumerate(rpts):
            if len(rpt) != N:
                raise DimensionMismatch(f"rpt {b + 1} has length {len(rpt)}, expected {N}")
            r = np.asarray(rpt, dtype=np.int64)
            if (r < 0).any():
                raise ArgumentError(f"rpt {b + 1} contains a negative entry")
            if int(r.sum()) != total:
                raise ArgumentError(f"rpt {b + 1} sums to {int(r.sum())}, but rpt 1 sums to "
                                    f"{total}; a batch shares one plan, so every pattern must "
                                    "have the same total")
        R = np.stack([np.asarray(r, dtype=np.int64) for r in rpts])
    if check_symmetric:
        _check_symmetric(A)
    if total == 0:
        return np.ones(B, T)
    if total % 2:
        return np.zeros(B, T)
    _check_method(method, total)
    if method in ("unrolled", "sieve"):
        _route(method, total)
    # the batched repeated form sends every total up to DP_HARD_MAX to the device under :auto
    # (src/hafnian.jl:1348); A and the index lists go to the device, which gathers P itself
    # every row is 1..N with row i repeated R[b, i] times (_expanded_indices of each pattern)
    idxs = np.repeat(np.tile(np.arange(1, N + 1, dtype=np.int64), B), R.ravel()).reshape(B, total)
    return _haf_dp_backend_repeated(_backend_or_default(backend), A.astype(T, copy=False), idxs,
                                    _plan_for(total, backend), max_batch=max_batch)


def hafnian_exact(A, *, check_symmetric: bool = True, backend=None,
                  max_batch: int | None = None):
    """Exact hafnian of an integer matrix (or of each matrix of a batch) as Python ints.

    New entry point (the reference promotes integers to Float64, src/hafnian.jl:744).  The DP
    runs in 128-bit two's-complement integers, which is exact whenever the hafnian fits in 127
    bits -- for 0/1 adjacency matrices at every degree up to 64 with room to spare at N <= 40,
    where (N-1)!! < 2^79.  Entries must be integers that fit in int64.  A matrix whose hafnian
    is not provably below 2^127 in magnitude (bound: the smaller of (N-1)!! * max|a|^(N/2) and
    sqrt(prod_i sum_{j != i} |a_ij|)) raises OverflowError before anything runs.
    """
    batch = _is_batch(A)
    As = [np.asarray(M) for M in A] if batch else [np.asarray(A)]
    for M in As:
        if M.dtype.kind not in "biu":
            raise ArgumentError(f"hafnian_exact needs an integer matrix, got {M.dtype}")
    B = len(As)
    if B == 0:
        return []
    N = _checksquare(As[0])
    for b, M in enumerate(As):
        if M.shape != (N, N):
            raise DimensionMismatch(f"matrix {b + 1} is {M.shape[0]}×{M.shape[-1]}, expected "
                                    f"{N}×{N}; a batch shares one plan")
        if check_symmetric and not (M == M.T).all():
            iu = np.argwhere(M != M.T)[0]
            raise ArgumentError(f"matrix is not symmetric at ({iu[0] + 1}, {iu[1] + 1})")
    if N == 0:
        vals = [1] * B
    elif N % 2:
        vals = [0] * B
    else:
        _check_method("dp", N)
        stack = np.stack([M.astype(np.int64) for M in As])
        P = _gather_pairs_batch(np.int64, stack, None, N)
        vals = _haf_dp_backend(_backend_or_default(backend), P, _plan_for(N, backend),
                               max_batch=max_batch)
    return vals if batch else vals[0]


def hafnian_mixed(As, *, check_symmetric: bool = True, backend=None, max_batch: int | None = None):
    """Hafnians of matrices of different sizes, in input order, in one device pass.

    New entry point: the reference's batch form requires one size (src/hafnian.jl:1277-1280).
    Matrices are grouped by degree and every group shares its plan.  Groups of 256 or more
    matrices walk concurrently, each with its own tuned launch on a stream of its own; the
    smaller groups share one walk in which each launch advances all of them by a level.
    ``max_batch`` bounds the matrices of any one degree on the device at once.  Each value is
    bit-identical to ``hafnian`` of that matrix alone.
    Degrees above DP_MAX (Int64 layout) and up to it (Int32) run as two passes.
    """
    As = [_as_matrix(A) for A in As]
    if not As:
        return np.empty(0)
    T = _haf_eltype(np.result_type(*[A.dtype for A in As]))
    out = np.zeros(len(As), T)
    groups: dict = {}
    sizes = [_checksquare(A) for A in As]
    for i, N in enumerate(sizes):
        if N == 0:
            out[i] = 1
        elif N % 2 == 0:
            _check_method("dp", N)
            groups.setdefault(N, []).append(i)
    stacks = {N: np.stack([As[i] for i in idx]) for N, idx in groups.items()}
    if check_symmetric:
        # when every degree's stack equals its transpose, only the odd-sized matrices are left
        exact = all(np.array_equal(S, S.transpose(0, 2, 1)) for S in stacks.values())
        for A, N in zip(As, sizes):
            if not exact or N % 2:
                _check_symmetric(A)
    be = _backend_or_default(backend)
    for wide in (False, True):
        Ns = sorted(N for N in groups if (N > DP_MAX) == wide)
        if not Ns:
            continue
        Ps = [_gather_pairs_batch(T, stacks[N], None, N) for N in Ns]
        res = _haf_dp_backend_mixed(be, Ps, [_plan_for(N, backend) for N in Ns], max_batch=max_batch)
        for N, r in zip(Ns, res):
            out[groups[N]] = r
    return out


def _scatter_pair_grad(dp_row, idx, n: int, dtype) -> np.ndarray:
    """Map d haf / d P (one instance) back to the matrix through the gather of _gather_pairs.

    Returns the symmetric G with G[u, v] = G[v, u] = d haf / d a_uv, the derivative with respect
    to the symmetric pair variable a_uv = A[u, v] = A[v, u] (u != v) or the diagonal entry a_uu:
    the sum of dP over every gathered pair (i < j) with {idx[i], idx[j]} = {u, v}.  With
    distinct indices each G[u, v] is the hafnian of A with rows and columns u, v removed."""
    K = len(idx)
    iu, ju = _triu(K)
    r = np.asarray(idx, dtype=np.int64) - 1
    u, v = r[iu], r[ju]
    G = np.zeros((n, n), dtype)
    np.add.at(G, (u, v), dp_row)
    off = u != v
    np.add.at(G, (v[off], u[off]), dp_row[off])
    return G


def hafnian_grad(A, *, check_symmetric: bool = True, backend=None, max_batch: int | None = None):
    """Hafnian and its gradient, for one matrix or a batch (list or (B, N, N) array).

    Returns ``(value, G)`` (batch: arrays of values and of N x N gradients).  ``G`` is symmetric
    with G[i, j] = G[j, i] = d haf / d a_ij for the symmetric pair variable a_ij, which equals
    the hafnian of A with rows and columns i, j removed; the diagonal is zero.  Computed by a
    reverse sweep over the same plan on the device (new: the reference has no derivative).
    """
    batch = _is_batch(A)
    if batch and isinstance(A, np.ndarray):
        S = _as_matrix(A)                       # a (B, N, N) stack: one element type and shape
        As = S
        if S.shape[0] == 0:
            return np.empty(0), np.empty((0, 0, 0))
        T = _haf_eltype(S.dtype)
        N = _checksquare(S[0])
    else:
        As = [_as_matrix(M) for M in A] if batch else [_as_matrix(A)]
        if len(As) == 0:
            return np.empty(0), np.empty((0, 0, 0))
        T = _haf_eltype(np.result_type(*[M.dtype for M in As]))
        N = _checksquare(As[0])
        for b, M in enumerate(As):
            if M.shape != (N, N):
                raise DimensionMismatch(f"matrix {b + 1} is {M.shape[0]}×{M.shape[-1]}, "
                                        f"expected {N}×{N}; a batch shares one plan")
        S = np.stack(As)
    B = S.shape[0]
    if check_symmetric and not np.array_equal(S, S.transpose(0, 2, 1)):
        for M in As:
            _check_symmetric(M)
    if N % 2 or N == 0:
        vals = np.full(B, 1 if N == 0 else 0, T)
        grads = np.zeros((B, N, N), T)
    else:
        _check_method("dp", N)
        _route("dp", N)
        P = _gather_pairs_batch(T, S, None, N)
        vals, dP = _haf_dp_backend_grad(_backend_or_default(backend), P, _plan_for(N, backend),
                                        max_batch=max_batch)
        # identity gather: every pair (i < j) is its own matrix entry, written on both sides
        iu, ju = _triu(N)
        grads = np.zeros((B, N, N), T)
        grads[:, iu, ju] = dP
        grads[:, ju, iu] = dP
    return (vals, grads) if batch else (vals[0], grads[0])


def hafnian_repeated_grad(A, rpt, *, check_symmetric: bool = True, backend=None):
    """Hafnian of A with row/column i repeated rpt[i] times, and its gradient with respect to
    the entries of A (symmetric pair variables, diagonal included: repeated rows read it)."""
    A = _as_matrix(A)
    N = _checksquare(A)
    r = np.asarray(rpt, dtype=np.int64)
    if len(r) != N:
        raise DimensionMismatch(f"rpt has length {len(r)}, expected {N}")
    if (r < 0).any():
        raise ArgumentError("rpt must contain non-negative integers")
    T = _haf_eltype(A.dtype)
    total = int(r.sum())
    if check_symmetric:
        _check_symmetric(A)
    if total == 0 or total % 2:
        return T.type(1 if total == 0 else 0), np.zeros((N, N), T)
    _check_method("dp", total)
    idx = _expanded_indices(r, total)
    P = _gather_pairs_batch(T, A, [idx], total)
    vals, dP = _haf_dp_backend_grad(_backend_or_default(backend), P, _plan_for(total, backend))
    return vals[0], _scatter_pair_grad(dP[0], idx, N, T)


def rrule(f, A, *args, **kwargs):
    """ChainRulesCore-style reverse rule for ``hafnian``: returns ``(y, pullback)`` where
    ``pullback(dy)`` gives ``(None, dA)``, the cotangent of A for a function of a symmetric
    matrix: ``dA = dy * conj(G) / 2`` with G from :func:`hafnian_grad` (G[i, j] = G[j, i] =
    d haf / d a_ij for the pair variable a_ij; zero diagonal).

    Under the Frobenius pairing a symmetric tangent dA then gives <dA_bar, dA> =
    sum_{i<j} G_ij dA_ij, the directional derivative: the symmetric pair (i, j), (j, i) shares
    one derivative, so each half carries G_ij / 2 (returning G itself would double every
    symmetric chain such as X -> X + X^T).  For complex A the conjugate follows ChainRules'
    convention for holomorphic functions (the cotangent is dy * conj(f')); for real A it is a
    no-op."""
    if f is not hafnian:
        raise ArgumentError("rrule is defined for hafnian only")
    y, G = hafnian_grad(A, *args, **kwargs)
    Gbar = np.conj(G) / 2

    def pullback(dy):
        dy = np.asarray(dy)
        if Gbar.ndim == 3:
            return None, dy[:, None, None] * Gbar
        return None, dy * Gbar

    return y, pullback


def reduction(A, rpt) -> np.ndarray:
    """A with row/column i repeated rpt[i] times (src/hafnian.jl:1369-1374)."""
    A = np.asarray(A)
    if A.shape[0] != len(rpt):
        raise DimensionMismatch("rpt must have one entry per row of A")
    idx = np.repeat(np.arange(len(rpt)), np.asarray(rpt, dtype=np.int64))
    return A[np.ix_(idx, idx)]


__all__ = ["hafnian", "hafnian_repeated", "hafnian_exact", "hafnian_grad",
           "hafnian_repeated_grad", "hafnian_mixed", "rrule", "reduction", "default_backend",
           "dtype_code", "I128"]
