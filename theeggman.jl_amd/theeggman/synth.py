"""Seeded synthetic inputs, bit-reproducible in any language.

Each (config, matrix) pair owns a SplitMix64 stream started at state ``config << 32 | matrix``.
Uniform(-1, 1) is ``2 * (x >> 11) * 2^-53 - 1``; a standard normal is Box-Muller's cosine branch
on two draws, rounded to a 2^-30 grid so the value does not depend on the libm used.  Matrices
are filled upper triangle first (row-major, i < j), mirrored to the lower triangle, and have a
zero diagonal.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(state: int, n: int) -> np.ndarray:
    """The first n outputs of SplitMix64 started at ``state``."""
    with np.errstate(over="ignore"):
        s = np.uint64(state & 0xFFFFFFFFFFFFFFFF) + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = s
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(state: int, n: int) -> np.ndarray:
    x = splitmix64(state, n)
    return 2.0 * (x >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 - 1.0


def normal(state: int, n: int) -> np.ndarray:
    x = splitmix64(state, 2 * n)
    u1 = ((x[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (x[1::2] >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    z = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    return np.round(z * 2.0 ** 30) * 2.0 ** -30


def _symmetric(vals: np.ndarray, N: int) -> np.ndarray:
    A = np.zeros((N, N), vals.dtype)
    iu = np.triu_indices(N, 1)
    A[iu] = vals
    A.T[iu] = vals
    return A


def stream(config: int, matrix: int) -> int:
    return (int(config) << 32) | int(matrix)


def sym_uniform(N: int, config: int, matrix: int) -> np.ndarray:
    """Real symmetric, entries Uniform(-1, 1)."""
    return _symmetric(uniform(stream(config, matrix), N * (N - 1) // 2), N)


def sym_normal(N: int, config: int, matrix: int, scale: float = 1.0) -> np.ndarray:
    """Real symmetric, entries N(0, 1) * scale."""
    return _symmetric(normal(stream(config, matrix), N * (N - 1) // 2) * scale, N)


def sym_complex_normal(N: int, config: int, matrix: int) -> np.ndarray:
    """Complex symmetric, entries (g1 + i g2) / sqrt(2N)."""
    n = N * (N - 1) // 2
    g = normal(stream(config, matrix), 2 * n)
    return _symmetric((g[0::2] + 1j * g[1::2]) / np.sqrt(2.0 * N), N)


def gnp_adjacency(N: int, config: int, matrix: int, p_half: bool = True) -> np.ndarray:
    """0/1 adjacency matrix of a G(N, 1/2) random graph (top bit of each draw)."""
    x = splitmix64(stream(config, matrix), N * (N - 1) // 2)
    return _symmetric((x >> np.uint64(63)).astype(np.int64), N)


def pair_batch(As: np.ndarray) -> np.ndarray:
    """(npairs, B) array of the strict upper triangles, i.e. P in [pair][batch] memory order."""
    N = As.shape[-1]
    iu, ju = np.triu_indices(N, 1)
    return np.ascontiguousarray(As[:, iu, ju].T)
