"""Autotuner front end: choose launch parameters per (degree, element type, batch size).

The library times its candidate launch shapes -- lanes per state, LDS staging of the pair rows,
workgroups per CU, instances per pass over the levels, and eager per-level launches versus a
replayed HIP graph of the whole level sequence -- and keeps the fastest.  Every candidate computes
bit-identical results, so tuning only ever changes speed.  Tables can be saved to and loaded from
a JSON file so a tuned choice survives the process.
"""
from __future__ import annotations

import ctypes
import json

import numpy as np

from ._lib import check, lib
from .backend import HIPBackend, _device_plan, dtype_code
from .plan import DP_MAX, PlanRef, _dp_plan

_FIELDS = ("bt", "chunk", "lds", "occ", "graph", "fused")
_TUNED: dict = {}   # (K, index_bytes, dtype code, B) -> choice, for save_tuning


def _parse(line: str) -> dict:
    kv = dict(tok.split("=") for tok in line.split() if "=" in tok)
    return {k: (float(v) if k == "ms" else int(v)) for k, v in kv.items()}


def autotune(K: int, dtype=np.float64, B: int = 1024, reps: int = 3,
             backend: HIPBackend | None = None) -> dict:
    """Tune the forward walk for degree K, element type ``dtype`` and batch size B.  Returns the
    chosen parameters and the per-candidate timings (ms per forward pass)."""
    backend = backend or HIPBackend()
    I = np.int32 if K <= DP_MAX else np.int64
    plan = PlanRef(K, np.dtype(I).itemsize) if backend.plan_source == "generate" else _dp_plan(K, I)
    h = _device_plan(backend, plan)
    buf = ctypes.create_string_buffer(1 << 16)
    code = dtype_code(dtype)
    check(lib().eggman_autotune(h, code, B, reps, buf, len(buf)), "autotune")
    lines = buf.value.decode().strip().splitlines()
    cands = [_parse(ln) for ln in lines if not ln.startswith("best")]
    choice = _parse(lines[-1])
    _TUNED[(K, np.dtype(I).itemsize, code, B)] = {k: choice[k] for k in _FIELDS}
    return {"choice": choice, "candidates": cands}


def tune_get(K: int, dtype=np.float64, B: int = 1024):
    """The tuning entry the library uses for (K, dtype, B), or None."""
    ib = 4 if K <= DP_MAX else 8
    vals = [ctypes.c_int() for _ in range(5)]
    chunk = ctypes.c_int64()
    rc = lib().eggman_tune_get(K, ib, dtype_code(dtype), B, ctypes.byref(vals[0]),
                               ctypes.byref(chunk), ctypes.byref(vals[1]), ctypes.byref(vals[2]),
                               ctypes.byref(vals[3]), ctypes.byref(vals[4]))
    if rc != 0:
        return None
    return {"bt": vals[0].value, "chunk": chunk.value, "lds": vals[1].value,
            "occ": vals[2].value, "graph": vals[3].value, "fused": vals[4].value}


def tune_set(K: int, dtype, B: int, bt: int, chunk: int, lds: int, occ: int, graph: int = 0,
             fused: int = 0) -> None:
    ib = 4 if K <= DP_MAX else 8
    code = dtype_code(dtype)
    check(lib().eggman_tune_set(K, ib, code, B, bt, chunk, lds, occ, graph, fused), "tune_set")
    _TUNED[(K, ib, code, B)] = {"bt": bt, "chunk": chunk, "lds": lds, "occ": occ, "graph": graph,
                                "fused": fused}


def save_tuning(path: str) -> None:
    rows = [{"K": K, "index_bytes": ib, "dtype": code, "B": B, **c}
            for (K, ib, code, B), c in sorted(_TUNED.items())]
    with open(path, "w") as f:
        json.dump(rows, f, indent=1)


def load_tuning(path: str) -> int:
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        check(lib().eggman_tune_set(r["K"], r["index_bytes"], r["dtype"], r["B"], r["bt"],
                                    r["chunk"], r["lds"], r["occ"], r.get("graph", 0),
                                    r.get("fused", 0)), "tune_set")
        _TUNED[(r["K"], r["index_bytes"], r["dtype"], r["B"])] = {k: r.get(k, 0) for k in _FIELDS}
    return len(rows)
