"""Autotuner front end: choose launch parameters per (degree, element type, batch size).

The library times its candidate launch shapes -- lanes per state, LDS staging of the pair rows,
8 or 16 waves per workgroup, workgroups per CU, 64- or 16-instance tiles with 8- or 16-byte
loads, instances per pass over the levels, one fused
LDS-resident kernel for small degrees, and eager per-level launches versus a replayed HIP graph
of the whole level sequence -- and keeps the fastest.  Every candidate computes bit-identical
results, so tuning only ever changes speed.  Tables can be saved to and loaded from a JSON file
so a tuned choice survives the process; the package loads ``tuning_mi355x.json`` beside it at
import when present.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from ._lib import check, lib
from .backend import HIPBackend, _device_plan, dtype_code
from .plan import DP_MAX, PlanRef, _dp_plan

LDS, GRAPH, FUSED, WIDE, CKPT, NARROW, VEC, DEEP, PAIRS = 1, 2, 4, 8, 16, 32, 64, 128, 256
_NAMES = {"lds": LDS, "graph": GRAPH, "fused": FUSED, "wide": WIDE, "ckpt": CKPT, "narrow": NARROW,
          "vec": VEC, "deep": DEEP, "pairs": PAIRS}   # EGGMAN_TUNE_*
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
_TUNED: dict = {}   # (K, index_bytes, dtype code, B) -> {"bt", "chunk", "flags", "occ"}


def _parse(line: str) -> dict:
    kv = dict(tok.split("=") for tok in line.split() if "=" in tok)
    d = {k: (float(v) if k == "ms" else int(v)) for k, v in kv.items()}
    for name, bit in _NAMES.items():
        d[name] = int(bool(d.get("flags", 0) & bit))
    return d


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
    _TUNED[(K, np.dtype(I).itemsize, code, B)] = {k: choice[k] for k in ("bt", "chunk", "flags", "occ")}
    return {"choice": choice, "candidates": cands}


def tune_get(K: int, dtype=np.float64, B: int = 1024):
    """The tuning entry the library uses for (K, dtype, B), or None."""
    ib = 4 if K <= DP_MAX else 8
    bt, flags, occ = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    chunk = ctypes.c_int64()
    rc = lib().eggman_tune_get(K, ib, dtype_code(dtype), B, ctypes.byref(bt), ctypes.byref(chunk),
                               ctypes.byref(flags), ctypes.byref(occ))
    if rc != 0:
        return None
    d = {"bt": bt.value, "chunk": chunk.value, "flags": flags.value, "occ": occ.value}
    for name, bit in _NAMES.items():
        d[name] = int(bool(flags.value & bit))
    return d


def tune_set(K: int, dtype, B: int, bt: int = 64, chunk: int = 0, occ: int = 0, *,
             lds: bool = True, graph: bool = False, fused: bool = False, wide: bool = False,
             ckpt: bool = False, narrow: bool = False, vec: bool = False, deep: bool = False,
             pairs: bool = False, index_bytes: int | None = None) -> None:
    """Set the tuning entry for (K, dtype, B) in the plan layout of ``index_bytes`` (default: the
    layout degree K uses, Int32 up to K = 32)."""
    ib = index_bytes or (4 if K <= DP_MAX else 8)
    code = dtype_code(dtype)
    flags = ((LDS if lds else 0) | (GRAPH if graph else 0) | (FUSED if fused else 0) |
             (WIDE if wide else 0) | (CKPT if ckpt else 0) | (NARROW if narrow else 0) |
             (VEC if vec else 0) | (DEEP if deep else 0) | (PAIRS if pairs else 0))
    check(lib().eggman_tune_set(K, ib, code, B, bt, chunk, flags, occ), "tune_set")
    _TUNED[(K, ib, code, B)] = {"bt": bt, "chunk": chunk, "flags": flags, "occ": occ}


def save_tuning(path: str) -> None:
    rows = [{"K": K, "index_bytes": ib, "dtype": code, "B": B, **c}
            for (K, ib, code, B), c in sorted(_TUNED.items())]
    with open(path, "w") as f:
        json.dump(rows, f, indent=1)


def load_tuning(path: str = DEFAULT_TABLE) -> int:
    """Load a saved table into the library (for the current device).  Returns the entries read."""
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        check(lib().eggman_tune_set(r["K"], r["index_bytes"], r["dtype"], r["B"], r["bt"],
                                    r["chunk"], r["flags"], r["occ"]), "tune_set")
        _TUNED[(r["K"], r["index_bytes"], r["dtype"], r["B"])] = {k: r[k] for k in ("bt", "chunk", "flags", "occ")}
    return len(rows)
