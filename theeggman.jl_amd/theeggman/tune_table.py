"""Regenerate the launch-choice table for this GPU.

    python theeggman.jl_amd/theeggman/tune_table.py [--out PATH] [--quick] [--missing]

Runs the autotuner over the degrees, element types and batch sizes the benchmarks and tests use
and writes the table (JSON) that ``theeggman`` loads at import (``tuning_mi355x.json`` beside
this file by default).  Every launch choice gives bit-identical results; the table only changes
speed.
"""
from __future__ import annotations

import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning_mi355x.json"))
    ap.add_argument("--quick", action="store_true", help="only the N=28 benchmark shapes")
    ap.add_argument("--missing", action="store_true",
                    help="keep the loaded table's entries and tune only the shapes it lacks")
    args = ap.parse_args()
    import numpy as np
    import theeggman as te

    shapes = [(28, np.float64, 1024), (28, np.float64, 1), (28, np.complex128, 1)]
    if not args.quick:
        for K in (16, 20, 24, 28, 32):
            for B in (1, 64, 256, 1024, 4096):
                shapes.append((K, np.float64, B))
        for K in (20, 24, 28, 32):
            shapes.append((K, np.complex128, 1))
            shapes.append((K, np.complex128, 256))
        for K in (24, 28, 32):
            shapes.append((K, np.float32, 1024))
            shapes.append((K, np.complex64, 256))
        for K in (36, 40):
            shapes.append((K, np.complex128, 1))
        shapes.append((40, np.int64, 64))
        # degrees between the benchmark ones, and larger complex and FP32 batches
        for B in (64, 256, 1024, 4096):
            shapes.append((26, np.float64, B))
        for B in (64, 256, 1024):
            shapes.append((30, np.float64, B))
        shapes += [(28, np.complex128, 1024), (28, np.complex64, 1024), (24, np.float32, 4096),
                   (28, np.float32, 4096)]
        # batches beyond 4096 for the small degrees (one pass of the whole batch fits)
        shapes += [(16, np.float64, 16384), (16, np.float64, 65536), (20, np.float64, 16384),
                   (24, np.float64, 16384)]
    from theeggman import tuning
    from theeggman.backend import dtype_code
    seen = set()
    for K, T, B in shapes:
        if (K, np.dtype(T).str, B) in seen:
            continue
        seen.add((K, np.dtype(T).str, B))
        ib = 4 if K <= te.DP_MAX else 8
        if args.missing and (K, ib, dtype_code(T), B) in tuning._TUNED:
            continue
        backend = te.HIPBackend(plan_source="generate") if K > 32 else te.HIPBackend()
        r = te.autotune(K, T, B, reps=3 if B <= 1024 else 2, backend=backend)
        c = r["choice"]
        print(f"K={K} {np.dtype(T).name} B={B}: bt={c['bt']} chunk={c['chunk']} flags={c['flags']} "
              f"occ={c['occ']} {c['ms']:.4f} ms", flush=True)
    te.save_tuning(args.out)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
