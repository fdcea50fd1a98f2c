"""Writes tests/golden/n40_exact.json: exact perfect-matching counts of G(40, 1/2) graphs.

The graphs are those of the 64-graph exact workload (synthetic config 52, matrices 0..63,
``theeggman.synth.gnp_adjacency``); eight of them, spread over the set, are counted here by the
C oracle's 128-bit walk (a restatement of _build_dp_plan and _haf_dp, src/dp.jl:223-304 and
:408-442) over the oracle's own Int64-layout plan, built on the host by its DFS / sort / rank-map
restatement -- independent of the product library's plan generation and kernels.  Every count is
below 39!! < 2^79, so the 128-bit walk is exact.  One graph is counted a second time after a
random relabeling of its vertices, which must give the same integer.

Needs about 40 GB of host memory (the K=40 plan is 22.7 GB) and a few minutes on 8 threads.

    python tests/golden/make_n40_exact.py
"""
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "theeggman.jl_amd"))
import oracle  # noqa: E402
from theeggman import synth  # noqa: E402  (the seeded generator only; no device code)

CONFIG = 52
GRAPHS = [0, 9, 18, 27, 36, 45, 54, 63]


def count(A, nthreads):
    P = oracle.gather_pairs(A).astype(np.int64).reshape(-1, 1)
    return oracle.haf_dp(P, 40, oracle.I128, nthreads, index_bytes=8)[0]


def main():
    nthreads = min(8, oracle.max_threads())
    t0 = time.time()
    oracle.build_plan(40, 8)
    print(f"K=40 plan built in {time.time() - t0:.0f} s", flush=True)
    counts = {}
    for b in GRAPHS:
        t = time.time()
        counts[b] = count(synth.gnp_adjacency(40, CONFIG, b), nthreads)
        print(f"graph {b}: {counts[b]} ({time.time() - t:.0f} s)", flush=True)
        assert 0 <= counts[b] < math.prod(range(39, 0, -2))
    A = synth.gnp_adjacency(40, CONFIG, GRAPHS[0])
    pi = np.random.default_rng(40).permutation(40)
    again = count(A[np.ix_(pi, pi)], nthreads)
    assert again == counts[GRAPHS[0]], (again, counts[GRAPHS[0]])
    out = {"what": "perfect matchings of G(40, 1/2) graphs synth.gnp_adjacency(40, 52, b), "
                   "by the C oracle's 128-bit walk over its own Int64 plan",
           "config": CONFIG, "degree": 40, "relabel_checked": GRAPHS[0],
           "counts": {str(b): c for b, c in counts.items()}}
    with open(os.path.join(HERE, "n40_exact.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote n40_exact.json", flush=True)


if __name__ == "__main__":
    main()
