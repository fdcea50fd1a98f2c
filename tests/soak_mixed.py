"""Randomized mixed-degree soak on the GPU (not collected by pytest; run by hand):

    python tests/soak_mixed.py SECONDS

One to four degrees of 16..28 with 1..699 matrices each, shuffled into one hafnian_mixed call
(with or without max_batch); the values must be the bits of the single-degree calls.  Groups of
256 or more walk on side streams, each with its passes two at a time on a second stream.
"""
import sys, os, time
root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(root, "theeggman.jl_amd"))
import numpy as np
import theeggman as te
from theeggman import synth
BE = te.HIPBackend()
rng = np.random.default_rng(3)
t_end = time.time() + float(sys.argv[1])
n = bad = 0
while time.time() < t_end:
    degs = rng.choice([16, 18, 20, 22, 24, 26, 28], size=int(rng.integers(1, 5)), replace=False)
    mats, want = [], []
    for d in degs:
        c = int(rng.integers(1, 700))
        As = [synth.sym_uniform(int(d), 500 + n, b) for b in range(c)]
        mats += As
        want.append(te.hafnian(np.stack(As), backend=BE))
    order = rng.permutation(len(mats))
    shuffled = [mats[i] for i in order]
    mb = None if rng.integers(2) else int(rng.integers(1, 800))
    got = te.hafnian_mixed(shuffled, backend=BE, max_batch=mb)
    flat = np.concatenate(want)
    ok = np.asarray(got).tobytes() == flat[order].tobytes()
    n += 1
    if not ok:
        bad += 1; print("MISMATCH", degs, mb, flush=True)
    if n % 10 == 0: print(f"{n} mixed calls, {bad} mismatches", flush=True)
print(f"DONE {n} mixed calls, {bad} mismatches", flush=True)
sys.exit(1 if bad else 0)
