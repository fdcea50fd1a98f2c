"""Randomized chunking soak on the GPU (not collected by pytest; run by hand):

    python tests/soak_chunking.py SEED SECONDS

Random degree (16..28), element type (FP64, FP32, complex64), batch size and max_batch; each
forward is compared bit for bit with the same batch in one call, in chunks of max_batch and again
in one call, and each forward + gradient with the same batch in chunks of max_batch.  Every launch
choice, pass split and stream split gives the same bits, so any mismatch is a bug.
"""
import sys, os, time
root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(root, "theeggman.jl_amd"))
import numpy as np
import theeggman as te
from theeggman import synth
BE = te.HIPBackend()
rng = np.random.default_rng(int(sys.argv[1]) if len(sys.argv) > 1 else 1)
t_end = time.time() + float(sys.argv[2]) if len(sys.argv) > 2 else time.time() + 60
n = 0; bad = 0
while time.time() < t_end:
    K = int(rng.choice([16, 20, 22, 24, 26, 28]))
    kind = rng.integers(3)
    T = [np.float64, np.float32, np.complex64][kind]
    grad = bool(rng.integers(2))
    B = int(rng.integers(1, 1500 if not grad else (600 if K >= 26 else 1500)))
    As = np.stack([synth.sym_uniform(K, 900 + n, b) for b in range(B)]).astype(T) if kind < 2 else \
         np.stack([synth.sym_complex_normal(K, 900 + n, b) for b in range(B)]).astype(T)
    P = te._gather_pairs_batch(T, As, None, K)
    plan = te._dp_plan(K)
    mb = int(rng.integers(1, B + 1))
    if grad:
        v1, g1 = te._haf_dp_backend_grad(BE, P, plan)
        v2, g2 = te._haf_dp_backend_grad(BE, P, plan, max_batch=mb)
        ok = np.asarray(v1).tobytes() == np.asarray(v2).tobytes() and np.asarray(g1).tobytes() == np.asarray(g2).tobytes()
    else:
        v1 = te._haf_dp_backend(BE, P, plan)
        v2 = te._haf_dp_backend(BE, P, plan, max_batch=mb)
        v3 = te._haf_dp_backend(BE, P, plan)
        ok = v1.tobytes() == v2.tobytes() == v3.tobytes()
    n += 1
    if not ok:
        bad += 1
        print(f"MISMATCH K={K} T={np.dtype(T)} B={B} mb={mb} grad={grad}", flush=True)
    if n % 20 == 0:
        print(f"{n} cases, {bad} mismatches", flush=True)
print(f"DONE {n} cases, {bad} mismatches", flush=True)
sys.exit(1 if bad else 0)
