"""Build libeggman_hip.so in place (gfx950 code objects, no JIT).

    python theeggman.jl_amd/theeggman/build.py

eggman.hip is compiled as six parts in parallel -- part 0 (shared state, the C ABI) and one part
per element type (its level drivers and kernels) -- together with the host plan builder, and the
objects are linked into one shared library.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "csrc")
OUT = os.path.join(HERE, "libeggman_hip.so")
PARTS = [0, 1, 2, 3, 4, 5]          # 0: state + C ABI; 1..5: EGGMAN_F32 .. EGGMAN_I128

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no fused multiply-add anywhere: each state's sum must round exactly as the reference's
    "-ffp-contract=off",
    # never emit packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
NOISE = "packed-fp32-ops"


def _run_all(cmds, verbose):
    procs = []
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        msgs = [ln for ln in out.splitlines() if NOISE not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(" ".join(cmd[-3:]) + ":\n" + "\n".join(msgs))
        elif verbose and msgs:
            print("\n".join(msgs))
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(failed))


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory(prefix="eggman_build_") as tmp:
        objs, cmds = [], []
        for part in PARTS:
            o = os.path.join(tmp, f"eggman_part{part}.o")
            cmds.append([hipcc, *HIPCC_FLAGS, f"-DEGG_PART={part}", "-c", "-o", o,
                         os.path.join(CSRC, "eggman.hip")])
            objs.append(o)
        o = os.path.join(tmp, "plan_host.o")
        cmds.append([hipcc, *HIPCC_FLAGS, "-c", "-o", o, os.path.join(CSRC, "plan_host.cpp")])
        objs.append(o)
        _run_all(cmds, verbose)
        tmp_so = OUT + ".tmp"
        _run_all([[hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_so, *objs,
                   "-lpthread"]], verbose)
        os.replace(tmp_so, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
    sys.exit(0)
