"""The host plan builders (plan_host.cpp: the plan, the reverse plan and the fused reverse step's
tables) under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU: any out-of-bounds
access, overflow or other undefined behaviour aborts the check.  Host code only -- the sanitizer
flags apply to the host compilation, no device code is built."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "theeggman.jl_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SANITIZE = "-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=all"


@pytest.mark.skipif(not os.path.exists(HIPCC) and shutil.which("hipcc") is None, reason="no hipcc")
def test_plan_builders_are_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "host_plan_check")
    cmd = [HIPCC, "-O1", "-g", "-std=c++17", *SANITIZE.split(), "-I", CSRC, "-o", exe,
           os.path.join(ROOT, "tests", "host_plan_check.cpp"), os.path.join(CSRC, "plan_host.cpp"),
           "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe] + [str(K) for K in range(4, 27, 2)], env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = [ln for ln in r.stdout.splitlines() if ln.startswith("K=")]
    assert len(rows) == 12
    for ln in rows:
        f = dict(kv.split("=") for kv in ln.split())
        assert f["rc"] == "0" and f["entries"] == f["terms"] == f["ntrans"], ln
