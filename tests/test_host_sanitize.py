"""The host shell (parser, problem registry, writers, checkpoints, indicator
weights reader) under AddressSanitizer and UndefinedBehaviorSanitizer, on the
CPU: tests/host_sanitize.cpp is built with g++ from the host sources and run
through malformed inputs, every registered problem and damaged files.  Any
sanitizer report aborts the program (-fno-sanitize-recover, halt_on_error)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "how-mhd-kokkos_amd"
HOST_SRC = ["host/params.cpp", "host/problems.cpp", "host/output.cpp", "host/checkpoint.cpp"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_host_shell_under_sanitizers(tmp_path):
    exe = tmp_path / "host_sanitize"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fopenmp", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", f"-I{ROOT / 'include'}",
           *[str(PKG / s) for s in HOST_SRC], str(ROOT / "tests" / "host_sanitize.cpp"), "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-3000:]
    scratch = tmp_path / "out"
    scratch.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(scratch), str(PKG / "weights" / "indicator_mlp.txt")],
                       capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    print(out[-2000:])
    assert "Sanitizer" not in out and "runtime error" not in out, out[-3000:]
    assert r.returncode == 0, out[-3000:]
    assert "host_sanitize: 0 failures" in out
