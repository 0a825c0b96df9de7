"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
server's device-memory arena, the batch's slot planner, the DEM generator,
reader and writer, and the CPU oracle, built from their sources with g++/gcc
and -fsanitize=address,undefined by tests/native/Makefile (no device code)
and exercised by tests/native/host_checks.cpp; any sanitizer finding aborts
the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")


def test_host_code_under_sanitizers(tmp_path):
    if shutil.which("g++") is None or shutil.which("gcc") is None:
        pytest.skip("no host compiler")
    out = tmp_path / "build"
    subprocess.run(["make", "-s", "-C", NATIVE, f"OUT={out}"], check=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([str(out / "host_checks")], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host checks ok" in r.stdout
