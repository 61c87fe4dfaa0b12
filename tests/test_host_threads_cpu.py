"""The library's host threading (csrc/host_threads.h: the ranges of the j pack of
gpunb_send_ and of the list assembly of gpunb_regf_) built with g++ under
ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer, and run
on the CPU (tests/sanitize/host_threads_check.cpp).  The HIP runtime does not start
under ThreadSanitizer, so the library itself is race-checked through this header."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent
SRC = HERE / "sanitize" / "host_threads_check.cpp"
INC = HERE.parent / "nbody6ppgpu-beijing_amd" / "csrc"


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
@pytest.mark.parametrize("san", ["thread", "address,undefined"])
def test_host_threads_under_sanitizer(tmp_path, san):
    exe = tmp_path / "check"
    cc = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-pthread", f"-fsanitize={san}",
                         "-fno-sanitize-recover=all", f"-I{INC}", str(SRC), "-o", str(exe)],
                        capture_output=True, text=True, timeout=240)
    assert cc.returncode == 0, cc.stderr
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0 and "host_threads_check: ok" in r.stdout, r.stdout + r.stderr
    assert "Sanitizer" not in r.stderr, r.stderr
