"""The host-side code of the path under AddressSanitizer and
UndefinedBehaviorSanitizer: tests/sanitize/san_host.cpp links the O(N) host
stages of the product (rcwa-cuda_amd/csrc/host_math.cpp) and the CPU oracle,
built with -fsanitize=address,undefined (g++, no device code), and runs them
over seeded random problems and edge shapes.  A sanitizer report aborts it;
its numerical checks tie the host tail to the oracle (per-order R, T to
1e-12 on the oracle's own S-matrix columns)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_and_ubsan():
    b = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(HERE, "san_host")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
