"""The host C++ of the oracle and of the library's sample selection, built
with g++ under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/cpp/sanitize_check.cpp) and run on small 2-D and 3-D problems: every
oracle entry point, dense and sparse sampling, masks, affine, B-spline and
composite transforms, the filters, the samplers and degenerate inputs.  No
sanitizer report and every call's expected status."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_host_code_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", CPP, "sanitize_check"], check=True, timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(CPP, "sanitize_check")], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
