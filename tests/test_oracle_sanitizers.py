"""The oracle's C++ (oracle/oracle.cpp) built with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and driven by
tests/native/oracle_check.cpp over every tree kind, leaf size and a set of
degenerate inputs (coincident, collinear, single body, empty), with topology
consistency checks and the theta-0 walk against the FP64 direct sum.  A
memory or UB error in the checker would silently corrupt every parity claim
built on it.  Host code only: runs without a GPU."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_check"
    # the oracle's own flags (oracle/Makefile: IEEE binary32, no contraction)
    # plus the sanitizers; OpenMP stays on so its pragmas compile as shipped
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
                    os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(ROOT, "tests", "native", "oracle_check.cpp"),
                    "-o", str(exe)], check=True)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "all checks passed" in r.stdout
