"""CPU check of the serving loop's admission policy (backfill with a bounded
number of bypasses, n-body_amd/csrc/admission.hpp), compiled with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_admission_policy_under_sanitizers(tmp_path):
    exe = tmp_path / "admission_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-I", os.path.join(ROOT, "n-body_amd", "csrc"),
                    os.path.join(ROOT, "tests", "native", "admission_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
