"""The oracle's pins catch plausible mistakes: each case builds oracle.cpp with
one deliberate mistake planted (-DORC_MUTATE=k; the lines sit beside the
ones they change) and runs the pin that covers that function against it,
which must fail, while the same pin passes on the build without a mistake.

  k  planted mistake                                       pin
  1  octree child offset side/2 instead of (side/2)/2       test_octree_structure (child centres, containment)
  2  linear-BVH leaf COM summed in reverse order            test_com_summation_order_restated
  3  FMM exclusion regions [c - s/2, c + s/2]               test_fmm_recommend_truth_table
  4  FMM descend-side rule with the two sides swapped       test_fmm_recommend_truth_table
  5  FMM particle-node cell closed above                    test_fmm_particle_far_truth_table
  6  FMM node-node distance from the active centre, not COM test_fmm_recommend_truth_table
  7  octree internal COM loop dropping the last child       test_com_summation_order_restated
"""

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "oracle.cpp")

CASES = {
    1: "test_octree_structure",
    2: "test_com_summation_order_restated",
    3: "test_fmm_recommend_truth_table",
    4: "test_fmm_recommend_truth_table",
    5: "test_fmm_particle_far_truth_table",
    6: "test_fmm_recommend_truth_table",
    7: "test_com_summation_order_restated",
}


def _build(k, out_dir):
    lib = os.path.join(out_dir, f"liboracle_m{k}.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-shared", f"-DORC_MUTATE={k}", "-o", lib, SRC], check=True)
    return lib


def _run_pin(lib, pin):
    env = dict(os.environ, ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(ROOT, "tests", "test_oracle.py"), "-k", pin],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout[-2000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("k", sorted(CASES))
def test_pin_catches_planted_mistake(tmp_path, k):
    pin = CASES[k]
    rc, out = _run_pin(_build(k, str(tmp_path)), pin)
    assert rc == 1, f"mistake {k} was not caught by {pin}:\n{out}"
    assert "failed" in out


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pins_pass_without_a_mistake(tmp_path):
    lib = _build(0, str(tmp_path))
    for pin in sorted(set(CASES.values())):
        rc, out = _run_pin(lib, pin)
        assert rc == 0, f"{pin} fails on the unmutated oracle:\n{out}"
