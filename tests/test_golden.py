"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py):
small Plummer spheres and the reference's three-body scenario, both solvers
and their quadrupole variants, both key widths.  The CPU tests check the files' hashes and that the oracle
still reproduces every array; the GPU tests check the device's ORDERED pass
against the stored arrays without the oracle."""

import glob
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))


def _load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def test_fixture_hashes():
    with open(os.path.join(HERE, "SHA256SUMS.json")) as f:
        sums = json.load(f)
    assert sorted(sums) == CASES and len(CASES) >= 7
    for name in CASES:
        h = hashlib.sha256(open(os.path.join(HERE, name + ".npz"), "rb").read()).hexdigest()
        assert h == sums[name], name


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_fixture(orc, name):
    d = _load(name)
    solver = int(d["solver"])
    t = orc.Tree(solver, d["x"], d["y"], d["z"], d["m"], theta=float(d["theta"]), leaf=int(d["leaf"]),
                 bits=int(d["bits"]), eps2=float(d["eps2"]))
    keys, perm = t.order()
    assert np.array_equal(t.bbox(), d["bbox"])
    assert np.array_equal(perm, d["perm"])
    if solver in (1, 4):
        assert np.array_equal(keys, d["keys"])
    nd = t.nodes()
    if "quad6" in d:
        q = t.quads()
        assert np.array_equal(np.isnan(q), np.isnan(d["quad6"]))
        assert np.array_equal(np.nan_to_num(q), np.nan_to_num(d["quad6"]))
    for k, fk in (("first", "first"), ("count", "count"), ("leaf", "leafflag"), ("depth", "depth"),
                  ("parent", "parent"), ("geo6", "geo6")):
        assert np.array_equal(nd[k], d[fk]), k
    assert np.array_equal(np.isnan(nd["com4"]), np.isnan(d["com4"]))
    assert np.array_equal(np.nan_to_num(nd["com4"]), np.nan_to_num(d["com4"]))
    acc, pp, npc = t.accel(d["x"], d["y"], d["z"], counts=True)
    assert np.array_equal(acc, d["acc"]) and np.array_equal(pp, d["pp"]) and np.array_equal(npc, d["np_"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_device_reproduces_fixture(nbg, gpu_ctx, name):
    d = _load(name)
    solver = int(d["solver"])
    n = len(d["x"])
    p = nbg.default_params()
    p.solver, p.theta, p.leaf_size, p.morton_bits = solver, float(d["theta"]), int(d["leaf"]), int(d["bits"])
    p.eps2, p.sum_mode = float(d["eps2"]), nbg.SUM_ORDERED
    s = nbg.System(gpu_ctx, p, n)
    s.set_state(nbg.Bodies(d["x"], d["y"], d["z"], d["m"], d["vx"], d["vy"], d["vz"]))
    acc = s.update_accelerations()
    assert np.array_equal(acc.view(np.uint32), d["acc"].view(np.uint32))
    assert s.counts() == (int(d["pp"].sum()), int(d["np_"].sum()))
    assert np.array_equal(s.debug_bbox(), d["bbox"])
    keys, perm = s.debug_order()
    if solver in (1, 4):
        assert np.array_equal(keys, d["keys"]) and np.array_equal(perm, d["perm"])
    s.close()
