"""Writes the golden fixtures in this directory from the oracle.

Each case is one .npz (numpy, no pickles): the input bodies, the oracle's
bounding box, body order, sorted keys (linear BVH), tree nodes, and the
accelerations with per-body interaction counts for every body.  The cases are
small Plummer spheres (the generator is nbg_generate_plummer) plus the bodies
of the reference's scenarios/threebody.json (read from the reference checkout
when it is present; its particles' mass, position and velocity are stored in
the fixture).

    python tests/golden/make_golden.py            # rewrites tests/golden/*.npz
    python tests/golden/make_golden.py NAME ...   # rewrites the named cases only

Solvers 3 and 4 are the quadrupole octree / linear BVH; their fixtures also
hold every node's second moment (quad6).
"""

import hashlib
import importlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SCENARIO = "/root/reference/scenarios/threebody.json"


def threebody():
    """(mass, position, velocity) of every particle with a mass."""
    with open(SCENARIO) as f:
        parts = json.load(f)["particles"]
    out = []
    for q in parts:
        if "mass" not in q:
            continue
        pos = q.get("position", {"x": 0.0, "y": 0.0, "z": 0.0})
        vel = q.get("velocity", {"x": 0.0, "y": 0.0, "z": 0.0})
        out.append((q["mass"], (pos["x"], pos["y"], pos["z"]), (vel["x"], vel["y"], vel["z"])))
    return out

# (name, solver, n or "threebody", leaf, bits, theta, eps2)
CASES = [
    ("lbvh63_n3", 1, 3, 1, 63, 0.5, 1e-4),
    ("lbvh30_n64", 1, 64, 4, 30, 0.5, 1e-4),
    ("lbvh63_n1000", 1, 1000, 16, 63, 0.5, 1e-4),
    ("lbvh30_n1000", 1, 1000, 16, 30, 0.3, 1e-4),
    ("octree_n3", 0, 3, 1, 63, 0.5, 1e-4),
    ("octree_n64", 0, 64, 4, 63, 0.7, 1e-4),
    ("octree_n1000", 0, 1000, 16, 63, 0.5, 1e-4),
    ("lbvh63_threebody", 1, "threebody", 64, 63, 0.3, 1e-5),
    ("octree_threebody", 0, "threebody", 64, 63, 0.3, 1e-5),
    ("lbvh30q_n64", 4, 64, 4, 30, 0.5, 1e-4),
    ("lbvh63q_n1000", 4, 1000, 16, 63, 0.72578126, 1e-4),
    ("octreeq_n1000", 3, 1000, 16, 63, 0.41289064, 1e-4),
    ("octreeq_threebody", 3, "threebody", 64, 63, 0.3, 1e-5),
]


def bodies_for(nbg, n):
    if n == "threebody":
        tb = threebody()
        m = np.array([t[0] for t in tb], np.float32)
        p = np.array([t[1] for t in tb], np.float32)
        v = np.array([t[2] for t in tb], np.float32)
        return nbg.Bodies(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m, v[:, 0].copy(), v[:, 1].copy(),
                          v[:, 2].copy())
    return nbg.plummer(n, seed=4242 + n)


def main():
    import oracle

    nbg = importlib.import_module("n-body_amd")
    index = {}
    only = set(sys.argv[1:])
    for name, solver, n, leaf, bits, theta, eps2 in CASES:
        if only and name not in only:
            continue
        if n == "threebody" and not os.path.exists(SCENARIO):
            print(f"skipping {name}: {SCENARIO} not present")
            continue
        b = bodies_for(nbg, n)
        t = oracle.Tree(solver, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, bits=bits, eps2=eps2)
        keys, perm = t.order()
        nd = t.nodes()
        acc, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(
            path, solver=solver, leaf=leaf, bits=bits, theta=np.float32(theta), eps2=np.float32(eps2),
            x=b.x, y=b.y, z=b.z, m=b.m, vx=b.vx, vy=b.vy, vz=b.vz,
            bbox=t.bbox(), perm=perm, keys=keys if solver in (1, 4) else np.zeros(0, np.uint64),
            first=nd["first"], count=nd["count"], leafflag=nd["leaf"], depth=nd["depth"], parent=nd["parent"],
            com4=nd["com4"], geo6=nd["geo6"], acc=acc, pp=pp, np_=npc,
            **({"quad6": t.quads()} if solver in (3, 4) else {}))
        index[name] = hashlib.sha256(open(path, "rb").read()).hexdigest()
    old = {}
    if os.path.exists(os.path.join(HERE, "SHA256SUMS.json")):
        with open(os.path.join(HERE, "SHA256SUMS.json")) as f:
            old = json.load(f)
    old.update(index)
    index = old
    with open(os.path.join(HERE, "SHA256SUMS.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
