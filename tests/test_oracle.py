"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each test names what fixes the expected value: a known-answer test from the
reference's own suite, a closed form, an invariant, or brute force on tiny
inputs.  Reference tests cited as src/NBody/... paths.
"""

import math

import numpy as np
import pytest

from treecheck import com_restatement, lbvh_check, octree_check, same_bits, summaries_check


def _rand_bodies(n, seed, spread=1.0):
    r = np.random.default_rng(seed)
    x, y, z = (r.normal(0, spread, n).astype(np.float32) for _ in range(3))
    m = r.uniform(0.5, 1.5, n).astype(np.float32)
    return x, y, z, m


# ------------------------------------------------------------------ reference KATs

def test_kat_particle_particle(orc):
    """src/NBody/Physics/test/gravity.cpp:13-26: pA=(0,0,0), mA=2, pB=(2,0,0),
    G=2, default epsilon 1e-5  =>  |a - (-1,0,0)| < 5e-4."""
    a = orc.particle_particle([0, 0, 0], 2.0, [2, 0, 0], 2.0, 1e-5)
    assert np.linalg.norm(a - np.array([-1, 0, 0], np.float32)) < 5e-4
    # closed form: -G m R / (R^2 + eps)^(3/2) = -8 / 4.00001^1.5
    exact = -8.0 / (4.00001 ** 1.5)
    assert abs(float(a[0]) - exact) < 4 * np.spacing(np.float32(1.0))
    assert a[1] == 0 and a[2] == 0


def test_kat_self_and_coincident_skip(orc):
    """Gravity.h:203: a body exactly at the passive position contributes 0."""
    assert np.all(orc.particle_particle([1, 2, 3], 5.0, [1, 2, 3], 1.0, 0.0) == 0)


def test_kat_center_of_mass(orc):
    """src/NBody/Physics/test/centerOfMassSummary.cpp:12-47: unit masses at
    z = -1 and +1; the monopole at (0, 20, 0) is within 0.5 % of the sum."""
    c = orc.com(np.zeros(2), np.zeros(2), np.array([-1, 1.0]), np.ones(2))
    assert np.array_equal(c, np.array([0, 0, 0, 2], np.float32))
    p = [0, 20, 0]
    direct = orc.particle_particle([0, 0, -1], 1.0, p, 1.0, 1e-5) + orc.particle_particle([0, 0, 1], 1.0, p, 1.0, 1e-5)
    approx = orc.particle_particle(c[:3], c[3], p, 1.0, 1e-5)
    assert np.linalg.norm(direct - approx) / np.linalg.norm(direct) < 0.005


def test_kat_box_intersection(orc):
    """src/NBody/Physics/test/boundingBox.cpp:12-32 truth table."""
    box = [0, 0, 0, 1, 1, 1]
    yes = [[0.5, 0.5, 0.5, 1.5, 1.5, 1.5], [-0.5, -0.5, -0.5, 0.5, 0.5, 0.5],
           [0.25, 0.25, -0.5, 0.75, 0.75, 1.5], [0.25, 0.25, 0.25, 0.75, 0.75, 0.75]]
    no = [[0, 0, 10, 1, 1, 11], [10, 10, 0, 11, 11, 1]]
    for b in yes:
        assert orc.boxes_intersect(box, b)
    assert orc.boxes_intersect(yes[3], box)
    for b in no:
        assert not orc.boxes_intersect(box, b)


# ------------------------------------------------------------------ Morton keys

@pytest.mark.parametrize("bits", [30, 63])
def test_expand_bits_matches_reference_constants(orc, bits):
    """The oracle's per-bit spread equals the reference's magic-constant
    expandBits10/21 (MortonSort.h:27-44) on edges and random values."""
    per = 10 if bits == 30 else 21
    r = np.random.default_rng(1)
    vals = [0, 1, 2, (1 << per) - 1, (1 << per) - 2, 1 << (per - 1)] + r.integers(0, 1 << per, 2000).tolist()
    for v in vals:
        assert orc.lib().orc_expand_bits(v, bits) == orc.lib().orc_expand_bits_magic(v, bits)


def _morton_np(x, y, z, bb, bits):
    """Closed form of toMortonCode in numpy float32 (IEEE, no FMA)."""
    mn = bb[:3].astype(np.float32)
    dims = (bb[3:] - bb[:3]).astype(np.float32)
    scale = np.float32(max(max(dims[0], dims[1]), dims[2]))
    per, mul, hi = (10, 1024.0, 1023.0) if bits == 30 else (21, 2097152.0, 2097151.0)
    key = np.zeros(len(x), dtype=object)
    for shift, c, m0 in ((2, x, mn[0]), (1, y, mn[1]), (0, z, mn[2])):
        v = (np.float32(c) - m0) / scale if np.isscalar(c) else (c.astype(np.float32) - m0) / scale
        v = (v * np.float32(mul)).astype(np.float32)
        v = np.where(v < 0, np.float32(0), v)
        v = np.where(np.float32(hi) < v, np.float32(hi), v).astype(np.float32)
        q = np.where(np.isnan(v), np.float32(0), v).astype(np.uint64)  # zero-extent box: NaN -> 0
        for b in range(per):
            key = key + (((q >> np.uint64(b)) & np.uint64(1)).astype(object) << (3 * b + shift))
    return np.array(key.tolist(), dtype=np.uint64)


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_closed_form(orc, bits):
    x, y, z, _ = _rand_bodies(3000, 7)
    bb = orc.bbox(x, y, z)
    assert np.array_equal(bb[:3], [x.min(), y.min(), z.min()]) and np.array_equal(bb[3:], [x.max(), y.max(), z.max()])
    k = orc.morton(x, y, z, bits, bb)
    assert np.array_equal(k, _morton_np(x, y, z, bb, bits))


@pytest.mark.parametrize("bits", [30, 63])
def test_morton_corners_and_axis_order(orc, bits):
    """Min corner -> 0, max corner -> all ones (clamped), x is the most
    significant axis of each triplet (4x + 2y + z)."""
    x = np.array([0, 1, 0, 0, 1], np.float32)
    y = np.array([0, 1, 0, 1, 0], np.float32)
    z = np.array([0, 1, 1, 0, 0], np.float32)
    bb = orc.bbox(x, y, z)
    k = orc.morton(x, y, z, bits, bb)
    per = 10 if bits == 30 else 21
    assert k[0] == 0
    assert k[1] == (1 << (3 * per)) - 1
    zbits = sum(1 << (3 * b) for b in range(per))
    assert k[2] == zbits and k[3] == zbits << 1 and k[4] == zbits << 2


# ------------------------------------------------------------------ sort

@pytest.mark.parametrize("n", [0, 1, 2, 1000, 20000])
def test_sort_is_stable_key_order(orc, n):
    r = np.random.default_rng(n)
    keys = r.integers(0, 50, n).astype(np.uint64)  # many ties
    perm = orc.sort(keys)
    assert np.array_equal(perm, np.argsort(keys, kind="stable").astype(np.uint32))


# ------------------------------------------------------------------ LBVH

@pytest.mark.parametrize("n,leaf,bits", [(1, 16, 63), (2, 1, 63), (3, 1, 30), (64, 4, 63), (1000, 16, 30),
                                         (1000, 1, 63), (10000, 16, 30), (10000, 64, 63)])
def test_lbvh_topology_and_summaries(orc, nbg, n, leaf, bits):
    b = nbg.plummer(n, seed=n)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, leaf=leaf, bits=bits)
    keys, perm = t.order()
    assert np.all(keys[1:] >= keys[:-1])
    bb = orc.bbox(b.x, b.y, b.z)
    assert np.array_equal(keys, _morton_np(b.x, b.y, b.z, bb, bits)[perm])
    nd = t.nodes()
    lbvh_check(keys, nd["first"], nd["count"], nd["leaf"], leaf)
    pos = np.stack([b.x, b.y, b.z], 1)[perm]
    summaries_check(pos, b.m[perm], nd["first"], nd["count"], nd["com4"], nd["geo6"])


def test_lbvh_duplicate_keys_make_oversized_leaves(orc):
    """Leaf rule 'all keys equal' (LinearBVH.h:114-119): 40 coincident bodies
    stay in one leaf even with leaf size 4."""
    x = np.concatenate([np.zeros(40), np.linspace(1, 2, 10)]).astype(np.float32)
    y = np.zeros(50, np.float32)
    z = np.zeros(50, np.float32)
    t = orc.Tree(1, x, y, z, np.ones(50, np.float32), leaf=4, bits=30)
    nd = t.nodes()
    keys, _ = t.order()
    lbvh_check(keys, nd["first"], nd["count"], nd["leaf"], 4)
    assert nd["count"][nd["leaf"]].max() == 40


# ------------------------------------------------------------------ octree

@pytest.mark.parametrize("n,leaf,max_depth", [(1, 16, 32), (9, 1, 32), (1000, 16, 32), (5000, 8, 3), (20000, 64, 32)])
def test_octree_structure(orc, nbg, n, leaf, max_depth):
    b = nbg.plummer(n, seed=100 + n)
    t = orc.Tree(0, b.x, b.y, b.z, b.m, leaf=leaf, max_depth=max_depth)
    _, perm = t.order()
    nd = t.nodes()
    pos = np.stack([b.x, b.y, b.z], 1)
    octree_check(pos, perm, nd, leaf, max_depth)
    # root cube: centre min + diag/2, side maxSide * 1.1 (Octree.h:189-190)
    bb = orc.bbox(b.x, b.y, b.z)
    diag = (bb[3:] - bb[:3]).astype(np.float32)
    assert np.array_equal(nd["geo6"][0, :3], (bb[:3] + diag / np.float32(2)).astype(np.float32))
    assert nd["geo6"][0, 3] == np.float32(diag.max() * np.float32(1.1))
    # leaf contents are in ascending original index (stable partitions)
    for f, c, lf in zip(nd["first"], nd["count"], nd["leaf"]):
        if lf and c > 1:
            assert np.all(np.diff(perm[f:f + c].astype(np.int64)) > 0)
    # An empty level-1 octant is still split and summarised from 8 empty
    # children: 0/0 = NaN COM, which poisons the root's COM (Tree.h:196-216,
    # CenterOfMassSummary.h:44-55).  The root is never accepted by the MAC.
    lvl1_empty = np.any((nd["depth"] == 1) & (nd["count"] == 0))
    assert np.isnan(nd["com4"][0, 0]) == bool(lvl1_empty)
    # summaries: COM/mass of non-empty nodes
    populated = (nd["count"] > 0) & ~np.isnan(nd["com4"][:, 0])
    keep = {k: v[populated] for k, v in nd.items()}
    pos_sorted = pos[perm]
    # octree nodes index the leaf order, which is contiguous per subtree
    summaries_check(pos_sorted, b.m[perm], keep["first"], keep["count"], keep["com4"])


# ------------------------------------------------------------------ walk

def _py_walk_lbvh(nd, pos, mass, p, theta2, G, eps2):
    """Brute-force restatement of activeTree over the LBVH in numpy float32,
    written from the reference text (activeTree.h:24-79, Gravity.h:199-209,
    DiagonalOverCenterDistance.h:18-22) independently of the oracle."""
    f32 = np.float32
    first, count, leaf, depth, com4, geo = nd["first"], nd["count"], nd["leaf"], nd["depth"], nd["com4"], nd["geo6"]
    kids = {}
    stack = []
    for i, d in enumerate(depth.tolist()):
        while stack and depth[stack[-1]] >= d:
            stack.pop()
        if stack:
            kids.setdefault(stack[-1], []).append(i)
        stack.append(i)

    def pp(a, m, q):
        if a[0] == q[0] and a[1] == q[1] and a[2] == q[2]:
            return np.zeros(3, f32)
        R = (q - a).astype(f32)
        r = f32(np.sqrt(f32(f32(f32(R[0] * R[0]) + f32(R[1] * R[1])) + f32(R[2] * R[2])) + f32(eps2)))
        r3 = f32(r * f32(r * r))
        return (((-R) * f32(m)).astype(f32) * f32(G)).astype(f32) / r3

    def walk(i):
        mn, mx = geo[i, :3], geo[i, 3:]
        c = ((mx + mn) / f32(2)).astype(f32)
        dg = (mx - mn).astype(f32)
        diag2 = f32(f32(f32(dg[0] * dg[0]) + f32(dg[1] * dg[1])) + f32(dg[2] * dg[2]))
        R = (p - c).astype(f32)
        d2 = f32(f32(f32(R[0] * R[0]) + f32(R[1] * R[1])) + f32(R[2] * R[2]))
        with np.errstate(divide="ignore", invalid="ignore"):
            if f32(diag2 / d2) < theta2:
                return pp(com4[i, :3], com4[i, 3], p)
        if leaf[i]:
            acc = np.zeros(3, f32)
            for j in range(first[i], first[i] + count[i]):
                acc = (acc + pp(pos[j], mass[j], p)).astype(f32)
            return acc
        r0, r1 = walk(kids[i][0]), walk(kids[i][1])
        return ((np.zeros(3, f32) + r0).astype(f32) + r1).astype(f32)

    return walk(0)


@pytest.mark.parametrize("n,leaf,theta", [(3, 1, 0.5), (64, 4, 0.5), (64, 1, 0.9), (200, 8, 0.3)])
def test_lbvh_walk_bruteforce_bitexact(orc, nbg, n, leaf, theta):
    b = nbg.plummer(n, seed=3 + n)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, bits=63, eps2=1e-4)
    keys, perm = t.order()
    nd = t.nodes()
    pos = np.stack([b.x, b.y, b.z], 1)
    acc = t.accel(b.x, b.y, b.z)
    theta2 = np.float32(np.float32(theta) * np.float32(theta))
    for i in range(n):
        ref = _py_walk_lbvh(nd, pos[perm], b.m[perm], pos[i], theta2, 1.0, 1e-4)
        assert np.array_equal(acc[i], ref), f"body {i}"


def _py_walk_octree(nd, pos, mass, p, theta, G, eps2):
    """Brute-force restatement of activeTree over the octree in numpy float32
    from the reference text (activeTree.h:24-79 with libstdc++'s 4-way
    transform_reduce grouping over 8 children, SideLengthOverDistance.h:17-22,
    DescentCriterionType.h:81-85, BoundingBox.h:61-64), independent of the
    oracle's walk; it reads only the oracle's tree."""
    f32 = np.float32
    first, count, leaf, com4, geo, parent = (nd["first"], nd["count"], nd["leaf"], nd["com4"], nd["geo6"],
                                            nd["parent"])
    kids = {}
    for i, q in enumerate(parent.tolist()):
        if q >= 0:
            kids.setdefault(q, []).append(i)
    z = np.zeros(3, f32)

    def pp(a, m, q):
        if a[0] == q[0] and a[1] == q[1] and a[2] == q[2]:
            return z.copy()
        R = (q - a).astype(f32)
        r = f32(np.sqrt(f32(f32(f32(R[0] * R[0]) + f32(R[1] * R[1])) + f32(R[2] * R[2])) + f32(eps2)))
        r3 = f32(r * f32(r * r))
        return (((-R) * f32(m)).astype(f32) * f32(G)).astype(f32) / r3

    def add(a, b):
        return (a + b).astype(f32)

    def walk(i):
        if count[i] == 0:
            return z.copy()
        c, side = geo[i, :3], f32(geo[i, 3])
        R = (p - com4[i, :3]).astype(f32)
        dist = f32(np.sqrt(f32(f32(f32(R[0] * R[0]) + f32(R[1] * R[1])) + f32(R[2] * R[2]))))
        lo, hi = (c - side).astype(f32), (c + side).astype(f32)
        inside = bool(np.all(lo <= p) and np.all(p < hi))
        with np.errstate(divide="ignore", invalid="ignore"):
            if f32(side / dist) < f32(theta) and not inside:
                return pp(com4[i, :3], com4[i, 3], p)
        if leaf[i]:
            acc = z.copy()
            for j in range(first[i], first[i] + count[i]):
                acc = add(acc, pp(pos[j], mass[j], p))
            return acc
        ks = kids[i]
        assert len(ks) == 8
        r = [walk(k) for k in ks]
        init = add(z, add(add(r[0], r[1]), add(r[2], r[3])))
        return add(init, add(add(r[4], r[5]), add(r[6], r[7])))

    return walk(0)


@pytest.mark.parametrize("n,leaf,theta", [(3, 1, 0.5), (64, 4, 0.5), (64, 1, 0.9), (300, 8, 0.3), (300, 2, 1.5)])
def test_octree_walk_bruteforce_bitexact(orc, nbg, n, leaf, theta):
    b = nbg.plummer(n, seed=5 + n)
    t = orc.Tree(0, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, eps2=1e-4)
    _, perm = t.order()
    nd = t.nodes()
    pos = np.stack([b.x, b.y, b.z], 1)
    acc, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    if theta >= 0.9:
        assert npc.sum() > 0, "the case must exercise accepted nodes"
    for i in range(n):
        ref = _py_walk_octree(nd, pos[perm], b.m[perm], pos[i], theta, 1.0, 1e-4)
        assert np.array_equal(acc[i], ref), f"body {i}"


@pytest.mark.parametrize("solver", [0, 1])
def test_walk_theta_zero_is_direct_sum(orc, nbg, solver):
    """theta = 0 never accepts a node, so every body meets every body: the
    walk equals the direct sum and counts n pp per body, 0 np."""
    n = 2000
    b = nbg.plummer(n, seed=11)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.0, leaf=16, eps2=1e-4)
    acc, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert np.all(pp == n) and np.all(npc == 0)
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    err = np.linalg.norm(acc - ref, axis=1)
    scale = np.linalg.norm(ref, axis=1)
    assert np.median(err / scale) < 1e-6
    assert np.all(err <= 1e-4 * scale + 1e-5 * np.sqrt(np.mean(scale ** 2)))


@pytest.mark.parametrize("solver", [0, 1])
def test_walk_f64_is_the_exact_interaction_set(orc, nbg, solver):
    """accel_f64 takes the FP32 walk's opening decisions and sums each
    interaction in FP64.  At theta = 0 it is the FP64 direct sum (every body
    meets every body); at theta 0.5 / 0.3 its counts equal the FP32 walk's,
    and the FP32 walk (pairwise tree-ordered sums) stays within FP32
    rounding of it."""
    n = 2000
    eps2 = float(np.float32(1e-4))  # the tree's eps2 is FP32
    b = nbg.plummer(n, seed=11)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.0, leaf=16, eps2=1e-4)
    a64, pp, npc = t.accel_f64(b.x, b.y, b.z, counts=True)
    assert np.all(pp == n) and np.all(npc == 0)
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, eps2)
    assert np.all(np.linalg.norm(a64 - ref, axis=1) <= 1e-12 * np.linalg.norm(ref, axis=1))
    b = nbg.plummer(20000, seed=5)
    sel = np.arange(0, 20000, 7)
    for theta in (0.5, 0.3):
        t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=theta, leaf=16, eps2=1e-4)
        a32, pp32, np32 = t.accel(b.x[sel], b.y[sel], b.z[sel], counts=True)
        a64, pp64, np64 = t.accel_f64(b.x[sel], b.y[sel], b.z[sel], counts=True)
        assert np.array_equal(pp32, pp64) and np.array_equal(np32, np64)
        assert np64.sum() > 0
        mag = np.linalg.norm(a64, axis=1)
        err = np.linalg.norm(a32 - a64, axis=1) / np.maximum(mag, np.sqrt(np.mean(mag ** 2)))
        assert err.max() < 5e-6
        # the FP64 set is not the FP64 direct sum: the approximation error
        # is orders of magnitude above the rounding
        d = orc.direct_f64(b.x, b.y, b.z, b.m, b.x[sel], b.y[sel], b.z[sel], 1.0, eps2)
        assert np.median(np.linalg.norm(a64 - d, axis=1) / np.linalg.norm(d, axis=1)) > 1e-5


def test_walk_f64_refuses_quadrupole_trees(orc, nbg):
    b = nbg.plummer(100, seed=1)
    t = orc.Tree(4, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=1e-4)
    with pytest.raises(ValueError):
        t.accel_f64(b.x, b.y, b.z)


@pytest.mark.parametrize("solver,theta,tol", [(0, 0.5, 3e-3), (1, 0.5, 5e-3), (1, 0.3, 1e-3)])
def test_walk_force_error_vs_fp64(orc, nbg, solver, theta, tol):
    """Monopole BH error shrinks with theta; median error at the stated
    bound against the FP64 direct sum."""
    n = 20000
    b = nbg.plummer(n, seed=5)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=theta, leaf=16, eps2=1e-4)
    sel = np.arange(0, n, 10)
    acc = t.accel(b.x[sel], b.y[sel], b.z[sel])
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x[sel], b.y[sel], b.z[sel], 1.0, 1e-4)
    rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(rel) < tol


def test_interaction_counts_fall_with_theta(orc, nbg):
    b = nbg.plummer(5000, seed=9)
    tot = []
    for th in (0.2, 0.4, 0.8):
        t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=th, leaf=16)
        _, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
        tot.append(int(pp.sum() + npc.sum()))
    assert tot[0] > tot[1] > tot[2]


# ------------------------------------------------------------------ quadrupole (solvers 3, 4)

def _moment_f64(pos, m):
    """FP64 centre of mass and raw second moment sum m (x - c)(x - c)^T."""
    pos = np.asarray(pos, np.float64)
    m = np.asarray(m, np.float64)
    c = (pos * m[:, None]).sum(0) / m.sum()
    o = pos - c
    return c, np.einsum("k,ki,kj->ij", m, o, o)


def _q6(Q):
    return np.array([Q[0, 0], Q[0, 1], Q[0, 2], Q[1, 1], Q[1, 2], Q[2, 2]])


def test_kat_quadrupole_three_bodies(orc):
    """src/NBody/Physics/test/multipoleMomentSummary.cpp:12-67: masses 1, 1, 2
    at (0,0,-1), (0,0,1), (0,3,1), target (0,10,0).  Closed form: COM
    (0, 1.5, 0.5) and moment yy = 9, yz = 3, zz = 3 (every product exact in
    FP32); the quadrupole force is nearer the direct sum than the monopole
    and within the test's absolute bound 0.005."""
    x = np.array([0, 0, 0], np.float32)
    y = np.array([0, 0, 3], np.float32)
    z = np.array([-1, 1, 1], np.float32)
    m = np.array([1, 1, 2], np.float32)
    t = orc.Tree(4, x, y, z, m, theta=0.5, leaf=16, bits=30, eps2=1e-5)
    nd = t.nodes()
    assert len(nd["first"]) == 1 and nd["leaf"][0]
    assert np.array_equal(nd["com4"][0], np.array([0, 1.5, 0.5, 4], np.float32))
    assert np.array_equal(t.quads()[0], np.array([0, 0, 0, 9, 3, 3], np.float32))
    acc, pp, npc = t.accel([0.0], [10.0], [0.0], counts=True)
    assert npc[0] == 1 and pp[0] == 0          # diag2 / d2 = 13 / 72.25 < 0.25: the root is accepted
    direct = orc.direct_f64(x, y, z, m, [0.0], [10.0], [0.0], 1.0, 1e-5)[0]
    mono = orc.particle_particle([0, 1.5, 0.5], 4.0, [0, 10, 0], 1.0, 1e-5)
    err_q = np.linalg.norm(acc[0] - direct)
    err_m = np.linalg.norm(mono - direct)
    assert err_q < 0.25 * err_m
    assert err_q <= 0.005                      # the reference test's absolute bound


def test_kat_quadrupole_real_scenario(orc):
    """src/NBody/Physics/test/gravity.cpp:122-153 (its four masses and target):
    the node's quadrupole approximation beats the monopole and stays within
    5% of the direct sum."""
    P = np.array([[12.237058, 7.9972196, 2.7898774], [15.70352, 12.367721, 3.9934757],
                  [12.150897, 13.606152, 3.4104824], [16.167948, 7.70833, 6.0922756]], np.float32)
    m = np.array([0.0570347, 0.112188, 0.371711, 0.144094], np.float32)
    p = np.array([0.41169, 1.1282315, 19.398197], np.float32)
    c, Q = _moment_f64(P, m)
    direct = orc.direct_f64(P[:, 0], P[:, 1], P[:, 2], m, [p[0]], [p[1]], [p[2]], 1.0, 1e-5)[0]
    quad = orc.node_particle_quad(c, float(m.sum()), _q6(Q), p, 1.0, 1e-5)
    mono = orc.particle_particle(c, float(m.sum()), p, 1.0, 1e-5)
    eq = np.linalg.norm(quad - direct) / np.linalg.norm(direct)
    em = np.linalg.norm(mono - direct) / np.linalg.norm(direct)
    assert eq < em and eq < 0.05


def test_quadrupole_isotropic_moment_adds_nothing(orc):
    """Closed form: D3 contracted with the identity is g3 R r^2 + 5 g2 R,
    which vanishes at eps2 = 0 (1/r is harmonic), so Q = s I gives the
    monopole force; a traceless moment does not."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.normal(0, 1, 3).astype(np.float32)
        p = (c + rng.normal(0, 4, 3)).astype(np.float32)
        s = float(rng.uniform(0.1, 2.0))
        iso = orc.node_particle_quad(c, 3.0, [s, 0, 0, s, 0, s], p, 1.5, 0.0)
        mono = orc.particle_particle(c, 3.0, p, 1.5, 0.0)
        assert np.allclose(iso, mono, rtol=2e-5, atol=1e-7 * np.linalg.norm(mono))
        tl = orc.node_particle_quad(c, 3.0, [s, 0, 0, -s, 0, 0], p, 1.5, 0.0)
        assert np.linalg.norm(tl - mono) > 1e-4 * np.linalg.norm(mono) * s / np.dot(p - c, p - c)


def test_quadrupole_is_second_order_taylor(orc):
    """The node-particle rule is the second-order expansion of the exact sum
    about the COM: at eps2 = 0 halving the cluster size h cuts the quadrupole
    error ~8x (next term h^3), the monopole error ~4x (h^2).  A wrong sign or
    factor on either contraction term leaves an h^2 error and fails."""
    rng = np.random.default_rng(8)
    base = rng.normal(0, 1, (12, 3))
    m = rng.uniform(0.5, 1.5, 12).astype(np.float32)
    p = np.array([0.3, -0.2, 1.0]) * 4.0
    errs_q, errs_m = [], []
    for h in (0.2, 0.1, 0.05):
        P = (base * h).astype(np.float32)
        c, Q = _moment_f64(P, m)
        direct = orc.direct_f64(P[:, 0], P[:, 1], P[:, 2], m, [p[0]], [p[1]], [p[2]], 1.0, 0.0)[0]
        q = orc.node_particle_quad(c, float(m.astype(np.float64).sum()), _q6(Q), p, 1.0, 0.0)
        mo = orc.particle_particle(c, float(m.astype(np.float64).sum()), p, 1.0, 0.0)
        errs_q.append(np.linalg.norm(q - direct))
        errs_m.append(np.linalg.norm(mo - direct))
    for k in range(2):
        assert 3.0 < errs_m[k] / errs_m[k + 1] < 5.5
        assert errs_q[k] / errs_q[k + 1] > 6.0
    assert errs_q[-1] < 0.1 * errs_m[-1]


@pytest.mark.parametrize("solver,leaf", [(3, 8), (4, 8), (4, 1)])
def test_quadrupole_node_moments_vs_fp64(orc, nbg, solver, leaf):
    """Every node's moment (leaf sums about the leaf COM, internal nodes by
    the parallel-axis combine of their children) equals the FP64 second
    moment of the node's bodies about its COM."""
    b = nbg.plummer(3000, seed=21)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf)
    nd = t.nodes()
    q = t.quads()
    _, perm = t.order()
    pos = np.stack([b.x, b.y, b.z], 1)
    checked = 0
    for i in range(len(nd["first"])):
        f, k = int(nd["first"][i]), int(nd["count"][i])
        if k == 0:
            assert np.all(q[i] == 0)
            continue
        sel = perm[f:f + k]
        c, Q = _moment_f64(pos[sel], b.m[sel])
        ref = _q6(Q)
        # FP32 offsets carry ~1 ulp of the positions (a one-body node's COM
        # need not round back to the body)
        ulp = 2.0 ** -23 * float(np.abs(pos[sel]).max())
        tol = 2e-4 * np.abs(ref).max() + 10.0 * float(b.m[sel].sum()) * ulp * ulp
        assert np.all(np.abs(q[i] - ref) <= tol), (i, q[i], ref)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("mono,quad,theta", [(0, 3, 0.5), (1, 4, 0.5), (1, 4, 0.7)])
def test_quadrupole_walk_same_decisions_smaller_error(orc, nbg, mono, quad, theta):
    """The quadrupole solvers take the monopole solvers' opening decisions
    (same MAC, same tree: equal per-body counts) and cut the median force
    error against the FP64 direct sum."""
    n = 20000
    b = nbg.plummer(n, seed=5)
    sel = np.arange(0, n, 10)
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x[sel], b.y[sel], b.z[sel], 1.0, 1e-4)
    out = {}
    for s in (mono, quad):
        t = orc.Tree(s, b.x, b.y, b.z, b.m, theta=theta, leaf=16, eps2=1e-4)
        acc, pp, npc = t.accel(b.x[sel], b.y[sel], b.z[sel], counts=True)
        rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
        out[s] = (np.median(rel), pp, npc)
    assert np.array_equal(out[mono][1], out[quad][1]) and np.array_equal(out[mono][2], out[quad][2])
    assert out[quad][0] < 0.5 * out[mono][0]


# ------------------------------------------------------------------ implicit FMM (solver 5)

@pytest.mark.parametrize("n,theta,leaf", [(1, 0.4, 16), (3, 0.4, 1), (2000, 0.4, 16), (20000, 0.4, 16),
                                          (20000, 0.8, 8), (5000, 0.3, 64)])
def test_fmm_every_source_reaches_every_target_once(orc, nbg, n, theta, leaf):
    """Interaction-set invariant of the dual traversal: run with each
    interaction adding its source mass (FP64; a node its bodies' mass, a local
    field the mass it gathered), every body receives exactly the total mass,
    i.e. every source body reaches it once, by one route (P2P, node-particle,
    particle-to-local or node-to-local).  A missed or doubled list entry fails."""
    b = nbg.plummer(n, seed=3 + n)
    b.m[:] = np.random.default_rng(n).uniform(0.5, 1.5, n).astype(np.float32)
    t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, eps2=1e-4)
    got = t.fmm_mass_received()
    total = float(b.m.astype(np.float64).sum())
    assert np.allclose(got, total, rtol=1e-12, atol=0)


def test_fmm_theta_zero_is_direct_sum(orc, nbg):
    """theta = 0 approximates nothing: every pair goes through a leaf P2P,
    so pp = n^2, the other counts are 0 and the result is the direct sum."""
    n = 3000
    b = nbg.plummer(n, seed=14)
    t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=0.0, leaf=16, eps2=1e-4)
    acc, c = t.fmm_accel(counts=True)
    assert c == (n * n, 0, 0, 0)
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    rel = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4


def test_fmm_force_error_and_counts_fall_with_theta(orc, nbg):
    """Against the FP64 direct sum the implicit FMM is as accurate as
    Barnes-Hut at the same theta (its near field is the same walk); a larger
    theta costs accuracy and saves interactions, P2P first."""
    n = 20000
    b = nbg.plummer(n, seed=5)
    sel = np.arange(0, n, 10)
    ref = orc.direct_f64(b.x, b.y, b.z, b.m, b.x[sel], b.y[sel], b.z[sel], 1.0, 1e-4)
    res = {}
    for th in (0.4, 0.7):
        t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=th, leaf=16, eps2=1e-4)
        acc, c = t.fmm_accel(counts=True)
        rel = np.linalg.norm(acc[sel] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        res[th] = (float(np.median(rel)), c)
    assert res[0.4][0] < 1.5e-3 and res[0.4][0] < res[0.7][0] < 1e-2
    assert res[0.7][1][0] < res[0.4][1][0] and sum(res[0.7][1]) < sum(res[0.4][1])


def test_fmm_local_expansion_is_first_order_taylor(orc):
    """A source's local field at c evaluated at c + h (vector plus gradient
    tensor) is the first-order expansion of its force: the error falls ~4x
    when h halves.  A wrong sign or factor on the gradient leaves an O(h)
    error (~2x)."""
    s = np.array([3.0, -2.0, 5.0], np.float32)
    c = np.array([0.2, 0.1, -0.3], np.float32)
    u = np.array([0.6, -0.3, 0.74], np.float64)
    errs = []
    for h in (0.2, 0.1, 0.05):
        p = (c + h * u).astype(np.float32)
        exact = orc.direct_f64([s[0]], [s[1]], [s[2]], [2.0], [p[0]], [p[1]], [p[2]], 1.5, 0.0)[0]
        got = orc.fmm_p2l(s, 2.0, c, p, 1.5, 0.0)
        errs.append(np.linalg.norm(got - exact))
    for k in range(2):
        assert 3.0 < errs[k] / errs[k + 1] < 5.5


def test_fmm_starting_nodes_partition_the_bodies(orc, nbg):
    """Tree::loadBalancedBreak(256): between 256 and 262 queued nodes whose
    ranges are disjoint and cover every body."""
    b = nbg.plummer(30000, seed=8)
    t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=0.4, leaf=16, eps2=1e-4)
    st = t.fmm_starts()
    nd = t.nodes()
    assert 256 <= len(st) <= 262
    cover = np.zeros(30000, np.int32)
    for s in st:
        f, k = int(nd["first"][s]), int(nd["count"][s])
        cover[f:f + k] += 1
    assert np.all(cover == 1)


# ------------------------------------------------------------------ integrator, direct sum, energy

def test_integrator_is_kick_then_drift(orc, nbg):
    """Solver::step (Solver.h:103-139): v' = v + a dt, p' = p + v' dt in FP32."""
    b = nbg.plummer(500, seed=2)
    dt = np.float32(1e-3)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=8, eps2=1e-4)
    a = t.accel(b.x, b.y, b.z)
    (x, y, z, vx, vy, vz), _ = orc.step(1, (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz), 1, theta=0.5, leaf=8, eps2=1e-4,
                                        dt=1e-3)
    ev = (b.vx + a[:, 0] * dt).astype(np.float32)
    assert np.array_equal(vx, ev)
    assert np.array_equal(x, (b.x + ev * dt).astype(np.float32))


def test_direct_f64_matches_numpy(orc, nbg):
    b = nbg.plummer(300, seed=4)
    got = orc.direct_f64(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    P = np.stack([b.x, b.y, b.z], 1).astype(np.float64)
    R = P[:, None, :] - P[None, :, :]
    r2 = (R ** 2).sum(-1) + 1e-4
    w = b.m.astype(np.float64)[None, :] / (np.sqrt(r2) * r2)
    np.fill_diagonal(w, 0.0)
    ref = -(R * w[:, :, None]).sum(1)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_plummer_energy_closed_form(orc, nbg):
    """Plummer W = -3 pi G M^2 / (32 a) (untruncated; r < 10a keeps 99.9 %
    of the mass) and virial 2K / |W| ~ 1."""
    n = 4000
    b = nbg.plummer(n, seed=21)
    W = orc.potential_f64(b.x, b.y, b.z, b.m, 1.0, 0.0)
    K = 0.5 * float(np.sum(b.m.astype(np.float64) * (b.vx.astype(np.float64) ** 2 + b.vy ** 2 + b.vz ** 2)))
    assert abs(W / (-3 * math.pi / 32) - 1) < 0.06
    assert abs(2 * K / abs(W) - 1) < 0.08


def test_direct_f32_is_the_sequential_reference_sum(orc, nbg):
    """NaiveSolver (NaiveSolver.h:22-52): acceleration reset to zero, then
    += Gravity::particleParticle over the actives in order; restated here in
    numpy float32 independently of the oracle."""
    b = nbg.plummer(40, seed=12)
    f32 = np.float32
    got = orc.direct_f32(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    for i in range(b.n):
        p = np.array([b.x[i], b.y[i], b.z[i]], f32)
        acc = np.zeros(3, f32)
        for j in range(b.n):
            a = np.array([b.x[j], b.y[j], b.z[j]], f32)
            if np.array_equal(a, p):
                r = np.zeros(3, f32)
            else:
                R = (p - a).astype(f32)
                rr = f32(np.sqrt(f32(f32(f32(R[0] * R[0]) + f32(R[1] * R[1])) + f32(R[2] * R[2])) + f32(1e-4)))
                r = ((((-R) * b.m[j]).astype(f32) * f32(1.0)).astype(f32) / f32(rr * f32(rr * rr))).astype(f32)
            acc = (acc + r).astype(f32)
        assert np.array_equal(got[i], acc), i


def test_fmm_threads_do_not_change_results(orc, nbg):
    """The oracle runs the starting nodes' traversals on OpenMP threads (as the
    reference's parallel_for_each); every body belongs to one starting node,
    so results and counts are independent of the thread count."""
    b = nbg.plummer(20000, seed=44)
    t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=0.4, leaf=16, eps2=1e-4)
    a1, c1 = t.fmm_accel(counts=True, threads=1)
    a4, c4 = t.fmm_accel(counts=True, threads=4)
    assert c1 == c4 and np.array_equal(a1, a4)


# ------------------------------------------------------------------ pins added for
# the octree cells, the COM summation order and the implicit FMM's criteria

@pytest.mark.parametrize("solver,bits", [(1, 63), (1, 30), (0, 63)])
@pytest.mark.parametrize("n,leaf", [(1, 16), (9, 1), (1000, 16), (5000, 8), (20000, 64)])
def test_com_summation_order_restated(orc, nbg, solver, bits, n, leaf):
    """Every node's COM and mass equal, bit for bit, an independent numpy
    float32 restatement of CenterOfMassSummary's leaf and child loops
    (CenterOfMassSummary.h:30-55) in the tree's own order."""
    b = nbg.plummer(n, seed=100 + n)
    if n > 1:
        b.m[::5] = np.float32(0.0)  # massless bodies inside leaves
        b.m[1::5] *= np.float32(3.0)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, leaf=leaf, bits=bits)
    nd = t.nodes()
    _, perm = t.order()
    pos = np.stack([b.x, b.y, b.z], 1)[perm]
    want = com_restatement(nd["first"], nd["count"], nd["leaf"], nd["parent"], nd["depth"], pos, b.m[perm],
                           solver == 0)
    assert same_bits(nd["com4"], want)


def _cell(c, side):
    return np.array([c[0], c[1], c[2], side], np.float32)


def test_fmm_recommend_truth_table(orc):
    """SideLengthOverDistance's node-node form (SideLengthOverDistance.h:33-56):
    approximate (0) iff 4 max(s_a^2, s_p^2) / |com_a - c_p|^2 < theta^2 and the
    exclusion regions [c - s, c + s] do not meet (closed intervals); else
    descend the passive node (2) when s_a^2 <= 4 s_p^2 and the active node (1)
    when s_p^2 <= 4 s_a^2."""
    f32 = np.float32
    z3 = np.zeros(3, f32)
    up = np.nextafter(f32(4.0), f32(5.0))
    # far apart, equal sides: approximate
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((10, 0, 0), 1.0), 0.5) == 0
    # quotient exactly theta^2 (4 * 1 / 16 = 0.25): the test is strict, so descend both
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((4, 0, 0), 1.0), 0.5) == 3
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((up, 0, 0), 1.0), 0.5) == 0
    # touching exclusion regions ([-1, 1] and [1, 3]) meet; a gap of one ulp does not
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((2, 0, 0), 1.0), 1.5) == 3
    two_up = np.nextafter(f32(2.0), f32(3.0))
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((two_up, 0, 0), 1.0), 1.5) == 0
    # the regions are [c - s, c + s], not the cells [c - s/2, c + s/2]: cells 3 apart
    # with sides 2 touch as regions ([-2, 2] and [1, 5]) though the cells do not
    assert orc.fmm_recommend(_cell(z3, 2.0), z3, _cell((3, 0, 0), 2.0), 4.0) == 3
    # side ratios (near pair, no approximation): exactly 2 descends both,
    # beyond 2 only the larger node, 4 only the larger node
    near = _cell((1, 0, 0), 1.0)
    assert orc.fmm_recommend(_cell(z3, 2.0), z3, near, 0.5) == 3
    assert orc.fmm_recommend(_cell(z3, np.nextafter(f32(2.0), f32(3.0))), z3, near, 0.5) == 1
    assert orc.fmm_recommend(_cell(z3, 4.0), z3, near, 0.5) == 1
    assert orc.fmm_recommend(_cell(z3, 0.5), z3, near, 0.5) == 3
    assert orc.fmm_recommend(_cell(z3, 0.25), z3, near, 0.5) == 2
    # the distance runs from the active node's centre of mass, not its centre
    com = np.array([-8.0, 0, 0], np.float32)
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((4, 0, 0), 1.0), 0.5) == 3
    assert orc.fmm_recommend(_cell(z3, 1.0), com, _cell((4, 0, 0), 1.0), 0.5) == 0
    # and to the passive node's centre (the other node's COM is not an input)
    assert orc.fmm_recommend(_cell(z3, 1.0), z3, _cell((0, 0, 12), 1.0), 0.5) == 0


def test_fmm_particle_far_truth_table(orc):
    """SideLengthOverDistance's particle-node form (SideLengthOverDistance.h:25-29):
    2 s / |x - c| < theta and x outside the half-open cell [c - s/2, c + s/2)
    (OctreeNode::boundingBox, Octree.h:92-98; BoundingBox::contains :61-64)."""
    f32 = np.float32
    cell = _cell((0, 0, 0), 1.0)
    assert orc.fmm_particle_far((10, 0, 0), cell, 0.5)
    # quotient exactly theta (2 / 4 = 0.5) is not below it
    assert not orc.fmm_particle_far((4, 0, 0), cell, 0.5)
    assert orc.fmm_particle_far((np.nextafter(f32(4.0), f32(5.0)), 0, 0), cell, 0.5)
    # with a large theta only the cell decides: closed below, open above
    assert not orc.fmm_particle_far((-0.5, 0, 0), cell, 100.0)
    assert orc.fmm_particle_far((0.5, 0, 0), cell, 100.0)
    assert not orc.fmm_particle_far((np.nextafter(f32(0.5), f32(0.0)), 0, 0), cell, 100.0)
    assert orc.fmm_particle_far((0, 0, np.nextafter(f32(-0.5), f32(-1.0))), cell, 100.0)
    # the cell is [c - s/2, c + s/2): a point at 0.75 of the side is outside
    assert orc.fmm_particle_far((0.75, 0, 0), cell, 100.0)
