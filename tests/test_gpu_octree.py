"""GPU parity of the Barnes-Hut octree force pass against the oracle.

ORDERED mode must be bit-exact with the oracle: bounding box, the set of
non-empty cells (level, centre and side from the float centre recurrence),
their body counts, leaf flags, leaf contents in order, centres of mass and
masses, accelerations and interaction counts.  FAST mode must take the same
opening decisions (equal interaction counts) and agree within

    |a_gpu - a_ref| <= 1e-5 * max(|a_ref|, a_rms)        (per body).

The oracle keeps empty children as nodes (the reference does); the device
tree stores non-empty cells only, so the comparison is over non-empty cells.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _params(nbg, leaf, max_depth=32, theta=0.5, eps2=1e-4, mode=0, dt=1e-3):
    p = nbg.default_params()
    p.solver = nbg.BARNES_HUT
    p.leaf_size = leaf
    p.max_depth = max_depth
    p.theta = theta
    p.eps2 = eps2
    p.sum_mode = mode
    p.dt = dt
    return p


def _gpu_pass(nbg, ctx, b, **kw):
    s = nbg.System(ctx, _params(nbg, **kw), b.n)
    s.set_state(b)
    acc = s.update_accelerations()
    return s, acc


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _cells(depth, first, count, leaf, com4, geo6, perm):
    """Non-empty cells keyed by (level, centre bits, side bits)."""
    out = {}
    g = _bits(geo6[:, :4])
    for i in np.nonzero(count > 0)[0].tolist():
        key = (int(depth[i]),) + tuple(int(v) for v in g[i])
        assert key not in out, f"duplicate cell {key}"
        f, c = int(first[i]), int(count[i])
        out[key] = (c, bool(leaf[i]), com4[i].copy(), perm[f:f + c].copy())
    return out


def _gpu_depth(parent):
    d = np.zeros(len(parent), np.int64)
    for i in range(1, len(parent)):
        p = int(parent[i])
        assert 0 <= p < i, "parent slots precede their children"
        d[i] = d[p] + 1
    return d


def _same_f32(a, b):
    """Bitwise equality, every NaN equal to every NaN."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return bool(np.all(both_nan | (_bits(a) == _bits(b))))


def _check_tree(s, t, n):
    gt = s.debug_tree()
    _, gperm = s.debug_order()
    on = t.nodes()
    _, operm = t.order()
    gd = _gpu_depth(gt["parent"])
    G = _cells(gd, gt["first"], gt["count"], gt["leaf"], gt["com4"], gt["geo6"], gperm)
    O = _cells(on["depth"], on["first"], on["count"], on["leaf"], on["com4"], on["geo6"], operm)
    assert G.keys() == O.keys(), f"cell sets differ: {len(G)} device vs {len(O)} oracle"
    for k, (c, lf, com, body) in O.items():
        gc, glf, gcom, gbody = G[k]
        assert gc == c and glf == lf, f"count/leaf of cell {k}"
        assert _same_f32(gcom, com), f"COM/mass of cell {k}: {gcom} vs {com}"
        if lf:
            assert np.array_equal(gbody, body), f"leaf contents of cell {k}"
        else:
            assert np.array_equal(np.sort(gbody), np.sort(body)), f"bodies of cell {k}"
    assert sorted(gperm.tolist()) == list(range(n))
    return len(G)


CASES = [(1, 16, 32), (2, 1, 32), (3, 1, 32), (9, 1, 32), (64, 8, 32), (1000, 16, 32), (1000, 1, 32),
         (5000, 8, 3), (10000, 16, 32), (10000, 64, 32), (20000, 16, 19), (65537, 16, 32)]


@pytest.mark.parametrize("n,leaf,max_depth", CASES)
def test_octree_ordered_bitexact(nbg, orc, gpu_ctx, n, leaf, max_depth):
    print(f"[octree ordered] n={n} leaf={leaf} max_depth={max_depth}", flush=True)
    b = nbg.plummer(n, seed=2000 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=leaf, max_depth=max_depth)
    t = orc.Tree(0, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf, max_depth=max_depth, eps2=1e-4)
    assert np.array_equal(s.debug_bbox(), t.bbox())
    _check_tree(s, t, n)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert np.array_equal(_bits(acc), _bits(ref)), "accelerations must be bit-exact in ORDERED mode"
    assert s.counts() == (int(pp.sum()), int(npc.sum()))


@pytest.mark.parametrize("n,leaf,theta", [(1000, 16, 0.5), (10000, 16, 0.5), (10000, 16, 0.3), (50000, 8, 0.7),
                                          (50000, 64, 0.5)])
def test_octree_fast_within_tolerance(nbg, orc, gpu_ctx, n, leaf, theta):
    print(f"[octree fast] n={n} leaf={leaf} theta={theta}", flush=True)
    b = nbg.plummer(n, seed=99 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=leaf, theta=theta, mode=1)
    t = orc.Tree(0, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, eps2=1e-4)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert s.counts() == (int(pp.sum()), int(npc.sum())), "same opening decisions"
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2))
    err = np.linalg.norm(acc - ref, axis=1)
    worst = float(np.max(err / np.maximum(mag, floor)))
    print(f"  fast worst {worst:.3g}", flush=True)
    assert worst <= FAST_TOL


@pytest.mark.parametrize("mode", [0, 1])
def test_octree_steps_match_oracle(nbg, orc, gpu_ctx, mode):
    """Solver::step x5 on the device vs the oracle's step (bit-exact ORDERED)."""
    n = 3000
    b = nbg.plummer(n, seed=41)
    s = nbg.System(gpu_ctx, _params(nbg, 16, mode=mode, dt=1e-3), n)
    s.set_state(b)
    s.step(5)
    s.synchronize()
    st = s.get_state()
    (x, y, z, vx, vy, vz), _ = orc.step(0, (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz), 5, theta=0.5, leaf=16,
                                        eps2=1e-4, dt=1e-3)
    if mode == 0:
        for g, r in zip((st.x, st.y, st.z, st.vx, st.vy, st.vz), (x, y, z, vx, vy, vz)):
            assert np.array_equal(_bits(g), _bits(r))
    else:
        assert np.max(np.abs(st.x - x)) < 1e-6 and np.max(np.abs(st.vx - vx)) < 1e-4


def _degenerate_sets(nbg):
    rng = np.random.default_rng(5)
    out = []
    # every body at one point: zero-extent root, all paths equal -> depth-limit leaf
    p = np.full((100, 3), 0.25, np.float32)
    out.append(("coincident", p, np.full(100, 0.01, np.float32), 4))
    # clusters of exact duplicates larger than the leaf size
    base = rng.standard_normal((40, 3)).astype(np.float32)
    p = np.repeat(base, 7, axis=0)
    out.append(("duplicates", p, np.full(len(p), 1.0 / len(p), np.float32), 4))
    # an integer lattice: many bodies exactly on cell planes (p[d] == c[d])
    g = np.arange(-4, 5, dtype=np.float32)
    p = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    out.append(("lattice", p, np.full(len(p), 1.0, np.float32), 8))
    # zero masses inside a Plummer set: 0/0 COMs that the MAC never accepts
    b = nbg.plummer(2000, seed=8)
    m = b.m.copy()
    m[::3] = 0.0
    out.append(("zero-mass", np.stack([b.x, b.y, b.z], 1), m, 16))
    # one octant of the root empty: NaN root COM
    p = np.abs(rng.standard_normal((3000, 3)).astype(np.float32))
    p[:1000, 0] *= -1.0
    out.append(("empty-octant", p, np.full(len(p), 1e-3, np.float32), 16))
    return out


@pytest.mark.parametrize("mode", [0, 1])
def test_octree_degenerate_inputs(nbg, orc, gpu_ctx, mode):
    for name, p, m, leaf in _degenerate_sets(nbg):
        print(f"[octree degenerate] {name} n={len(p)} leaf={leaf} mode={mode}", flush=True)
        n = len(p)
        z = np.zeros(n, np.float32)
        b = nbg.Bodies(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m, z, z.copy(), z.copy())
        s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=leaf, mode=mode)
        t = orc.Tree(0, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf, eps2=1e-4)
        _check_tree(s, t, n)
        ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
        assert s.counts() == (int(pp.sum()), int(npc.sum())), name
        if mode == 0:
            assert np.array_equal(_bits(acc), _bits(ref)), name
        else:
            mag = np.linalg.norm(ref, axis=1)
            floor = max(np.sqrt(np.mean(mag ** 2)), 1e-30)
            assert np.max(np.linalg.norm(acc - ref, axis=1) / np.maximum(mag, floor)) <= FAST_TOL, name
        s.close()


def test_octree_empty_system(nbg, gpu_ctx):
    s = nbg.System(gpu_ctx, _params(nbg, 16), 0)
    s.set_state(nbg.Bodies.empty(0))
    assert s.update_accelerations().shape == (0, 3)
    s.step(2)
    s.synchronize()


def test_octree_solver_switch(nbg, orc, gpu_ctx):
    """set_params switches a live system between the two solvers."""
    b = nbg.plummer(5000, seed=12)
    s = nbg.System(gpu_ctx, _params(nbg, 16), b.n)
    s.set_state(b)
    a_oct = s.update_accelerations()
    p = _params(nbg, 16)
    p.solver = nbg.LINEAR_BVH
    p.morton_bits = 63
    s.params = p
    a_bvh = s.update_accelerations()
    t0 = orc.Tree(0, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=1e-4)
    t1 = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=63, eps2=1e-4)
    assert np.array_equal(a_oct, t0.accel(b.x, b.y, b.z))
    assert np.array_equal(a_bvh, t1.accel(b.x, b.y, b.z))


def test_octree_large_sample(nbg, orc, gpu_ctx):
    """1M-body Plummer sphere: device FAST and ORDERED passes against the
    oracle's walk for a sample of targets in the same tree; the device's
    interaction counts over all bodies equal the oracle's (its passive-point
    walk at every body's position meets the same nodes and bodies)."""
    n = 1 << 20
    print(f"[octree large] n={n}", flush=True)
    b = nbg.plummer(n, seed=1)
    t = orc.Tree(0, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=2.5e-5)
    sel = np.random.default_rng(3).choice(n, 8192, replace=False)
    ref = t.accel(b.x[sel], b.y[sel], b.z[sel])
    _, opp, onp = t.accel(b.x, b.y, b.z, counts=True)
    ocounts = (int(opp.sum()), int(onp.sum()))
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2))
    for mode in (1, 0):
        s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=16, eps2=2.5e-5, mode=mode)
        if mode == 0:
            assert np.array_equal(_bits(acc[sel]), _bits(ref))
        else:
            err = np.linalg.norm(acc[sel] - ref, axis=1)
            assert np.max(err / np.maximum(mag, floor)) <= FAST_TOL
        pp, npc = s.counts()
        print(f"  mode {mode}: pp={pp} np={npc}", flush=True)
        assert (pp, npc) == ocounts
        s.close()
