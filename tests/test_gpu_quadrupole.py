"""GPU parity of the quadrupole tree solvers (QuadrupoleLinearBVHSolver,
QuadrupoleBarnesHutSolver) against the oracle.

ORDERED mode must be bit-exact with the oracle: every node's second moment
(MultipoleMassSummary<2>), accelerations through Gravity::nodeParticle<2>,
interaction counts and trajectories.  FAST mode takes the same opening
decisions (equal counts) and agrees within

    |a_gpu - a_ref| <= 1e-5 * max(|a_ref|, a_rms)        (per body).

The device octree stores non-empty cells only (the oracle keeps empty
children), so octree moments are compared over non-empty cells.
"""

import numpy as np
import pytest

from test_gpu_octree import _bits, _cells, _gpu_depth, _same_f32

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _params(nbg, solver, leaf, bits=63, max_depth=32, theta=0.5, eps2=1e-4, mode=0, dt=1e-3):
    p = nbg.default_params()
    p.solver = solver
    p.leaf_size = leaf
    p.morton_bits = bits
    p.max_depth = max_depth
    p.theta = theta
    p.eps2 = eps2
    p.sum_mode = mode
    p.dt = dt
    return p


def _gpu_pass(nbg, ctx, b, solver, **kw):
    s = nbg.System(ctx, _params(nbg, solver, **kw), b.n)
    s.set_state(b)
    acc = s.update_accelerations()
    return s, acc


def _lbvh_moments(first, count, q):
    order = np.lexsort((count, first))
    return first[order], count[order], q[order]


def _check_lbvh_moments(s, t):
    gt = s.debug_tree()
    on = t.nodes()
    g = _lbvh_moments(gt["first"], gt["count"], s.debug_quads())
    o = _lbvh_moments(on["first"], on["count"], t.quads())
    assert np.array_equal(g[0], o[0]) and np.array_equal(g[1], o[1]), "node ranges"
    assert _same_f32(g[2], o[2]), "node second moments must be bit-exact"


def _check_octree_moments(s, t):
    gt = s.debug_tree()
    _, gperm = s.debug_order()
    on = t.nodes()
    _, operm = t.order()
    gq, oq = s.debug_quads(), t.quads()
    # key each cell's moment by the cell (level, centre, side); the COM slot
    # of _cells carries the moment here
    G = _cells(_gpu_depth(gt["parent"]), gt["first"], gt["count"], gt["leaf"], gq, gt["geo6"], gperm)
    O = _cells(on["depth"], on["first"], on["count"], on["leaf"], oq, on["geo6"], operm)
    assert G.keys() == O.keys()
    for k, v in O.items():
        assert _same_f32(G[k][2], v[2]), f"second moment of cell {k}: {G[k][2]} vs {v[2]}"
    return len(G)


LBVH_CASES = [(1, 16, 63), (2, 1, 30), (3, 1, 63), (17, 4, 30), (64, 8, 63), (1000, 16, 30), (1000, 1, 63),
              (4099, 16, 63), (10000, 64, 30), (30011, 16, 63)]


@pytest.mark.parametrize("n,leaf,bits", LBVH_CASES)
def test_quad_lbvh_ordered_bitexact(nbg, orc, gpu_ctx, n, leaf, bits):
    print(f"[quad lbvh ordered] n={n} leaf={leaf} bits={bits}", flush=True)
    b = nbg.plummer(n, seed=4000 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, nbg.LINEAR_BVH_QUADRUPOLE, leaf=leaf, bits=bits)
    t = orc.Tree(4, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf, bits=bits, eps2=1e-4)
    _check_lbvh_moments(s, t)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert np.array_equal(_bits(acc), _bits(ref)), "accelerations must be bit-exact in ORDERED mode"
    assert s.counts() == (int(pp.sum()), int(npc.sum()))


OCT_CASES = [(1, 16, 32), (2, 1, 32), (9, 1, 32), (64, 8, 32), (1000, 16, 32), (1000, 1, 32), (5000, 8, 3),
             (10000, 16, 32), (20000, 16, 19), (30011, 64, 32)]


@pytest.mark.parametrize("n,leaf,max_depth", OCT_CASES)
def test_quad_octree_ordered_bitexact(nbg, orc, gpu_ctx, n, leaf, max_depth):
    print(f"[quad octree ordered] n={n} leaf={leaf} max_depth={max_depth}", flush=True)
    b = nbg.plummer(n, seed=5000 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, nbg.BARNES_HUT_QUADRUPOLE, leaf=leaf, max_depth=max_depth)
    t = orc.Tree(3, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf, max_depth=max_depth, eps2=1e-4)
    _check_octree_moments(s, t)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert np.array_equal(_bits(acc), _bits(ref)), "accelerations must be bit-exact in ORDERED mode"
    assert s.counts() == (int(pp.sum()), int(npc.sum()))


@pytest.mark.parametrize("solver,n,leaf,theta", [(4, 1000, 16, 0.5), (4, 20000, 16, 0.7), (4, 50000, 8, 0.5),
                                                 (3, 1000, 16, 0.5), (3, 20000, 16, 0.4), (3, 50000, 64, 0.6)])
def test_quad_fast_within_tolerance(nbg, orc, gpu_ctx, solver, n, leaf, theta):
    print(f"[quad fast] solver={solver} n={n} leaf={leaf} theta={theta}", flush=True)
    b = nbg.plummer(n, seed=123 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=leaf, theta=theta, mode=1)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, eps2=1e-4)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert s.counts() == (int(pp.sum()), int(npc.sum())), "same opening decisions"
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2))
    err = np.linalg.norm(acc - ref, axis=1)
    worst = float(np.max(err / np.maximum(mag, floor)))
    print(f"  fast-vs-ordered worst {worst:.3g}, median rel {np.median(err / mag):.3g}", flush=True)
    assert worst <= FAST_TOL


@pytest.mark.parametrize("solver", [3, 4])
def test_quad_fast_zero_softening(nbg, orc, gpu_ctx, solver):
    """eps2 = 0 takes the guarded rsqrt path (r2 = 0 contributes nothing)."""
    b = nbg.plummer(5000, seed=17)
    b.x[1] = b.x[0]; b.y[1] = b.y[0]; b.z[1] = b.z[0]   # a coincident pair
    s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=16, eps2=0.0, mode=1)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=0.0)
    ref = t.accel(b.x, b.y, b.z)
    assert np.all(np.isfinite(acc))
    mag = np.linalg.norm(ref, axis=1)
    err = np.linalg.norm(acc - ref, axis=1)
    assert np.max(err / np.maximum(mag, np.sqrt(np.mean(mag ** 2)))) <= FAST_TOL


@pytest.mark.parametrize("solver,mode", [(3, 0), (4, 0), (3, 1), (4, 1)])
def test_quad_steps_match_oracle(nbg, orc, gpu_ctx, solver, mode):
    """Solver::step x4 with the quadrupole solvers (bit-exact ORDERED)."""
    n = 3000
    b = nbg.plummer(n, seed=41)
    s = nbg.System(gpu_ctx, _params(nbg, solver, 16, mode=mode), n)
    s.set_state(b)
    s.step(4)
    s.synchronize()
    st = s.get_state()
    (x, y, z, vx, vy, vz), _ = orc.step(solver, (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz), 4, theta=0.5, leaf=16,
                                        bits=63, eps2=1e-4, dt=1e-3)
    if mode == 0:
        for g, r in zip((st.x, st.y, st.z, st.vx, st.vy, st.vz), (x, y, z, vx, vy, vz)):
            assert np.array_equal(g, r)
    else:
        assert np.max(np.abs(st.x - x)) < 1e-6 and np.max(np.abs(st.vx - vx)) < 1e-4


@pytest.mark.parametrize("solver", [3, 4])
def test_quad_passive_targets(nbg, orc, gpu_ctx, solver):
    """Accelerations at points that are not bodies (the tree of the bodies,
    walked for other positions): bit-exact ORDERED, tolerance FAST."""
    b = nbg.plummer(8000, seed=3)
    rng = np.random.default_rng(4)
    tx, ty, tz = (rng.normal(0, 1.5, 3000).astype(np.float32) for _ in range(3))
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=1e-4)
    ref = t.accel(tx, ty, tz)
    for mode in (0, 1):
        s = nbg.System(gpu_ctx, _params(nbg, solver, 16, mode=mode), b.n)
        s.set_state(b)
        got = s.accelerations_at(tx, ty, tz)
        if mode == 0:
            assert np.array_equal(_bits(got), _bits(ref))
        else:
            mag = np.linalg.norm(ref, axis=1)
            err = np.linalg.norm(got - ref, axis=1)
            assert np.max(err / np.maximum(mag, np.sqrt(np.mean(mag ** 2)))) <= FAST_TOL


@pytest.mark.parametrize("mono,quad", [(0, 3), (1, 4)])
def test_quad_more_accurate_than_monopole(nbg, orc, gpu_ctx, mono, quad):
    """Against the FP64 direct sum the quadrupole solvers' median error is
    well below the monopole solvers' at the same theta (FAST mode)."""
    n = 100000
    b = nbg.plummer(n, seed=12)
    sel = np.arange(0, n, 50)
    ref = gpu_ctx.direct_f64(b, b.x[sel], b.y[sel], b.z[sel], 1.0, 1e-4)
    med = {}
    for solver in (mono, quad):
        s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=16, theta=0.6, mode=1)
        rel = np.linalg.norm(acc[sel] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        med[solver] = float(np.median(rel))
    print(f"  median error monopole {med[mono]:.3g}, quadrupole {med[quad]:.3g}", flush=True)
    assert med[quad] < 0.5 * med[mono]


@pytest.mark.parametrize("solver", [3, 4])
def test_quad_degenerate_sets(nbg, orc, gpu_ctx, solver):
    # zero bodies
    s = nbg.System(gpu_ctx, _params(nbg, solver, 16), 0)
    s.set_state(nbg.Bodies.empty(0))
    assert s.update_accelerations().shape == (0, 3)
    # every body at one point: zero moment, zero force
    n = 300
    b = nbg.Bodies.empty(n)
    b.x[:] = 1.5
    b.m[:] = 1.0
    for mode in (0, 1):
        s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=8, mode=mode)
        assert np.all(acc == 0)
    # two clusters of exact duplicates (ORDERED, bit-exact)
    b = nbg.Bodies.empty(200)
    b.x[100:] = 1.0
    b.m[:] = 0.5
    s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=4)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=4, eps2=1e-4)
    assert np.array_equal(_bits(acc), _bits(t.accel(b.x, b.y, b.z)))


def test_quad_solver_switch(nbg, orc, gpu_ctx):
    """One system switched between the monopole and quadrupole kinds keeps
    giving each kind's oracle result (buffers follow the solver)."""
    b = nbg.plummer(6000, seed=61)
    s = nbg.System(gpu_ctx, _params(nbg, 1, 16), b.n)
    s.set_state(b)
    for solver in (1, 4, 0, 3, 4, 1):
        s.params = _params(nbg, solver, 16)
        acc = s.update_accelerations()
        t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, eps2=1e-4)
        assert np.array_equal(_bits(acc), _bits(t.accel(b.x, b.y, b.z))), f"solver {solver}"


@pytest.mark.parametrize("solver", [1, 0, 4, 3])
def test_large_system_summaries_bitexact(nbg, orc, gpu_ctx, solver):
    """Above ~1.8M bodies the summaries add a many-workgroup middle pass;
    ORDERED accelerations of a sample stay the oracle's bits (every node
    summary, moment and MAC record on the sample's paths included)."""
    n = 2200000
    b = nbg.plummer(n, seed=77)
    print(f"[large summaries] solver={solver} n={n}", flush=True)
    theta = {1: 0.5, 0: 0.5, 4: 0.72578126, 3: 0.41289064}[solver]
    s, acc = _gpu_pass(nbg, gpu_ctx, b, solver, leaf=16, theta=theta)
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=theta, leaf=16, eps2=1e-4)
    sel = np.random.default_rng(5).choice(n, 3000, replace=False)
    ref = t.accel(b.x[sel], b.y[sel], b.z[sel])
    assert np.array_equal(_bits(acc[sel]), _bits(ref))
    s.close()
