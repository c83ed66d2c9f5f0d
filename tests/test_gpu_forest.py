"""Batched forest: independent systems built and walked together.

Every system's result must be the bits it gets alone (an nbg_sys with the
same parameters, same sum mode) — in both sum modes, because each
wavefront serves one system and each system's tree, summaries and
traversal order are its own — and, in ORDERED mode, the oracle's bits.
This holds for all four tree solvers a batch runs: the linear BVH with
30-bit keys (solver 1) and the Barnes-Hut octree (solver 0, BarnesHutSolver,
Solvers/BarnesHutSolver.h:14-32), each also in its quadrupole form (4, 3),
whose node second moments are summarized with the forest."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(nbg, theta=0.5, leaf=16, eps2=1e-4, G=1.0, dt=1e-3, mode=1, solver=1):
    p = nbg.default_params()
    p.solver, p.morton_bits = solver, 30
    p.theta, p.leaf_size, p.eps2, p.G, p.dt, p.sum_mode = theta, leaf, eps2, G, dt, mode
    return p


def _systems(nbg):
    """Uneven sizes (wave-boundary cases included) and per-system parameters."""
    spec = [(1, 0.5, 16, 1e-4, 1.0, 1e-3), (63, 0.5, 8, 1e-4, 1.0, 1e-3), (64, 0.3, 16, 1e-4, 2.0, 5e-4),
            (65, 0.7, 4, 1e-3, 1.0, 1e-3), (1000, 0.5, 16, 1e-4, 1.0, 1e-3), (5003, 0.5, 32, 2.5e-5, 0.5, 2e-3),
            (20000, 0.4, 16, 1e-4, 1.0, 1e-3), (3, 0.5, 1, 0.0, 1.0, 1e-3)]
    out = []
    for k, (n, th, lf, e2, G, dt) in enumerate(spec):
        out.append((nbg.plummer(n, seed=500 + k, a=1.0 + 0.1 * k), dict(theta=th, leaf=lf, eps2=e2, G=G, dt=dt)))
    return out


def _alone(nbg, ctx, b, kw, mode, steps=0, solver=1):
    s = nbg.System(ctx, _params(nbg, mode=mode, solver=solver, **kw), b.n)
    s.set_state(b)
    if steps:
        s.step(steps)
        s.synchronize()
        st = s.get_state()
        a = s.accelerations()
    else:
        a = s.update_accelerations()
        st = None
    c = s.counts()
    s.close()
    return a, st, c


@pytest.mark.parametrize("solver", [1, 4, 0, 3])
@pytest.mark.parametrize("mode", [0, 1])
def test_forest_pass_equals_alone_and_oracle(nbg, orc, gpu_ctx, mode, solver):
    systems = _systems(nbg)
    B = nbg.Batch(gpu_ctx, 40000, 16, sum_mode=mode)
    ids = [B.add(_params(nbg, mode=mode, solver=solver, **kw), b) for b, kw in systems]
    B.update_accelerations()
    tot = [0, 0]
    for sid, (b, kw) in zip(ids, systems):
        print(f"[forest pass] solver={solver} mode={mode} n={b.n} {kw}", flush=True)
        _, acc = B.get(sid, accelerations=True)
        a1, _, c = _alone(nbg, gpu_ctx, b, kw, mode, solver=solver)
        tot[0] += c[0]
        tot[1] += c[1]
        assert np.array_equal(acc.view(np.uint32), a1.view(np.uint32)), f"n={b.n}: batch differs from alone"
        if mode == 0:
            t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=kw["theta"], leaf=kw["leaf"], bits=30, G=kw["G"],
                         eps2=kw["eps2"])
            assert np.array_equal(acc, t.accel(b.x, b.y, b.z)), f"n={b.n}: batch differs from the oracle"
    assert B.counts() == tuple(tot)
    B.close()


@pytest.mark.parametrize("mode,solver", [(0, 1), (1, 1), (0, 4), (1, 4), (0, 0), (1, 0), (1, 3)])
def test_forest_membership_changes(nbg, gpu_ctx, mode, solver):
    """Systems join and leave between steps; the others' trajectories are
    unaffected (same bits as alone with the same number of steps)."""
    systems = _systems(nbg)
    B = nbg.Batch(gpu_ctx, 60000, 16, sum_mode=mode)
    ids = [B.add(_params(nbg, mode=mode, solver=solver, **kw), b) for b, kw in systems[:6]]
    steps = {sid: 0 for sid in ids}
    data = {sid: systems[k] for k, sid in enumerate(ids)}
    B.step(3)
    for sid in ids:
        steps[sid] += 3
    B.remove(ids[2])           # a middle system leaves
    del steps[ids[2]]
    B.remove(ids[0])           # and the first
    del steps[ids[0]]
    late = B.add(_params(nbg, mode=mode, solver=solver, **systems[6][1]), systems[6][0])  # a new one joins
    steps[late] = 0
    data[late] = systems[6]
    B.step(2)
    for sid in steps:
        steps[sid] += 2
    B.synchronize()
    assert B.info() == (5, sum(data[s][0].n for s in steps))
    for sid, k in steps.items():
        b, kw = data[sid]
        st = B.get(sid)
        _, ref, _ = _alone(nbg, gpu_ctx, b, kw, mode, steps=k, solver=solver)
        for g, r in zip((st.x, st.y, st.z, st.vx, st.vy, st.vz), (ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz)):
            assert np.array_equal(g, r), f"system {sid} (n={b.n}) after {k} steps"
    B.close()


@pytest.mark.parametrize("solver", [1, 4, 0, 3])
def test_forest_graph_replay_equals_eager(nbg, gpu_ctx, solver):
    systems = _systems(nbg)[:6]
    runs = []
    for graph in (False, True):
        B = nbg.Batch(gpu_ctx, 40000, 8)
        ids = [B.add(_params(nbg, solver=solver, **kw), b) for b, kw in systems]
        B.use_graph(graph)
        snaps = []
        for _ in range(4):     # replays with host synchronisation in between
            B.step(1)
            B.synchronize()
            snaps.append([B.get(sid).x.copy() for sid in ids])
        runs.append(snaps)
        B.close()
    for r_eager, r_graph in zip(*runs):
        for a, b in zip(r_eager, r_graph):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("solver", [1, 0])
def test_forest_autotune(nbg, gpu_ctx, solver):
    systems = _systems(nbg)
    B = nbg.Batch(gpu_ctx, 40000, 16)
    ids = [B.add(_params(nbg, solver=solver, **kw), b) for b, kw in systems]
    before = [B.get(sid).x.copy() for sid in ids]
    best, ms = B.autotune((8, 16, 32, 64), repeats=2)
    print(f"[forest autotune] best leaf {best}, ms {ms}", flush=True)
    assert best in (8, 16, 32, 64) and all(v > 0 for v in ms.values())
    for sid, x0 in zip(ids, before):
        assert np.array_equal(B.get(sid).x, x0), "autotuning must not move bodies"
    B.update_accelerations()
    b, kw = systems[4]
    _, acc = B.get(ids[4], accelerations=True)
    kw = dict(kw, leaf=best)
    a1, _, _ = _alone(nbg, gpu_ctx, b, kw, 1, solver=solver)
    assert np.array_equal(acc, a1)
    B.close()


@pytest.mark.parametrize("solver", [1, 4])
def test_forest_targets_per_lane_same_bits(nbg, gpu_ctx, solver):
    """The FAST linear-BVH forest walk with 2 or 4 targets per lane (128 or
    256 bodies per wavefront, waves aligned to systems at that width): each
    system keeps the bits of one target per lane, and of itself alone."""
    systems = _systems(nbg)
    runs = []
    for tpl in (1, 2, 4):
        B = nbg.Batch(gpu_ctx, 40000, 16, sum_mode=1)
        ids = [B.add(_params(nbg, mode=1, solver=solver, **kw), b) for b, kw in systems]
        B.set_targets_per_lane(tpl)
        B.update_accelerations()
        runs.append(([B.get(sid, accelerations=True)[1] for sid in ids], B.counts()))
        B.close()
    for tpl, (accs, c) in zip((2, 4), runs[1:]):
        assert c == runs[0][1]
        for (b, _), a, a1 in zip(systems, accs, runs[0][0]):
            assert np.array_equal(a.view(np.uint32), a1.view(np.uint32)), f"{tpl} targets per lane, n={b.n}"
    b, kw = systems[6]
    a1, _, _ = _alone(nbg, gpu_ctx, b, kw, 1, solver=solver)
    assert np.array_equal(runs[2][0][6], a1)


def test_forest_autotune_grid(nbg, gpu_ctx):
    """Leaf size x bodies per wavefront: every pair is timed, the fastest is
    kept, no body moves, and the batch then gives each system the bits it
    gets alone with the chosen leaf size; an octree batch times one target
    per lane only."""
    systems = _systems(nbg)
    B = nbg.Batch(gpu_ctx, 40000, 16)
    ids = [B.add(_params(nbg, **kw), b) for b, kw in systems]
    before = [B.get(sid).x.copy() for sid in ids]
    leaf, tpl, table = B.autotune_grid((8, 16), (1, 2, 4), repeats=1)
    print(f"[forest autotune grid] leaf {leaf}, targets per lane {tpl}: {table}", flush=True)
    assert set(table) == {(l, t) for l in (8, 16) for t in (1, 2, 4)} and all(v > 0 for v in table.values())
    assert (leaf, tpl) == min(table, key=table.get)
    for sid, x0 in zip(ids, before):
        assert np.array_equal(B.get(sid).x, x0)
    B.update_accelerations()
    b, kw = systems[6]
    a1, _, _ = _alone(nbg, gpu_ctx, b, dict(kw, leaf=leaf), 1)
    assert np.array_equal(B.get(ids[6], accelerations=True)[1], a1)
    B.close()
    B = nbg.Batch(gpu_ctx, 40000, 16)
    for b, kw in systems:
        B.add(_params(nbg, solver=0, **kw), b)
    leaf, tpl, table = B.autotune_grid((8, 16), (1, 2, 4), repeats=1)
    assert set(table) == {(8, 1), (16, 1)} and tpl == 1
    B.close()


def test_forest_capacity_and_arguments(nbg, gpu_ctx):
    B = nbg.Batch(gpu_ctx, 100, 2)
    b = nbg.plummer(60, seed=1)
    B.add(_params(nbg), b)
    with pytest.raises(nbg.NbgError):
        B.add(_params(nbg), b)            # 120 > 100 bodies
    p = _params(nbg)
    p.solver = nbg.BARNES_HUT
    with pytest.raises(nbg.NbgError):
        B.add(p, nbg.plummer(10, seed=2))  # members share one solver (octree and linear BVH)
    p.solver = nbg.IMPLICIT_FMM
    with pytest.raises(nbg.NbgError):
        B.add(p, nbg.plummer(10, seed=2))  # the implicit FMM is not batched
    with pytest.raises(nbg.NbgError):
        B.add(_params(nbg, solver=4), nbg.plummer(10, seed=3))  # members share one solver
    B.add(_params(nbg), nbg.plummer(10, seed=3))
    with pytest.raises(nbg.NbgError):
        B.add(_params(nbg), nbg.plummer(10, seed=4))  # 3 > 2 systems
    B.close()


def test_forest_solver_kind_changes(nbg, orc, gpu_ctx):
    """A batch's solver follows its members: a sole member switched between
    the linear-BVH and octree, monopole and quadrupole kinds, and a batch
    emptied and refilled with another kind, give each kind's bits alone and
    the oracle's."""
    b1, b2 = nbg.plummer(3000, seed=71), nbg.plummer(2000, seed=72, a=1.5)
    kw = dict(theta=0.5, leaf=16, eps2=1e-4, G=1.0, dt=1e-3)
    B = nbg.Batch(gpu_ctx, 8000, 4, sum_mode=0)

    def check(sid, b, solver):
        B.update_accelerations()
        _, acc = B.get(sid, accelerations=True)
        a1, _, _ = _alone(nbg, gpu_ctx, b, kw, 0, solver=solver)
        assert np.array_equal(acc.view(np.uint32), a1.view(np.uint32)), f"solver {solver}"
        t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=30, eps2=1e-4)
        assert np.array_equal(acc, t.accel(b.x, b.y, b.z)), f"solver {solver}: oracle"

    sid = B.add(_params(nbg, mode=0, **kw), b1)
    check(sid, b1, 1)
    for solver in (4, 0, 3, 1, 4):
        B.set_params(sid, _params(nbg, mode=0, solver=solver, **kw))
        check(sid, b1, solver)
    other = B.add(_params(nbg, mode=0, solver=4, **kw), b2)
    with pytest.raises(nbg.NbgError):
        B.set_params(sid, _params(nbg, mode=0, solver=1, **kw))  # would mix the kinds
    check(other, b2, 4)
    B.remove(sid)
    B.remove(other)
    sid = B.add(_params(nbg, mode=0, solver=1, **kw), b2)
    check(sid, b2, 1)
    B.close()


@pytest.mark.parametrize("mode,solver", [(0, 1), (1, 1), (0, 4), (0, 0), (1, 0)])
def test_forest_large_equals_alone(nbg, gpu_ctx, mode, solver):
    """A forest above 2M bodies (the summaries then run their middle,
    many-workgroup pass, whose ranges cut through systems): every system
    still gets its bits alone."""
    sizes = [300000, 70001, 450000, 129, 500000, 260000, 333333, 70000, 65, 300000]
    systems = [(nbg.plummer(n, seed=900 + k, a=0.7 + 0.2 * k), dict(theta=0.5, leaf=16, eps2=1e-4, G=1.0, dt=1e-3))
               for k, n in enumerate(sizes)]
    total = sum(b.n for b, _ in systems)
    assert total > 8 * 4096 * 64
    B = nbg.Batch(gpu_ctx, total, len(systems), sum_mode=mode)
    ids = [B.add(_params(nbg, mode=mode, solver=solver, **kw), b) for b, kw in systems]
    B.update_accelerations()
    for sid, (b, kw) in zip(ids, systems):
        print(f"[forest large] solver={solver} mode={mode} n={b.n}", flush=True)
        _, acc = B.get(sid, accelerations=True)
        a1, _, _ = _alone(nbg, gpu_ctx, b, kw, mode, solver=solver)
        assert np.array_equal(acc.view(np.uint32), a1.view(np.uint32))
    B.close()
