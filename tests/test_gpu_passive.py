"""Passive bodies (no Mass): the reference builds the tree from the actives
and walks every body that has an Acceleration (ActiveTreeSolver.h:33-72).
nbg_sys_accelerations_at walks arbitrary points through the tree of the
system's state; ORDERED results must be the oracle's walk at those points,
bit for bit, with equal interaction counts."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _points(b, n, seed):
    rng = np.random.default_rng(seed)
    pts = rng.standard_normal((n, 3)).astype(np.float32) * np.float32(2.0)  # inside and outside the box
    pts[: n // 10] = np.stack([b.x, b.y, b.z], 1)[: n // 10]                # coincident with actives
    pts[n // 10: n // 10 + 5] = np.float32(50.0)                             # far outside
    return pts


@pytest.mark.parametrize("solver", [0, 1])
@pytest.mark.parametrize("mode", [0, 1])
def test_passive_points_match_oracle(nbg, orc, gpu_ctx, solver, mode):
    b = nbg.plummer(6000, seed=21)
    pts = _points(b, 4001, 5)
    p = nbg.default_params()
    p.solver, p.theta, p.leaf_size, p.morton_bits, p.eps2, p.sum_mode = solver, 0.5, 16, 63, 1e-4, mode
    s = nbg.System(gpu_ctx, p, b.n)
    s.set_state(b)
    acc = s.accelerations_at(pts[:, 0], pts[:, 1], pts[:, 2])
    t = orc.Tree(solver, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=63, eps2=1e-4)
    ref, pp, npc = t.accel(pts[:, 0], pts[:, 1], pts[:, 2], counts=True)
    assert s.counts() == (int(pp.sum()), int(npc.sum()))
    if mode == 0:
        assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))
    else:
        mag = np.linalg.norm(ref, axis=1)
        floor = np.sqrt(np.mean(mag ** 2))
        assert np.max(np.linalg.norm(acc - ref, axis=1) / np.maximum(mag, floor)) <= 1e-5
    # the bodies' own pass on the same tree, then passives again (tree reused)
    own = s.update_accelerations()
    again = s.accelerations_at(pts[:, 0], pts[:, 1], pts[:, 2])
    assert np.array_equal(again, acc)
    if mode == 0:
        assert np.array_equal(own, t.accel(b.x, b.y, b.z))
    # after a step the tree is rebuilt from the moved bodies
    s.step(1)
    s.synchronize()
    st = s.get_state()
    moved = s.accelerations_at(pts[:, 0], pts[:, 1], pts[:, 2])
    t2 = orc.Tree(solver, st.x, st.y, st.z, b.m, theta=0.5, leaf=16, bits=63, eps2=1e-4)
    ref2 = t2.accel(pts[:, 0], pts[:, 1], pts[:, 2])
    if mode == 0:
        assert np.array_equal(moved, ref2)
    s.close()


def test_passive_without_actives(nbg, gpu_ctx):
    s = nbg.System(gpu_ctx, nbg.default_params(), 0)
    s.set_state(nbg.Bodies.empty(0))
    a = s.accelerations_at(np.zeros(3), np.ones(3), np.zeros(3))
    assert a.shape == (3, 3) and not a.any()


def test_passive_scratch_reused_across_sizes(nbg, orc, gpu_ctx):
    """The passive-point walk keeps its device scratch between calls: target
    counts that grow, shrink and grow again give the oracle's bits each time."""
    b = nbg.plummer(5000, seed=23)
    p = nbg.default_params()
    p.solver, p.theta, p.leaf_size, p.morton_bits, p.eps2, p.sum_mode = 1, 0.5, 16, 63, 1e-4, 0
    s = nbg.System(gpu_ctx, p, b.n)
    s.set_state(b)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=63, eps2=1e-4)
    for k, nt in enumerate((3000, 17, 9001, 64, 1)):
        pts = _points(b, nt, 40 + k) if nt >= 20 else np.random.default_rng(k).standard_normal((nt, 3)).astype(
            np.float32)
        acc = s.accelerations_at(pts[:, 0], pts[:, 1], pts[:, 2])
        ref, pp, npc = t.accel(pts[:, 0], pts[:, 1], pts[:, 2], counts=True)
        assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32)), f"nt={nt}"
        assert s.counts() == (int(pp.sum()), int(npc.sum()))
    s.close()
