"""GPU NaiveSolver (Solvers/NaiveSolver.h:22-52): direct summation in body
order.  ORDERED is bit-exact with the oracle's sequential FP32 sum (itself
pinned by an independent numpy restatement in test_oracle.py); FAST agrees
within 1e-5 relative with the RMS floor; interaction counts are n per
body (SimpleTrackingRule counts the self call too)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(nbg, mode, eps2=1e-4, dt=1e-3):
    p = nbg.default_params()
    p.solver, p.eps2, p.dt, p.sum_mode = nbg.NAIVE, eps2, dt, mode
    return p


@pytest.mark.parametrize("n", [1, 2, 100, 3001, 20000])
def test_naive_matches_oracle(nbg, orc, gpu_ctx, n):
    b = nbg.plummer(n, seed=70 + n)
    ref = orc.direct_f32(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    for mode in (0, 1):
        s = nbg.System(gpu_ctx, _params(nbg, mode), n)
        s.set_state(b)
        acc = s.update_accelerations()
        assert s.counts() == (n * n, 0)
        if mode == 0:
            assert np.array_equal(acc.view(np.uint32), ref.view(np.uint32))
        else:
            mag = np.linalg.norm(ref, axis=1)
            floor = max(float(np.sqrt(np.mean(mag ** 2))), 1e-30)
            assert np.max(np.linalg.norm(acc - ref, axis=1) / np.maximum(mag, floor)) <= 1e-5
        s.close()


def test_naive_step_and_passive_points(nbg, orc, gpu_ctx):
    n = 2000
    b = nbg.plummer(n, seed=3)
    s = nbg.System(gpu_ctx, _params(nbg, 0), n)
    s.set_state(b)
    s.step(1)
    s.synchronize()
    st = s.get_state()
    a = orc.direct_f32(b.x, b.y, b.z, b.m, b.x, b.y, b.z, 1.0, 1e-4)
    f32, dt = np.float32, np.float32(1e-3)
    vx = (b.vx + a[:, 0] * dt).astype(f32)
    x = (b.x + vx * dt).astype(f32)
    assert np.array_equal(st.vx, vx) and np.array_equal(st.x, x)
    pts = np.random.default_rng(1).standard_normal((500, 3)).astype(f32)
    s.set_state(b)
    got = s.accelerations_at(pts[:, 0], pts[:, 1], pts[:, 2])
    ref = orc.direct_f32(b.x, b.y, b.z, b.m, pts[:, 0], pts[:, 1], pts[:, 2], 1.0, 1e-4)
    assert np.array_equal(got, ref)
    assert s.counts() == (n * 500, 0)
    s.close()
