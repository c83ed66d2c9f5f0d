"""GPU parity of the implicit FMM (ImplicitFMMSolver = MultipoleImplicitFMMSolver<1>,
Solvers/ImplicitFMMSolver.h:14-36) against the oracle's restatement of the
reference's dual traversal.

The GPU runs the traversal breadth-first (one launch per traversal step) and
takes every decision the reference takes (the same starting nodes, lockstep
partitions, particle and node criteria in FP32), so the four SimpleTrackingRule
counters
(pp, np, pn, nn) must be equal.  It sums in FAST mode (rsqrt + FMA, partial
local fields per lane), so accelerations agree within

    |a_gpu - a_ref| <= 1e-5 * max(|a_ref|, a_rms)        (per body).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _params(nbg, leaf, theta=0.4, eps2=1e-4, max_depth=32, dt=1e-3):
    p = nbg.default_params()
    p.solver = nbg.IMPLICIT_FMM
    p.leaf_size = leaf
    p.max_depth = max_depth
    p.theta = theta
    p.eps2 = eps2
    p.sum_mode = nbg.SUM_FAST
    p.dt = dt
    return p


def _gpu(nbg, ctx, b, leaf, **kw):
    s = nbg.System(ctx, _params(nbg, leaf, **kw), b.n)
    s.set_state(b)
    return s, s.update_accelerations()


def _oracle(orc, b, leaf, theta=0.4, eps2=1e-4, max_depth=32):
    t = orc.Tree(5, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, max_depth=max_depth, eps2=eps2)
    return t.fmm_accel(counts=True)


def _worst(acc, ref):
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2)) if len(mag) else 0.0
    err = np.linalg.norm(acc - ref, axis=1)
    return float(np.max(err / np.maximum(np.maximum(mag, floor), 1e-30))) if len(mag) else 0.0


CASES = [(1, 16, 0.4, 32), (2, 1, 0.4, 32), (9, 1, 0.4, 32), (64, 8, 0.4, 32), (1000, 16, 0.4, 32),
         (1000, 1, 0.4, 32), (5000, 8, 0.4, 3), (10000, 16, 0.4, 32), (20000, 16, 0.25, 32), (30011, 64, 0.4, 32),
         (50000, 16, 0.6, 32), (20000, 4, 0.8, 19)]


@pytest.mark.parametrize("n,leaf,theta,max_depth", CASES)
def test_fmm_matches_oracle(nbg, orc, gpu_ctx, n, leaf, theta, max_depth):
    print(f"[fmm] n={n} leaf={leaf} theta={theta} max_depth={max_depth}", flush=True)
    b = nbg.plummer(n, seed=7000 + n)
    s, acc = _gpu(nbg, gpu_ctx, b, leaf, theta=theta, max_depth=max_depth)
    ref, c = _oracle(orc, b, leaf, theta=theta, max_depth=max_depth)
    assert s.counts4() == c, "same interaction set (pp, np, pn, nn)"
    assert s.counts() == c[:2]
    w = _worst(acc, ref)
    print(f"  counts {c}, worst {w:.3g}", flush=True)
    assert w <= FAST_TOL


def test_fmm_matches_oracle_200k(nbg, orc, gpu_ctx):
    """A larger case against the oracle (about 7 s of CPU for the oracle)."""
    n = 200000
    b = nbg.plummer(n, seed=21)
    s, acc = _gpu(nbg, gpu_ctx, b, 16)
    ref, c = _oracle(orc, b, 16)
    assert s.counts4() == c
    w = _worst(acc, ref)
    print(f"  counts {c}, worst {w:.3g}", flush=True)
    assert w <= FAST_TOL


def test_fmm_two_clusters(nbg, orc, gpu_ctx):
    """Two Plummer spheres far apart (the colliding-spheres configuration at a
    small size): deep subtrees under a shallow root, far fields across the gap."""
    n = 20000
    a1, a2 = nbg.plummer(n, seed=31), nbg.plummer(n, seed=32)
    b = nbg.Bodies.empty(2 * n)
    for f in ("x", "y", "z", "vx", "vy", "vz"):
        getattr(b, f)[:n] = getattr(a1, f)
        getattr(b, f)[n:] = getattr(a2, f)
    b.x[:n] -= 10.0; b.x[n:] += 10.0
    b.y[:n] += 1.0; b.y[n:] -= 1.0
    b.m[:] = 0.5 / n
    s, acc = _gpu(nbg, gpu_ctx, b, 16, eps2=4e-6)
    ref, c = _oracle(orc, b, 16, eps2=4e-6)
    assert s.counts4() == c
    assert _worst(acc, ref) <= FAST_TOL


def test_fmm_small_theta_grows_the_pool(nbg, orc, gpu_ctx):
    """theta = 0.1: long close lists outgrow the list pool's first size (32
    words per body plus 128K), so the pass is repeated with a larger pool; the
    result is the oracle's, and the next pass starts with the larger pool."""
    n = 30000
    b = nbg.plummer(n, seed=23)
    s, acc = _gpu(nbg, gpu_ctx, b, 16, theta=0.1)
    ref, c = _oracle(orc, b, 16, theta=0.1)
    ws = s.walk_stats()
    print(f"  counts {c}, traversal {ws}", flush=True)
    assert ws["pool_capacity"] > 32 * n + (1 << 17) and ws["pool_peak"] <= ws["pool_capacity"]
    assert s.counts4() == c
    assert _worst(acc, ref) <= FAST_TOL
    again = s.update_accelerations()
    assert np.array_equal(again, acc) and s.walk_stats()["pool_capacity"] == ws["pool_capacity"]


def test_fmm_uniform_lattice(nbg, orc, gpu_ctx):
    """A cubic lattice of equal masses: many nodes hold equal body counts, so
    the load-balanced break resolves heap ties (std::push_heap/pop_heap) and
    many distances tie at the criteria's thresholds."""
    k = 24
    g = (np.arange(k, dtype=np.float32) + 0.5) / k
    x, y, z = (v.ravel().astype(np.float32) for v in np.meshgrid(g, g, g, indexing="ij"))
    b = nbg.Bodies.empty(len(x))
    b.x[:], b.y[:], b.z[:] = x, y, z
    b.m[:] = 1.0 / len(x)
    s, acc = _gpu(nbg, gpu_ctx, b, 8)
    ref, c = _oracle(orc, b, 8)
    assert s.counts4() == c
    assert _worst(acc, ref) <= FAST_TOL


def test_fmm_zero_softening(nbg, orc, gpu_ctx):
    """eps2 = 0: coincident bodies contribute nothing to each other."""
    b = nbg.plummer(5000, seed=17)
    b.x[1] = b.x[0]; b.y[1] = b.y[0]; b.z[1] = b.z[0]
    s, acc = _gpu(nbg, gpu_ctx, b, 16, eps2=0.0)
    ref, c = _oracle(orc, b, 16, eps2=0.0)
    bad = np.flatnonzero(~np.all(np.isfinite(acc), axis=1))
    if len(bad):
        print(f"  non-finite bodies {bad[:8]} (of {len(bad)}): device {acc[bad[:4]]}, oracle {ref[bad[:4]]}, "
              f"positions {np.stack([b.x, b.y, b.z], 1)[bad[:4]]}, counts {s.counts4()} vs {c}", flush=True)
    assert len(bad) == 0
    assert s.counts4() == c
    assert _worst(acc, ref) <= FAST_TOL


def test_fmm_theta_zero_is_direct_sum(nbg, gpu_ctx):
    """theta = 0 approximates nothing: every pair is a particle-particle call
    and the result is the direct sum."""
    n = 3000
    b = nbg.plummer(n, seed=5)
    s, acc = _gpu(nbg, gpu_ctx, b, 16, theta=0.0)
    pp, np_, pn, nn = s.counts4()
    assert (pp, pn, nn) == (n * n, 0, 0)
    ref = gpu_ctx.direct_f64(b, b.x, b.y, b.z, 1.0, 1e-4)
    assert _worst(acc, ref) <= 1e-5


def test_fmm_accuracy_at_one_million(nbg, gpu_ctx):
    """At the full Plummer size the FMM stays close to the FP64 direct sum
    (sampled), and a second pass repeats the first exactly."""
    n = 1 << 20
    b = nbg.plummer(n, seed=1)
    s, acc = _gpu(nbg, gpu_ctx, b, 16)
    assert np.all(np.isfinite(acc))
    sel = np.arange(0, n, 997)
    ref = gpu_ctx.direct_f64(b, b.x[sel], b.y[sel], b.z[sel], 1.0, 1e-4)
    rel = np.linalg.norm(acc[sel] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"  median {np.median(rel):.3g}, p99 {np.quantile(rel, 0.99):.3g}, max {rel.max():.3g}", flush=True)
    assert np.median(rel) < 2e-3 and np.quantile(rel, 0.99) < 2e-2
    pp, np_, pn, nn = s.counts4()
    assert pp > 0 and np_ > 0 and pn > 0 and nn > 0
    ws = s.walk_stats()
    print(f"  counts {(pp, np_, pn, nn)}, traversal {ws}", flush=True)
    assert 0 < ws["pool_peak"] <= ws["pool_capacity"] and 0 < ws["frontier_peak"] <= n
    again = s.update_accelerations()
    assert np.array_equal(again, acc)
    assert s.counts4() == (pp, np_, pn, nn)


@pytest.mark.slow
def test_fmm_colliding_spheres_16M(nbg, gpu_ctx):
    """The 16.7M-body colliding-spheres configuration (deep trees, a large
    density contrast): the pass completes within its list capacities, and a
    sample agrees with the FP64 direct sum as well as at 1M bodies."""
    from test_gpu_configs import _two_spheres

    b = _two_spheres(nbg, 8388608)
    n = b.n
    print(f"[fmm two spheres] n={n}", flush=True)
    s, acc = _gpu(nbg, gpu_ctx, b, 16, eps2=4e-6)
    assert np.all(np.isfinite(acc))
    ws = s.walk_stats()
    sel = np.random.default_rng(9).choice(n, 4096, replace=False)
    ref = gpu_ctx.direct_f64(b, b.x[sel], b.y[sel], b.z[sel], 1.0, 4e-6)
    rel = np.linalg.norm(acc[sel] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"  median {np.median(rel):.3g}, p99 {np.quantile(rel, 0.99):.3g}; traversal {ws}, counts {s.counts4()}",
          flush=True)
    assert np.median(rel) < 2e-3 and np.quantile(rel, 0.99) < 2e-2
    assert ws["pool_peak"] <= ws["pool_capacity"]
    s.close()


def test_fmm_steps_match_oracle(nbg, orc, gpu_ctx):
    """Solver::step with the FMM's accelerations (FAST tolerance)."""
    n = 3000
    b = nbg.plummer(n, seed=41)
    s = nbg.System(gpu_ctx, _params(nbg, 16), n)
    s.set_state(b)
    s.step(3)
    s.synchronize()
    st = s.get_state()
    (x, y, z, vx, vy, vz), _ = orc.step(5, (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz), 3, theta=0.4, leaf=16,
                                        bits=63, eps2=1e-4, dt=1e-3)
    assert np.max(np.abs(st.x - x)) < 1e-6 and np.max(np.abs(st.vx - vx)) < 1e-4


def test_fmm_rejects_unsupported_use(nbg, gpu_ctx):
    p = _params(nbg, 16)
    p.sum_mode = nbg.SUM_ORDERED
    with pytest.raises(nbg.NbgError, match="FAST"):
        nbg.System(gpu_ctx, p, 100)
    b = nbg.plummer(100, seed=2)
    s, _ = _gpu(nbg, gpu_ctx, b, 16)
    with pytest.raises(nbg.NbgError, match="passive"):
        s.accelerations_at(b.x[:3], b.y[:3], b.z[:3])


def test_fmm_degenerate_sets(nbg, orc, gpu_ctx):
    s = nbg.System(gpu_ctx, _params(nbg, 16), 0)
    s.set_state(nbg.Bodies.empty(0))
    assert s.update_accelerations().shape == (0, 3)
    # every body at one point: zero-size cells, 0/0 in both criteria
    b = nbg.Bodies.empty(300)
    b.m[:] = 1.0
    for theta in (0.4, 0.5, 0.7, 0.3):
        s, acc = _gpu(nbg, gpu_ctx, b, 16, max_depth=4, theta=theta)
        ref, c = _oracle(orc, b, 16, max_depth=4, theta=theta)
        assert s.counts4() == c, theta
        assert np.array_equal(acc, np.zeros_like(acc)) and np.array_equal(ref, np.zeros_like(ref))


def test_fmm_facade(nbg, orc, gpu_ctx):
    """The Python facade: ImplicitFMMSolver through MultiSolver-style select."""
    sim = nbg.Simulation(nbg.plummer(2000, seed=9))
    solver = nbg.select("gpu-implicit-fmm-2p")(sim, nbg.Gravity(1.0, 1e-4))
    assert solver.theta() == pytest.approx(0.4)
    solver.setMaxLeafSize(16)
    solver.updateAccelerations()
    b = sim.bodies
    ref, c = _oracle(orc, b, 16)
    assert solver.counts4() == c
    assert _worst(sim.acceleration, ref) <= FAST_TOL
