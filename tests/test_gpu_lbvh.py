"""GPU parity of the linear-BVH force pass against the oracle.

ORDERED mode must be bit-exact with the oracle at every stage: bounding box,
Morton keys, sort permutation (ties by original index), tree topology,
summaries, accelerations and interaction counts.  FAST mode must take the
same opening decisions (equal interaction counts) and agree within the
stated floating-point tolerance:

    |a_gpu - a_ref| <= 1e-5 * max(|a_ref|, a_rms)        (per body)

where a_rms is the RMS acceleration over the bodies.  The floor a_rms keeps
the bound meaningful at bodies whose net force cancels (|a_ref| -> 0).
"""

import numpy as np
import pytest

from treecheck import lbvh_check

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _params(nbg, leaf, bits, theta=0.5, eps2=1e-4, mode=0, dt=1e-3):
    p = nbg.default_params()
    p.solver = nbg.LINEAR_BVH
    p.leaf_size = leaf
    p.morton_bits = bits
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


def _node_table(first, count, leaf, com4, geo6):
    order = np.lexsort((count, first))
    return first[order], count[order], leaf[order], com4[order], geo6[order]


CASES = [(1, 16, 63), (2, 1, 30), (3, 1, 63), (17, 4, 30), (64, 8, 63), (1000, 16, 30), (1000, 1, 63),
         (4099, 16, 63), (10000, 16, 30), (10000, 64, 63), (65537, 16, 30)]


@pytest.mark.parametrize("n,leaf,bits", CASES)
def test_lbvh_ordered_bitexact(nbg, orc, gpu_ctx, n, leaf, bits):
    print(f"[lbvh ordered] n={n} leaf={leaf} bits={bits}", flush=True)
    b = nbg.plummer(n, seed=1000 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=leaf, bits=bits)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=leaf, bits=bits, eps2=1e-4)

    assert np.array_equal(s.debug_bbox(), t.bbox())
    gk, gp = s.debug_order()
    ok, op = t.order()
    assert np.array_equal(gk, ok), "sorted Morton keys"
    assert np.array_equal(gp, op), "sort permutation (ties by original index)"

    gt = s.debug_tree()
    on = t.nodes()
    lbvh_check(gk, gt["first"], gt["count"], gt["leaf"], leaf)
    g = _node_table(gt["first"], gt["count"], gt["leaf"], gt["com4"], gt["geo6"])
    o = _node_table(on["first"], on["count"], on["leaf"], on["com4"], on["geo6"])
    for a_, b_, what in zip(g, o, ("first", "count", "leaf", "com/mass", "bbox")):
        assert np.array_equal(a_, b_), f"node {what}"

    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert np.array_equal(acc, ref), "accelerations must be bit-exact in ORDERED mode"
    assert s.counts() == (int(pp.sum()), int(npc.sum()))


@pytest.mark.parametrize("n,leaf,bits,theta", [(1000, 16, 30, 0.5), (10000, 16, 30, 0.5), (10000, 16, 63, 0.3),
                                               (50000, 8, 63, 0.7)])
def test_lbvh_fast_within_tolerance(nbg, orc, gpu_ctx, n, leaf, bits, theta):
    print(f"[lbvh fast] n={n} leaf={leaf} bits={bits} theta={theta}", flush=True)
    b = nbg.plummer(n, seed=77 + n)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=leaf, bits=bits, theta=theta, mode=1)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, bits=bits, eps2=1e-4)
    ref, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert s.counts() == (int(pp.sum()), int(npc.sum())), "same opening decisions"
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2))
    err = np.linalg.norm(acc - ref, axis=1)
    worst = float(np.max(err / np.maximum(mag, floor)))
    print(f"  fast-vs-ordered worst {worst:.3g}, median rel {np.median(err / mag):.3g}", flush=True)
    assert worst <= FAST_TOL


@pytest.mark.parametrize("n,bits,quad", [(20037, 30, False), (50000, 63, False), (20037, 30, True), (130, 30, False),
                                          (1, 30, False), (100003, 30, False), (100003, 30, True)])
def test_lbvh_targets_per_lane_same_bits(nbg, gpu_ctx, n, bits, quad):
    """FAST with 2 or 4 targets per lane (128 or 256 bodies per wavefront):
    every 64-target slot keeps its own masks and visits its nodes in the
    same depth-first order as a wavefront of its own, so accelerations and
    counts are the bits of one target per lane (sizes not a multiple of the
    wavefront's targets included)."""
    b = nbg.plummer(n, seed=41)
    res = []
    for tpl in (1, 2, 4):
        p = _params(nbg, 16, bits, mode=1)
        if quad:
            p.solver = nbg.LINEAR_BVH_QUADRUPOLE
        p.targets_per_lane = tpl
        s = nbg.System(gpu_ctx, p, n)
        s.set_state(b)
        res.append((s.update_accelerations(), s.counts()))
        s.close()
    for tpl, (acc, c) in zip((2, 4), res[1:]):
        assert c == res[0][1], f"{tpl} targets per lane: counts"
        assert np.array_equal(acc.view(np.uint32), res[0][0].view(np.uint32)), f"{tpl} targets per lane: bits"


def test_targets_per_lane_is_checked(nbg, gpu_ctx):
    p = _params(nbg, 16, 30, mode=1)
    p.targets_per_lane = 3
    with pytest.raises(nbg.NbgError):
        nbg.System(gpu_ctx, p, 100)


@pytest.mark.parametrize("mode", [0, 1])
def test_lbvh_steps_match_oracle(nbg, orc, gpu_ctx, mode):
    """Solver::step x5 on the device vs the oracle's step (bit-exact ORDERED)."""
    n = 3000
    b = nbg.plummer(n, seed=31)
    s = nbg.System(gpu_ctx, _params(nbg, 16, 63, mode=mode, dt=1e-3), n)
    s.set_state(b)
    s.step(5)
    s.synchronize()
    st = s.get_state()
    (x, y, z, vx, vy, vz), _ = orc.step(1, (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz), 5, theta=0.5, leaf=16, bits=63,
                                        eps2=1e-4, dt=1e-3)
    if mode == 0:
        for g, r in zip((st.x, st.y, st.z, st.vx, st.vy, st.vz), (x, y, z, vx, vy, vz)):
            assert np.array_equal(g, r)
    else:
        assert np.max(np.abs(st.x - x)) < 1e-6 and np.max(np.abs(st.vx - vx)) < 1e-4


def test_lbvh_empty_and_degenerate(nbg, orc, gpu_ctx):
    # zero bodies
    s = nbg.System(gpu_ctx, _params(nbg, 16, 63), 0)
    s.set_state(nbg.Bodies.empty(0))
    assert s.update_accelerations().shape == (0, 3)
    # all bodies coincident: zero-extent box -> every key 0 -> one leaf
    n = 300
    b = nbg.Bodies.empty(n)
    b.x[:] = 1.5
    b.m[:] = 1.0
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=8, bits=30)
    gt = s.debug_tree()
    assert len(gt["first"]) == 1 and gt["leaf"][0] and gt["count"][0] == n
    assert np.all(acc == 0)
    assert s.counts() == (n * n, 0)
    # two clusters of exact duplicates
    b = nbg.Bodies.empty(200)
    b.x[100:] = 1.0
    b.m[:] = 0.5
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=4, bits=63)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, leaf=4, bits=63, eps2=1e-4)
    assert np.array_equal(acc, t.accel(b.x, b.y, b.z))


def test_lbvh_zero_softening_and_zero_mass(nbg, orc, gpu_ctx):
    b = nbg.plummer(2000, seed=8)
    b.m[::7] = 0.0
    for mode in (0, 1):
        s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=16, bits=63, eps2=0.0, mode=mode)
        t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=63, eps2=0.0)
        ref = t.accel(b.x, b.y, b.z)
        assert np.all(np.isfinite(acc))
        if mode == 0:
            assert np.array_equal(acc, ref)


def test_lbvh_large_properties(nbg, orc, gpu_ctx):
    """1,048,576-body Plummer, 30-bit keys, theta 0.5: sort/tree equal to the
    oracle's; ORDERED accelerations bit-exact on a target sample; force
    error against the FP64 direct sum on a 16,384-body sample."""
    n = 1 << 20
    print(f"[lbvh large] n={n}", flush=True)
    b = nbg.plummer(n, seed=1)
    s, acc = _gpu_pass(nbg, gpu_ctx, b, leaf=16, bits=30, theta=0.5, eps2=2.5e-5)
    t = orc.Tree(1, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=30, eps2=2.5e-5)
    gk, gp = s.debug_order()
    ok, op = t.order()
    assert np.array_equal(gk, ok) and np.array_equal(gp, op)
    gt = s.debug_tree()
    on = t.nodes()
    g = _node_table(gt["first"], gt["count"], gt["leaf"], gt["com4"], gt["geo6"])
    o = _node_table(on["first"], on["count"], on["leaf"], on["com4"], on["geo6"])
    for a_, b_ in zip(g, o):
        assert np.array_equal(a_, b_)
    sel = np.random.default_rng(0).choice(n, 4096, replace=False)
    ref = t.accel(b.x[sel], b.y[sel], b.z[sel])
    assert np.array_equal(acc[sel], ref)
    sub = np.random.default_rng(1).choice(n, 16384, replace=False)
    d = gpu_ctx.direct_f64(b, b.x[sub], b.y[sub], b.z[sub], 1.0, 2.5e-5)
    rel = np.linalg.norm(acc[sub] - d, axis=1) / np.linalg.norm(d, axis=1)
    print(f"  force error vs FP64: median {np.median(rel):.3g} p99 {np.quantile(rel, 0.99):.3g}", flush=True)
    assert np.median(rel) < 5e-3


def test_direct_f64_gpu_matches_oracle(nbg, orc, gpu_ctx):
    b = nbg.plummer(3000, seed=12)
    d = gpu_ctx.direct_f64(b, b.x[:500], b.y[:500], b.z[:500], 1.0, 1e-4)
    r = orc.direct_f64(b.x, b.y, b.z, b.m, b.x[:500], b.y[:500], b.z[:500], 1.0, 1e-4)
    assert np.allclose(d, r, rtol=1e-11, atol=0)


@pytest.mark.parametrize("n", [2, 3, 63, 65, 130, 4000])
def test_potential_gpu_matches_oracle(nbg, orc, gpu_ctx, n):
    """The device potential (each pair once, refined FP64 rsqrt, several
    launches for large n) against the oracle's FP64 double sum."""
    b = nbg.plummer(n, seed=13 + n)
    s = nbg.System(gpu_ctx, _params(nbg, 16, 63), b.n)
    s.set_state(b)
    w = s.potential_f64()
    # the system's eps2 is an FP32 parameter: the oracle gets the same value
    r = orc.potential_f64(b.x, b.y, b.z, b.m, 1.0, float(np.float32(1e-4)))
    assert abs(w - r) <= 1e-12 * abs(r)


def test_potential_gpu_closed_form(nbg, gpu_ctx):
    """Two bodies at distance 3 with eps2 = 16: W = -G m1 m2 / 5."""
    b = nbg.Bodies.empty(2)
    b.x[1] = 3.0
    b.m[:] = (2.0, 0.5)
    p = _params(nbg, 16, 63, eps2=16.0)
    p.G = 4.0
    s = nbg.System(gpu_ctx, p, 2)
    s.set_state(b)
    assert abs(s.potential_f64() - (-4.0 * 2.0 * 0.5 / 5.0)) <= 1e-15


def test_phase_timing(nbg, gpu_ctx):
    b = nbg.plummer(20000, seed=2)
    s = nbg.System(gpu_ctx, _params(nbg, 16, 30, mode=1), b.n)
    s.set_state(b)
    s.enable_timing(True)
    s.step(3)
    t = s.phase_times()
    assert set(t) == set(nbg.PHASES) and t["walk"] > 0
