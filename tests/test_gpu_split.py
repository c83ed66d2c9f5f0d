"""The FAST tree walks (linear BVH and octree, monopole and quadrupole)
split over several waves per group of 64 targets: every split (1 to 16
waves, and the size-chosen default) gives the same bits and the same
interaction counts, and those counts are the oracle's.  The same bits as
the forest and the several-targets-per-lane walks are checked where those
are tested against a system alone."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FAST_TOL = 1e-5


def _params(nbg, solver, theta=0.5, leaf=16, bits=63, eps2=1e-4):
    p = nbg.default_params()
    p.solver, p.theta, p.leaf_size, p.morton_bits, p.eps2 = solver, theta, leaf, bits, eps2
    p.sum_mode = nbg.SUM_FAST
    return p


CASES = [(1, "plummer"), (2, "cube"), (63, "plummer"), (65, "cube"), (1000, "plummer"), (4099, "cube"),
         (10000, "plummer"), (50000, "plummer"), (65536, "cube")]


@pytest.mark.parametrize("tree", ["lbvh", "octree"])
@pytest.mark.parametrize("quad", [False, True])
@pytest.mark.parametrize("n,kind", CASES)
def test_split_walk_same_bits(nbg, orc, gpu_ctx, n, kind, quad, tree):
    print(f"[split] n={n} {kind} quad={quad} {tree}", flush=True)
    b = nbg.plummer(n, seed=3000 + n) if kind == "plummer" else nbg.cube(n, seed=3000 + n, side=2.0)
    if tree == "lbvh":
        solver = nbg.LINEAR_BVH_QUADRUPOLE if quad else nbg.LINEAR_BVH
    else:
        solver = nbg.BARNES_HUT_QUADRUPOLE if quad else nbg.BARNES_HUT
    s = nbg.System(gpu_ctx, _params(nbg, solver), n)
    s.set_state(b)
    ref = None
    for k in (1, 2, 4, 8, 16, 0):
        s.set_walk_split(k)
        acc = s.update_accelerations()
        counts = s.counts()
        if ref is None:
            ref = (acc, counts)
            continue
        assert np.array_equal(acc, ref[0]), f"split {k}"
        assert counts == ref[1], f"split {k}"
    s.close()
    if quad:
        return
    t = orc.Tree(1 if tree == "lbvh" else 0, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=63, eps2=1e-4)
    want, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    assert ref[1] == (int(pp.sum()), int(npc.sum()))
    mag = np.linalg.norm(want, axis=1)
    floor = np.sqrt(np.mean(mag ** 2)) if n > 1 else 1.0
    err = np.linalg.norm(ref[0] - want, axis=1) / np.maximum(mag, floor)
    assert float(err.max()) <= FAST_TOL


def test_split_argument(nbg, gpu_ctx):
    s = nbg.System(gpu_ctx, _params(nbg, nbg.LINEAR_BVH), 100)
    with pytest.raises(nbg.NbgError):
        s.set_walk_split(3)
    s.close()
    B = nbg.Batch(gpu_ctx, 1000, 4)
    with pytest.raises(nbg.NbgError):
        B.set_walk_split(32)
    B.close()


@pytest.mark.parametrize("tree", ["lbvh", "octree"])
@pytest.mark.parametrize("quad", [False, True])
def test_forest_member_equals_alone_on_both_sides_of_the_split_limit(nbg, gpu_ctx, tree, quad):
    """Systems of 65,536 bodies or fewer (block form, split alone) and larger
    (one block: the plain walk alone) in one forest: each member's
    accelerations are the bits of the same system alone."""
    if tree == "lbvh":
        solver = nbg.LINEAR_BVH_QUADRUPOLE if quad else nbg.LINEAR_BVH
    else:
        solver = nbg.BARNES_HUT_QUADRUPOLE if quad else nbg.BARNES_HUT
    systems = [nbg.plummer(70001, seed=81), nbg.cube(9000, seed=82, side=2.0), nbg.plummer(65536, seed=83)]
    p = _params(nbg, solver, bits=30)
    B = nbg.Batch(gpu_ctx, 150000, 8, sum_mode=nbg.SUM_FAST)
    ids = [B.add(p, b) for b in systems]
    B.update_accelerations()
    for sid, b in zip(ids, systems):
        _, got = B.get(sid, accelerations=True)
        s = nbg.System(gpu_ctx, p, b.n)
        s.set_state(b)
        want = s.update_accelerations()
        s.close()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"n={b.n}"


@pytest.mark.parametrize("tree", ["lbvh", "octree"])
@pytest.mark.parametrize("quad", [False, True])
def test_small_forest_split_same_bits(nbg, gpu_ctx, tree, quad):
    """A FAST forest of at most 65,536 bodies in all splits each group over
    several waves (the member's frontier blocks per wave, the sums met in
    one reduce): every member keeps its bits and the counts for 1 to 16
    waves, and equals the same system alone."""
    if tree == "lbvh":
        solver = nbg.LINEAR_BVH_QUADRUPOLE if quad else nbg.LINEAR_BVH
    else:
        solver = nbg.BARNES_HUT_QUADRUPOLE if quad else nbg.BARNES_HUT
    sizes = [1, 2, 65, 700, 5000, 20000, 129, 31000]
    systems = [nbg.plummer(m, seed=90 + i) if i % 2 else nbg.cube(m, seed=90 + i, side=2.0)
               for i, m in enumerate(sizes)]
    assert sum(sizes) <= 65536
    p = _params(nbg, solver, bits=30)
    runs = []
    for k in (1, 2, 4, 8, 16, 0):
        print(f"[forest split] {tree} quad={quad} waves={k}", flush=True)
        B = nbg.Batch(gpu_ctx, 65536, 16, sum_mode=nbg.SUM_FAST)
        ids = [B.add(p, b) for b in systems]
        B.set_walk_split(k)
        B.update_accelerations()
        runs.append(([B.get(sid, accelerations=True)[1] for sid in ids], B.counts()))
        B.close()
    for k, (accs, c) in zip((2, 4, 8, 16, 0), runs[1:]):
        assert c == runs[0][1], f"split {k}"
        for b, a, a1 in zip(systems, accs, runs[0][0]):
            assert np.array_equal(a.view(np.uint32), a1.view(np.uint32)), f"split {k}, n={b.n}"
    for b, a in zip(systems, runs[-1][0]):
        s = nbg.System(gpu_ctx, p, b.n)
        s.set_state(b)
        want = s.update_accelerations()
        s.close()
        assert np.array_equal(a.view(np.uint32), want.view(np.uint32)), f"alone, n={b.n}"


@pytest.mark.parametrize("tree", ["lbvh", "octree"])
@pytest.mark.parametrize("quad", [False, True])
@pytest.mark.parametrize("n,kind,bits", [(1, "plummer", 63), (65, "cube", 30), (1000, "plummer", 63),
                                         (4099, "cube", 30), (10000, "plummer", 63), (16384, "plummer", 30),
                                         (65536, "cube", 63)])
def test_ordered_split_bitexact(nbg, orc, gpu_ctx, n, kind, bits, quad, tree):
    """The ORDERED walks split over 1 to 16 waves per group (each frontier
    subtree's result computed from zero, as the reference's recursion
    returns it, and used where the walk meets it) give the oracle's bits and
    counts for every split."""
    print(f"[ordered split] n={n} {kind} bits={bits} quad={quad} {tree}", flush=True)
    b = nbg.plummer(n, seed=5000 + n) if kind == "plummer" else nbg.cube(n, seed=5000 + n, side=2.0)
    if tree == "lbvh":
        solver, kind_id = (nbg.LINEAR_BVH_QUADRUPOLE, 4) if quad else (nbg.LINEAR_BVH, 1)
    else:
        solver, kind_id = (nbg.BARNES_HUT_QUADRUPOLE, 3) if quad else (nbg.BARNES_HUT, 0)
    p = _params(nbg, solver, bits=bits)
    p.sum_mode = nbg.SUM_ORDERED
    t = orc.Tree(kind_id, b.x, b.y, b.z, b.m, theta=0.5, leaf=16, bits=bits, eps2=1e-4)
    want, pp, npc = t.accel(b.x, b.y, b.z, counts=True)
    s = nbg.System(gpu_ctx, p, n)
    s.set_state(b)
    for k in (1, 2, 4, 8, 16, 0):
        s.set_walk_split(k)
        acc = s.update_accelerations()
        assert np.array_equal(acc.view(np.uint32), want.view(np.uint32)), f"split {k}"
        assert s.counts() == (int(pp.sum()), int(npc.sum())), f"split {k}"
    s.close()
