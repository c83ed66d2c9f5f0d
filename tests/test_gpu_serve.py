"""Continuous-batching server: jobs arriving while others run are admitted
as capacity frees up, run together in one forest, and each result is the
same bits as the same system stepped alone (nbg_sys, same parameters)."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(nbg, theta, leaf=16, eps2=1e-4, dt=1e-3, mode=1):
    p = nbg.default_params()
    p.solver, p.morton_bits = nbg.LINEAR_BVH, 30
    p.theta, p.leaf_size, p.eps2, p.dt, p.sum_mode = theta, leaf, eps2, dt, mode
    return p


def _jobs(nbg, count, nmin, nmax, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        if k % 2:
            b = nbg.cube(n, seed=900 + k, side=2.0)
        else:
            b = nbg.plummer(n, seed=900 + k)
        theta = float(rng.choice([0.3, 0.5, 0.7]))
        steps = int(rng.integers(2, 13))
        out.append((b, theta, steps))
    return out


def _alone(nbg, ctx, b, theta, steps, mode):
    s = nbg.System(ctx, _params(nbg, theta, mode=mode), b.n)
    s.set_state(b)
    s.step(steps)
    s.synchronize()
    st = s.get_state()
    s.close()
    return st


@pytest.mark.parametrize("mode", [1, 0])
def test_serving_results_equal_alone(nbg, gpu_ctx, mode):
    jobs = _jobs(nbg, 12, 200, 15000, 3)
    budget = nbg.Batch.bytes_for(40000, 64)
    srv = nbg.Server(gpu_ctx, budget, max_systems=64, sum_mode=mode)
    cap, _ = srv.capacity()
    assert 39000 <= cap <= 40000
    ids = []
    for b, theta, steps in jobs:
        ids.append(srv.submit(_params(nbg, theta, mode=mode), b, steps))
        time.sleep(0.002)  # later jobs arrive while earlier ones run
    results = [srv.wait(j, timeout=120.0) for j in ids]
    st = srv.stats()
    print(f"[serve] mode={mode} stats {st}", flush=True)
    assert st["jobs_done"] == len(jobs) and st["peak_bodies"] <= cap and st["peak_systems"] >= 2
    for (b, theta, steps), (out, t) in zip(jobs, results):
        ref = _alone(nbg, gpu_ctx, b, theta, steps, mode)
        for g, r in zip((out.x, out.y, out.z, out.vx, out.vy, out.vz), (ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz)):
            assert np.array_equal(g, r), f"job n={b.n} theta={theta} steps={steps}"
        assert t["latency"] >= t["device"] >= 0.0
    srv.close()


def test_serving_budget_admission(nbg, gpu_ctx):
    """Jobs that do not fit together wait for capacity; a job larger than the
    whole budget is refused at submission."""
    budget = nbg.Batch.bytes_for(5000, 8)
    srv = nbg.Server(gpu_ctx, budget, max_systems=8)
    cap, _ = srv.capacity()
    with pytest.raises(nbg.NbgError):
        srv.submit(_params(nbg, 0.5), nbg.plummer(cap + 1, seed=1), 2)
    jobs = [(nbg.plummer(3000, seed=10 + k), 0.5, 3) for k in range(5)]   # 15000 bodies > capacity
    ids = [srv.submit(_params(nbg, th), b, s) for b, th, s in jobs]
    times = [srv.wait(j, timeout=120.0)[1] for j in ids]
    st = srv.stats()
    assert st["jobs_done"] == 5 and st["peak_bodies"] <= cap and st["peak_systems"] == 1
    assert max(t["queued"] for t in times) > 0.0
    srv.close()


def test_serving_backfill(nbg, gpu_ctx):
    """A job that does not fit waits while smaller jobs behind it are admitted
    into the room left (backfill), and it still runs once room frees."""
    budget = nbg.Batch.bytes_for(10000, 8)
    srv = nbg.Server(gpu_ctx, budget, max_systems=8)
    cap, _ = srv.capacity()
    first = srv.submit(_params(nbg, 0.5), nbg.plummer(cap // 2 + 100, seed=40), 3000)
    time.sleep(0.2)  # the first job is running
    big = srv.submit(_params(nbg, 0.5), nbg.plummer(cap // 2 + 100, seed=41), 2)   # does not fit beside it
    small = srv.submit(_params(nbg, 0.5), nbg.plummer(500, seed=42), 2)            # fits beside it
    _, t_small = srv.wait(small, timeout=120.0)
    _, t_big = srv.wait(big, timeout=120.0)
    srv.wait(first, timeout=120.0)
    print(f"[serve backfill] small {t_small} big {t_big}", flush=True)
    # the small job was admitted while the big one still waited
    assert t_small["latency"] < t_big["queued"]
    srv.close()


def test_serving_nonfinite_job_fails_alone(nbg, gpu_ctx):
    """A job whose state becomes infinite (eps2 = 0 and a pair 1e-18 apart:
    the acceleration overflows) fails with its own error; the jobs beside it
    finish with the bits they get alone.  A non-finite input is refused at
    submission."""
    budget = nbg.Batch.bytes_for(40000, 16)
    srv = nbg.Server(gpu_ctx, budget, max_systems=16)
    bad_in = nbg.plummer(100, seed=50)
    bad_in.x[7] = np.nan
    with pytest.raises(nbg.NbgError):
        srv.submit(_params(nbg, 0.5), bad_in, 3)
    good = [(nbg.plummer(3000 + 500 * k, seed=60 + k), 0.5, 6) for k in range(3)]
    blow = nbg.Bodies.empty(2)
    blow.x[1] = np.float32(1e-18)
    blow.m[:] = 1.0
    ids = [srv.submit(_params(nbg, th), b, st) for b, th, st in good[:2]]
    bad = srv.submit(_params(nbg, 0.5, eps2=0.0), blow, 6)
    ids.append(srv.submit(_params(nbg, good[2][1]), good[2][0], good[2][2]))
    with pytest.raises(nbg.NbgError, match="non-finite"):
        srv.wait(bad, timeout=120.0)
    for (b, theta, steps), j in zip(good, ids):
        out, _ = srv.wait(j, timeout=120.0)
        ref = _alone(nbg, gpu_ctx, b, theta, steps, 1)
        for g, r in zip((out.x, out.y, out.z, out.vx, out.vy, out.vz), (ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz)):
            assert np.array_equal(g, r)
    srv.close()


def test_serving_destroy_releases_waiters(nbg, gpu_ctx):
    """Destroying a server with jobs queued and running ends them with an
    error; a client blocked in wait() returns instead of hanging."""
    import threading

    budget = nbg.Batch.bytes_for(20000, 4)
    srv = nbg.Server(gpu_ctx, budget, max_systems=4)
    cap, _ = srv.capacity()
    running = srv.submit(_params(nbg, 0.5), nbg.plummer(cap - 10, seed=70), 100000)
    queued = srv.submit(_params(nbg, 0.5), nbg.plummer(cap - 10, seed=71), 5)
    result = {}

    def waiter():
        try:
            srv.wait(queued)
            result["status"] = "finished"
        except nbg.NbgError as e:
            result["status"] = str(e)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.5)
    srv.close()
    th.join(timeout=60.0)
    assert not th.is_alive()
    assert "destroyed" in result["status"]
    del running


@pytest.mark.parametrize("mode", [1, 0])
def test_serving_express_lane_results_equal_alone(nbg, gpu_ctx, mode):
    """With an express lane, jobs up to its size run in a second forest that
    steps concurrently with the bulk forest; every job, in either lane, ends
    with the bits it gets alone.  A long large job in the bulk lane does not
    hold the small jobs back: they finish while it runs."""
    budget = nbg.Batch.bytes_for(300000, 16)
    xbytes = nbg.Batch.bytes_for(40000, 32)
    srv = nbg.Server(gpu_ctx, budget + xbytes, max_systems=16, sum_mode=mode, express_max_bodies=8192,
                     express_bytes=xbytes, express_systems=32)
    cap, slots = srv.capacity()
    assert 299000 <= cap <= 300000 and slots == 48
    big_b = nbg.plummer(200000, seed=80)
    big = srv.submit(_params(nbg, 0.5, mode=mode), big_b, 150)
    time.sleep(0.05)
    small = _jobs(nbg, 10, 300, 8192, 5)          # express lane
    mid = _jobs(nbg, 3, 9000, 30000, 6)           # bulk lane, beside the big job
    ids = [srv.submit(_params(nbg, th, mode=mode), b, st) for b, th, st in small + mid]
    results = [srv.wait(j, timeout=300.0) for j in ids]
    big_out, t_big = srv.wait(big, timeout=300.0)
    st = srv.stats()
    print(f"[serve express] mode={mode} stats {st} big {t_big}", flush=True)
    assert st["jobs_done"] == len(ids) + 1
    for (b, theta, steps), (out, t) in zip(small + mid, results):
        ref = _alone(nbg, gpu_ctx, b, theta, steps, mode)
        for g, r in zip((out.x, out.y, out.z, out.vx, out.vy, out.vz), (ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz)):
            assert np.array_equal(g, r), f"job n={b.n} theta={theta} steps={steps}"
    ref = _alone(nbg, gpu_ctx, big_b, 0.5, 150, mode)
    for g, r in zip((big_out.x, big_out.vx), (ref.x, ref.vx)):
        assert np.array_equal(g, r)
    # the small jobs (at most 12 steps each) finished while the 150-step
    # 200k-body job was still on the device
    assert max(t["latency"] for _, t in results[:len(small)]) < t_big["latency"]
    srv.close()


def test_serving_express_lane_arguments(nbg, gpu_ctx):
    budget = nbg.Batch.bytes_for(20000, 4)
    with pytest.raises(nbg.NbgError):   # express budget not inside the server's
        nbg.Server(gpu_ctx, budget, max_systems=4, express_max_bodies=1000, express_bytes=budget, express_systems=4)
    with pytest.raises(nbg.NbgError):   # express lane cannot hold one job of its size
        nbg.Server(gpu_ctx, budget, max_systems=4, express_max_bodies=10 ** 6,
                   express_bytes=nbg.Batch.bytes_for(1000, 4), express_systems=4)


def _same_bits(out, ref, what):
    for g, r in zip((out.x, out.y, out.z, out.vx, out.vy, out.vz), (ref.x, ref.y, ref.z, ref.vx, ref.vy, ref.vz)):
        assert np.array_equal(g, r), what


@pytest.mark.parametrize("mode", [1, 0])
def test_serving_snapshot_resume_equals_uninterrupted(nbg, gpu_ctx, mode):
    """Job snapshots: a resident job checkpointed (it runs on), a resident job
    suspended, and a queued job suspended, each resumed from its snapshot on
    this server or on a second one, end with the bits of the same system
    stepped alone without interruption.  A queued job's snapshot holds its
    submitted state; a suspended job's id is forgotten."""
    budget = nbg.Batch.bytes_for(60000, 16)
    srv = nbg.Server(gpu_ctx, budget, max_systems=16, sum_mode=mode)
    srv2 = nbg.Server(gpu_ctx, budget, max_systems=16, sum_mode=mode)
    a = (nbg.plummer(40000, seed=51), 0.5, 200)
    b = (nbg.cube(15000, seed=52, side=2.0), 0.3, 150)
    c = (nbg.plummer(30000, seed=53), 0.7, 10)     # 85k bodies with a: queued
    ia, ib, ic = [srv.submit(_params(nbg, th, mode=mode), body, st) for body, th, st in (a, b, c)]
    t0 = time.time()
    while srv.stats()["steps_run"] < 3 and time.time() - t0 < 60.0:
        time.sleep(0.001)
    snap_c = srv.suspend(ic)
    info = nbg.snapshot_info(snap_c)
    assert info["steps_done"] == 0 and info["steps_total"] == 10 and info["n"] == 30000
    arrays = np.frombuffer(snap_c, dtype=np.float32, offset=112).reshape(7, 30000)
    for k, v in enumerate((c[0].x, c[0].y, c[0].z, c[0].m, c[0].vx, c[0].vy, c[0].vz)):
        assert np.array_equal(arrays[k], v)
    srv._n[ic] = c[0].n
    with pytest.raises(nbg.NbgError, match="unknown job"):
        srv.wait(ic, timeout=0.0)
    snap_a = srv.checkpoint(ia)
    snap_b = srv.suspend(ib)
    ia_info, ib_info = nbg.snapshot_info(snap_a), nbg.snapshot_info(snap_b)
    print(f"[serve snapshot] mode={mode} a {ia_info} b {ib_info}", flush=True)
    assert 0 < ia_info["steps_done"] < 200 and 0 < ib_info["steps_done"] < 150
    ra = srv2.resume(snap_a)
    rb = srv.resume(snap_b)
    rc = srv2.resume(snap_c)
    _same_bits(srv.wait(ia, timeout=300.0)[0], _alone(nbg, gpu_ctx, a[0], a[1], a[2], mode), "checkpointed job")
    _same_bits(srv2.wait(ra, timeout=300.0)[0], _alone(nbg, gpu_ctx, a[0], a[1], a[2], mode), "resumed checkpoint")
    _same_bits(srv.wait(rb, timeout=300.0)[0], _alone(nbg, gpu_ctx, b[0], b[1], b[2], mode), "resumed suspension")
    _same_bits(srv2.wait(rc, timeout=300.0)[0], _alone(nbg, gpu_ctx, c[0], c[1], c[2], mode), "resumed queued job")
    bad = bytearray(snap_b)
    bad[1000] ^= 0x10
    with pytest.raises(nbg.NbgError, match="checksum"):
        srv.resume(bytes(bad))
    srv.close()
    srv2.close()
