#!/usr/bin/env python3
"""Benchmark of the MI355X tree-code force pass.

Workload (default): Plummer sphere, N = 1,048,576 (a = 1, M = 1, G = 1,
r < 10 a), FP32, eps = 0.005 (eps2 = 2.5e-5), theta = 0.5, leaf <= 16,
leapfrog dt = 5e-4, seed 1.  A "step" is Solver::step(): tree build + walk +
kick/drift, all on the device with the state resident in HBM.

The headline line is the linear-BVH solver with 30-bit Morton keys over the
bounding cube; the same run measures the octree Barnes-Hut solver on the
same bodies and reports it in the "barnes_hut" block, and the quadrupole
variants of both trees ("linear_bvh_4p", "barnes_hut_4p") at the opening
angles the reference's benchmark tunes them to (src/Benchmark/profile.cpp:
47-53: 0.72578126 and 0.41289064), and the implicit FMM ("implicit_fmm") at
its own criterion's angle (ImplicitDualTraversalSolver.h:22: 0.4).  Each
solver carries:
bodies*steps/s over the timed steps, interactions/s, the walk kernel's FP32
roofline fraction (HIP events on the system stream), the force error against
an FP64 direct sum on a 16,384-body sample, the energy drift over the run and
the oracle's CPU time for the same solver on this host.
"""

import argparse
import copy
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bodies·steps/s at N=1M θ=0.5; batched-serving p50/p99 latency at fixed force err"
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (packed) peak
FLOPS_PER_INTERACTION = 20
SOLVER_NAMES = {0: "barnes-hut", 1: "linear-bvh", 3: "barnes-hut-4p", 4: "linear-bvh-4p", 5: "implicit-fmm"}
ORDERED_WALK_KERNELS = {0: "k_walk_oct_ordered", 1: "k_walk_lbvh_ordered", 3: "k_walk_oct_ordered",
                        4: "k_walk_lbvh_ordered"}
WALK_KERNELS = {0: "k_walk_oct_fast", 1: "k_walk_lbvh_plain", 3: "k_walk_oct_fast", 4: "k_walk_lbvh_plain",
                5: "k_fmm_groups + k_fmm_step (the whole traversal)"}
# FP32 operations of one body or node into a local field (difference, r^2,
# rsqrt and powers, the vector's 3 and the gradient's 6 accumulating FMAs).
FLOPS_PER_LOCAL = 40
# FP32 operations of one quadrupole node interaction in the FAST walk
# (difference, r^2, rsqrt and powers, Q d, d.Q.d, the two coefficients and
# the 6 accumulating FMAs); monopole node and body interactions count 20.
FLOPS_PER_QUAD_NODE = 50
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_traffic_plummer1M.txt")


_T0 = time.perf_counter()


def log(msg):
    print(f"[{time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def load_traffic():
    """HBM bytes per walk launch measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (scaled by a copy kernel of known size), kept as text beside
    the profiles; keyed by solver name."""
    out = {}
    if not os.path.exists(TRAFFIC_FILE):
        return out
    with open(TRAFFIC_FILE) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2 and not line.startswith("#"):
                try:
                    out[parts[0]] = float(parts[1])
                except ValueError:
                    pass
    return out


def host_info():
    """CPU model and core counts of this host (the CPU baseline's machine)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def error_stats(acc, ref, m):
    """Force error of `acc` against the FP64 reference `ref` (N x 3): median,
    p99 and max of |da| / |a_ref|, and the reference benchmark's
    constitutional error (src/Benchmark/Graders/ConstitutionalGrader.h:50-64):
    100 * max_i |F_i - F_ref,i| / min(|F_ref,i|, F_rms), F = m a, F_rms the
    RMS reference force over the bodies compared."""
    acc = np.asarray(acc, np.float64)
    ref = np.asarray(ref, np.float64)
    m = np.asarray(m, np.float64)
    d = np.linalg.norm(acc - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    rel = d / mag
    F, dF = mag * m, d * m
    f_rms = np.sqrt(np.mean(F ** 2))
    return {
        "force_err_vs_fp64_median": float(np.median(rel)),
        "force_err_vs_fp64_p99": float(np.quantile(rel, 0.99)),
        "force_err_vs_fp64_max": float(rel.max()),
        "constitutional_error_pct": float(100.0 * np.max(dF / np.minimum(F, f_rms))),
        "sample": int(len(rel)),
    }


def tolerance_stats(acc, ref):
    """FAST accelerations against the reference's own FP32 arithmetic (the
    ORDERED pass, bit-exact with the oracle) over every body: the largest
    |da| / max(|a_ref|, a_rms), the largest strict |da| / |a_ref|, and the
    fraction of bodies with |da| > 1e-5 |a_ref|."""
    acc = np.asarray(acc, np.float64)
    ref = np.asarray(ref, np.float64)
    d = np.linalg.norm(acc - ref, axis=1)
    mag = np.linalg.norm(ref, axis=1)
    floor = np.sqrt(np.mean(mag ** 2))
    nz = mag > 0
    return {
        "max_rel_floored": float(np.max(d / np.maximum(mag, floor))),
        "max_rel_strict": float(np.max(d[nz] / mag[nz])) if nz.any() else 0.0,
        "frac_above_1e-5_strict": float(np.mean(d > 1e-5 * mag)),
        "bodies": int(len(d)),
    }


def energy(m, vx, vy, vz, a, dt, W):
    """Total energy with velocities synchronised to the positions: the
    reference's step (Solver.h:103-139: v += a dt, then x += v dt) carries
    v_{n-1/2}, so v(t_n) = v + a(x_n) dt / 2 at both ends of a run."""
    m = np.asarray(m, np.float64)
    v = np.stack([vx, vy, vz], 1).astype(np.float64) + np.asarray(a, np.float64) * (dt / 2)
    return 0.5 * np.sum(m * (v ** 2).sum(1)) + W


def solver_theta(args, solver):
    return {3: args.theta_4p_bh, 4: args.theta_4p_lbvh, 5: args.theta_fmm}.get(solver, args.theta)


def cpu_baseline(b, args, solver, threads):
    """The oracle's restatement of the reference CPU path for `solver`, timed
    on this host: a full tree build plus the walk for a target sample,
    scaled to N, plus the kick/drift passes."""
    import oracle

    n = b.n
    t0 = time.perf_counter()
    t = oracle.Tree(solver, b.x, b.y, b.z, b.m, theta=solver_theta(args, solver), leaf=args.leaf, bits=args.bits,
                    eps2=args.eps2)
    t_build = time.perf_counter() - t0
    if solver == 5:
        # the implicit FMM has no per-target walk: time the traversals from a
        # sample of its load-balanced starting nodes (one per thread at a time,
        # as the reference's tbb::parallel_for_each), scaled by the bodies held
        t0 = time.perf_counter()
        covered = t.fmm_sample(args.cpu_fmm_starts, threads=threads)
        t_walk = (time.perf_counter() - t0) * (n / max(1, covered))
        step_s = t_build + t_walk
        del t
        return {
            "value": n / step_s,
            "unit": "bodies·steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"oracle implicit FMM on {threads} threads: one build over all {n} bodies + the dual "
                      f"traversals from {args.cpu_fmm_starts} of its starting nodes ({covered} bodies), scaled by "
                      f"N/{covered} (build {t_build:.2f} s, traversal {t_walk:.2f} s extrapolated)",
        }
    sample = min(n, args.cpu_sample)
    sel = np.linspace(0, n - 1, sample).astype(np.int64)
    t0 = time.perf_counter()
    t.accel(b.x[sel], b.y[sel], b.z[sel], threads=threads)
    t_walk = (time.perf_counter() - t0) * (n / sample)
    x = b.x.copy()
    t0 = time.perf_counter()
    x += b.vx * np.float32(args.dt)
    t_update = (time.perf_counter() - t0) * 6  # v and p, 3 components each
    del t
    step_s = t_build + t_walk + t_update
    return {
        "value": n / step_s,
        "unit": "bodies·steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle {SOLVER_NAMES[solver]}: one build over all {n} bodies + the walk for {sample} evenly "
                  f"spaced targets, scaled by N/sample (build {t_build:.2f} s, walk {t_walk:.2f} s extrapolated)",
    }


def measure(nbg, ctx, b, args, solver, mode=None):
    mode = mode or args.mode
    p = nbg.default_params()
    p.solver = solver
    p.theta, p.leaf_size, p.morton_bits = solver_theta(args, solver), args.leaf, args.bits
    p.eps2, p.dt = args.eps2, args.dt
    p.sum_mode = nbg.SUM_FAST if mode == "fast" else nbg.SUM_ORDERED
    p.targets_per_lane = getattr(args, "tpl", 1)
    s = nbg.System(ctx, p, args.n)
    s.set_state(b)
    a0 = s.update_accelerations()  # force pass at t = 0

    log(f"[bench] {SOLVER_NAMES[solver]}: warmup {args.warmup} steps")
    s.step(args.warmup)
    s.synchronize()
    s.set_state(b)  # timed run starts from the initial conditions
    s.enable_counts(False)  # production steps do not count interactions
    s.enable_timing(True)
    log(f"[bench] {SOLVER_NAMES[solver]}: timing {args.steps} steps")
    s.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    phases = s.phase_times()
    s.enable_timing(False)
    # interaction counts and schedule statistics from one extra, untimed force
    # pass on the final state of the timed run
    s.enable_counts(True)
    s.enable_walk_stats(True)
    s.update_accelerations()
    pp, npc, pn, nn = s.counts4()
    ws = s.walk_stats()
    s.enable_walk_stats(False)

    # ORDERED walks always count (counts come with the bit-exact sums); the
    # FAST walk's schedule statistics do not apply
    if mode != "fast":
        ws = {"node_visits": 0, "active_lanes": 0, "leaf_visits": 0, "leaf_lanes": 0}
    ms_per_step = 1e3 * elapsed / args.steps
    inter = pp + npc + pn + nn
    walk_ms = phases["walk"]
    quad = solver in (3, 4)
    fmm = solver == 5
    flops = (FLOPS_PER_INTERACTION * pp + (FLOPS_PER_QUAD_NODE if quad else FLOPS_PER_INTERACTION) * npc
             + FLOPS_PER_LOCAL * (pn + nn))
    achieved = flops / (walk_ms * 1e-3) / 1e12
    waves = max(1, (args.n + 63) // 64)
    # the traffic file keys the ORDERED walks' rows "<solver>-ordered"
    traffic = load_traffic().get(SOLVER_NAMES[solver] + ("" if mode == "fast" else "-ordered"))
    out = {
        "value": args.n * args.steps / elapsed,
        "ms_per_step": ms_per_step,
        "sum_mode": mode,
        "theta": p.theta,
        "interactions_per_step": inter,
        "interactions_per_s": inter / (ms_per_step * 1e-3),
        "phases_ms": phases,
        "interaction_counts": {"pp": pp, "np": npc, "pn": pn, "nn": nn},
        "walk_schedule": {
            "node_visits_per_wave": ws["node_visits"] / waves,
            "lane_utilisation_nodes": ws["active_lanes"] / max(1, 64 * ws["node_visits"]),
            "leaf_visits_per_wave": ws["leaf_visits"] / waves,
            "lane_utilisation_leaves": ws["leaf_lanes"] / max(1, 64 * ws["leaf_visits"]),
        } if not fmm and ws["node_visits"] else (ws if fmm else None),
        "roofline": {
            "bound": "valu",
            "kernel": WALK_KERNELS[solver] if mode == "fast" else ORDERED_WALK_KERNELS[solver],
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "note": (f"{FLOPS_PER_INTERACTION} flops per body interaction and "
                     f"{FLOPS_PER_QUAD_NODE if quad else FLOPS_PER_INTERACTION} per node interaction"
                     + (f", {FLOPS_PER_LOCAL} per body or node into a local field" if fmm else "")
                     + " (SimpleTrackingRule counting) / "
                     + ("the traversal's device time (all its launches) " if fmm else "mean walk-kernel time ")
                     + "from HIP events on the system stream; traffic = HBM bytes per "
                     + ("traversal" if fmm else "walk launch")
                     + " from rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated"),
        },
    }
    if not args.no_accuracy:
        log(f"[bench] {SOLVER_NAMES[solver]}: accuracy (FP64 direct sum on {len(args.sub)} bodies, energy drift)")
        st = s.get_state()
        end = nbg.Bodies(st.x, st.y, st.z, b.m, st.vx, st.vy, st.vz)
        s2 = nbg.System(ctx, p, args.n)
        s2.set_state(end)
        aN = s2.update_accelerations()
        out["accuracy"] = error_stats(a0[args.sub], args.direct, b.m[args.sub])
        if getattr(args, "W0", None) is not None:
            log(f"[bench] {SOLVER_NAMES[solver]}: FP64 potential of the final state")
            WN = s2.potential_f64()
            E0 = energy(b.m, b.vx, b.vy, b.vz, a0, args.dt, args.W0)
            EN = energy(b.m, end.vx, end.vy, end.vz, aN, args.dt, WN)
            out["accuracy"].update({"energy_drift": float(abs(EN - E0) / abs(E0)), "E0": float(E0),
                                    "E_end": float(EN), "drift_steps": args.steps})
        else:
            out["accuracy"]["energy_drift"] = None
        s2.close()
        if mode == "fast" and solver != 5:
            # the same pass in the reference's FP32 arithmetic (ORDERED, the
            # oracle's bits) over every body
            po = type(p).from_buffer_copy(p)
            po.sum_mode = nbg.SUM_ORDERED
            so = nbg.System(ctx, po, args.n)
            so.set_state(b)
            out["accuracy"]["vs_reference_fp32"] = tolerance_stats(a0, so.update_accelerations())
            so.close()
    s.close()
    return out


def splitmix64(x):
    """SplitMix64 output function (Steele, Lea, Flood 2014) on a 64-bit state."""
    m = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def ensemble(nbg, nsys, nmin, nmax, seed):
    """Batched-ensemble inputs: N_i = floor(exp U[ln nmin, ln nmax]), scale
    radius a_i ~ U[0.5, 2], M = 1, per-system seed SplitMix64(seed + i)."""
    rng = np.random.default_rng(seed)
    ns = np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), nsys))).astype(np.int64)
    aa = rng.uniform(0.5, 2.0, nsys)
    return [(nbg.plummer(int(n), seed=splitmix64(seed + i), a=float(a)), float(a)) for i, (n, a) in
            enumerate(zip(ns, aa))]


def bench_forest(nbg, ctx, args):
    """Batched ensemble: many independent Plummer systems stepped together
    (one build + one walk launch per step, replayed from a HIP graph), leaf
    size picked by the autotuner, against the same systems stepped one at a
    time through nbg_sys (a sample, scaled by bodies)."""
    log(f"[bench] forest: generating {args.forest_systems} systems")
    systems = ensemble(nbg, args.forest_systems, 256, 65536, args.forest_seed)
    total = sum(b.n for b, _ in systems)
    p = nbg.default_params()
    solver = args.forest_solver
    quad = solver.endswith("quadrupole")
    octree = solver.startswith("barnes-hut")
    p.solver = {"linear-bvh": nbg.LINEAR_BVH, "linear-bvh-quadrupole": nbg.LINEAR_BVH_QUADRUPOLE,
                "barnes-hut": nbg.BARNES_HUT, "barnes-hut-quadrupole": nbg.BARNES_HUT_QUADRUPOLE}[solver]
    p.morton_bits, p.theta, p.leaf_size, p.max_depth = 30, 0.5, 16, 32
    p.eps2, p.dt, p.sum_mode = 1e-4, 1e-3, nbg.SUM_FAST
    B = nbg.Batch(ctx, total, len(systems), sum_mode=nbg.SUM_FAST)
    ids = [B.add(p, b) for b, _ in systems]
    B.enable_counts(False)
    log(f"[bench] forest ({solver}): {len(systems)} systems, {total} bodies; autotuning leaf size x bodies per "
        f"wavefront")
    best, best_tpl, tune = B.autotune_grid((8, 16, 32, 64), (1, 2, 4), repeats=2)
    tune_ms = {f"leaf {l}, {64 * t} bodies/wavefront": v for (l, t), v in tune.items()}
    p.leaf_size = best
    p.targets_per_lane = best_tpl
    # accuracy sample: each sampled system's first pass and energy, alone
    # (the forest gives each system the bits it gets alone)
    acc_sample = np.random.default_rng(3).choice(len(systems), min(args.forest_accuracy_sample, len(systems)),
                                                 replace=False)
    start = {}
    for k in acc_sample:
        b, _ = systems[int(k)]
        s = nbg.System(ctx, p, b.n)
        s.set_state(b)
        a0 = s.update_accelerations()
        start[int(k)] = (a0, s.potential_f64())
        s.close()
    B.use_graph(True)
    B.step(1)            # capture + warm-up
    B.synchronize()
    t0 = time.perf_counter()
    B.step(args.forest_steps)
    B.synchronize()
    el = time.perf_counter() - t0
    B.enable_counts(True)
    B.update_accelerations()
    pp, npc = B.counts()
    errs, drifts = [], []
    for k in acc_sample:
        b, _ = systems[int(k)]
        a0, W0 = start[int(k)]
        d = ctx.direct_f64(b, b.x, b.y, b.z, 1.0, p.eps2)
        errs.append(error_stats(a0, d, b.m))
        end, aN = B.get(ids[int(k)], accelerations=True)
        s = nbg.System(ctx, p, b.n)
        s.set_state(nbg.Bodies(end.x, end.y, end.z, b.m, end.vx, end.vy, end.vz))
        WN = s.potential_f64()
        s.close()
        E0 = energy(b.m, b.vx, b.vy, b.vz, a0, p.dt, W0)
        EN = energy(b.m, end.vx, end.vy, end.vz, aN, p.dt, WN)
        drifts.append(abs(EN - E0) / abs(E0))
    B.close()
    value = total * args.forest_steps / el
    # one system at a time through nbg_sys, on a sample of the same systems
    sample = np.random.default_rng(1).choice(len(systems), min(args.forest_alone_sample, len(systems)),
                                             replace=False)
    t_alone, bodies_alone = 0.0, 0
    for k in sample:
        b, _ = systems[int(k)]
        s = nbg.System(ctx, p, b.n)
        s.set_state(b)
        s.enable_counts(False)
        s.step(1)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(args.forest_steps)
        s.synchronize()
        t_alone += time.perf_counter() - t0
        bodies_alone += b.n
        s.close()
    alone = bodies_alone * args.forest_steps / t_alone
    return {
        "workload": f"{len(systems)} Plummer systems, N log-uniform in [256, 65536] (total {total}), a ~ U[0.5, 2], "
                    f"M = 1, eps2 = 1e-4, theta = 0.5, dt = 1e-3, {args.forest_steps} steps, "
                    f"{'quadrupole ' if quad else ''}{'Barnes-Hut octree' if octree else 'linear BVH 30-bit keys'}, "
                    f"FAST sums, seed {args.forest_seed}",
        "value": value,
        "unit": "bodies·steps/s",
        "ms_per_step": 1e3 * el / args.forest_steps,
        "systems": len(systems),
        "bodies": total,
        "leaf_size_autotuned": best,
        "bodies_per_wavefront_autotuned": 64 * best_tpl,
        "autotune_ms_per_pass": tune_ms,
        "interactions_per_step": pp + npc,
        "interactions_per_s": (pp + npc) * args.forest_steps / el,
        # the per-interaction flop count is the monopole kernel's
        "roofline_frac_of_fp32_peak": None if quad else FLOPS_PER_INTERACTION * (pp + npc) * args.forest_steps
                                      / el / 1e12 / FP32_PEAK_TFLOPS,
        "one_at_a_time": {"value": alone, "unit": "bodies·steps/s",
                          "sample": f"{len(sample)} of the systems stepped alone through nbg_sys "
                                    f"(same parameters and leaf size), bodies-weighted"},
        "batched_over_one_at_a_time": value / alone,
        "accuracy": {
            "sample": f"{len(acc_sample)} systems: first force pass against the FP64 direct sum over all their bodies, "
                      f"energy drift over the warm-up step and the {args.forest_steps} timed steps",
            "force_err_vs_fp64_median": float(np.median([e["force_err_vs_fp64_median"] for e in errs])),
            "force_err_vs_fp64_p99": float(np.max([e["force_err_vs_fp64_p99"] for e in errs])),
            "force_err_vs_fp64_max": float(np.max([e["force_err_vs_fp64_max"] for e in errs])),
            "constitutional_error_pct_max": float(np.max([e["constitutional_error_pct"] for e in errs])),
            "energy_drift_median": float(np.median(drifts)),
            "energy_drift_max": float(np.max(drifts)),
        },
    }


def two_spheres(nbg, n_each):
    """Two Plummer spheres of n_each bodies (a = 1, M = 0.5 each, seeds 31 and
    32), centres +-(10, 1, 0), bulk velocities -+(0.3, 0, 0)."""
    a = nbg.plummer(n_each, seed=31, M=0.5)
    c = nbg.plummer(n_each, seed=32, M=0.5)
    f32 = np.float32
    return nbg.Bodies(np.concatenate([a.x + f32(10.0), c.x - f32(10.0)]),
                      np.concatenate([a.y + f32(1.0), c.y - f32(1.0)]), np.concatenate([a.z, c.z]),
                      np.concatenate([a.m, c.m]), np.concatenate([a.vx - f32(0.3), c.vx + f32(0.3)]),
                      np.concatenate([a.vy, c.vy]), np.concatenate([a.vz, c.vz]))


def bench_two_spheres(nbg, ctx, args, threads):
    """The colliding-spheres configuration (2 x 8,388,608 bodies, eps2 4e-6,
    theta 0.5, leaf 16, 63-bit keys, dt 1e-3, 50 steps) for the linear BVH
    and the octree, FAST sums: throughput, the force error against the FP64
    direct sum on 16,384 bodies, the error against the reference's FP32
    arithmetic over every body, the energy drift over the 50 steps (FP64
    potential, each pair once, in launches of about 1e12 pairs), and the
    oracle's CPU time for one step (build + a walk sample)."""
    log("[bench] two spheres: generating 2 x 8388608 bodies")
    b = two_spheres(nbg, 8388608)
    a2 = copy.copy(args)
    a2.n, a2.eps2, a2.dt, a2.bits, a2.theta, a2.leaf = b.n, 4e-6, 1e-3, 63, 0.5, 16
    a2.steps, a2.warmup, a2.no_accuracy = args.two_spheres_steps, 1, False
    a2.sub = np.random.default_rng(0).choice(b.n, 16384, replace=False)
    log("[bench] two spheres: FP64 direct sum on 16384 bodies")
    a2.direct = ctx.direct_f64(b, b.x[a2.sub], b.y[a2.sub], b.z[a2.sub], 1.0, a2.eps2)
    t_pot = None
    a2.W0 = None
    if args.two_spheres_drift:
        # the exact FP64 potential is O(N^2): 1.4e14 pairs, minutes per evaluation
        log("[bench] two spheres: FP64 potential of the initial state")
        p0 = nbg.default_params()
        p0.eps2 = a2.eps2
        s0 = nbg.System(ctx, p0, b.n)
        s0.set_state(b)
        t0 = time.perf_counter()
        a2.W0 = s0.potential_f64()
        t_pot = time.perf_counter() - t0
        s0.close()
        log(f"[bench] two spheres: potential took {t_pot:.1f} s")
    out = {"workload": f"two Plummer spheres of 8388608 bodies (a = 1, M = 0.5 each, seeds 31 and 32), centres "
                       f"+-(10, 1, 0), bulk velocities -+(0.3, 0, 0); eps2 4e-6, theta 0.5, leaf 16, 63-bit keys, "
                       f"dt 1e-3, {a2.steps} steps, FAST sums",
           "n_bodies": b.n, "potential_eval_s": t_pot,
           "energy_drift_note": None if args.two_spheres_drift else
           "not evaluated in the default run: the exact FP64 potential of 16.7M bodies (1.4e14 pairs) takes "
           "minutes per evaluation; bench.py --two-spheres --two-spheres-drift measures it"}
    for sv in (1, 0):
        m = measure(nbg, ctx, b, a2, sv)
        m["roofline"]["traffic"] = None  # the profiled walk traffic is the 1M workload's
        if not args.no_cpu:
            a3 = copy.copy(a2)
            a3.cpu_sample = args.two_spheres_cpu_sample
            log(f"[bench] two spheres: CPU baseline ({SOLVER_NAMES[sv]}) with {threads} threads")
            m["cpu_baseline"] = cpu_baseline(b, a3, sv, threads)
            m["gpu_over_cpu"] = m["value"] / m["cpu_baseline"]["value"]
        out[SOLVER_NAMES[sv].replace("-", "_")] = m
    return out


def bench_c1(nbg, ctx, args, threads):
    """The 10,000-body configuration: Plummer N = 10,000 (seed 0), eps2 1e-4,
    theta 0.5, leaf 16, 63-bit keys (the reference's default key width,
    MortonSort.h:25), dt 1e-3, 100 steps, both trees and both sum modes,
    timed over the whole run and reported per step as the reference's
    timedRun does (src/Benchmark/benchmark.h:29-39), beside the oracle's
    100 steps on this host's cores (the reference's CPU path), with the force
    error of the first pass against the FP64 direct sum over every body and
    the energy drift of each run."""
    import oracle

    n, steps, eps2, dt, theta, leaf, bits = 10000, 100, 1e-4, 1e-3, 0.5, 16, 63
    b = nbg.plummer(n, seed=0)
    direct = ctx.direct_f64(b, b.x, b.y, b.z, 1.0, eps2)

    def potential(x, y, z):
        p0 = nbg.default_params()
        p0.eps2 = eps2
        s0 = nbg.System(ctx, p0, n)
        s0.set_state(nbg.Bodies(x, y, z, b.m, b.vx, b.vy, b.vz))
        w = s0.potential_f64()
        s0.close()
        return w

    W0 = potential(b.x, b.y, b.z)
    out = {"workload": f"Plummer N = {n} (a = 1, M = 1, G = 1, r < 10, seed 0), eps2 {eps2}, theta {theta}, "
                       f"leaf {leaf}, {bits}-bit keys, dt {dt}, {steps} steps"}
    for sv in (1, 0):
        name = SOLVER_NAMES[sv]
        log(f"[bench] 10k bodies: {name}")
        blk = {}
        for mode in ("fast", "ordered"):
            p = nbg.default_params()
            p.solver, p.theta, p.leaf_size, p.morton_bits, p.eps2, p.dt = sv, theta, leaf, bits, eps2, dt
            p.sum_mode = nbg.SUM_FAST if mode == "fast" else nbg.SUM_ORDERED
            s = nbg.System(ctx, p, n)
            s.set_state(b)
            a0 = s.update_accelerations()
            s.enable_counts(False)
            s.step(5)
            s.synchronize()
            s.set_state(b)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(steps)
            s.synchronize()
            el = time.perf_counter() - t0
            aN = s.update_accelerations()
            st = s.get_state()
            s.close()
            E0 = energy(b.m, b.vx, b.vy, b.vz, a0, dt, W0)
            EN = energy(b.m, st.vx, st.vy, st.vz, aN, dt, potential(st.x, st.y, st.z))
            acc = error_stats(a0, direct, b.m)
            acc["energy_drift"] = float(abs(EN - E0) / abs(E0))
            blk[mode] = {"value": n * steps / el, "unit": "bodies·steps/s", "ms_per_step": 1e3 * el / steps,
                         "accuracy": acc}
        # the oracle: all 100 steps, timed, with its own force error and drift
        bodies = (b.x, b.y, b.z, b.m, b.vx, b.vy, b.vz)
        t0 = time.perf_counter()
        (x, y, z, vx, vy, vz), _ = oracle.step(sv, bodies, steps, theta=theta, leaf=leaf, bits=bits, eps2=eps2,
                                              dt=dt, threads=threads)
        el = time.perf_counter() - t0
        a0c = oracle.Tree(sv, b.x, b.y, b.z, b.m, theta=theta, leaf=leaf, bits=bits, eps2=eps2).accel(b.x, b.y, b.z)
        aNc = oracle.Tree(sv, x, y, z, b.m, theta=theta, leaf=leaf, bits=bits, eps2=eps2).accel(x, y, z)
        E0 = energy(b.m, b.vx, b.vy, b.vz, a0c, dt, W0)
        EN = energy(b.m, vx, vy, vz, aNc, dt, potential(x, y, z))
        cacc = error_stats(a0c, direct, b.m)
        cacc["energy_drift"] = float(abs(EN - E0) / abs(E0))
        blk["cpu_baseline"] = {"value": n * steps / el, "unit": "bodies·steps/s", "cores": threads, "kind": "port",
                               "sample": f"the oracle's {name} run, all {steps} steps", "accuracy": cacc}
        blk["gpu_over_cpu"] = blk["fast"]["value"] / blk["cpu_baseline"]["value"]
        blk["ordered_gpu_over_cpu"] = blk["ordered"]["value"] / blk["cpu_baseline"]["value"]
        out[name.replace("-", "_")] = blk
    return out


def bench_serving(nbg, ctx, args):
    """Serving: a Poisson stream of jobs (Plummer sphere or uniform cube, N
    log-uniform in [1024, 2097152], steps uniform in [10, 500], theta in
    {0.3, 0.5, 0.7}, eps2 = 1e-4) submitted at their arrival times to a
    continuous-batching server with a 200 GB HBM budget.  The arrival rate
    offers `serve_load` x the batched capacity measured on a sample of the
    same jobs.  Reports latency percentiles, aggregate throughput, and the
    same jobs run one at a time through nbg_sys (a sample, steps capped,
    scaled)."""
    rng = np.random.default_rng(args.serve_seed)
    specs = []
    for k in range(args.serve_jobs):
        n = int(np.floor(np.exp(rng.uniform(np.log(1024), np.log(2097152)))))
        kind = "plummer" if rng.random() < 0.5 else "cube"
        steps = int(rng.integers(10, 501))
        theta = float(rng.choice([0.3, 0.5, 0.7]))
        specs.append((kind, n, steps, theta, splitmix64(args.serve_seed * 1000003 + k)))
    work = np.array([n * st for _, n, st, _, _ in specs], np.float64)
    bodies = [nbg.plummer(n, seed=sd) if kind == "plummer" else nbg.cube(n, seed=sd, side=2.0)
              for kind, n, _, _, sd in specs]

    def params(theta):
        p = nbg.default_params()
        p.solver, p.morton_bits, p.theta, p.leaf_size = nbg.LINEAR_BVH, 30, theta, 16
        p.eps2, p.dt, p.sum_mode = 1e-4, 1e-3, nbg.SUM_FAST
        return p

    # one job at a time through nbg_sys, and the same sample batched, to
    # calibrate the offered load (at most 10 timed steps each)
    sample = np.random.default_rng(2).choice(len(specs), min(args.serve_alone_sample, len(specs)), replace=False)
    t_tot, w_tot = 0.0, 0.0
    for k in sample:
        kind, n, steps, theta, _ = specs[int(k)]
        s = nbg.System(ctx, params(theta), n)
        s.set_state(bodies[int(k)])
        s.enable_counts(False)
        s.step(1)
        s.synchronize()
        k_steps = min(steps, 10)
        t1 = time.perf_counter()
        s.step(k_steps)
        s.synchronize()
        t_tot += (time.perf_counter() - t1) * steps / k_steps
        w_tot += n * steps
        s.close()
    alone = w_tot / t_tot
    nb = sum(specs[int(k)][1] for k in sample)
    B = nbg.Batch(ctx, nb, len(sample))
    for k in sample:
        B.add(params(specs[int(k)][3]), bodies[int(k)])
    B.enable_counts(False)
    B.step(1)
    B.synchronize()
    t1 = time.perf_counter()
    B.step(5)
    B.synchronize()
    capacity = nb * 5 / (time.perf_counter() - t1)
    B.close()
    rate = args.serve_load * capacity / work.mean()
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(specs)))
    log(f"[bench] serving: {len(specs)} jobs, {work.sum():.3g} body-steps, batched capacity {capacity:.3g}/s, "
        f"arrival rate {rate:.3g}/s (~{arrivals[-1]:.0f} s of arrivals)")

    # jobs of at most serve_express bodies run in an express forest (8 GB of
    # the budget) that steps concurrently with the bulk forest
    xmax = args.serve_express
    srv = nbg.Server(ctx, 200 * 10 ** 9, max_systems=4096, sum_mode=nbg.SUM_FAST, express_max_bodies=xmax,
                     express_bytes=8 * 10 ** 9 if xmax > 0 else 0, express_systems=4096 if xmax > 0 else 0)
    cap_bodies, _ = srv.capacity()
    ids = []
    t0 = time.perf_counter()
    for k, ((kind, n, steps, theta, _), b, t_arr) in enumerate(zip(specs, bodies, arrivals)):
        now = time.perf_counter() - t0
        if t_arr > now:
            time.sleep(t_arr - now)
        ids.append(srv.submit(params(theta), b, steps))
        if (k + 1) % 10 == 0:
            st = srv.stats()
            log(f"[bench] serving: t={time.perf_counter() - t0:.1f}s submitted {k + 1}, done {int(st['jobs_done'])}, "
                f"waiting {int(st['waiting'])}")
    lat, que = [], []
    for k, j in enumerate(ids):
        while True:
            try:
                _, t = srv.wait(j, timeout=30.0)
                break
            except nbg.NbgError as e:
                if "not finished" not in str(e):
                    raise
                st = srv.stats()
                log(f"[bench] serving: t={time.perf_counter() - t0:.1f}s waiting for job {k}; "
                    f"done {int(st['jobs_done'])}, steps {int(st['steps_run'])}")
        lat.append(t["latency"])
        que.append(t["queued"])
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    served = work.sum() / wall
    lat = np.array(lat)
    # latency by job size: small jobs should not wait behind large ones
    sizes = np.array([n for _, n, _, _, _ in specs])
    classes = {}
    for name, lo, hi in (("N < 16k", 0, 16384), ("16k <= N < 256k", 16384, 262144), ("N >= 256k", 262144, 1 << 40)):
        sel = (sizes >= lo) & (sizes < hi)
        if sel.any():
            classes[name] = {"jobs": int(sel.sum()), "latency_p50_s": float(np.quantile(lat[sel], 0.5)),
                             "latency_max_s": float(lat[sel].max()),
                             "latency_per_Mbody_step_p50_s": float(np.quantile(lat[sel] / (work[sel] / 1e6), 0.5))}
    # force error of the served jobs: a served job's bits are its bits alone
    # (tests/test_gpu_serve.py), so the first force pass of a sample of the
    # jobs, alone, against the FP64 direct sum on up to 4,096 of its bodies
    err_sample = [int(k) for k in np.argsort(sizes)[::max(1, len(sizes) // 8)][:8]]
    errs = []
    for k in err_sample:
        kind, n, steps, theta, _ = specs[k]
        b = bodies[k]
        s = nbg.System(ctx, params(theta), n)
        s.set_state(b)
        a0 = s.update_accelerations()
        s.close()
        sub = np.random.default_rng(k).choice(n, min(n, 4096), replace=False)
        d = ctx.direct_f64(b, b.x[sub], b.y[sub], b.z[sub], 1.0, 1e-4)
        e = error_stats(a0[sub], d, b.m[sub])
        errs.append({"job": k, "kind": kind, "n": n, "theta": theta, "force_err_vs_fp64_median": e[
            "force_err_vs_fp64_median"], "force_err_vs_fp64_p99": e["force_err_vs_fp64_p99"]})
    return {
        "workload": f"{len(specs)} jobs (Plummer or cube side 2, N log-uniform [1024, 2097152], steps U[10, 500], "
                    f"theta in {{0.3, 0.5, 0.7}}, eps2 1e-4, dt 1e-3, linear BVH 30-bit, FAST), Poisson arrivals at "
                    f"{rate:.3g}/s = {args.serve_load} x the batched capacity {capacity:.3g} bodies*steps/s measured "
                    f"on {len(sample)} of the jobs, 200 GB HBM budget ({cap_bodies} bodies), "
                    + (f"express lane for jobs of <= {xmax} bodies (8 GB of the budget), " if xmax > 0 else "")
                    + f"seed {args.serve_seed}",
        "value": served,
        "unit": "bodies·steps/s",
        "latency_p50_s": float(np.quantile(lat, 0.5)),
        "latency_p99_s": float(np.quantile(lat, 0.99)),
        "queue_p50_s": float(np.quantile(que, 0.5)),
        "queue_p99_s": float(np.quantile(que, 0.99)),
        "wall_s": wall,
        "device_busy_s": st["busy_s"],
        "forest_steps": int(st["steps_run"]),
        "peak_bodies_resident": int(st["peak_bodies"]),
        "peak_systems_resident": int(st["peak_systems"]),
        "device_busy_throughput": work.sum() / st["busy_s"],
        "one_at_a_time": {"value": alone, "unit": "bodies·steps/s",
                          "sample": f"{len(sample)} of the jobs run alone through nbg_sys, at most 10 timed steps "
                                    f"each, scaled to their step counts"},
        "batched_same_sample": {"value": capacity, "unit": "bodies·steps/s",
                                "sample": f"the same {len(sample)} jobs stepped together in one forest (5 steps)"},
        "batched_over_one_at_a_time": capacity / alone,
        "served_over_one_at_a_time": served / alone,
        "latency_by_size": classes,
        "force_error_by_job": errs,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--leaf", type=int, default=16)
    ap.add_argument("--bits", type=int, default=30)
    ap.add_argument("--eps2", type=float, default=2.5e-5)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--mode", choices=["fast", "ordered"], default="fast")
    ap.add_argument("--tpl", type=int, default=1, choices=(1, 2, 4),
                    help="targets per lane of the FAST linear-BVH walk (64 x this many bodies per wavefront)")
    ap.add_argument("--solvers", default="linear-bvh,barnes-hut,linear-bvh-4p,barnes-hut-4p,implicit-fmm",
                    help="comma list; the first is the headline line")
    ap.add_argument("--theta-4p-lbvh", type=float, default=0.72578126,
                    help="quadrupole LBVH opening angle (the reference benchmark's tuned value)")
    ap.add_argument("--theta-4p-bh", type=float, default=0.41289064,
                    help="quadrupole octree opening angle (the reference benchmark's tuned value)")
    ap.add_argument("--theta-fmm", type=float, default=0.4,
                    help="implicit FMM opening angle (its descent criterion's default)")
    ap.add_argument("--theta-secondary", type=float, default=0.3,
                    help="second opening angle measured for the monopole trees (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--cpu-fmm-starts", type=int, default=64,
                    help="starting nodes whose traversals time the oracle's implicit FMM")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-forest", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--forest-systems", type=int, default=4096)
    ap.add_argument("--forest-steps", type=int, default=20)
    ap.add_argument("--forest-seed", type=int, default=7)
    ap.add_argument("--forest-solver", choices=("linear-bvh", "linear-bvh-quadrupole", "barnes-hut",
                                                "barnes-hut-quadrupole"), default="linear-bvh",
                    help="solver of every system in the batched ensemble")
    ap.add_argument("--forest-alone-sample", type=int, default=64)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--two-spheres", action="store_true",
                    help="only measure the 16.7M-body colliding-spheres configuration (one JSON line)")
    ap.add_argument("--no-two-spheres", action="store_true",
                    help="skip the colliding-spheres block of the default run")
    ap.add_argument("--two-spheres-steps", type=int, default=50)
    ap.add_argument("--two-spheres-drift", action="store_true",
                    help="also the energy drift of the colliding spheres (two exact FP64 potentials, minutes each)")
    ap.add_argument("--two-spheres-cpu-sample", type=int, default=32768)
    ap.add_argument("--no-c1", action="store_true", help="skip the 10,000-body block")
    ap.add_argument("--no-ordered", action="store_true", help="skip the ORDERED (bit-exact) blocks")
    ap.add_argument("--ordered-steps", type=int, default=5)
    ap.add_argument("--forest-accuracy-sample", type=int, default=32)
    ap.add_argument("--serve-jobs", type=int, default=60)
    ap.add_argument("--serve-load", type=float, default=0.8)
    ap.add_argument("--serve-express", type=int, default=131072,
                    help="largest job (bodies) served by the express forest; 0: one forest for every job")
    ap.add_argument("--serve-seed", type=int, default=11)
    ap.add_argument("--serve-alone-sample", type=int, default=16)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device tensor copy three times (the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (systems shard by body "
              "or by system with no communication; see DESIGN.md).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    ids = {v: k for k, v in SOLVER_NAMES.items()}
    solvers = [ids[x.strip()] for x in args.solvers.split(",") if x.strip()]

    nbg = importlib.import_module("n-body_amd")
    if nbg.device_count() == 0:
        raise SystemExit("no GPU visible")
    ctx = nbg.Context(0)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    host = host_info()
    if args.two_spheres:
        print(json.dumps(bench_two_spheres(nbg, ctx, args, threads)), flush=True)
        ctx.close()
        return
    log(f"[bench] generating Plummer N={args.n} seed={args.seed}")
    b = nbg.plummer(args.n, seed=args.seed)
    if not args.no_accuracy:
        log("[bench] FP64 references: direct sum on a 16384-body sample, initial potential")
        args.sub = np.random.default_rng(0).choice(args.n, min(16384, args.n), replace=False)
        args.direct = ctx.direct_f64(b, b.x[args.sub], b.y[args.sub], b.z[args.sub], 1.0, args.eps2)
        p0 = nbg.default_params()
        p0.eps2 = args.eps2
        s0 = nbg.System(ctx, p0, args.n)
        s0.set_state(b)
        args.W0 = s0.potential_f64()
        s0.close()

    res = {sv: measure(nbg, ctx, b, args, sv) for sv in solvers}
    # the 1M configuration's second opening angle, for the monopole trees
    extra = {}
    if args.theta_secondary > 0:
        for sv in (1, 0):
            if sv in solvers:
                a2 = copy.copy(args)
                a2.theta = args.theta_secondary
                log(f"[bench] {SOLVER_NAMES[sv]} at theta {a2.theta:g}")
                m2 = measure(nbg, ctx, b, a2, sv)
                m2["roofline"]["traffic"] = None  # the profiled walk traffic is theta 0.5's
                extra[f"{SOLVER_NAMES[sv].replace('-', '_')}_theta{a2.theta:g}"] = m2
    # the bit-exact path (the reference's FP32 arithmetic) at the same configuration
    ordered = {}
    if not args.no_ordered:
        for sv in (1, 0):
            if sv in solvers:
                a2 = copy.copy(args)
                a2.steps, a2.warmup = args.ordered_steps, 1
                log(f"[bench] {SOLVER_NAMES[sv]} ORDERED")
                m2 = measure(nbg, ctx, b, a2, sv, mode="ordered")  # traffic: the "-ordered" rows
                ordered[sv] = m2
                extra[f"{SOLVER_NAMES[sv].replace('-', '_')}_ordered"] = m2
    if not args.no_cpu:
        for sv in solvers:
            log(f"[bench] CPU baseline ({SOLVER_NAMES[sv]}) with {threads} threads")
            cb = cpu_baseline(b, args, sv, threads)
            cb["host"] = host
            res[sv]["cpu_baseline"] = cb
            res[sv]["gpu_over_cpu"] = res[sv]["value"] / cb["value"]
            if sv in ordered:
                ordered[sv]["cpu_baseline"] = cb
                ordered[sv]["gpu_over_cpu"] = ordered[sv]["value"] / cb["value"]

    c1 = None
    if not args.no_c1:
        try:
            c1 = bench_c1(nbg, ctx, args, threads)
        except Exception as e:  # reported, not fatal for the headline line
            c1 = {"error": f"{type(e).__name__}: {e}"}

    c3 = None
    if not args.no_two_spheres:
        try:
            c3 = bench_two_spheres(nbg, ctx, args, threads)
        except Exception as e:
            c3 = {"error": f"{type(e).__name__}: {e}"}

    forest = forest_oct = None
    if not args.no_forest:
        try:
            forest = bench_forest(nbg, ctx, args)
        except Exception as e:  # reported, not fatal for the headline line
            forest = {"error": f"{type(e).__name__}: {e}"}
        if args.forest_solver == "linear-bvh":
            # the same ensemble as Barnes-Hut octree members
            a2 = copy.copy(args)
            a2.forest_solver = "barnes-hut"
            try:
                forest_oct = bench_forest(nbg, ctx, a2)
            except Exception as e:
                forest_oct = {"error": f"{type(e).__name__}: {e}"}

    serving = None
    if not args.no_serving:
        try:
            serving = bench_serving(nbg, ctx, args)
        except Exception as e:
            serving = {"error": f"{type(e).__name__}: {e}"}

    head = res[solvers[0]]
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "bodies·steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": f"plummer-{args.n} {SOLVER_NAMES[solvers[0]]} step "
                        f"(theta={solver_theta(args, solvers[0]):.8g}, leaf={args.leaf}, "
                        f"{args.bits}-bit keys, eps2={args.eps2}, dt={args.dt}, seed={args.seed})",
            "n_bodies": args.n, "theta": solver_theta(args, solvers[0]), "leaf_size": args.leaf,
            "morton_bits": args.bits,
            "eps2": args.eps2, "dt": args.dt, "sum_mode": args.mode, "solver": SOLVER_NAMES[solvers[0]],
        },
    }
    for k, v in head.items():
        if k not in ("value", "ms_per_step"):
            out[k] = v
    for sv in solvers[1:]:
        out[SOLVER_NAMES[sv].replace("-", "_")] = res[sv]
    out.update(extra)
    if c1 is not None:
        out["plummer_10k"] = c1
    if c3 is not None:
        out["two_spheres"] = c3
    if forest is not None:
        out["batched_ensemble"] = forest
    if forest_oct is not None:
        out["batched_ensemble_barnes_hut"] = forest_oct
    if serving is not None:
        out["serving"] = serving
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
