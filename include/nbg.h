/*
 * nbg.h — C ABI of the MI355X tree-code force pass (HIP, gfx950).
 *
 * This library is a drop-in for the reference simulator's CPU tree solvers
 * (JacksonCampolattaro/n-body):  BarnesHutSolver (octree) and LinearBVHSolver,
 * i.e. ActiveTreeSolver::updateAccelerations() plus the Solver::step()
 * integrator around it.  Every entry point names the reference interface it
 * replaces (paths relative to the reference's src/NBody/).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative nbg_status on error;
 *     no C++ exception crosses the ABI.  nbg_last_error() returns the message
 *     of the last failure on the calling thread.
 *   - Plain pointers and sizes only.  Host arrays are SoA float32, caller
 *     owned, and copied; device state is owned by the nbg_sys.
 *   - One HIP stream per system.  Calls on one nbg_sys must be serialised by
 *     the caller; distinct systems may be driven from different threads.
 *     Every call makes the system's device current itself (the reference
 *     drives Solver::step from a background std::thread, Solver.h:155-186).
 */
#ifndef NBG_H
#define NBG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nbg_status {
    NBG_OK = 0,
    NBG_ERR_ARG = -1,       /* bad argument (null pointer, size, parameter range) */
    NBG_ERR_HIP = -2,       /* HIP runtime error (message in nbg_last_error) */
    NBG_ERR_NO_DEVICE = -3, /* no GPU / no HIP runtime */
    NBG_ERR_DEVICE = -4,    /* a kernel reported a failure (e.g. tree deeper than supported) */
    NBG_ERR_BUDGET = -5,    /* memory budget exhausted (batched / serving forms) */
    NBG_ERR_STATE = -6      /* call out of order (e.g. counts before a force pass) */
};

/* Solver kind.  BARNES_HUT = BarnesHutSolver (Solvers/BarnesHutSolver.h:14-32,
 * id "barnes-hut"); LINEAR_BVH = LinearBVHSolver (Solvers/LinearBVHSolver.h:18-28,
 * id "linear-bvh"); NAIVE = NaiveSolver (Solvers/NaiveSolver.h:22-52, id "naive"):
 * direct summation over every body in body order, no tree (theta, leaf_size,
 * max_depth and morton_bits are ignored).  The QUADRUPOLE kinds are the
 * same trees whose nodes also carry the raw second mass moment
 * (MultipoleMassSummary<2>, Physics/Summaries/MultipoleMassSummary.h) and
 * interact through Gravity::nodeParticle<2> (Physics/Rules/Gravity.h:238-259):
 * QuadrupoleBarnesHutSolver (BarnesHutSolver.h:34-66, id "barnes-hut-4p")
 * and QuadrupoleLinearBVHSolver (LinearBVHSolver.h:30-53, id "linear-bvh-4p"). */
enum nbg_solver_kind {
    NBG_BARNES_HUT = 0,
    NBG_LINEAR_BVH = 1,
    NBG_NAIVE = 2,
    NBG_BARNES_HUT_QUADRUPOLE = 3,
    NBG_LINEAR_BVH_QUADRUPOLE = 4,
    /* ImplicitFMMSolver = MultipoleImplicitFMMSolver<1> (Solvers/ImplicitFMMSolver.h:14-47,
     * id "implicit-fmm-2p"): the lockstep dual traversal of ImplicitDualTraversalSolver
     * (Solvers/ImplicitDualTraversalSolver.h:17-70) over the Barnes-Hut octree, monopole
     * sources and local fields with their gradient; theta is its SideLengthOverDistance
     * (reference default 0.4).  FAST sum mode only.  Its passive points are its bodies:
 * massless bodies are not in the reference's dual tree (CenterOfMassSummary::context,
 * Physics/Summaries/CenterOfMassSummary.h:22-24) and keep the zero acceleration the
 * solver resets them to (ImplicitDualTraversalSolver.h:50-54), so
 * nbg_sys_accelerations_at refuses this kind. */
    NBG_IMPLICIT_FMM = 5
};

/* Per-body sum order.  ORDERED reproduces the reference's arithmetic exactly:
 * IEEE sqrt/div, no FMA, and the std::transform_reduce bracketing of child
 * results (Descent/activeTree.h:44-52,66-76).  FAST keeps the reference's
 * interaction set (the same, exact opening decisions) but sums each body's
 * interactions into one accumulator, in the walk's own order, with rsqrt
 * and FMA. */
enum nbg_sum_mode { NBG_SUM_ORDERED = 0, NBG_SUM_FAST = 1 };

typedef struct nbg_params {
    int32_t solver;      /* nbg_solver_kind */
    float theta;         /* opening angle; ActiveTreeSolver default 0.3 (ActiveTreeSolver.h:21) */
    int32_t leaf_size;   /* max leaf size; Octree/LinearBVH default 64 (Octree.h:168, LinearBVH.h:135) */
    int32_t max_depth;   /* octree refine depth below the 2 forced levels; default 32 (Octree.h:167) */
    int32_t morton_bits; /* LBVH key width, 30 or 63; default 63 (MortonSort.h:25) */
    float G;             /* Gravity G, default 1 (Physics/Rules/Gravity.h:195) */
    float eps2;          /* Gravity "epsilon", added to r^2; default 1e-5 (Gravity.h:195,207) */
    float dt;            /* Solver::timeStep(), default 1e-3 (Simulation/Solver.h:34,101) */
    int32_t sum_mode;    /* nbg_sum_mode */
    int32_t targets_per_lane; /* FAST walks: bodies each lane walks the tree for, 1, 2 or 4 (0 reads as 1);
                                 a wavefront serves 64 x this many bodies (no reference counterpart:
                                 a schedule choice that leaves every result's interaction set unchanged) */
} nbg_params;

typedef struct nbg_ctx nbg_ctx;
typedef struct nbg_sys nbg_sys;

/* Fills p with the reference's defaults (see the field comments). */
void nbg_params_default(nbg_params *p);

/* Message of the last failed call on this thread ("" if none). */
const char *nbg_last_error(void);

/* Library version string. */
const char *nbg_version(void);

/* Number of visible HIP devices (0 if none or no runtime). */
int nbg_device_count(void);

/* A context binds one GPU.  Replaces nothing in the reference (which is
 * CPU-only); it owns the device selection that Solver::maxThreadCount()
 * (Solver.h:97-99) expresses for TBB. */
int nbg_ctx_create(int device, nbg_ctx **out);
int nbg_ctx_destroy(nbg_ctx *ctx);

/* A system is one Simulation's bodies plus one solver's state
 * (ActiveTreeSolver(Simulation&, Rule&), Solvers/ActiveTreeSolver.h:25-27).
 * Capacity n is fixed at creation; device buffers are allocated here. */
int nbg_sys_create(nbg_ctx *ctx, const nbg_params *p, int64_t n, nbg_sys **out);
int nbg_sys_destroy(nbg_sys *sys);

/* theta() / maxLeafSize() / maxDepth() / Gravity / timeStep() setters
 * (ActiveTreeSolver.h:78-80, BarnesHutSolver.h:25-31, Gravity.h:195,
 * Solver.h:101).  The solver kind and key width may change too. */
int nbg_sys_set_params(nbg_sys *sys, const nbg_params *p);
int nbg_sys_get_params(const nbg_sys *sys, nbg_params *p);

/* Copies host SoA state in (the Simulation's Position, Mass and Velocity
 * components; Physics/Position.h:17-29, Mass.h:10-28, Velocity.h:16-24).
 * vx/vy/vz may be NULL (velocities set to 0). */
int nbg_sys_set_state(nbg_sys *sys, const float *x, const float *y, const float *z, const float *m,
                      const float *vx, const float *vy, const float *vz);

/* Copies device state out; any pointer may be NULL. */
int nbg_sys_get_state(nbg_sys *sys, float *x, float *y, float *z, float *vx, float *vy, float *vz);

/* ActiveTreeSolver::updateAccelerations() (Solvers/ActiveTreeSolver.h:33-72):
 * builds the tree from the current positions and writes every body's
 * acceleration.  ax/ay/az (host) may all be NULL to keep them on the device. */
int nbg_sys_update_accelerations(nbg_sys *sys, float *ax, float *ay, float *az);

/* Copies the accelerations of the last force pass out (Acceleration component,
 * Physics/Acceleration.h:16-42). */
int nbg_sys_get_accelerations(nbg_sys *sys, float *ax, float *ay, float *az);

/* Accelerations at nt passive points (bodies without Mass): the reference
 * builds its tree from the actives, view<const Position, const Mass>, and walks
 * every body with an Acceleration, view<const Position, Acceleration>
 * (Solvers/ActiveTreeSolver.h:33-72; Simulation.cpp:208-211 gives massless
 * scenario bodies no Mass).  Uses the tree of the system's current state
 * (building it first if the state or parameters changed since the last force
 * pass) and writes ax/ay/az in the targets' order.  Counts (nbg_sys_counts)
 * then describe this walk.  Host arrays; blocking. */
int nbg_sys_accelerations_at(nbg_sys *sys, int64_t nt, const float *tx, const float *ty, const float *tz,
                             float *ax, float *ay, float *az);

/* The same walk as nbg_sys_accelerations_at, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc); it has no counterpart in the reference.  It calls
 * no hipMalloc, hipFree or allocating library routine: its scratch is one
 * static device array per GPU (about 17 MB, held until the process ends),
 * so at most NBG_NOALLOC_MAX_POINTS targets per call (more: NBG_ERR_ARG).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several threads, on any systems of that device) race. */
#define NBG_NOALLOC_MAX_POINTS 262144
int nbg_sys_accelerations_at_noalloc(nbg_sys *sys, int64_t nt, const float *tx, const float *ty, const float *tz,
                                     float *ax, float *ay, float *az);

/* Solver::step() repeated nsteps times (Simulation/Solver.h:103-139):
 * updateAccelerations, then v = v + a*dt, then p = p + v*dt, on the device,
 * with no host copies.  Asynchronous: returns after enqueueing. */
int nbg_sys_step(nbg_sys *sys, int32_t nsteps);

/* Waits for the system's stream and reports any device-side failure. */
int nbg_sys_synchronize(nbg_sys *sys);

/* Interaction counts of the last force pass as SimpleTrackingRule counts
 * them (Physics/Rules/SimpleTrackingRule.h:29-45): pp = particle-particle
 * calls (self pairs included), np = node-particle calls.  64-bit.  The
 * reference counts with a separate rule; here counting is a switch of the
 * walk (on by default; FAST passes skip it while it is off, ORDERED passes
 * always count). */
int nbg_sys_counts(nbg_sys *sys, uint64_t *pp, uint64_t *np);
/* All four SimpleTrackingRule counters (Physics/Rules/SimpleTrackingRule.h:29-61):
 * out4 = {pp, np, pn, nn}; pn (particle-node: a body into a local field) and
 * nn (node-node: a node into a local field) are nonzero for the implicit FMM
 * only.  Empty octants, which the reference keeps as body-less nodes, are not
 * counted. */
int nbg_sys_counts4(nbg_sys *sys, uint64_t *out4);
int nbg_sys_enable_counts(nbg_sys *sys, int32_t on);

/* Schedule statistics of the last FAST force pass (no reference counterpart;
 * diagnostics of the wavefront walk): out4 = {wave-level node visits, active
 * lanes summed over those visits, wave-level leaf visits, lanes interacting
 * with a leaf summed over leaf visits}.  Counted only while enabled (it
 * costs scalar instructions in the walk); zero after an ORDERED pass.
 * For the implicit FMM (always available) out4 = {largest list-pool use of
 * one traversal step in words, largest frontier of a step in nodes,
 * traversal steps launched, list-pool capacity per step in words}. */
int nbg_sys_enable_walk_stats(nbg_sys *sys, int32_t on);
int nbg_sys_walk_stats(nbg_sys *sys, uint64_t *out4);

/* Waves per group of 64 targets in the tree walks (FAST: linear BVH with one
 * target per lane, octree; ORDERED: linear BVH; no reference counterpart: a
 * schedule choice).  0 (the default) chooses by size: systems of at most
 * 65,536 bodies fill the GPU with up to 16 waves per group, within 4,096
 * waves in all (FAST: linear BVH 4, octree 8), larger ones use 1.  Results are the same bits for every
 * choice: FAST sums each target's interactions per fixed block of subtrees and
 * adds the blocks in one order (the forest and the several-targets-per-lane
 * walks use the same blocks); ORDERED computes each frontier subtree's result
 * from zero, as the reference's recursion does, and adds it where the walk
 * meets it.  A split walk keeps 408 (FAST) or 1,536 (ORDERED) bytes per body
 * of sums in HBM. */
int nbg_sys_set_walk_split(nbg_sys *sys, int32_t waves);

/* Per-phase device time (ms) of the last force pass, measured with HIP events
 * on the system's stream when timing is enabled.  Phases, in order:
 * 0 bbox, 1 keys, 2 sort, 3 topology, 4 gather, 5 summaries, 6 walk,
 * 7 update.  Returns the number of phases written (<= cap). */
int nbg_sys_enable_timing(nbg_sys *sys, int32_t on);
int nbg_sys_phase_times(nbg_sys *sys, float *ms, int32_t cap);

/* ---- parity hooks: read back intermediate results of the last force pass ---- */

/* outerBoundingBox (Physics/Summaries/SummaryType.h:73-81): min xyz, max xyz. */
int nbg_sys_debug_bbox(nbg_sys *sys, float *out6);

/* LBVH: sorted Morton keys (setMortonCodes, Solvers/MortonSort.h:63-73; keys
 * widened to 64 bit) and the sort permutation (recursiveParallelMortonSort,
 * MortonSort.h:141-183), ties by original index.  Octree: keys_or_null is
 * ignored and perm is the body order of the leaves. */
int nbg_sys_debug_order(nbg_sys *sys, uint64_t *keys_or_null, uint32_t *perm);

/* Tree topology and summaries.  Call with arrays NULL to get the node count.
 * Per node: first body (sorted position), body count, leaf flag, parent
 * (-1 for the root), centre of mass + total mass (CenterOfMassSummary.h:14-63)
 * and geometry: LBVH bbox min/max (BoundingBoxSummary.h:14-74); octree
 * centre xyz, side, 0, 0 (Octree.h:73-74). */
int nbg_sys_debug_tree(nbg_sys *sys, int64_t *num_nodes, int64_t *first, int64_t *count, int32_t *leaf,
                       int32_t *parent, float *com4, float *geo6);

/* Quadrupole solvers: each node's raw second moment sum m (x - c)(x - c)^T
 * about its centre of mass as (xx, xy, xz, yy, yz, zz)
 * (MultipoleMassSummary::moment(), Summaries/MultipoleMassSummary.h:99-101),
 * in nbg_sys_debug_tree's node order; q6 holds cap nodes. */
int nbg_sys_debug_quads(nbg_sys *sys, int64_t cap, float *q6);

/* ---- scenario files (the reference's formats) ----
 *
 * Simulation::from_json (src/NBody/Simulation/Simulation.cpp:154-228) and
 * Simulation::from_tipsy (:273-395), read into host arrays in file order.
 * Per particle, flags: NBG_SCENARIO_POSITION / _VELOCITY (given in the file),
 * _ACTIVE (has a Mass: a source, goes to nbg_sys_set_state) and _PASSIVE (has
 * an Acceleration: receives forces; a passive particle without a mass goes
 * to nbg_sys_accelerations_at).  JSON: "passive": false drops the
 * Acceleration unless the file gives one (Simulation.cpp:208-211).  Tipsy:
 * the stars, or the gas when there are no stars, never the dark matter; all
 * active and passive; either byte order. */
#define NBG_SCENARIO_POSITION 1
#define NBG_SCENARIO_VELOCITY 2
#define NBG_SCENARIO_ACTIVE 4
#define NBG_SCENARIO_PASSIVE 8
typedef struct nbg_scenario nbg_scenario;
int nbg_scenario_load_json(const char *path, nbg_scenario **out);
int nbg_scenario_load_tipsy(const char *path, nbg_scenario **out);
int64_t nbg_scenario_count(const nbg_scenario *s);
/* Copies the particles out; any array may be NULL (n entries each). */
int nbg_scenario_get(const nbg_scenario *s, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *m,
                     uint8_t *flags);
int nbg_scenario_free(nbg_scenario *s);

/* ---- direct summation (NaiveSolver, Solvers/NaiveSolver.h:22-52) ---- */

/* FP64 direct sum on the device: accelerations at nt target points from ns
 * sources, the reference's softening (eps2 added to r^2) and self-skip by
 * exact position equality.  Host arrays in/out. */
int nbg_direct_f64(nbg_ctx *ctx, int64_t ns, const float *sx, const float *sy, const float *sz,
                   const float *sm, int64_t nt, const float *tx, const float *ty, const float *tz,
                   double G, double eps2, double *ax, double *ay, double *az);

/* FP64 softened potential energy -sum_{i<j} G m_i m_j / sqrt(r^2 + eps2)
 * of a system's current device state. */
int nbg_sys_potential_f64(nbg_sys *sys, double *W);

/* ---- batched forest: many independent systems stepped together ----
 *
 * No reference counterpart (the reference steps one Simulation).  Each system
 * is a linear-BVH Solver (Solvers/LinearBVHSolver.h:18-28), or the quadrupole
 * form of it (NBG_LINEAR_BVH_QUADRUPOLE; all members of a batch run the same
 * one), with its own theta, leaf size, G, eps2 and dt; every pass builds all
 * systems' trees in one segmented build and walks them in one launch.  A
 * system's results are the same bits it gets alone in an nbg_sys with the
 * same parameters (30-bit Morton keys, the batch's sum mode).  Systems can join and leave
 * between steps (continuous batching); state stays in HBM. */
typedef struct nbg_batch nbg_batch;

/* HBM bytes a batch of this capacity allocates (for admission budgets). */
int64_t nbg_batch_bytes(int64_t max_bodies, int32_t max_systems);

/* Capacity is fixed at creation (NBG_ERR_BUDGET if the device cannot hold it). */
int nbg_batch_create(nbg_ctx *ctx, int64_t max_bodies, int32_t max_systems, int32_t sum_mode, nbg_batch **out);
int nbg_batch_destroy(nbg_batch *b);

/* Adds a system of n >= 1 bodies (host SoA copied in; vx/vy/vz may be NULL).
 * p->solver must be NBG_LINEAR_BVH or NBG_LINEAR_BVH_QUADRUPOLE, the same as
 * the other members', and p->morton_bits 30.  *id receives the system's
 * handle.  NBG_ERR_BUDGET when the batch is full (the first quadrupole member
 * also allocates 64 bytes per body of capacity for the node moments). */
int nbg_batch_add(nbg_batch *b, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
                  const float *m, const float *vx, const float *vy, const float *vz, int32_t *id);

/* Removes a system; the others keep their state (moved on the device). */
int nbg_batch_remove(nbg_batch *b, int32_t id);

/* Changes one system's parameters (same constraints as nbg_batch_add). */
int nbg_batch_set_params(nbg_batch *b, int32_t id, const nbg_params *p);

/* Number of systems and bodies currently in the batch. */
int nbg_batch_info(nbg_batch *b, int32_t *nsys, int64_t *nbodies);

/* Copies one system's state out (any pointer may be NULL); accelerations are
 * those of the last pass and need a pass since the last membership change. */
int nbg_batch_get(nbg_batch *b, int32_t id, float *x, float *y, float *z, float *vx, float *vy, float *vz,
                  float *ax, float *ay, float *az);

/* Non-finite guard (no reference counterpart: the reference lets NaN
 * propagate): finds the systems whose positions or velocities hold a NaN or
 * an infinity.  Writes up to cap of their ids to ids and their number to
 * *count.  Removing them before the next step keeps their values out of the
 * shared build. */
int nbg_batch_nonfinite(nbg_batch *b, int32_t *ids, int32_t cap, int32_t *count);

/* Force pass for every system (ActiveTreeSolver::updateAccelerations each). */
int nbg_batch_update_accelerations(nbg_batch *b);

/* nsteps Solver::step()s of every system; asynchronous. */
int nbg_batch_step(nbg_batch *b, int32_t nsteps);
int nbg_batch_synchronize(nbg_batch *b);

/* Interaction totals of the last pass over all systems (SimpleTrackingRule
 * counting; a switch as for nbg_sys). */
int nbg_batch_counts(nbg_batch *b, uint64_t *pp, uint64_t *np);
int nbg_batch_enable_counts(nbg_batch *b, int32_t on);

/* Replays each step from a captured HIP graph (captured again after any
 * membership or parameter change). */
int nbg_batch_use_graph(nbg_batch *b, int32_t on);

/* Leaf-size autotuner: times `repeats` force passes of the current batch for
 * each candidate leaf size (state untouched), then sets every system's leaf
 * size to the fastest.  *best receives it; ms_out[k] (optional) the mean
 * pass time of candidate k. */
int nbg_batch_autotune(nbg_batch *b, const int32_t *leaf_sizes, int32_t k, int32_t repeats, int32_t *best,
                       float *ms_out);

/* The autotuner over leaf size x bodies per wavefront: times every pair
 * (leaf_sizes[i], targets per lane tpls[j] in {1, 2, 4}; a wavefront serves
 * 64 x tpl bodies) as above and keeps the fastest pair.  ms_out[i * kt + j]
 * (optional) receives each mean pass time, -1 for pairs the batch's walk
 * does not run (more than one target per lane exists for the FAST
 * linear-BVH walk only).  The traversal stacks need no tuning: each walk's
 * stack holds the deepest path its trees can have. */
int nbg_batch_autotune_grid(nbg_batch *b, const int32_t *leaf_sizes, int32_t kl, const int32_t *tpls, int32_t kt,
                            int32_t repeats, int32_t *best_leaf, int32_t *best_tpl, float *ms_out);

/* Targets per lane of the batch's FAST linear-BVH walk (1, 2 or 4); the
 * autotuner sets it too.  Results do not depend on it. */
int nbg_batch_set_targets_per_lane(nbg_batch *b, int32_t tpl);

/* Waves per group of 64 targets in the batch's FAST walks (no reference
 * counterpart: a schedule choice, as nbg_sys_set_walk_split).  0 (the
 * default) chooses by size: a forest of at most 65,536 bodies in all, walked
 * with one target per lane and with at least half of its bodies in systems of
 * 4,096 bodies or more, splits each group over up to 4 waves (octree: 8);
 * 1 never splits.  Each member's results are the same bits for every
 * choice.  A FAST batch keeps 408 bytes per body, for up to 65,536 bodies,
 * of split sums in HBM. */
int nbg_batch_set_walk_split(nbg_batch *b, int32_t waves);

/* ---- serving: continuous batching over a forest ----
 *
 * No reference counterpart.  Jobs (a linear-BVH system, 30-bit keys, plus a
 * step count) are submitted from any thread; a worker thread admits them in
 * arrival order whenever the forest's capacity allows, steps all admitted
 * systems together, and retires each when its step count is reached.  The
 * forest's capacity is the largest that fits hbm_budget_bytes
 * (nbg_batch_bytes).  A job's result is the same bits as the same system
 * stepped alone in an nbg_sys. */
typedef struct nbg_server nbg_server;

int nbg_server_create(nbg_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_systems, int32_t sum_mode,
                      nbg_server **out);
/* The same server with an express lane: jobs of at most express_max_bodies
 * bodies go to a second forest (express_budget_bytes of the budget,
 * express_max_systems slots) with its own worker thread and stream, which
 * steps concurrently with the bulk forest.  A small job then advances at the
 * pace of a forest of small systems rather than one step per step of the
 * largest resident job.  nbg_server_capacity reports the largest job either
 * lane accepts and the slots of both.  No reference counterpart. */
int nbg_server_create_express(nbg_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_systems, int32_t sum_mode,
                              int64_t express_max_bodies, int64_t express_budget_bytes, int32_t express_max_systems,
                              nbg_server **out);
/* Stops the worker.  Jobs still queued or running end with NBG_ERR_STATE;
 * clients blocked in nbg_server_wait return that status, and destroy returns
 * once they have left.  A finished job's result is held until it is waited
 * for (or until destroy). */
int nbg_server_destroy(nbg_server *s);
int nbg_server_capacity(nbg_server *s, int64_t *max_bodies, int32_t *max_systems);

/* Queues a job (host arrays copied; vx/vy/vz may be NULL); *job receives its
 * id.  NBG_ERR_BUDGET if the job alone exceeds the capacity. */
int nbg_server_submit(nbg_server *s, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
                      const float *m, const float *vx, const float *vy, const float *vz, int32_t nsteps,
                      int64_t *job);

/* Waits for a job (timeout_s < 0: no limit; NBG_ERR_STATE on time-out), then
 * copies its final state out (any pointer may be NULL) and its times in
 * seconds: times3 = {queued, on the device, submit-to-done latency}.  A job's
 * result can be collected once. */
int nbg_server_wait(nbg_server *s, int64_t job, double timeout_s, float *x, float *y, float *z, float *vx, float *vy,
                    float *vz, double *times3);

/* out8 = {jobs finished, forest steps run, body-steps, device-busy seconds,
 * seconds since creation, peak bodies resident, peak systems resident, jobs
 * waiting}. */
int nbg_server_stats(nbg_server *s, double *out8);

/* ---- job snapshots ----
 *
 * No reference counterpart: the reference saves whole scenarios as JSON
 * (src/NBody/Simulation/Simulation.cpp:53-63,154-228) with no step counter.
 * A snapshot is one contiguous little-endian buffer holding a job between
 * two steps:
 *   [0, 64)    "NBGSNAP1", u32 version 1, u32 sizeof(nbg_params), i64 n,
 *              i32 steps_total, i32 steps_done, u64 checksum (FNV-1a 64 of
 *              every byte from offset 64 on), 24 zero bytes
 *   [64, 112)  the job's nbg_params, zero padded
 *   [112, 112 + 28 n)  float32 x, y, z, m, vx, vy, vz, n values each
 * A step reads only positions, velocities and masses, so a job resumed from
 * a snapshot ends with the same bits as the job run uninterrupted, on a
 * server of the same sum mode.  Buffers must be 4-byte aligned. */

/* Copies a job's state into buf (cap bytes) and lets the job run on;
 * *size receives the snapshot's size, and buf NULL only asks for it.  A
 * resident job is copied at its forest's next step boundary (the call waits
 * for at most one forest step); a queued job gives its submitted state, a
 * finished one its final state (steps_done == steps_total), a failed one its
 * error. */
int nbg_server_checkpoint(nbg_server *s, int64_t job, void *buf, int64_t cap, int64_t *size);
/* The same, and the job leaves the server: its forest slot frees, clients
 * waiting on it return NBG_ERR_STATE and its id is forgotten.  A finished
 * job is collected with nbg_server_wait instead (NBG_ERR_STATE here). */
int nbg_server_suspend(nbg_server *s, int64_t job, void *buf, int64_t cap, int64_t *size);
/* Queues the job a snapshot holds for its remaining steps, under a new id.
 * NBG_ERR_ARG for a damaged snapshot (checksum), another format version or
 * parameter layout, a size that does not match n, or no steps left. */
int nbg_server_resume(nbg_server *s, const void *buf, int64_t size, int64_t *job);

/* ---- synthetic inputs (host, CPU) ---- */

/* Plummer sphere: scale radius a, total mass M, gravitational constant G,
 * radius truncated at rmax*a; positions by inverse CDF, velocities by the
 * Aarseth-Henon-Wielen rejection; std::mt19937_64(seed) with 53-bit
 * doubles; generated in FP64, rounded to FP32.  m_i = M / n. */
int nbg_generate_plummer(uint64_t seed, int64_t n, double a, double M, double G, double rmax, float *x,
                         float *y, float *z, float *m, float *vx, float *vy, float *vz);

/* Uniform cube of side `side` centred at the origin, zero velocity, total
 * mass M. */
int nbg_generate_cube(uint64_t seed, int64_t n, double side, double M, float *x, float *y, float *z,
                      float *m, float *vx, float *vy, float *vz);

#ifdef __cplusplus
}
#endif

#endif /* NBG_H */
