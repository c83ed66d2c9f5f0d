// C ABI of the MI355X tree-code force pass.  Host code only: it owns the
// device buffers of each system, enqueues the build / walk / update kernels
// on the system's stream and reports errors as status codes.

#include <algorithm>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/nbg.h"
#include "host_util.hpp"
#include "launch.hpp"

using namespace nbg;

namespace {

constexpr int kPhases = 8;

bool is_octree(int solver) {
    return solver == NBG_BARNES_HUT || solver == NBG_BARNES_HUT_QUADRUPOLE || solver == NBG_IMPLICIT_FMM;
}
// Implicit FMM: waves of a traversal step, the seed queues' capacity, and
// the list pool of a pass (words per body plus a floor).
constexpr int kFmmWaves = 5120, kFmmSeedCap = 16384;
static_assert(kFmmWaves <= kFmmGroupWaves, "the step waves share the leaf-group waves' close-node lists");
constexpr int64_t kFmmPoolPerBody = 32, kFmmPoolFloor = (int64_t)1 << 17, kFmmPoolMax = (int64_t)1 << 31;
bool is_quad(int solver) { return solver == NBG_BARNES_HUT_QUADRUPOLE || solver == NBG_LINEAR_BVH_QUADRUPOLE; }
constexpr int kEventSets = 256;

}  // namespace

struct nbg_sys {
    nbg_ctx *ctx = nullptr;
    nbg_params prm{};
    int64_t cap = 0;
    int64_t n = 0;
    hipStream_t stream = nullptr;
    bool have_state = false;
    bool have_forces = false;
    bool timing = false;
    bool walk_stats = false;  // FAST walk counts its schedule statistics
    bool counting = true;     // FAST walk counts interactions (SimpleTrackingRule)
    bool tree_valid = false;  // the tree matches the current state and parameters
    // Timing: one set of kPhases+1 events per force pass, recorded on the
    // system stream; folded into per-phase sums when the stream is synchronised.
    std::vector<hipEvent_t> ev;    // kEventSets * (kPhases + 1)
    int ev_used = 0;               // sets recorded since the last fold
    double phase_sum[kPhases] = {};
    int64_t phase_n = 0;           // passes folded into phase_sum

    // state, original order
    float *x = nullptr, *y = nullptr, *z = nullptr, *m = nullptr;
    float *vx = nullptr, *vy = nullptr, *vz = nullptr;
    float *ax = nullptr, *ay = nullptr, *az = nullptr;
    // build work
    void *keys = nullptr, *keys_alt = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;
    float4 *bodies = nullptr;
    int32_t *kfirst = nullptr, *klast = nullptr, *kgamma = nullptr, *kparent = nullptr;
    uint32_t *ksplit = nullptr, *krank = nullptr;
    int32_t *leaf_start = nullptr;  // leaf slot by first body (n)
    uint32_t *glist = nullptr;      // straddling parents queued by the summaries (2 queues of 2n + 64)
    uint32_t *scan_ws = nullptr;
    uint32_t *sort_status = nullptr;
    // tree (capacity 2n nodes)
    NodeRec *nodes = nullptr;
    float4 *bmin = nullptr, *bmax = nullptr;
    uint32_t *flags = nullptr;
    // octree work and nodes (allocated when the Barnes-Hut solver is selected)
    OctBuffers ob;
    // second moments, 2 float4 per node slot (allocated for the quadrupole solvers)
    float4 *quad = nullptr;
    int64_t quad_slots = 0;
    // implicit FMM work (allocated for that solver): start flags by slot,
    // starting nodes, leaf list, counters, per-wave list scratch
    // implicit FMM work (allocated for that solver): frontier items and list
    // pool (two halves each), seed queues, counters, starting nodes
    uint32_t *fmm_items = nullptr, *fmm_pool = nullptr, *fmm_seed = nullptr, *fmm_misc = nullptr;
    uint32_t *fmm_ulist = nullptr;
    int32_t *fmm_starts = nullptr;
    int64_t fmm_item_cap = 0, fmm_pool_cap = 0;
    int fmm_seed_cap = 0;
    int fmm_shift = 0;  // capacity doublings after passes that outgrew the pool
    // split FAST linear-BVH walks of small systems: per-target sums and
    // per-group block counts (see WalkArgs::split); split_req 0 = automatic
    double *split_part = nullptr;
    int32_t *split_nb = nullptr;
    int64_t split_cap = 0;
    int64_t split_words = 0;  // doubles per target in split_part
    int split_req = 0;
    // passive-point walks (nbg_sys_accelerations_at): one buffer that grows
    // to the largest target count seen and is kept for later calls
    void *pts = nullptr;
    size_t pts_bytes = 0;
    // small
    float *bbox_partials = nullptr, *bbox = nullptr;
    // [1] err, [2] node count, [4..16) counts: pp, np, then the FAST walk's
    // wave-level node visits, active lanes summed over visits, leaf visits
    // and open lanes summed over leaf visits (6 x u64).
    uint32_t *misc = nullptr;

    uint32_t *err_word() { return misc + 1; }
    uint32_t *node_count() { return misc + 2; }
    unsigned long long *counts() { return (unsigned long long *)(misc + 4); }
};

namespace {

int check_params(const nbg_params *p) {
    if (!p) return fail(NBG_ERR_ARG, "params is NULL");
    if (p->solver < NBG_BARNES_HUT || p->solver > NBG_IMPLICIT_FMM) return fail(NBG_ERR_ARG, "unknown solver kind");
    if (!(p->theta >= 0.0f)) return fail(NBG_ERR_ARG, "theta must be >= 0");
    if (p->leaf_size < 1) return fail(NBG_ERR_ARG, "leaf_size must be >= 1");
    if (p->max_depth < 0 || p->max_depth > 32) return fail(NBG_ERR_ARG, "max_depth must be in [0, 32]");
    if (p->morton_bits != 30 && p->morton_bits != 63) return fail(NBG_ERR_ARG, "morton_bits must be 30 or 63");
    if (!(p->eps2 >= 0.0f)) return fail(NBG_ERR_ARG, "eps2 must be >= 0");
    if (p->sum_mode != NBG_SUM_ORDERED && p->sum_mode != NBG_SUM_FAST) return fail(NBG_ERR_ARG, "unknown sum mode");
    if (p->solver == NBG_IMPLICIT_FMM && p->sum_mode != NBG_SUM_FAST)
        return fail(NBG_ERR_ARG, "the implicit FMM runs in the FAST sum mode only");
    if (p->targets_per_lane != 0 && p->targets_per_lane != 1 && p->targets_per_lane != 2 && p->targets_per_lane != 4)
        return fail(NBG_ERR_ARG, "targets_per_lane must be 1, 2 or 4");
    return NBG_OK;
}

void free_oct(OctBuffers &b) {
    void *ptrs[] = {b.key_hi, b.key_lo, b.lo_sorted, b.nown, b.nbase, b.nfirst, b.ncount, b.nparent, b.nmeta,
                    b.occ, b.cb, b.flags, b.glist, b.scan_ws, b.nodes, b.centre, b.thr, b.misc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    b = OctBuffers{};
}

// Allocates the octree buffers for the system's size and parameters (no-op
// when the current ones are large enough).
int ensure_octree(nbg_sys *s) {
    if (!is_octree(s->prm.solver)) return NBG_OK;
    const int64_t n = s->n;
    const int64_t cap = octree_capacity(n, s->prm.leaf_size, s->prm.max_depth);
    if (cap >= ((int64_t)1 << 28))
        return fail(NBG_ERR_ARG, "octree node bound exceeds 2^28 slots; raise leaf_size or lower n");
    if (s->ob.cap >= cap && s->ob.n == n) return NBG_OK;
    if (s->stream) NBG_HIP(hipStreamSynchronize(s->stream));
    free_oct(s->ob);
    OctBuffers &b = s->ob;
    int r = NBG_OK;
    auto A = [&](auto **ptr, int64_t count) {
        if (r == NBG_OK) r = dalloc(ptr, count);
    };
    A(&b.key_hi, n); A(&b.key_lo, n); A(&b.lo_sorted, n);
    A(&b.nown, n); A(&b.nbase, n);
    A(&b.nfirst, cap); A(&b.ncount, cap); A(&b.nparent, cap); A(&b.nmeta, cap);
    A(&b.occ, cap); A(&b.cb, cap); A(&b.flags, cap); A(&b.glist, 2 * (cap + 64));
    A(&b.scan_ws, 128 + (cap + kScanTile - 1) / kScanTile);
    A(&b.nodes, cap); A(&b.centre, cap); A(&b.thr, 64); A(&b.misc, 64);
    if (r != NBG_OK) {
        free_oct(b);
        return r;
    }
    b.cap = cap;
    b.n = n;
    NBG_HIP(hipMemset(b.misc, 0, 64 * sizeof(uint32_t)));
    return NBG_OK;
}

// Allocates the quadrupole moment array for the current tree's node slots
// (no-op for monopole solvers or when the current array is large enough).
int ensure_quad(nbg_sys *s) {
    if (!is_quad(s->prm.solver)) return NBG_OK;
    const int64_t slots = is_octree(s->prm.solver) ? s->ob.cap : 2 * s->cap;
    if (s->quad_slots >= slots) return NBG_OK;
    if (s->stream) NBG_HIP(hipStreamSynchronize(s->stream));
    if (s->quad) (void)hipFree(s->quad);
    s->quad = nullptr;
    s->quad_slots = 0;
    if (int r = dalloc(&s->quad, 2 * (slots > 0 ? slots : 1))) return r;
    s->quad_slots = slots;
    return NBG_OK;
}

void free_fmm(nbg_sys *s) {
    void *ptrs[] = {s->fmm_items, s->fmm_pool, s->fmm_seed, s->fmm_misc, s->fmm_starts, s->fmm_ulist};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    s->fmm_items = s->fmm_pool = s->fmm_seed = s->fmm_misc = s->fmm_ulist = nullptr;
    s->fmm_starts = nullptr;
    s->fmm_item_cap = s->fmm_pool_cap = 0;
}

// Allocates the implicit FMM's work buffers (no-op for other solvers or when
// the current ones are large enough).
int ensure_fmm(nbg_sys *s) {
    if (s->prm.solver != NBG_IMPLICIT_FMM) return NBG_OK;
    // a step's frontier holds disjoint non-empty subtrees: at most one item per body
    const int64_t n = s->cap > 0 ? s->cap : 1;
    const int64_t item_cap = n + 64;
    const int64_t pool_cap = std::min(kFmmPoolMax, (kFmmPoolPerBody * n + kFmmPoolFloor) << s->fmm_shift);
    const int seed_cap = kFmmSeedCap << std::min(s->fmm_shift, 4);
    if (s->fmm_items && s->fmm_item_cap >= item_cap && s->fmm_pool_cap >= pool_cap && s->fmm_seed_cap >= seed_cap)
        return NBG_OK;
    if (s->stream) NBG_HIP(hipStreamSynchronize(s->stream));
    free_fmm(s);
    int r = NBG_OK;
    auto A = [&](auto **ptr, int64_t count) {
        if (r == NBG_OK) r = dalloc(ptr, count);
    };
    A(&s->fmm_items, 3 * item_cap * kFmmItemWords); A(&s->fmm_pool, pool_cap);
    A(&s->fmm_seed, (int64_t)kFmmSeedWaves * 3 * seed_cap); A(&s->fmm_misc, 64);
    A(&s->fmm_starts, kFmmStartCap); A(&s->fmm_ulist, (int64_t)kFmmGroupWaves * kFmmUlistCap);
    if (r != NBG_OK) {
        free_fmm(s);
        return r;
    }
    s->fmm_item_cap = item_cap;
    s->fmm_pool_cap = pool_cap;
    s->fmm_seed_cap = seed_cap;
    return NBG_OK;
}

bool is_lbvh(int solver) { return solver == NBG_LINEAR_BVH || solver == NBG_LINEAR_BVH_QUADRUPOLE; }

// Split waves per group for the system's tree walk (1 for the other
// solvers and several targets per lane).  The FAST octree's frontier holds
// at most 8 blocks, so it uses at most 8 waves; by size the FAST linear-BVH
// walk uses at most 4 (measured fastest from 2,000 to 65,536 bodies).
int split_of(const nbg_sys *s) {
    const bool lbvh = is_lbvh(s->prm.solver);
    const bool tree = lbvh || s->prm.solver == NBG_BARNES_HUT || s->prm.solver == NBG_BARNES_HUT_QUADRUPOLE;
    if (s->prm.sum_mode != NBG_SUM_FAST) {
        if (!tree || s->n > kOrderedSplitMaxBodies) return 1;
        return s->split_req > 0 ? s->split_req : split_waves(s->n);
    }
    if (!lbvh && !(is_octree(s->prm.solver) && s->prm.solver != NBG_IMPLICIT_FMM)) return 1;
    if (lbvh && s->prm.targets_per_lane > 1) return 1;
    if (s->split_req > 0) return lbvh ? s->split_req : std::min(s->split_req, 8);
    return std::min(split_waves(s->n), lbvh ? kSplitAutoLbvh : 8);
}

void free_split(nbg_sys *s) {
    if (s->split_part) (void)hipFree(s->split_part);
    if (s->split_nb) (void)hipFree(s->split_nb);
    s->split_part = nullptr;
    s->split_nb = nullptr;
    s->split_cap = 0;
    s->split_words = 0;
}

// Allocates the split walk's sums when the system's walk splits (no-op
// otherwise or when the current ones are large enough).
// Doubles of split sums per target: the FAST blocks' FP64 sums, or the
// ORDERED frontier results (FP32, two to a double).
int64_t split_words(const nbg_sys *s) {
    return s->prm.sum_mode == NBG_SUM_FAST ? (int64_t)kSplitSlots * 3 : (int64_t)kOrderedSplitSlots * 3 / 2;
}

int ensure_split(nbg_sys *s) {
    if (split_of(s) <= 1 || (s->split_cap >= s->n && s->split_words >= split_words(s))) return NBG_OK;
    if (s->stream) NBG_HIP(hipStreamSynchronize(s->stream));
    free_split(s);
    const int64_t n = s->n > 0 ? s->n : 1;
    int r = dalloc(&s->split_part, split_words(s) * n);
    if (r == NBG_OK) r = dalloc(&s->split_nb, (n + 63) / 64);
    if (r != NBG_OK) {
        free_split(s);
        return r;
    }
    s->split_cap = n;
    s->split_words = split_words(s);
    return NBG_OK;
}

int ensure_buffers(nbg_sys *s) {
    if (int r = ensure_octree(s)) return r;
    if (int r = ensure_fmm(s)) return r;
    if (int r = ensure_split(s)) return r;
    return ensure_quad(s);
}

void free_sys(nbg_sys *s) {
    free_oct(s->ob);
    free_fmm(s);
    free_split(s);
    if (s->quad) (void)hipFree(s->quad);
    if (s->pts) (void)hipFree(s->pts);
    void *ptrs[] = {s->x, s->y, s->z, s->m, s->vx, s->vy, s->vz, s->ax, s->ay, s->az, s->keys, s->keys_alt,
                    s->idx, s->idx_alt, s->bodies, s->kfirst, s->klast, s->kgamma, s->kparent, s->ksplit,
                    s->krank, s->leaf_start, s->glist, s->scan_ws, s->sort_status, s->nodes, s->bmin,
                    s->bmax, s->flags,
                    s->bbox_partials, s->bbox, s->misc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (auto &e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
}

// Records event `phase` of the current pass's set (no-op when timing is off
// or every set is in use until the next synchronisation).
void mark(nbg_sys *s, int phase) {
    if (!s->timing || s->ev_used >= kEventSets) return;
    (void)hipEventRecord(s->ev[(size_t)s->ev_used * (kPhases + 1) + phase], s->stream);
    if (phase == kPhases) ++s->ev_used;
}

// Folds the recorded sets into the per-phase sums (stream must be idle).
void fold_timing(nbg_sys *s) {
    for (int k = 0; k < s->ev_used; ++k) {
        hipEvent_t *e = &s->ev[(size_t)k * (kPhases + 1)];
        for (int i = 0; i < kPhases; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e[i], e[i + 1]) == hipSuccess) s->phase_sum[i] += ms;
        }
    }
    s->phase_n += s->ev_used;
    s->ev_used = 0;
}

WalkArgs walk_args(nbg_sys *s);
void launch_walk(nbg_sys *s, const WalkArgs &wa, int stats);

// ActiveTreeSolver::updateAccelerations (Solvers/ActiveTreeSolver.h:33-72):
// build the tree (LinearBVH::build, Trees/LinearBVH.h:106-133, or
// Octree::build, Trees/Octree.h:185-205), then walk it for every body.
// Enqueue only.
int enqueue_force_pass(nbg_sys *s) {
    const int64_t n = s->n;
    const nbg_params &p = s->prm;
    hipStream_t st = s->stream;
    const bool octree = is_octree(p.solver);
    if (int r = ensure_buffers(s)) return r;
    float4 *quad = is_quad(p.solver) ? s->quad : nullptr;
    launch_fill_u32(s->misc + 4, 0u, 12, st);  // interaction counters
    if (p.solver == NBG_NAIVE) {
        // NaiveSolver: no tree; every body sums every body in body order.
        // SimpleTrackingRule counts n calls per body, np = 0.
        for (int k = 0; k <= 5; ++k) mark(s, k);
        launch_fill_u32(s->node_count(), 0u, 1, st);
        launch_pack(s->x, s->y, s->z, s->m, n, s->bodies, st);
        const uint64_t pp = (uint64_t)n * (uint64_t)n;
        launch_fill_u32(s->misc + 4, (uint32_t)pp, 1, st);
        launch_fill_u32(s->misc + 5, (uint32_t)(pp >> 32), 1, st);
        mark(s, 6);
        launch_direct_f32(s->bodies, n, s->bodies, n, nullptr, p.G, p.eps2, p.sum_mode == NBG_SUM_ORDERED, s->ax,
                          s->ay, s->az, st);
        mark(s, 7);
        NBG_HIP(hipGetLastError());
        s->tree_valid = true;
        return NBG_OK;
    }
    mark(s, 0);
    launch_bbox(s->x, s->y, s->z, n, s->bbox_partials, s->bbox, st);
    mark(s, 1);
    if (octree) {
        OctBuffers &ob = s->ob;
        uint64_t *keys = (uint64_t *)s->keys, *keys_alt = (uint64_t *)s->keys_alt;
        launch_octree_keys(s->x, s->y, s->z, n, s->bbox, p.max_depth, ob, keys, s->idx, s->idx_alt, st);
        mark(s, 2);
        launch_octree_sort(n, p.max_depth, ob, keys, s->idx, keys_alt, s->idx_alt, s->sort_status, s->err_word(), st);
        mark(s, 3);
        launch_octree_topology(n, s->bbox, p.leaf_size, p.max_depth, p.theta, ob, keys, s->idx_alt, s->idx,
                               s->leaf_start, s->err_word(), st);
        mark(s, 4);
        launch_gather(s->x, s->y, s->z, s->m, s->idx, n, s->bodies, st);
        mark(s, 5);
        launch_octree_summaries(s->bodies, n, ob, s->leaf_start, s->err_word(), st, quad);
    } else {
        launch_morton(s->x, s->y, s->z, n, s->bbox, p.morton_bits, s->keys, s->idx, st);
        mark(s, 2);
        launch_radix_sort(s->keys, s->idx, s->keys_alt, s->idx_alt, n, p.morton_bits, s->sort_status, s->err_word(),
                          st);
        mark(s, 3);
        launch_lbvh_topology(s->keys, p.morton_bits, n, p.leaf_size, s->kfirst, s->klast, s->kgamma, s->ksplit,
                             s->krank, s->kparent, s->scan_ws, s->nodes, s->leaf_start, s->node_count(), st);
        mark(s, 4);
        launch_gather(s->x, s->y, s->z, s->m, s->idx, n, s->bodies, st);
        mark(s, 5);
        launch_lbvh_summaries(s->bodies, n, s->nodes, s->leaf_start, 2 * n, s->flags, s->glist, s->bmin, s->bmax,
                              p.theta, s->err_word(), st, nullptr, nullptr, 0, quad);
    }
    mark(s, 6);
    if (p.solver == NBG_IMPLICIT_FMM) {
        FmmArgs fa;
        fa.n = n;
        fa.bodies = s->bodies;
        fa.nodes = s->ob.nodes;
        fa.centre = s->ob.centre;
        fa.misc = s->ob.misc;
        fa.max_nodes = s->ob.cap;
        fa.perm = s->idx;
        fa.ax = s->ax; fa.ay = s->ay; fa.az = s->az;
        fa.theta = p.theta; fa.G = p.G; fa.eps2 = p.eps2;
        fa.starts = s->fmm_starts;
        fa.ctr = s->fmm_misc;
        fa.pool_ctr = s->fmm_misc + 12;
        fa.group_ctr = s->fmm_misc + 13;
        fa.peak = s->fmm_misc + 16;
        fa.nstarts = s->fmm_misc + 20;
        fa.items[0] = s->fmm_items;
        fa.items[1] = s->fmm_items + s->fmm_item_cap * kFmmItemWords;
        fa.groups = s->fmm_items + 2 * s->fmm_item_cap * kFmmItemWords;
        fa.pool = s->fmm_pool;
        fa.item_cap = s->fmm_item_cap;
        fa.pool_cap = s->fmm_pool_cap;
        fa.seed_scratch = s->fmm_seed;
        fa.ulist = s->fmm_ulist;
        fa.ulist_cap = kFmmUlistCap;
        fa.cap_a = s->fmm_seed_cap;
        fa.waves = kFmmWaves;
        fa.steps = p.max_depth + 4;  // a step per tree level below the shallowest starting node
        fa.counts = s->counts();
        fa.err = s->err_word();
        launch_fmm(fa, st);
    } else {
        WalkArgs wa = walk_args(s);
        const int k = split_of(s);
        if (k > 1 && s->split_part && s->split_cap >= s->n && s->split_words >= split_words(s)) {
            wa.split = k;
            wa.split_part = s->split_part;
            wa.split_part_f = reinterpret_cast<float *>(s->split_part);
            wa.split_nb = s->split_nb;
            wa.split_stride = s->n;
        }
        const int stats = s->walk_stats ? 2 : (s->counting ? 1 : 0);
        launch_walk(s, wa, stats);
    }
    mark(s, 7);
    NBG_HIP(hipGetLastError());
    s->tree_valid = true;
    return NBG_OK;
}

// The walk's arguments for the system's own bodies as targets.
WalkArgs walk_args(nbg_sys *s) {
    const int64_t n = s->n;
    const nbg_params &p = s->prm;
    const bool octree = is_octree(p.solver);
    WalkArgs wa;
    wa.n = n;
    wa.bodies = s->bodies;
    wa.nt = n;
    wa.targets = s->bodies;
    wa.perm = s->idx;
    wa.nodes = s->nodes;
    wa.onodes = s->ob.nodes;
    wa.levels = p.max_depth + 2;
    wa.theta = p.theta;
    wa.theta2 = p.theta * p.theta;
    mac_threshold(p.theta, &wa.mac_mid, &wa.mac_tie);
    wa.G = p.G;
    wa.eps2 = p.eps2;
    wa.max_nodes = octree ? s->ob.cap : 2 * n;
    wa.ax = s->ax;
    wa.ay = s->ay;
    wa.az = s->az;
    wa.counts = s->counts();
    wa.err = s->err_word();
    wa.quad = is_quad(p.solver) ? s->quad : nullptr;
    wa.tpl = p.targets_per_lane > 1 ? p.targets_per_lane : 1;
    return wa;
}

void launch_walk(nbg_sys *s, const WalkArgs &wa, int stats) {
    const nbg_params &p = s->prm;
    if (p.solver == NBG_NAIVE) {
        launch_direct_f32(s->bodies, s->n, wa.targets, wa.nt, wa.perm, p.G, p.eps2, p.sum_mode == NBG_SUM_ORDERED,
                          wa.ax, wa.ay, wa.az, s->stream);
        const uint64_t pp = (uint64_t)s->n * (uint64_t)wa.nt;
        launch_fill_u32(s->misc + 4, (uint32_t)pp, 1, s->stream);
        launch_fill_u32(s->misc + 5, (uint32_t)(pp >> 32), 1, s->stream);
        return;
    }
    if (is_octree(p.solver))
        launch_walk_oct(wa, p.sum_mode == NBG_SUM_ORDERED, stats, s->stream);
    else
        launch_walk_lbvh(wa, p.sum_mode == NBG_SUM_ORDERED, p.morton_bits == 30 ? 32 : 64, stats, s->stream);
}

// Clears the device error word (stream idle) and reports the failure its
// bits name.
int report_device_error(nbg_sys *s, uint32_t err) {
    uint32_t zero = 0;
    (void)hipMemcpy(s->err_word(), &zero, sizeof zero, hipMemcpyHostToDevice);
    if (err & 1u) return fail(NBG_ERR_DEVICE, "radix sort look-back exceeded its spin limit");
    if (err & 2u) return fail(NBG_ERR_DEVICE, "tree traversal exceeded its stack, depth or index limits");
    if (err & 4u) return fail(NBG_ERR_DEVICE, "tree summaries found an invalid node index");
    if (err & 8u) return fail(NBG_ERR_DEVICE, "octree build found an inconsistent node index");
    if (err & 16u) return fail(NBG_ERR_DEVICE, "implicit FMM interaction lists exceeded their capacity");
    return fail(NBG_ERR_DEVICE, "device error flag set");
}

int sync_and_check(nbg_sys *s) {
    NBG_HIP(hipStreamSynchronize(s->stream));
    uint32_t err = 0;
    NBG_HIP(hipMemcpy(&err, s->err_word(), sizeof err, hipMemcpyDeviceToHost));
    if (err) return report_device_error(s, err);
    fold_timing(s);
    return NBG_OK;
}

// A force pass.  The implicit FMM's list pool is sized by an estimate; a pass
// that outgrows it (error bit 16 alone) is repeated with twice the capacity
// (bodies and velocities are untouched until the update), at the cost of one
// synchronisation per pass.
int force_pass(nbg_sys *s) {
    if (s->prm.solver != NBG_IMPLICIT_FMM) return enqueue_force_pass(s);
    for (;;) {
        if (int r = enqueue_force_pass(s)) return r;
        NBG_HIP(hipStreamSynchronize(s->stream));
        uint32_t err = 0;
        NBG_HIP(hipMemcpy(&err, s->err_word(), sizeof err, hipMemcpyDeviceToHost));
        if (err == 0) return NBG_OK;
        // a failed pass is reported here, before the caller enqueues an update
        // that would move the bodies with partial accelerations
        if (err != 16u || s->fmm_pool_cap >= kFmmPoolMax) return report_device_error(s, err);
        const uint32_t zero = 0;
        NBG_HIP(hipMemcpy(s->err_word(), &zero, sizeof zero, hipMemcpyHostToDevice));
        ++s->fmm_shift;
    }
}

// Device scratch of a passive-point walk of nt targets, carved from one
// buffer: positions (3 nt floats), results (3 nt floats), Morton keys and
// their sort's alternate buffer (2 nt u64), the permutation and its
// alternate (2 nt u32), the targets in walk order (nt float4) and the sort's
// status words.  Offsets are 256-byte aligned.
struct PointsLayout {
    size_t t3, a3, k2, i2, t4, status, bytes;
};
constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr PointsLayout points_layout(int64_t nt) {
    PointsLayout L{};
    size_t o = 0;
    L.t3 = o; o = align256(o + (size_t)nt * 3 * sizeof(float));
    L.a3 = o; o = align256(o + (size_t)nt * 3 * sizeof(float));
    L.k2 = o; o = align256(o + (size_t)nt * 2 * sizeof(uint64_t));
    L.i2 = o; o = align256(o + (size_t)nt * 2 * sizeof(uint32_t));
    L.t4 = o; o = align256(o + (size_t)nt * sizeof(float4));
    L.status = o; o = align256(o + (size_t)sort_status_words(nt, 64) * sizeof(uint32_t));
    L.bytes = o;
    return L;
}

struct PointsBuf {
    float *t3, *a3;
    uint64_t *k2;
    uint32_t *i2, *status;
    float4 *t4;
};

PointsBuf points_carve(void *base, int64_t nt) {
    const PointsLayout L = points_layout(nt);
    char *b = (char *)base;
    PointsBuf p;
    p.t3 = (float *)(b + L.t3);
    p.a3 = (float *)(b + L.a3);
    p.k2 = (uint64_t *)(b + L.k2);
    p.i2 = (uint32_t *)(b + L.i2);
    p.t4 = (float4 *)(b + L.t4);
    p.status = (uint32_t *)(b + L.status);
    return p;
}

// Scratch of nbg_sys_accelerations_at_noalloc: a file-scope device array
// (points_layout of NBG_NOALLOC_MAX_POINTS targets, about 17 MB) that every
// GPU the library runs on holds until the process ends.  In host code its
// name is not a device address: nbg_ctx_create looks up each device's own
// copy with hipGetSymbolAddress, with that device current, and host code
// uses only that pointer (kernel arguments, copies, fills).
static __device__ __align__(256) unsigned char scratch[points_layout(NBG_NOALLOC_MAX_POINTS).bytes];

constexpr int kMaxNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxNoAllocDevices] = {};

// This device's scratch pointer, looked up once per device (the caller has
// made the device current).
int noalloc_scratch_of(int device, void **out) {
    *out = nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (device < kMaxNoAllocDevices && g_noalloc_ptr[device]) {
        *out = g_noalloc_ptr[device];
        return NBG_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(NBG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    if (device < kMaxNoAllocDevices) g_noalloc_ptr[device] = p;
    *out = p;
    return NBG_OK;
}

int check_points(nbg_sys *s, int64_t nt, const float *tx, const float *ty, const float *tz, float *ax, float *ay,
                 float *az) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (nt < 0 || nt > kMaxBodies) return fail(NBG_ERR_ARG, "target count out of range");
    if (nt > 0 && (!tx || !ty || !tz || !ax || !ay || !az)) return fail(NBG_ERR_ARG, "NULL array");
    if (!s->have_state) return fail(NBG_ERR_STATE, "set_state has not been called");
    if (s->prm.solver == NBG_IMPLICIT_FMM)
        return fail(NBG_ERR_STATE, "the implicit FMM's passive points are its own bodies");
    return NBG_OK;
}

// The walk of nt > 0 passive points through the device scratch `pb` (laid
// out by points_layout(nt)).  Uses the tree of the system's current state,
// building it first when the state or parameters changed since the last
// force pass.  Targets are walked in Morton order over the system's bounding
// box (63-bit keys, clamped outside it) so that a wavefront's targets are
// spatially adjacent; results return in the caller's order.  Blocking.
int walk_points(nbg_sys *s, int64_t nt, const float *tx, const float *ty, const float *tz, float *ax, float *ay,
                float *az, const PointsBuf &pb) {
    if (s->n == 0) {  // no actives: every passive acceleration is zero
        for (int64_t i = 0; i < nt; ++i) ax[i] = ay[i] = az[i] = 0.0f;
        return NBG_OK;
    }
    if (!s->tree_valid) {
        // builds the tree (and, as a side effect, the bodies' own accelerations)
        if (int r = enqueue_force_pass(s)) return r;
        mark(s, kPhases);
        if (int r = sync_and_check(s)) return r;
        s->have_forces = true;
    }
    hipStream_t st = s->stream;
    const size_t b = (size_t)nt * sizeof(float);
    NBG_HIP(hipMemcpyAsync(pb.t3, tx, b, hipMemcpyHostToDevice, st));
    NBG_HIP(hipMemcpyAsync(pb.t3 + nt, ty, b, hipMemcpyHostToDevice, st));
    NBG_HIP(hipMemcpyAsync(pb.t3 + 2 * nt, tz, b, hipMemcpyHostToDevice, st));
    launch_morton(pb.t3, pb.t3 + nt, pb.t3 + 2 * nt, nt, s->bbox, 63, pb.k2, pb.i2, st);
    launch_radix_sort(pb.k2, pb.i2, pb.k2 + nt, pb.i2 + nt, nt, 63, pb.status, s->err_word(), st);
    launch_gather(pb.t3, pb.t3 + nt, pb.t3 + 2 * nt, pb.t3, pb.i2, nt, pb.t4, st);
    launch_fill_u32(s->misc + 4, 0u, 12, st);
    WalkArgs wa = walk_args(s);
    wa.nt = nt;
    wa.targets = pb.t4;
    wa.perm = pb.i2;
    wa.ax = pb.a3;
    wa.ay = pb.a3 + nt;
    wa.az = pb.a3 + 2 * nt;
    launch_walk(s, wa, s->counting || s->prm.sum_mode == NBG_SUM_ORDERED ? 1 : 0);
    NBG_HIP(hipGetLastError());
    if (int r = sync_and_check(s)) return r;
    NBG_HIP(hipMemcpy(ax, pb.a3, b, hipMemcpyDeviceToHost));
    NBG_HIP(hipMemcpy(ay, pb.a3 + nt, b, hipMemcpyDeviceToHost));
    NBG_HIP(hipMemcpy(az, pb.a3 + 2 * nt, b, hipMemcpyDeviceToHost));
    return NBG_OK;
}

}  // namespace

extern "C" {

void nbg_params_default(nbg_params *p) {
    if (!p) return;
    p->solver = NBG_LINEAR_BVH;
    p->theta = 0.3f;
    p->leaf_size = 64;
    p->max_depth = 32;
    p->morton_bits = 63;
    p->G = 1.0f;
    p->eps2 = 0.00001f;
    p->dt = 0.001f;
    p->sum_mode = NBG_SUM_ORDERED;
    p->targets_per_lane = 1;
}

const char *nbg_last_error(void) { return g_err.c_str(); }

const char *nbg_version(void) { return "nbg 0.1 (gfx950)"; }

int nbg_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int nbg_ctx_create(int device, nbg_ctx **out) {
    if (!out) return fail(NBG_ERR_ARG, "out is NULL");
    *out = nullptr;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c <= 0) return fail(NBG_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= c) return fail(NBG_ERR_ARG, "device index out of range");
    NBG_HIP(hipSetDevice(device));
    void *scr = nullptr;
    if (int r = noalloc_scratch_of(device, &scr)) return r;
    nbg_ctx *ctx = new nbg_ctx;
    ctx->device = device;
    ctx->noalloc_scratch = scr;
    *out = ctx;
    return NBG_OK;
}

int nbg_ctx_destroy(nbg_ctx *ctx) {
    delete ctx;
    return NBG_OK;
}

int nbg_sys_create(nbg_ctx *ctx, const nbg_params *p, int64_t n, nbg_sys **out) {
    if (!ctx || !out) return fail(NBG_ERR_ARG, "ctx or out is NULL");
    *out = nullptr;
    if (int r = check_params(p)) return r;
    if (n < 0 || n > kMaxBodies) return fail(NBG_ERR_ARG, "body count out of range");
    DeviceGuard g(ctx->device);
    nbg_sys *s = new nbg_sys;
    s->ctx = ctx;
    s->prm = *p;
    s->cap = n;
    s->n = n;
    int r = NBG_OK;
    auto A = [&](auto **ptr, int64_t count) {
        if (r == NBG_OK) r = dalloc(ptr, count);
    };
    A(&s->x, n); A(&s->y, n); A(&s->z, n); A(&s->m, n);
    A(&s->vx, n); A(&s->vy, n); A(&s->vz, n);
    A(&s->ax, n); A(&s->ay, n); A(&s->az, n);
    A((uint64_t **)&s->keys, n); A((uint64_t **)&s->keys_alt, n);
    A(&s->idx, n); A(&s->idx_alt, n);
    A(&s->bodies, n);
    A(&s->kfirst, n); A(&s->klast, n); A(&s->kgamma, n); A(&s->kparent, n);
    A(&s->ksplit, n); A(&s->krank, n);
    A(&s->leaf_start, n); A(&s->glist, 2 * (2 * n + 64));
    A(&s->scan_ws, 64 + (n + kScanTile - 1) / kScanTile + 64);
    A(&s->sort_status, sort_status_words(n, 64));
    A(&s->nodes, 2 * n); A(&s->bmin, 2 * n); A(&s->bmax, 2 * n);
    A(&s->flags, 2 * n);
    A(&s->bbox_partials, 6 * kBBoxBlocks);
    A(&s->bbox, 8);
    A(&s->misc, 32);
    if (r == NBG_OK) {
        hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
        if (e != hipSuccess) r = fail(NBG_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (r == NBG_OK) {
        hipError_t e = hipMemsetAsync(s->misc, 0, 32 * sizeof(uint32_t), s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) r = fail(NBG_ERR_HIP, std::string("init: ") + hipGetErrorString(e));
    }
    if (r == NBG_OK) r = ensure_buffers(s);
    if (r != NBG_OK) {
        free_sys(s);
        delete s;
        return r;
    }
    *out = s;
    return NBG_OK;
}

int nbg_sys_destroy(nbg_sys *s) {
    if (!s) return NBG_OK;
    DeviceGuard g(s->ctx->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    free_sys(s);
    delete s;
    return NBG_OK;
}

int nbg_sys_set_params(nbg_sys *s, const nbg_params *p) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (int r = check_params(p)) return r;
    DeviceGuard g(s->ctx->device);
    const nbg_params old = s->prm;
    s->prm = *p;
    if (int r = ensure_buffers(s)) {
        s->prm = old;
        return r;
    }
    s->tree_valid = false;
    return NBG_OK;
}

int nbg_sys_get_params(const nbg_sys *s, nbg_params *p) {
    if (!s || !p) return fail(NBG_ERR_ARG, "sys or params is NULL");
    *p = s->prm;
    return NBG_OK;
}

int nbg_sys_set_state(nbg_sys *s, const float *x, const float *y, const float *z, const float *m, const float *vx,
                      const float *vy, const float *vz) {
    if (!s || !x || !y || !z || !m) return fail(NBG_ERR_ARG, "positions and masses are required");
    DeviceGuard g(s->ctx->device);
    const size_t b = (size_t)s->n * sizeof(float);
    if (s->n > 0) {
        NBG_HIP(hipMemcpyAsync(s->x, x, b, hipMemcpyHostToDevice, s->stream));
        NBG_HIP(hipMemcpyAsync(s->y, y, b, hipMemcpyHostToDevice, s->stream));
        NBG_HIP(hipMemcpyAsync(s->z, z, b, hipMemcpyHostToDevice, s->stream));
        NBG_HIP(hipMemcpyAsync(s->m, m, b, hipMemcpyHostToDevice, s->stream));
        float *dv[3] = {s->vx, s->vy, s->vz};
        const float *hv[3] = {vx, vy, vz};
        for (int d = 0; d < 3; ++d) {
            if (hv[d]) NBG_HIP(hipMemcpyAsync(dv[d], hv[d], b, hipMemcpyHostToDevice, s->stream));
            else launch_fill_u32((uint32_t *)dv[d], 0u, s->n, s->stream);
        }
        NBG_HIP(hipStreamSynchronize(s->stream));
    }
    s->have_state = true;
    s->have_forces = false;
    s->tree_valid = false;
    return NBG_OK;
}

int nbg_sys_get_state(nbg_sys *s, float *x, float *y, float *z, float *vx, float *vy, float *vz) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    DeviceGuard g(s->ctx->device);
    const size_t b = (size_t)s->n * sizeof(float);
    float *h[6] = {x, y, z, vx, vy, vz};
    float *d[6] = {s->x, s->y, s->z, s->vx, s->vy, s->vz};
    NBG_HIP(hipStreamSynchronize(s->stream));
    for (int i = 0; i < 6; ++i)
        if (h[i] && s->n > 0) NBG_HIP(hipMemcpy(h[i], d[i], b, hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_sys_update_accelerations(nbg_sys *s, float *ax, float *ay, float *az) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (!s->have_state) return fail(NBG_ERR_STATE, "set_state has not been called");
    DeviceGuard g(s->ctx->device);
    if (int r = force_pass(s)) return r;
    mark(s, kPhases);  // a force pass alone has an empty update phase
    if (int r = sync_and_check(s)) return r;
    s->have_forces = true;
    if (ax || ay || az) return nbg_sys_get_accelerations(s, ax, ay, az);
    return NBG_OK;
}

int nbg_sys_get_accelerations(nbg_sys *s, float *ax, float *ay, float *az) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    const size_t b = (size_t)s->n * sizeof(float);
    if (s->n > 0) {
        if (ax) NBG_HIP(hipMemcpy(ax, s->ax, b, hipMemcpyDeviceToHost));
        if (ay) NBG_HIP(hipMemcpy(ay, s->ay, b, hipMemcpyDeviceToHost));
        if (az) NBG_HIP(hipMemcpy(az, s->az, b, hipMemcpyDeviceToHost));
    }
    return NBG_OK;
}

int nbg_sys_accelerations_at(nbg_sys *s, int64_t nt, const float *tx, const float *ty, const float *tz, float *ax,
                             float *ay, float *az) {
    if (int r = check_points(s, nt, tx, ty, tz, ax, ay, az)) return r;
    if (nt == 0) return NBG_OK;
    DeviceGuard g(s->ctx->device);
    const size_t need = points_layout(nt).bytes;
    if (need > s->pts_bytes) {
        NBG_HIP(hipStreamSynchronize(s->stream));
        if (s->pts) (void)hipFree(s->pts);
        s->pts = nullptr;
        s->pts_bytes = 0;
        NBG_HIP(hipMalloc(&s->pts, need));
        s->pts_bytes = need;
    }
    return walk_points(s, nt, tx, ty, tz, ax, ay, az, points_carve(s->pts, nt));
}

int nbg_sys_accelerations_at_noalloc(nbg_sys *s, int64_t nt, const float *tx, const float *ty, const float *tz,
                                     float *ax, float *ay, float *az) {
    if (int r = check_points(s, nt, tx, ty, tz, ax, ay, az)) return r;
    if (nt > NBG_NOALLOC_MAX_POINTS)
        return fail(NBG_ERR_ARG, "nbg_sys_accelerations_at_noalloc: more than NBG_NOALLOC_MAX_POINTS (262144) "
                                 "targets; use nbg_sys_accelerations_at");
    if (nt == 0) return NBG_OK;
    if (!s->ctx->noalloc_scratch) return fail(NBG_ERR_STATE, "the context has no device scratch");
    DeviceGuard g(s->ctx->device);
    return walk_points(s, nt, tx, ty, tz, ax, ay, az, points_carve(s->ctx->noalloc_scratch, nt));
}

int nbg_sys_step(nbg_sys *s, int32_t nsteps) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (nsteps < 0) return fail(NBG_ERR_ARG, "nsteps must be >= 0");
    if (!s->have_state) return fail(NBG_ERR_STATE, "set_state has not been called");
    DeviceGuard g(s->ctx->device);
    for (int i = 0; i < nsteps; ++i) {
        if (int r = force_pass(s)) return r;
        launch_update(s->x, s->y, s->z, s->vx, s->vy, s->vz, s->ax, s->ay, s->az, s->prm.dt, s->n, s->stream);
        mark(s, kPhases);
        s->tree_valid = false;  // positions moved after the tree was built
    }
    NBG_HIP(hipGetLastError());
    if (nsteps > 0) s->have_forces = true;
    return NBG_OK;
}

int nbg_sys_synchronize(nbg_sys *s) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    DeviceGuard g(s->ctx->device);
    return sync_and_check(s);
}

int nbg_sys_counts(nbg_sys *s, uint64_t *pp, uint64_t *np) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    unsigned long long c[2];
    NBG_HIP(hipMemcpy(c, s->counts(), sizeof c, hipMemcpyDeviceToHost));
    if (pp) *pp = c[0];
    if (np) *np = c[1];
    return NBG_OK;
}

int nbg_sys_counts4(nbg_sys *s, uint64_t *out4) {
    if (!s || !out4) return fail(NBG_ERR_ARG, "sys or out is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    unsigned long long c[4];
    NBG_HIP(hipMemcpy(c, s->counts(), sizeof c, hipMemcpyDeviceToHost));
    const bool fmm = s->prm.solver == NBG_IMPLICIT_FMM;
    out4[0] = c[0];
    out4[1] = c[1];
    out4[2] = fmm ? c[2] : 0;
    out4[3] = fmm ? c[3] : 0;
    return NBG_OK;
}

int nbg_sys_walk_stats(nbg_sys *s, uint64_t *out4) {
    if (!s || !out4) return fail(NBG_ERR_ARG, "sys or out is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    if (s->prm.solver == NBG_IMPLICIT_FMM) {
        uint32_t pk[4] = {0, 0, 0, 0};
        if (s->n > 0 && s->fmm_misc) NBG_HIP(hipMemcpy(pk, s->fmm_misc + 16, sizeof pk, hipMemcpyDeviceToHost));
        out4[0] = pk[0];
        out4[1] = pk[1];
        out4[2] = pk[2];
        out4[3] = (uint64_t)s->fmm_pool_cap;
        return NBG_OK;
    }
    unsigned long long c[6];
    NBG_HIP(hipMemcpy(c, s->counts(), sizeof c, hipMemcpyDeviceToHost));
    for (int i = 0; i < 4; ++i) out4[i] = c[2 + i];
    return NBG_OK;
}

int nbg_sys_enable_counts(nbg_sys *s, int32_t on) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    s->counting = on != 0;
    return NBG_OK;
}

int nbg_sys_enable_walk_stats(nbg_sys *s, int32_t on) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    s->walk_stats = on != 0;
    return NBG_OK;
}

int nbg_sys_set_walk_split(nbg_sys *s, int32_t waves) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4 && waves != 8 && waves != 16)
        return fail(NBG_ERR_ARG, "waves must be 0 (automatic), 1, 2, 4, 8 or 16");
    DeviceGuard g(s->ctx->device);
    const int old = s->split_req;
    s->split_req = waves;
    if (int r = ensure_split(s)) {
        s->split_req = old;
        return r;
    }
    return NBG_OK;
}

int nbg_sys_enable_timing(nbg_sys *s, int32_t on) {
    if (!s) return fail(NBG_ERR_ARG, "sys is NULL");
    DeviceGuard g(s->ctx->device);
    if (on && s->ev.empty()) {
        s->ev.assign((size_t)kEventSets * (kPhases + 1), nullptr);
        for (auto &e : s->ev)
            NBG_HIP(hipEventCreate(&e));
    }
    NBG_HIP(hipStreamSynchronize(s->stream));
    s->timing = on != 0;
    s->ev_used = 0;
    s->phase_n = 0;
    for (double &v : s->phase_sum) v = 0.0;
    return NBG_OK;
}

// Mean per-phase milliseconds over every pass since timing was enabled.
int nbg_sys_phase_times(nbg_sys *s, float *ms, int32_t cap) {
    if (!s || !ms) return fail(NBG_ERR_ARG, "sys or ms is NULL");
    DeviceGuard g(s->ctx->device);
    if (int r = sync_and_check(s)) return r;
    if (!s->timing || s->phase_n == 0) return fail(NBG_ERR_STATE, "timing is not enabled or no pass has run");
    int k = cap < kPhases ? cap : kPhases;
    for (int i = 0; i < k; ++i) ms[i] = (float)(s->phase_sum[i] / (double)s->phase_n);
    return k;
}

int nbg_sys_debug_bbox(nbg_sys *s, float *out6) {
    if (!s || !out6) return fail(NBG_ERR_ARG, "sys or out is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    NBG_HIP(hipMemcpy(out6, s->bbox, 6 * sizeof(float), hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_sys_debug_order(nbg_sys *s, uint64_t *keys, uint32_t *perm) {
    if (!s || !perm) return fail(NBG_ERR_ARG, "sys or perm is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    if (s->prm.solver == NBG_NAIVE) return fail(NBG_ERR_STATE, "the naive solver builds no tree");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    const int64_t n = s->n;
    if (n == 0) return NBG_OK;
    NBG_HIP(hipMemcpy(perm, s->idx, (size_t)n * 4, hipMemcpyDeviceToHost));
    if (keys) {
        if (s->prm.morton_bits == 30) {
            std::vector<uint32_t> k32(n);
            NBG_HIP(hipMemcpy(k32.data(), s->keys, (size_t)n * 4, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; ++i) keys[i] = k32[i];
        } else {
            NBG_HIP(hipMemcpy(keys, s->keys, (size_t)n * 8, hipMemcpyDeviceToHost));
        }
    }
    return NBG_OK;
}

int nbg_sys_debug_tree(nbg_sys *s, int64_t *num_nodes, int64_t *first, int64_t *count, int32_t *leaf,
                       int32_t *parent, float *com4, float *geo6) {
    if (!s || !num_nodes) return fail(NBG_ERR_ARG, "sys or num_nodes is NULL");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    DeviceGuard g(s->ctx->device);
    NBG_HIP(hipStreamSynchronize(s->stream));
    uint32_t nn = 0;
    const bool octree = is_octree(s->prm.solver);
    if (octree && s->n > 0) NBG_HIP(hipMemcpy(&nn, s->ob.misc, 4, hipMemcpyDeviceToHost));
    else if (!octree) NBG_HIP(hipMemcpy(&nn, s->node_count(), 4, hipMemcpyDeviceToHost));
    *num_nodes = nn;
    if (!first && !count && !leaf && !parent && !com4 && !geo6) return NBG_OK;
    if (octree) {
        std::vector<OctRec> nd(nn);
        std::vector<float4> cs(nn);
        if (nn) {
            NBG_HIP(hipMemcpy(nd.data(), s->ob.nodes, nn * sizeof(OctRec), hipMemcpyDeviceToHost));
            NBG_HIP(hipMemcpy(cs.data(), s->ob.centre, nn * sizeof(float4), hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < nn; ++i) {
            const int4 t = nd[i].topo;
            const float4 c = nd[i].com;
            if (first) first[i] = t.x;
            if (count) count[i] = t.y;
            if (leaf) leaf[i] = t.z < 0;
            if (parent) parent[i] = t.w;
            if (com4) { com4[4 * i] = c.x; com4[4 * i + 1] = c.y; com4[4 * i + 2] = c.z; com4[4 * i + 3] = c.w; }
            if (geo6) {
                float gg[6] = {cs[i].x, cs[i].y, cs[i].z, cs[i].w, 0.0f, 0.0f};
                std::memcpy(geo6 + 6 * i, gg, sizeof gg);
            }
        }
        return NBG_OK;
    }
    std::vector<NodeRec> nd(nn);
    std::vector<float4> lo(nn), hi(nn);
    if (nn) {
        NBG_HIP(hipMemcpy(nd.data(), s->nodes, nn * sizeof(NodeRec), hipMemcpyDeviceToHost));
        NBG_HIP(hipMemcpy(lo.data(), s->bmin, nn * sizeof(float4), hipMemcpyDeviceToHost));
        NBG_HIP(hipMemcpy(hi.data(), s->bmax, nn * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for (uint32_t i = 0; i < nn; ++i) {
        const int4 t = nd[i].topo;
        const float4 c = nd[i].com;
        if (first) first[i] = t.x;
        if (count) count[i] = t.y;
        if (leaf) leaf[i] = t.z < 0;
        if (parent) parent[i] = t.w;
        if (com4) { com4[4 * i] = c.x; com4[4 * i + 1] = c.y; com4[4 * i + 2] = c.z; com4[4 * i + 3] = c.w; }
        if (geo6) {
            float gg[6] = {lo[i].x, lo[i].y, lo[i].z, hi[i].x, hi[i].y, hi[i].z};
            std::memcpy(geo6 + 6 * i, gg, sizeof gg);
        }
    }
    return NBG_OK;
}

int nbg_sys_debug_quads(nbg_sys *s, int64_t cap, float *q6) {
    if (!s || !q6) return fail(NBG_ERR_ARG, "sys or q6 is NULL");
    if (!is_quad(s->prm.solver)) return fail(NBG_ERR_STATE, "the solver carries no quadrupole moments");
    if (!s->have_forces) return fail(NBG_ERR_STATE, "no force pass has run");
    int64_t nn = 0;
    if (int r = nbg_sys_debug_tree(s, &nn, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)) return r;
    if (cap < nn) return fail(NBG_ERR_ARG, "q6 holds fewer nodes than the tree has");
    if (nn == 0) return NBG_OK;
    std::vector<float4> h((size_t)(2 * nn));
    NBG_HIP(hipMemcpy(h.data(), s->quad, h.size() * sizeof(float4), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nn; ++i) {
        const float4 a = h[2 * i], b = h[2 * i + 1];
        const float v[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
        std::memcpy(q6 + 6 * i, v, sizeof v);
    }
    return NBG_OK;
}

int nbg_direct_f64(nbg_ctx *ctx, int64_t ns, const float *sx, const float *sy, const float *sz, const float *sm,
                   int64_t nt, const float *tx, const float *ty, const float *tz, double G, double eps2, double *ax,
                   double *ay, double *az) {
    if (!ctx || ns < 0 || nt < 0) return fail(NBG_ERR_ARG, "bad arguments");
    if ((ns > 0 && (!sx || !sy || !sz || !sm)) || (nt > 0 && (!tx || !ty || !tz || !ax || !ay || !az)))
        return fail(NBG_ERR_ARG, "NULL array");
    if (nt == 0) return NBG_OK;
    DeviceGuard g(ctx->device);
    std::vector<float4> hs(ns), ht(nt);
    for (int64_t i = 0; i < ns; ++i) hs[i] = make_float4(sx[i], sy[i], sz[i], sm[i]);
    for (int64_t i = 0; i < nt; ++i) ht[i] = make_float4(tx[i], ty[i], tz[i], 0.0f);
    float4 *ds = nullptr, *dt = nullptr;
    double *da = nullptr;
    hipStream_t st = nullptr;
    int r = NBG_OK;
    auto chk = [&](hipError_t e, const char *what) {
        if (r == NBG_OK && e != hipSuccess) r = fail(NBG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
    };
    chk(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
    chk(hipMalloc((void **)&ds, (size_t)(ns > 0 ? ns : 1) * sizeof(float4)), "hipMalloc");
    chk(hipMalloc((void **)&dt, (size_t)nt * sizeof(float4)), "hipMalloc");
    chk(hipMalloc((void **)&da, (size_t)nt * 3 * sizeof(double)), "hipMalloc");
    if (ns > 0) chk(hipMemcpyAsync(ds, hs.data(), ns * sizeof(float4), hipMemcpyHostToDevice, st), "H2D");
    chk(hipMemcpyAsync(dt, ht.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st), "H2D");
    if (r == NBG_OK) {
        launch_direct_f64(ds, ns, dt, nt, G, eps2, da, da + nt, da + 2 * nt, st);
        chk(hipGetLastError(), "launch");
    }
    chk(hipStreamSynchronize(st), "sync");
    std::vector<double> out(3 * nt);
    chk(hipMemcpy(out.data(), da, 3 * nt * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    if (r == NBG_OK)
        for (int64_t i = 0; i < nt; ++i) { ax[i] = out[i]; ay[i] = out[nt + i]; az[i] = out[2 * nt + i]; }
    if (ds) (void)hipFree(ds);
    if (dt) (void)hipFree(dt);
    if (da) (void)hipFree(da);
    if (st) (void)hipStreamDestroy(st);
    return r;
}

int nbg_sys_potential_f64(nbg_sys *s, double *W) {
    if (!s || !W) return fail(NBG_ERR_ARG, "sys or W is NULL");
    DeviceGuard g(s->ctx->device);
    const int64_t n = s->n;
    *W = 0.0;
    if (n < 2) return NBG_OK;
    const int64_t nw = potential_waves(n);
    double *part = nullptr;
    double4 *packed = nullptr;
    NBG_HIP(hipMalloc((void **)&part, (size_t)nw * sizeof(double)));
    hipError_t e = hipMalloc((void **)&packed, (size_t)n * sizeof(double4));
    if (e == hipSuccess) {
        launch_pack_f64(s->x, s->y, s->z, s->m, n, packed, s->stream);
        // Launches of about kPairsPerLaunch pairs each (wave w holds about
        // 64 (n - 64 w) of them), each waited for, so no single kernel runs
        // long even at tens of millions of bodies.
        constexpr double kPairsPerLaunch = 1e12;
        int64_t w0 = 0;
        while (w0 < nw && e == hipSuccess) {
            int64_t w1 = w0;
            double pairs = 0.0;
            while (w1 < nw && (w1 == w0 || pairs < kPairsPerLaunch)) {
                pairs += 64.0 * (double)(n - 64 * w1);
                ++w1;
            }
            launch_potential_f64(packed, n, (double)s->prm.eps2, w0, w1, part, s->stream);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
            w0 = w1;
        }
    }
    std::vector<double> h(nw);
    if (e == hipSuccess) e = hipMemcpy(h.data(), part, nw * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(part);
    if (packed) (void)hipFree(packed);
    if (e != hipSuccess) return fail(NBG_ERR_HIP, std::string("potential: ") + hipGetErrorString(e));
    double w = 0.0;
    for (double v : h) w += v;
    *W = -(double)s->prm.G * w;
    return NBG_OK;
}

}  // extern "C"
