// C ABI of the batched forest: many independent systems stepped together.
// Host code only.  The systems' states are concatenated in HBM in membership
// order; one pass builds every system's tree and walks every system in one
// launch, with wavefronts aligned to systems.  The members share one solver:
//   * the linear BVH (LinearBVHSolver): segmented boxes, one sort of
//     system-prefixed 30-bit keys, one radix-tree build, one summary pass;
//   * the Barnes-Hut octree (BarnesHutSolver): segmented boxes, each body's
//     octant path over its own system's cube and depth, one sort by (system,
//     path), one octree build in which systems are a digit above the root,
//     one summary pass;
// each monopole or quadrupole (second moments summarized with the tree,
// quadrupole walk).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nbg.h"
#include "host_util.hpp"
#include "launch.hpp"

using namespace nbg;

namespace {

struct Member {
    int32_t id;
    int64_t n;
    nbg_params p;
    int64_t first;  // offset in the concatenated state
};

// Per-body device arrays of a forest and the bytes each takes.
constexpr int64_t kStateArrays = 10;   // x y z m vx vy vz ax ay az
constexpr int64_t kStageArrays = 7;    // compaction staging for x y z m vx vy vz

}  // namespace

struct nbg_batch {
    nbg_ctx *ctx = nullptr;
    int64_t cap_n = 0;
    int32_t cap_s = 0;
    int32_t sum_mode = NBG_SUM_FAST;
    hipStream_t stream = nullptr;
    std::vector<Member> sys;
    int32_t next_id = 0;
    int64_t n = 0;
    int64_t nwaves = 0;
    bool tables_dirty = true;
    bool have_forces = false;
    bool counting = true;
    bool use_graph = false;
    bool all_soft = true;
    // every member runs the same solver: NBG_LINEAR_BVH or
    // NBG_LINEAR_BVH_QUADRUPOLE (set by the first member of an empty batch)
    int32_t solver = NBG_LINEAR_BVH;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_valid = false;

    float *st[kStateArrays] = {};
    float *stage[kStageArrays] = {};
    uint32_t *bsys = nullptr;
    uint64_t *keys = nullptr, *keys_alt = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;
    float4 *bodies = nullptr;
    int32_t *kfirst = nullptr, *klast = nullptr, *kgamma = nullptr, *kparent = nullptr;
    uint32_t *ksplit = nullptr, *krank = nullptr;
    int32_t *leaf_start = nullptr;
    uint32_t *glist = nullptr, *scan_ws = nullptr, *sort_status = nullptr;
    NodeRec *nodes = nullptr;
    float4 *bmin = nullptr, *bmax = nullptr;
    uint32_t *flags = nullptr;
    // per-system tables
    int64_t *sys_first = nullptr, *wave_first = nullptr, *old_first = nullptr;
    int32_t *sys_root = nullptr, *leaf_tab = nullptr;
    float4 *sys_prm = nullptr;
    MacParam *mp_tab = nullptr;
    float *bbox6 = nullptr;
    uint32_t *misc = nullptr;  // [1] err, [2] node count, [4..16) counters
    uint32_t *bad = nullptr;   // one word per system slot: the non-finite guard's flags
    // second moments of every node (2 float4 per node slot), allocated when
    // the first quadrupole member joins (linear BVH) or with the octree
    float4 *quad = nullptr;
    int64_t quad_slots = 0;
    // octree members: work and nodes sized for the members' node bound
    // (allocated when the tables are uploaded), levels per system
    OctBuffers ob;
    int32_t *lv_tab = nullptr;
    int max_levels = 0;
    // targets per lane of the FAST linear-BVH walk (the autotuner's
    // bodies-per-wavefront choice: 64 x tpl); the other walks use 1
    int32_t tpl = 1;
    // split FAST walks of small forests (several waves per group of 64
    // targets): per-target sums for up to kSplitMaxBodies bodies and one
    // block count per wave; split_req 0 = by size, 1 = off
    double *split_part = nullptr;
    int32_t *split_nb = nullptr;
    int64_t split_bodies = 0, split_nb_cap = 0;
    int32_t split_req = 0;

    float *x() { return st[0]; }
    float *y() { return st[1]; }
    float *z() { return st[2]; }
    float *m() { return st[3]; }
    float *vx() { return st[4]; }
    float *vy() { return st[5]; }
    float *vz() { return st[6]; }
    float *ax() { return st[7]; }
    float *ay() { return st[8]; }
    float *az() { return st[9]; }
    uint32_t *err_word() { return misc + 1; }
    uint32_t *node_count() { return misc + 2; }
    unsigned long long *counts() { return (unsigned long long *)(misc + 4); }
};

namespace {

int64_t bytes_for(int64_t cap_n, int64_t cap_s) {
    const int64_t n = cap_n > 0 ? cap_n : 1, s = cap_s > 0 ? cap_s : 1;
    int64_t b = 0;
    b += (kStateArrays + kStageArrays) * 4 * n;               // state + staging
    b += 4 * n;                                               // bsys
    b += 2 * 8 * n + 2 * 4 * n + 16 * n;                      // keys, idx, bodies
    b += 6 * 4 * n + 4 * n;                                   // Karras scratch, leaf_start
    b += 4 * 2 * (2 * n + 64) + 4 * (128 + n / kScanTile);    // glist (two queues), scan_ws
    b += 4 * sort_status_words(n, 64);                        // sort status
    b += (int64_t)sizeof(NodeRec) * 2 * n + 2 * 16 * 2 * n + 4 * 2 * n;  // nodes, bmin/bmax, flags
    b += 3 * 8 * (s + 1) + 2 * 4 * s + 16 * s + (int64_t)sizeof(MacParam) * s + 24 * s + 4 * 32 + 4 * s;
    const int64_t ns = std::min(n, kSplitMaxBodies);
    b += 8 * kSplitSlots * 3 * ns + 4 * (ns / 64 + s + 1);   // split walk sums, block counts
    return b;
}

bool is_octree_solver(int32_t solver) { return solver == NBG_BARNES_HUT || solver == NBG_BARNES_HUT_QUADRUPOLE; }

void free_octree(OctBuffers &o) {
    void *ptrs[] = {o.key_hi, o.key_lo, o.lo_sorted, o.nown, o.nbase, o.nfirst, o.ncount, o.nparent, o.nmeta,
                    o.occ, o.cb, o.flags, o.glist, o.scan_ws, o.nodes, o.centre, o.thr, o.misc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    o = OctBuffers{};
}

void free_batch(nbg_batch *b) {
    free_octree(b->ob);
    if (b->lv_tab) (void)hipFree(b->lv_tab);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    std::vector<void *> ptrs = {b->bsys, b->keys, b->keys_alt, b->idx, b->idx_alt, b->bodies, b->kfirst, b->klast,
                                b->kgamma, b->kparent, b->ksplit, b->krank, b->leaf_start, b->glist, b->scan_ws,
                                b->sort_status, b->nodes, b->bmin, b->bmax, b->flags, b->sys_first, b->wave_first,
                                b->old_first, b->sys_root, b->leaf_tab, b->sys_prm, b->mp_tab, b->bbox6, b->misc,
                                b->quad, b->bad, b->split_part, b->split_nb};
    for (float *p : b->st) ptrs.push_back(p);
    for (float *p : b->stage) ptrs.push_back(p);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

int check_member_params(const nbg_params *p) {
    if (!p) return fail(NBG_ERR_ARG, "params is NULL");
    const bool octree = is_octree_solver(p->solver);
    if (!octree && p->solver != NBG_LINEAR_BVH && p->solver != NBG_LINEAR_BVH_QUADRUPOLE)
        return fail(NBG_ERR_ARG, "the batched forest runs the Barnes-Hut octree and linear-BVH solvers "
                                 "(monopole or quadrupole)");
    if (!octree && p->morton_bits != 30)
        return fail(NBG_ERR_ARG, "the batched forest's linear BVH uses 30-bit Morton keys");
    if (octree && (p->max_depth < 0 || p->max_depth > 32)) return fail(NBG_ERR_ARG, "max_depth must be in [0, 32]");
    if (!(p->theta >= 0.0f)) return fail(NBG_ERR_ARG, "theta must be >= 0");
    if (p->leaf_size < 1) return fail(NBG_ERR_ARG, "leaf_size must be >= 1");
    if (!(p->eps2 >= 0.0f)) return fail(NBG_ERR_ARG, "eps2 must be >= 0");
    return NBG_OK;
}

// Members share one solver; `self` (or null) is the member being changed.
// A quadrupole batch allocates its node moments here, outside any pass.
// Targets per lane the walk of this batch uses.
int walk_tpl(const nbg_batch *b) {
    const bool lbvh = b->solver == NBG_LINEAR_BVH || b->solver == NBG_LINEAR_BVH_QUADRUPOLE;
    return lbvh && b->sum_mode == NBG_SUM_FAST ? b->tpl : 1;
}

// Waves per group of the forest's FAST walk: a forest of at most
// kSplitMaxBodies bodies in all (one target per lane) fills the chip by
// splitting each group's frontier blocks over several waves (by size up
// to 4 for the linear BVH, 8 for the octree), as a single small system does.  By size, only when at least half
// of the bodies are in systems of kSplitMinMember bodies or more: a tiny
// system's tree has too few frontier blocks to share, and its extra waves
// cost more than they save.  The block sums, and so the bits, do not depend
// on the split.
constexpr int64_t kSplitMinMember = 4096;

int forest_split(const nbg_batch *b) {
    if (b->sum_mode != NBG_SUM_FAST || !b->split_part || b->split_req == 1 || walk_tpl(b) != 1) return 1;
    if (b->n > b->split_bodies || b->nwaves > b->split_nb_cap) return 1;
    if (b->split_req == 0) {
        int64_t big = 0;
        for (const Member &m : b->sys)
            if (m.n >= kSplitMinMember) big += m.n;
        if (2 * big < b->n) return 1;
    }
    int k = b->split_req > 1 ? b->split_req : split_waves_groups(b->nwaves);
    if (is_octree_solver(b->solver)) k = std::min(k, 8);
    else if (b->split_req == 0) k = std::min(k, kSplitAutoLbvh);
    return k;
}

void set_split(const nbg_batch *b, WalkArgs &wa) {
    wa.split = forest_split(b);
    if (wa.split > 1) {
        wa.split_part = b->split_part;
        wa.split_nb = b->split_nb;
        wa.split_stride = b->n;
    }
}

int ensure_quad(nbg_batch *b, int64_t slots) {
    if (b->quad && b->quad_slots >= slots) return NBG_OK;
    NBG_HIP(hipStreamSynchronize(b->stream));
    if (b->quad) (void)hipFree(b->quad);
    b->quad = nullptr;
    b->quad_slots = 0;
    if (dalloc(&b->quad, 2 * slots) != NBG_OK) {
        b->quad = nullptr;
        (void)hipGetLastError();  // the failed allocation must not fail the next launch check
        return fail(NBG_ERR_BUDGET, "no device memory for the quadrupole moments");
    }
    b->quad_slots = slots;
    return NBG_OK;
}

int check_member_solver(nbg_batch *b, const nbg_params *p, const Member *self) {
    for (const Member &m : b->sys)
        if (&m != self && m.p.solver != p->solver)
            return fail(NBG_ERR_ARG, "the systems of a batch share one solver");
    // the octree's moments are sized with its nodes (upload_tables)
    if (p->solver == NBG_LINEAR_BVH_QUADRUPOLE) return ensure_quad(b, 2 * b->cap_n);
    return NBG_OK;
}

// Octree work and node arrays for the current members (their node bound is
// the sum of each system's), and the quadrupole moments with them.
int ensure_octree(nbg_batch *b) {
    int64_t cap = 0;
    for (const Member &m : b->sys) cap += octree_capacity(m.n, m.p.leaf_size, m.p.max_depth);
    if (cap >= ((int64_t)1 << 28))
        return fail(NBG_ERR_BUDGET, "the forest's octree node bound exceeds 2^28 slots; raise leaf sizes");
    OctBuffers &o = b->ob;
    if (o.cap < cap) {
        NBG_HIP(hipStreamSynchronize(b->stream));
        free_octree(o);
        const int64_t n = b->cap_n;
        int r = NBG_OK;
        auto A = [&](auto **ptr, int64_t count) {
            if (r == NBG_OK) r = dalloc(ptr, count);
        };
        A(&o.key_hi, n); A(&o.key_lo, n); A(&o.lo_sorted, n); A(&o.nown, n); A(&o.nbase, n);
        A(&o.nfirst, cap); A(&o.ncount, cap); A(&o.nparent, cap); A(&o.nmeta, cap);
        A(&o.occ, cap); A(&o.cb, cap); A(&o.flags, cap); A(&o.glist, 2 * (cap + 64));
        A(&o.scan_ws, 128 + (cap + kScanTile - 1) / kScanTile);
        A(&o.nodes, cap); A(&o.centre, cap); A(&o.thr, (int64_t)kOctThrStride * b->cap_s); A(&o.misc, 64);
        if (r != NBG_OK) {
            free_octree(o);
            (void)hipGetLastError();
            return fail(NBG_ERR_BUDGET, "no device memory for the forest's octree");
        }
        o.cap = cap;
        o.n = n;
        NBG_HIP(hipMemset(o.misc, 0, 64 * sizeof(uint32_t)));
    }
    if (b->solver == NBG_BARNES_HUT_QUADRUPOLE) return ensure_quad(b, o.cap);
    return NBG_OK;
}

// Uploads the per-system tables after a membership or parameter change.
int upload_tables(nbg_batch *b) {
    if (!b->tables_dirty) return NBG_OK;
    const size_t S = b->sys.size();
    std::vector<int64_t> first(S + 1), waves(S + 1);
    std::vector<int32_t> leaf(S), levels(S), roots(S);
    std::vector<float4> prm(S);
    std::vector<MacParam> mp(S);
    int64_t off = 0, w = 0;
    b->all_soft = true;
    b->max_levels = 0;
    if (S) b->solver = b->sys[0].p.solver;
    const int64_t width = 64 * (int64_t)walk_tpl(b);  // targets per wave
    for (size_t i = 0; i < S; ++i) {
        Member &m = b->sys[i];
        m.first = off;
        first[i] = off;
        waves[i] = w;
        off += m.n;
        w += (m.n + width - 1) / width;
        leaf[i] = m.p.leaf_size;
        levels[i] = m.p.max_depth + 2;
        if (levels[i] > b->max_levels) b->max_levels = levels[i];
        roots[i] = (int32_t)i;  // an octree forest puts system i's root in slot i
        prm[i] = make_float4(m.p.G, m.p.eps2, m.p.theta, m.p.dt);
        mac_threshold(m.p.theta, &mp[i].mid, &mp[i].tie);
        if (!(m.p.eps2 > 0.0f)) b->all_soft = false;
    }
    if (S && is_octree_solver(b->solver))
        if (int r = ensure_octree(b)) return r;
    first[S] = off;
    waves[S] = w;
    b->n = off;
    b->nwaves = w;
    hipStream_t s = b->stream;
    NBG_HIP(hipMemcpyAsync(b->sys_first, first.data(), (S + 1) * 8, hipMemcpyHostToDevice, s));
    NBG_HIP(hipMemcpyAsync(b->wave_first, waves.data(), (S + 1) * 8, hipMemcpyHostToDevice, s));
    if (S) {
        NBG_HIP(hipMemcpyAsync(b->leaf_tab, leaf.data(), S * 4, hipMemcpyHostToDevice, s));
        NBG_HIP(hipMemcpyAsync(b->sys_prm, prm.data(), S * sizeof(float4), hipMemcpyHostToDevice, s));
        NBG_HIP(hipMemcpyAsync(b->mp_tab, mp.data(), S * sizeof(MacParam), hipMemcpyHostToDevice, s));
        NBG_HIP(hipMemcpyAsync(b->lv_tab, levels.data(), S * 4, hipMemcpyHostToDevice, s));
        if (is_octree_solver(b->solver))
            NBG_HIP(hipMemcpyAsync(b->sys_root, roots.data(), S * 4, hipMemcpyHostToDevice, s));
    }
    for (size_t i = 0; i < S; ++i)
        launch_fill_u32(b->bsys + b->sys[i].first, (uint32_t)i, b->sys[i].n, s);
    NBG_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    b->tables_dirty = false;
    b->graph_valid = false;
    return NBG_OK;
}

// The octree form of enqueue_pass: every system's Barnes-Hut octree
// (BarnesHutSolver, Solvers/BarnesHutSolver.h:14-32) in one build, walked
// in one launch.
void enqueue_pass_octree(nbg_batch *b) {
    const int64_t n = b->n;
    const int S = (int)b->sys.size();
    hipStream_t s = b->stream;
    OctBuffers &ob = b->ob;
    launch_seg_bbox(b->x(), b->y(), b->z(), S, b->sys_first, b->bbox6, s);
    launch_octree_forest_build(b->x(), b->y(), b->z(), n, S, b->bsys, b->bbox6, b->lv_tab, b->max_levels, ob,
                               b->keys, b->idx, b->keys_alt, b->idx_alt, b->sort_status, b->err_word(), s);
    launch_octree_forest_topology(n, S, b->bbox6, b->sys_prm, b->lv_tab, b->leaf_tab, b->sys_first, b->max_levels,
                                  ob, b->keys, b->keys_alt, b->idx_alt, b->idx, b->leaf_start, b->err_word(), s);
    launch_gather(b->x(), b->y(), b->z(), b->m(), b->idx, n, b->bodies, s);
    float4 *quad = b->solver == NBG_BARNES_HUT_QUADRUPOLE ? b->quad : nullptr;
    launch_octree_summaries(b->bodies, n, ob, b->leaf_start, b->err_word(), s, quad);
    WalkArgs wa;
    wa.n = n;
    wa.bodies = b->bodies;
    wa.nt = n;
    wa.targets = b->bodies;
    wa.perm = b->idx;
    wa.nodes = nullptr;
    wa.onodes = ob.nodes;
    wa.levels = b->max_levels;
    wa.theta = 0.0f;
    wa.theta2 = 0.0f;
    wa.mac_mid = 0.0;
    wa.mac_tie = 0;
    wa.G = 1.0f;
    wa.eps2 = b->all_soft ? 1.0f : 0.0f;  // see enqueue_pass
    wa.max_nodes = ob.cap;
    wa.ax = b->ax();
    wa.ay = b->ay();
    wa.az = b->az();
    wa.counts = b->counts();
    wa.err = b->err_word();
    wa.nsys = S;
    wa.nwaves = b->nwaves;
    wa.wave_first = b->wave_first;
    wa.sys_first = b->sys_first;
    wa.sys_root = b->sys_root;
    wa.sys_prm = b->sys_prm;
    wa.quad = quad;
    set_split(b, wa);
    const bool ordered = b->sum_mode == NBG_SUM_ORDERED;
    launch_walk_oct(wa, ordered, (b->counting || ordered) ? 1 : 0, s);
}

// One force pass over every system (enqueue only).
void enqueue_pass(nbg_batch *b) {
    const int64_t n = b->n;
    const int S = (int)b->sys.size();
    hipStream_t s = b->stream;
    launch_fill_u32(b->misc + 4, 0u, 12, s);
    if (n == 0) return;
    if (is_octree_solver(b->solver)) {
        enqueue_pass_octree(b);
        return;
    }
    launch_forest_build(b->x(), b->y(), b->z(), n, S, b->sys_first, b->bsys, b->bbox6, b->keys, b->idx, b->keys_alt,
                        b->idx_alt, b->sort_status, b->err_word(), s);
    launch_lbvh_topology(b->keys, 63, n, 0, b->kfirst, b->klast, b->kgamma, b->ksplit, b->krank, b->kparent,
                         b->scan_ws, b->nodes, b->leaf_start, b->node_count(), s, 30, b->leaf_tab);
    launch_gather(b->x(), b->y(), b->z(), b->m(), b->idx, n, b->bodies, s);
    float4 *quad = b->solver == NBG_LINEAR_BVH_QUADRUPOLE ? b->quad : nullptr;
    launch_lbvh_summaries(b->bodies, n, b->nodes, b->leaf_start, 2 * n, b->flags, b->glist, b->bmin, b->bmax, 0.5f,
                          b->err_word(), s, b->mp_tab, b->keys, 30, quad);
    launch_forest_roots(b->nodes, b->node_count(), 2 * n, b->keys, n, S, b->sys_root, b->err_word(), s);
    WalkArgs wa;
    wa.n = n;
    wa.bodies = b->bodies;
    wa.nt = n;
    wa.targets = b->bodies;
    wa.perm = b->idx;
    wa.nodes = b->nodes;
    wa.onodes = nullptr;
    wa.levels = 0;
    wa.theta = 0.0f;
    wa.theta2 = 0.0f;
    wa.mac_mid = 0.0;
    wa.mac_tie = 0;
    wa.G = 1.0f;
    // the FAST walk is instantiated for eps2 > 0 (every pair has r2 > 0) or
    // with the r2 = 0 guard; the per-system values come from sys_prm
    wa.eps2 = b->all_soft ? 1.0f : 0.0f;
    wa.max_nodes = 2 * n;
    wa.ax = b->ax();
    wa.ay = b->ay();
    wa.az = b->az();
    wa.counts = b->counts();
    wa.err = b->err_word();
    wa.nsys = S;
    wa.nwaves = b->nwaves;
    wa.wave_first = b->wave_first;
    wa.sys_first = b->sys_first;
    wa.sys_root = b->sys_root;
    wa.sys_prm = b->sys_prm;
    wa.quad = quad;
    wa.tpl = walk_tpl(b);
    set_split(b, wa);
    const bool ordered = b->sum_mode == NBG_SUM_ORDERED;
    launch_walk_lbvh(wa, ordered, 32, (b->counting || ordered) ? 1 : 0, s);
}

void enqueue_step(nbg_batch *b) {
    enqueue_pass(b);
    launch_update_forest(b->x(), b->y(), b->z(), b->vx(), b->vy(), b->vz(), b->ax(), b->ay(), b->az(), b->bsys,
                         b->sys_prm, b->n, b->stream);
}

int sync_check(nbg_batch *b) {
    NBG_HIP(hipStreamSynchronize(b->stream));
    uint32_t err = 0;
    NBG_HIP(hipMemcpy(&err, b->err_word(), 4, hipMemcpyDeviceToHost));
    if (err) {
        uint32_t zero = 0;
        (void)hipMemcpy(b->err_word(), &zero, 4, hipMemcpyHostToDevice);
        return fail(NBG_ERR_DEVICE, "forest pass reported device error flags " + std::to_string(err));
    }
    return NBG_OK;
}

Member *find(nbg_batch *b, int32_t id, size_t *pos = nullptr) {
    for (size_t i = 0; i < b->sys.size(); ++i)
        if (b->sys[i].id == id) {
            if (pos) *pos = i;
            return &b->sys[i];
        }
    return nullptr;
}

}  // namespace

// Moves the bodies of surviving systems to their new offsets: new position
// i of system s takes old position old_first[s] + (i - new_first[s]).
namespace nbg {
void launch_forest_compact(float *const *dst, const float *const *src, int arrays, int64_t n, int nsys,
                           const int64_t *new_first, const int64_t *old_first, hipStream_t s);
}

extern "C" {

int64_t nbg_batch_bytes(int64_t max_bodies, int32_t max_systems) {
    if (max_bodies < 0 || max_systems < 0) return fail(NBG_ERR_ARG, "negative capacity");
    return bytes_for(max_bodies, max_systems);
}

int nbg_batch_create(nbg_ctx *ctx, int64_t max_bodies, int32_t max_systems, int32_t sum_mode, nbg_batch **out) {
    if (!ctx || !out) return fail(NBG_ERR_ARG, "ctx or out is NULL");
    *out = nullptr;
    if (max_bodies < 1 || max_bodies > kMaxBodies) return fail(NBG_ERR_ARG, "max_bodies out of range");
    if (max_systems < 1 || max_systems > (1 << 24)) return fail(NBG_ERR_ARG, "max_systems out of range");
    if (sum_mode != NBG_SUM_ORDERED && sum_mode != NBG_SUM_FAST) return fail(NBG_ERR_ARG, "unknown sum mode");
    DeviceGuard g(ctx->device);
    nbg_batch *b = new nbg_batch;
    b->ctx = ctx;
    b->cap_n = max_bodies;
    b->cap_s = max_systems;
    b->sum_mode = sum_mode;
    const int64_t n = max_bodies, S = max_systems;
    int r = NBG_OK;
    auto A = [&](auto **ptr, int64_t count) {
        if (r == NBG_OK) r = dalloc(ptr, count);
    };
    for (auto &p : b->st) A(&p, n);
    for (auto &p : b->stage) A(&p, n);
    A(&b->bsys, n);
    A(&b->keys, n); A(&b->keys_alt, n); A(&b->idx, n); A(&b->idx_alt, n); A(&b->bodies, n);
    A(&b->kfirst, n); A(&b->klast, n); A(&b->kgamma, n); A(&b->kparent, n); A(&b->ksplit, n); A(&b->krank, n);
    A(&b->leaf_start, n); A(&b->glist, 2 * (2 * n + 64)); A(&b->scan_ws, 128 + n / kScanTile);
    A(&b->sort_status, sort_status_words(n, 64));
    A(&b->nodes, 2 * n); A(&b->bmin, 2 * n); A(&b->bmax, 2 * n); A(&b->flags, 2 * n);
    A(&b->sys_first, S + 1); A(&b->wave_first, S + 1); A(&b->old_first, S + 1);
    A(&b->sys_root, S); A(&b->leaf_tab, S); A(&b->sys_prm, S); A(&b->mp_tab, S); A(&b->bbox6, 6 * S);
    A(&b->lv_tab, S);
    A(&b->misc, 32);
    A(&b->bad, S);
    if (sum_mode == NBG_SUM_FAST) {
        b->split_bodies = std::min(n, kSplitMaxBodies);
        b->split_nb_cap = b->split_bodies / 64 + S + 1;
        A(&b->split_part, (int64_t)kSplitSlots * 3 * b->split_bodies);
        A(&b->split_nb, b->split_nb_cap);
    }
    if (r == NBG_OK) {
        hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipMemsetAsync(b->misc, 0, 32 * 4, b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) r = fail(NBG_ERR_HIP, std::string("batch init: ") + hipGetErrorString(e));
    }
    if (r != NBG_OK) {
        free_batch(b);
        delete b;
        return r == NBG_ERR_HIP ? NBG_ERR_BUDGET : r;
    }
    *out = b;
    return NBG_OK;
}

int nbg_batch_destroy(nbg_batch *b) {
    if (!b) return NBG_OK;
    DeviceGuard g(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    delete b;
    return NBG_OK;
}

int nbg_batch_add(nbg_batch *b, const nbg_params *p, int64_t n, const float *x, const float *y, const float *z,
                  const float *m, const float *vx, const float *vy, const float *vz, int32_t *id) {
    if (!b || !id) return fail(NBG_ERR_ARG, "batch or id is NULL");
    if (int r = check_member_params(p)) return r;
    if (n < 1) return fail(NBG_ERR_ARG, "a system needs at least one body");
    if (!x || !y || !z || !m) return fail(NBG_ERR_ARG, "positions and masses are required");
    int64_t used = 0;
    for (const Member &mm : b->sys) used += mm.n;
    if ((int64_t)b->sys.size() >= b->cap_s) return fail(NBG_ERR_BUDGET, "batch system capacity reached");
    if (used + n > b->cap_n) return fail(NBG_ERR_BUDGET, "batch body capacity reached");
    DeviceGuard g(b->ctx->device);
    if (int r = check_member_solver(b, p, nullptr)) return r;
    hipStream_t s = b->stream;
    NBG_HIP(hipStreamSynchronize(s));
    const size_t bytes = (size_t)n * 4;
    const float *h[7] = {x, y, z, m, vx, vy, vz};
    for (int k = 0; k < 7; ++k) {
        if (h[k]) NBG_HIP(hipMemcpyAsync(b->st[k] + used, h[k], bytes, hipMemcpyHostToDevice, s));
        else launch_fill_u32((uint32_t *)(b->st[k] + used), 0u, n, s);
    }
    for (int k = 7; k < 10; ++k) launch_fill_u32((uint32_t *)(b->st[k] + used), 0u, n, s);
    NBG_HIP(hipStreamSynchronize(s));
    Member mm;
    mm.id = b->next_id++;
    mm.n = n;
    mm.p = *p;
    mm.first = used;
    b->sys.push_back(mm);
    b->tables_dirty = true;
    *id = mm.id;
    return NBG_OK;
}

int nbg_batch_remove(nbg_batch *b, int32_t id) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    size_t pos = 0;
    if (!find(b, id, &pos)) return fail(NBG_ERR_ARG, "unknown system id");
    DeviceGuard g(b->ctx->device);
    hipStream_t s = b->stream;
    NBG_HIP(hipStreamSynchronize(s));
    // old offsets of the survivors, in membership order
    std::vector<Member> keep;
    std::vector<int64_t> old_first, new_first;
    int64_t off = 0;
    for (size_t i = 0; i < b->sys.size(); ++i) {
        if (i == pos) continue;
        keep.push_back(b->sys[i]);
        old_first.push_back(b->sys[i].first);
        new_first.push_back(off);
        off += b->sys[i].n;
    }
    new_first.push_back(off);
    old_first.push_back(0);
    for (size_t i = 0; i < keep.size(); ++i) keep[i].first = new_first[i];
    b->sys = keep;
    if (!keep.empty() && pos < keep.size()) {  // systems after the removed one move down
        NBG_HIP(hipMemcpyAsync(b->old_first, old_first.data(), old_first.size() * 8, hipMemcpyHostToDevice, s));
        NBG_HIP(hipMemcpyAsync(b->sys_first, new_first.data(), new_first.size() * 8, hipMemcpyHostToDevice, s));
        launch_forest_compact(b->stage, b->st, kStageArrays, off, (int)keep.size(), b->sys_first, b->old_first, s);
        for (int k = 0; k < kStageArrays; ++k) std::swap(b->st[k], b->stage[k]);
        NBG_HIP(hipStreamSynchronize(s));
    }
    b->tables_dirty = true;
    b->have_forces = false;  // accelerations of the survivors were not moved
    return NBG_OK;
}

int nbg_batch_info(nbg_batch *b, int32_t *nsys, int64_t *nbodies) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    int64_t used = 0;
    for (const Member &mm : b->sys) used += mm.n;
    if (nsys) *nsys = (int32_t)b->sys.size();
    if (nbodies) *nbodies = used;
    return NBG_OK;
}

int nbg_batch_set_params(nbg_batch *b, int32_t id, const nbg_params *p) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    if (int r = check_member_params(p)) return r;
    Member *mm = find(b, id);
    if (!mm) return fail(NBG_ERR_ARG, "unknown system id");
    DeviceGuard g(b->ctx->device);
    if (int r = check_member_solver(b, p, mm)) return r;
    mm->p = *p;
    b->tables_dirty = true;
    return NBG_OK;
}

int nbg_batch_get(nbg_batch *b, int32_t id, float *x, float *y, float *z, float *vx, float *vy, float *vz, float *ax,
                  float *ay, float *az) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    Member *mm = find(b, id);
    if (!mm) return fail(NBG_ERR_ARG, "unknown system id");
    DeviceGuard g(b->ctx->device);
    if (int r = sync_check(b)) return r;
    if ((ax || ay || az) && !b->have_forces) return fail(NBG_ERR_STATE, "no force pass since the last change");
    float *h[9] = {x, y, z, vx, vy, vz, ax, ay, az};
    const int map[9] = {0, 1, 2, 4, 5, 6, 7, 8, 9};
    // offsets may be stale if membership changed after the last pass
    int64_t first = 0;
    for (const Member &o : b->sys) {
        if (o.id == id) break;
        first += o.n;
    }
    for (int k = 0; k < 9; ++k)
        if (h[k]) NBG_HIP(hipMemcpy(h[k], b->st[map[k]] + first, (size_t)mm->n * 4, hipMemcpyDeviceToHost));
    return NBG_OK;
}

int nbg_batch_update_accelerations(nbg_batch *b) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    DeviceGuard g(b->ctx->device);
    if (int r = upload_tables(b)) return r;
    enqueue_pass(b);
    NBG_HIP(hipGetLastError());
    if (int r = sync_check(b)) return r;
    b->have_forces = true;
    return NBG_OK;
}

int nbg_batch_step(nbg_batch *b, int32_t nsteps) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    if (nsteps < 0) return fail(NBG_ERR_ARG, "nsteps must be >= 0");
    if (nsteps == 0 || b->sys.empty()) return NBG_OK;
    DeviceGuard g(b->ctx->device);
    if (int r = upload_tables(b)) return r;
    if (b->use_graph) {
        if (!b->graph_valid) {
            if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
            if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
            NBG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            enqueue_step(b);
            hipGraph_t gr = nullptr;
            NBG_HIP(hipStreamEndCapture(b->stream, &gr));
            b->graph = gr;
            NBG_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
            b->graph_valid = true;
        }
        for (int i = 0; i < nsteps; ++i) NBG_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        for (int i = 0; i < nsteps; ++i) enqueue_step(b);
        NBG_HIP(hipGetLastError());
    }
    b->have_forces = true;
    return NBG_OK;
}

int nbg_batch_synchronize(nbg_batch *b) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    DeviceGuard g(b->ctx->device);
    return sync_check(b);
}

int nbg_batch_counts(nbg_batch *b, uint64_t *pp, uint64_t *np) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    DeviceGuard g(b->ctx->device);
    if (int r = sync_check(b)) return r;
    unsigned long long c[2];
    NBG_HIP(hipMemcpy(c, b->counts(), sizeof c, hipMemcpyDeviceToHost));
    if (pp) *pp = c[0];
    if (np) *np = c[1];
    return NBG_OK;
}

int nbg_batch_nonfinite(nbg_batch *b, int32_t *ids, int32_t cap, int32_t *count) {
    if (!b || !count || (cap > 0 && !ids)) return fail(NBG_ERR_ARG, "bad arguments");
    *count = 0;
    if (b->sys.empty()) return NBG_OK;
    DeviceGuard g(b->ctx->device);
    if (int r = upload_tables(b)) return r;  // bsys must describe the current membership
    const int64_t S = (int64_t)b->sys.size();
    hipStream_t s = b->stream;
    launch_fill_u32(b->bad, 0u, S, s);
    launch_forest_nonfinite(b->x(), b->y(), b->z(), b->vx(), b->vy(), b->vz(), b->bsys, b->n, b->bad, s);
    NBG_HIP(hipGetLastError());
    if (int r = sync_check(b)) return r;
    std::vector<uint32_t> h((size_t)S);
    NBG_HIP(hipMemcpy(h.data(), b->bad, (size_t)S * 4, hipMemcpyDeviceToHost));
    int32_t k = 0;
    for (int64_t i = 0; i < S; ++i)
        if (h[(size_t)i]) {
            if (k < cap) ids[k] = b->sys[(size_t)i].id;
            ++k;
        }
    *count = k;
    return NBG_OK;
}

int nbg_batch_enable_counts(nbg_batch *b, int32_t on) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    b->counting = on != 0;
    b->graph_valid = false;
    return NBG_OK;
}

int nbg_batch_use_graph(nbg_batch *b, int32_t on) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    b->use_graph = on != 0;
    b->graph_valid = false;
    return NBG_OK;
}

int nbg_batch_autotune_grid(nbg_batch *b, const int32_t *leaf_sizes, int32_t kl, const int32_t *tpls, int32_t kt,
                            int32_t repeats, int32_t *best_leaf, int32_t *best_tpl, float *ms_out) {
    if (!b || !leaf_sizes || kl < 1 || !tpls || kt < 1 || repeats < 1)
        return fail(NBG_ERR_ARG, "bad autotune arguments");
    for (int i = 0; i < kl; ++i)
        if (leaf_sizes[i] < 1) return fail(NBG_ERR_ARG, "leaf sizes must be >= 1");
    for (int j = 0; j < kt; ++j)
        if (tpls[j] != 1 && tpls[j] != 2 && tpls[j] != 4) return fail(NBG_ERR_ARG, "targets per lane: 1, 2 or 4");
    if (b->sys.empty()) return fail(NBG_ERR_STATE, "the batch is empty");
    DeviceGuard g(b->ctx->device);
    hipEvent_t e0, e1;
    NBG_HIP(hipEventCreate(&e0));
    NBG_HIP(hipEventCreate(&e1));
    const bool counting = b->counting;
    b->counting = false;
    float best_ms = 0.0f;
    bool have = false;
    int32_t bl = leaf_sizes[0], bt = b->tpl;
    int r = NBG_OK;
    for (int i = 0; i < kl && r == NBG_OK; ++i) {
        for (int j = 0; j < kt && r == NBG_OK; ++j) {
            if (ms_out) ms_out[i * kt + j] = -1.0f;
            b->tpl = tpls[j];
            // a walk without a multi-target form runs one target per lane:
            // its other candidates would time the same schedule again
            if (walk_tpl(b) != tpls[j]) continue;
            for (Member &mm : b->sys) mm.p.leaf_size = leaf_sizes[i];
            b->tables_dirty = true;
            if ((r = upload_tables(b))) break;
            enqueue_pass(b);  // warm-up
            (void)hipEventRecord(e0, b->stream);
            for (int t = 0; t < repeats; ++t) enqueue_pass(b);
            (void)hipEventRecord(e1, b->stream);
            if ((r = sync_check(b))) break;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= (float)repeats;
            if (ms_out) ms_out[i * kt + j] = ms;
            if (!have || ms < best_ms) { best_ms = ms; bl = leaf_sizes[i]; bt = tpls[j]; have = true; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->counting = counting;
    for (Member &mm : b->sys) mm.p.leaf_size = bl;
    b->tpl = bt;
    b->tables_dirty = true;
    b->have_forces = false;
    if (best_leaf) *best_leaf = bl;
    if (best_tpl) *best_tpl = bt;
    return r;
}

int nbg_batch_autotune(nbg_batch *b, const int32_t *leaf_sizes, int32_t k, int32_t repeats, int32_t *best,
                       float *ms_out) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    const int32_t tpl = b->tpl;
    return nbg_batch_autotune_grid(b, leaf_sizes, k, &tpl, 1, repeats, best, nullptr, ms_out);
}

int nbg_batch_set_walk_split(nbg_batch *b, int32_t waves) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    if (waves != 0 && waves != 1 && waves != 2 && waves != 4 && waves != 8 && waves != 16)
        return fail(NBG_ERR_ARG, "walk split: 0 (by size), 1, 2, 4, 8 or 16 waves");
    b->split_req = waves;
    b->tables_dirty = true;
    return NBG_OK;
}

int nbg_batch_set_targets_per_lane(nbg_batch *b, int32_t tpl) {
    if (!b) return fail(NBG_ERR_ARG, "batch is NULL");
    if (tpl != 1 && tpl != 2 && tpl != 4) return fail(NBG_ERR_ARG, "targets per lane: 1, 2 or 4");
    b->tpl = tpl;
    b->tables_dirty = true;
    return NBG_OK;
}

}  // extern "C"
