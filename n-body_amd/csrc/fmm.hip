// Implicit fast multipole method (MultipoleImplicitFMMSolver<1>,
// Solvers/ImplicitFMMSolver.h:14-36) on the octree of the Barnes-Hut build.
//
// The reference (ImplicitDualTraversalSolver::updateAccelerations,
// Solvers/ImplicitDualTraversalSolver.h:46-68) breaks the tree into about 256
// starting passive nodes (Tree::loadBalancedBreak, Trees/Tree.h:82-106) and,
// from each, runs a lockstep dual traversal: the active nodes of a passive
// node P are split into far ones (their monopole goes into P's local field,
// a vector and its gradient tensor), leaves and internal nodes; an internal
// P hands the bodies of the active leaves to a passive-tree traversal (far
// bodies go into local fields, close ones down the tree) and the children of
// the internal active nodes to each of its children, with the local field
// translated to the child's centre; a passive leaf walks its bodies through
// the remaining active nodes (Barnes-Hut from those nodes), adds the close
// bodies directly and evaluates its local field
// (Descent/balancedLockstepDualTreeImplicitField.h:14-45,
// Descent/lockstepDualTreeImplicitField.h:18-170,
// Descent/passiveTreeImplicitField.h:13-78).
//
// Here the traversal runs breadth-first, one launch per traversal step: a
// step takes the frontier of passive nodes (each with its local field, its
// active-node list and its list of close bodies) and produces the next
// frontier, the children of its internal nodes, which share their parent's
// two lists.  One wavefront serves one passive node; lanes classify 64 list
// entries at a time with the reference's criteria (SideLengthOverDistance,
// node-node and particle-node forms, FP32 in the reference's operation
// order), so every decision, and with it every interaction count, is the
// reference's.  Lists live in one pool for the pass; a wave reserves its
// output lists with one atomic add.  Leaf children are collected as leaf
// groups (up to 64 bodies) and served by a last launch over all of them, so
// the heavy near-field work is one load-balanced launch rather than the tail
// of every step.  Far-field contributions are per-lane partial local fields
// summed once per node.  Empty octants, which the reference keeps as
// body-less nodes, contribute nothing and have no record here.

#include <cmath>
#include <limits>

#include "device.hpp"
#include "launch.hpp"
#include "walk_common.hpp"

namespace nbg {

namespace {

// ---------------------------------------------------------------- starting nodes
// Tree::loadBalancedBreak with libstdc++'s binary-heap algorithms (push: sift
// the new element up while its parent compares less; pop: move the root to
// the end, walk the hole down along the larger child to the bottom, then sift
// the displaced last element up from there), so ties between equal body
// counts resolve as in the reference.
struct HeapEnt {
    int slot;   // node slot, -1 for an empty octant
    int count;  // bodies
};

__device__ void heap_sift_up(HeapEnt *h, int hole, int top, HeapEnt v) {
    int parent = (hole - 1) / 2;
    while (hole > top && h[parent].count < v.count) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = v;
}

__device__ void heap_push(HeapEnt *h, int len) { heap_sift_up(h, len - 1, 0, h[len - 1]); }

__device__ void heap_pop(HeapEnt *h, int len) {
    if (len < 2) return;
    const HeapEnt v = h[len - 1];
    h[len - 1] = h[0];
    const int n = len - 1;
    int hole = 0, child = 0;
    while (child < (n - 1) / 2) {
        child = 2 * (child + 1);
        if (h[child].count < h[child - 1].count) --child;
        h[hole] = h[child];
        hole = child;
    }
    if ((n & 1) == 0 && child == (n - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    heap_sift_up(h, hole, 0, v);
}

constexpr int kFmmStarts = 256;
constexpr int kHeapCap = kFmmStarts + 8;
static_assert(kHeapCap <= kFmmStartCap && kHeapCap <= kFmmSeedWaves, "starting-node buffers too small");

__global__ __launch_bounds__(64) void k_fmm_starts(const OctRec *__restrict__ nodes, const uint32_t *__restrict__ misc,
                                                   int32_t *starts, uint32_t *nstarts, uint32_t *err) {
    // lane 0 keeps the heap in LDS; the popped node's children are read by
    // one lane each, so a round costs two dependent loads, not nine
    __shared__ HeapEnt q[kHeapCap];
    const int lane = lane_id();
    const int64_t T = misc[0];
    int len = 1;
    bool bad = false;
    if (lane == 0) q[0] = HeapEnt{0, nodes[0].topo.y};
    __builtin_amdgcn_wave_barrier();
    while (len < kFmmStarts) {
        if (lane == 0) heap_pop(q, len);
        __builtin_amdgcn_wave_barrier();
        const int slot = uniform_i32(q[len - 1].slot);
        if (slot < 0) break;
        const OctRec r = nodes[slot];
        if (r.topo.z < 0) break;
        --len;
        const int c0 = uniform_i32(r.topo.z), nch = uniform_i32(oct_nchild(r));
        if (c0 <= slot || nch < 1 || nch > 8 || (int64_t)c0 + nch > T) {
            bad = true;
            break;
        }
        int oc = 8, cnt = 0;
        if (lane < nch) {
            const OctRec cr = nodes[c0 + lane];
            oc = oct_octant(cr);
            cnt = cr.topo.y;
        }
        int j = 0;
        for (int oct = 0; oct < 8; ++oct) {
            const int ocj = __shfl(oc, j & 7, kWave), cj = __shfl(cnt, j & 7, kWave);
            HeapEnt e{-1, 0};
            if (j < nch && ocj == oct) {
                e = HeapEnt{c0 + j, cj};
                ++j;
            }
            if (lane == 0) {
                q[len] = e;
                heap_push(q, len + 1);
            }
            ++len;
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (lane == 0) {
        if (bad) atomicOr(err, 8u);
        int k = 0;
        for (int i = 0; i < len; ++i)
            if (q[i].slot >= 0) starts[k++] = q[i].slot;
        *nstarts = (uint32_t)k;
    }
}

// ---------------------------------------------------------------- criteria
// SideLengthOverDistance, node-node (Descent/DescentCriterion/SideLengthOverDistance.h:33-56):
// 0 = Approximate, else DescendActiveNode (1) | DescendPassiveNode (2).
// fl(x / y) < t for FP32 x, y >= 0 and an FP32 t >= 0, without the
// division: with p the float below t and m = (p + t) / 2, fl(x / y) <= p
// exactly when x / y < m.  A quotient of floats never equals m: m's mantissa
// is an odd 25-bit integer, so x = m y would need a 25-bit mantissa (x = 0
// aside, where 0 / y is below t for y > 0 and NaN for y = 0, as 0 < m y
// says).  m y has at most 49 significant bits, so the FP64 product and
// comparison are exact; infinities compare as the quotient would.
struct QuotientBound {
    double mid;
};

__device__ __forceinline__ bool quotient_below(float x, float y, QuotientBound t) {
    return (double)x < t.mid * (double)y;
}

__device__ __forceinline__ int fmm_recommend(float4 acom, float4 ac, float4 pc, QuotientBound theta2) {
    const float sa2 = ac.w * ac.w, sp2 = pc.w * pc.w;
    const float dx = pc.x - acom.x, dy = pc.y - acom.y, dz = pc.z - acom.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float big = sa2 < sp2 ? sp2 : sa2;
    if (quotient_below(4.0f * big, d2, theta2)) {
        // exclusion regions [c - side, c + side] overlap (DescentCriterionType.h:81-85, BoundingBox.h:77-82)
        const bool meet = fminf(pc.x + pc.w, ac.x + ac.w) >= fmaxf(pc.x - pc.w, ac.x - ac.w) &&
                          fminf(pc.y + pc.w, ac.y + ac.w) >= fmaxf(pc.y - pc.w, ac.y - ac.w) &&
                          fminf(pc.z + pc.w, ac.z + ac.w) >= fmaxf(pc.z - pc.w, ac.z - ac.w);
        if (!meet) return 0;
    }
    return ((sa2 <= sp2 * 4.0f) ? 2 : 0) | ((sp2 <= sa2 * 4.0f) ? 1 : 0);
}

// SideLengthOverDistance, particle-node (SideLengthOverDistance.h:25-29):
// 2 side / |x - c| < theta and x outside the cell [c - side/2, c + side/2).
__device__ __forceinline__ bool fmm_particle_far(float x, float y, float z, float4 pc, QuotientBound theta) {
    const float dx = x - pc.x, dy = y - pc.y, dz = z - pc.z;
    const float dist = ieee_sqrt(dx * dx + dy * dy + dz * dz);
    if (!quotient_below(2.0f * pc.w, dist, theta)) return false;
    const float h = pc.w / 2.0f;
    const bool in = pc.x - h <= x && pc.y - h <= y && pc.z - h <= z && x < pc.x + h && y < pc.y + h && z < pc.z + h;
    return !in;
}

// A local field (MultipoleAcceleration<2>: vector and symmetric gradient
// xx xy xz yy yz zz), without the factor G.
struct LocalF {
    float ax, ay, az, t0, t1, t2, t3, t4, t5;
};

// Gravity::particleNode into the local field at c (Physics/Rules/Gravity.h:281-304)
// with rsqrt: a -= R m / r^3, T += m (3 R(x)R / r^5 - I / r^3), R = c - s;
// a source exactly at c contributes nothing.
__device__ __forceinline__ void fmm_to_local(float sx, float sy, float sz, float m, float4 c, float eps2, LocalF &F) {
    if (sx == c.x && sy == c.y && sz == c.z) return;
    const float Rx = c.x - sx, Ry = c.y - sy, Rz = c.z - sz;
    const float r2 = __builtin_fmaf(Rx, Rx, __builtin_fmaf(Ry, Ry, __builtin_fmaf(Rz, Rz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv2 = inv * inv, m3 = m * (inv * inv2), m5 = 3.0f * m3 * inv2;
    F.ax = __builtin_fmaf(-Rx, m3, F.ax);
    F.ay = __builtin_fmaf(-Ry, m3, F.ay);
    F.az = __builtin_fmaf(-Rz, m3, F.az);
    F.t0 = __builtin_fmaf(Rx * Rx, m5, F.t0 - m3);
    F.t1 = __builtin_fmaf(Rx * Ry, m5, F.t1);
    F.t2 = __builtin_fmaf(Rx * Rz, m5, F.t2);
    F.t3 = __builtin_fmaf(Ry * Ry, m5, F.t3 - m3);
    F.t4 = __builtin_fmaf(Ry * Rz, m5, F.t4);
    F.t5 = __builtin_fmaf(Rz * Rz, m5, F.t5 - m3);
}

// MultipoleAcceleration::at (MultipoleAcceleration.h:33-47) and ::translated
// (:49-70): a + T d.
__device__ __forceinline__ void fmm_at(const LocalF &F, float dx, float dy, float dz, float &x, float &y, float &z) {
    x = __builtin_fmaf(F.t0, dx, __builtin_fmaf(F.t1, dy, __builtin_fmaf(F.t2, dz, F.ax)));
    y = __builtin_fmaf(F.t1, dx, __builtin_fmaf(F.t3, dy, __builtin_fmaf(F.t4, dz, F.ay)));
    z = __builtin_fmaf(F.t2, dx, __builtin_fmaf(F.t4, dy, __builtin_fmaf(F.t5, dz, F.az)));
}

// Makes the wave's earlier global-memory list writes visible to its later
// reads by other lanes (one wave per workgroup).
__device__ __forceinline__ void wave_mem_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float wave_fsum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Appends v for every lane with `want`; writes stay below `cap` (a longer
// result sets `over`); returns the new (uniform) length, at most cap.
__device__ __forceinline__ int wave_append(uint32_t *list, int len, int cap, bool want, uint32_t v, bool &over) {
    const uint64_t b = __ballot(want);
    if (want) {
        const int k = len + __popcll(b & lanemask_lt());
        if (k < cap) list[k] = v;
    }
    const int nl = len + __popcll(b);
    if (nl > cap) over = true;
    return nl < cap ? nl : cap;
}

// Appends the children (contiguous slots c0 .. c0 + k - 1) of every lane's
// node; returns the new length.  A child range outside the node array (an
// inconsistent tree) sets `bad` and appends nothing for that lane.
__device__ __forceinline__ int wave_append_children(uint32_t *list, int len, int cap, int c0, int k, int64_t max_nodes,
                                                    bool &over, bool &bad) {
    if (k > 0 && (c0 < 1 || k > 8 || (int64_t)c0 + k > max_nodes)) {
        bad = true;
        k = 0;
    }
    const uint32_t inc = wave_incl_scan((uint32_t)k);
    const int base = len + (int)(inc - (uint32_t)k);
    for (int j = 0; j < k; ++j)
        if (base + j < cap) list[base + j] = (uint32_t)(c0 + j);
    const int nl = len + (int)__shfl(inc, kWave - 1, kWave);
    if (nl > cap) over = true;
    return nl < cap ? nl : cap;
}

// Reserves `words` of a pool (wave-uniform); -1 when it does not fit.
__device__ __forceinline__ int64_t pool_reserve(uint32_t *counter, int64_t words, int64_t cap, uint32_t *peak) {
    uint32_t off = 0;
    if (lane_id() == 0 && words > 0) {
        off = atomicAdd(counter, (uint32_t)words);
        atomicMax(peak, off + (uint32_t)words);
    }
    const int64_t o = (int64_t)__shfl(off, 0, kWave);
    return (words == 0 || o + words <= cap) ? o : -1;
}

// Descent::activeTree from `start` for the lanes in `lanes` (the octree's
// FAST group walk, counting interactions).
template <bool SOFT>
__device__ void fmm_walk_from(const FmmArgs &a, const OctRec &st, uint64_t lanes, FastWave &W, int *s_grp,
                              uint64_t *s_msk, bool &fail) {
    const OctRec *__restrict__ nodes = a.nodes;
    uint64_t mask = visit_oct<SOFT, 1, false>(st, nullptr, lanes, a.bodies, a.n, a.eps2, W, fail);
    int c = st.topo.z, nch = oct_nchild(st);
    int sp = 0;
    int64_t iters = 0;
    while (true) {
        if (mask == 0) {
            if (sp == 0) break;
            --sp;
            __builtin_amdgcn_wave_barrier();
            const int g = uniform_i32(s_grp[sp]);
            c = g & 0x0FFFFFFF;
            nch = (int)((uint32_t)g >> 28);
            mask = uniform_u64(s_msk[sp]);
        }
        c = uniform_i32(c);
        nch = uniform_i32(nch);
        if (++iters > a.max_nodes + 64 || c < 0 || nch < 1 || nch > 8 || (int64_t)c + nch > a.max_nodes) {
            fail = true;
            break;
        }
        uint64_t nm = 0;
        int nc = 0, nn = 0;
        for (int k = 0; k < nch; ++k) {
            const OctRec r = nodes[c + k];
            const uint64_t o = visit_oct<SOFT, 1, false>(r, nullptr, mask, a.bodies, a.n, a.eps2, W, fail);
            if (o) {
                if (nm) {
                    if (sp >= kStackOct) { fail = true; break; }
                    if (lane_id() == 0) { s_grp[sp] = nc | (nn << 28); s_msk[sp] = nm; }
                    __builtin_amdgcn_wave_barrier();
                    ++sp;
                }
                nm = o;
                nc = r.topo.z;
                nn = oct_nchild(r);
            }
        }
        mask = nm;
        c = nc;
        nch = nn;
    }
}

// A frontier item: a passive node, its active list and close-body list
// (offsets into the step's pool) and its local field at the node's centre.
// Flags: the active list holds child groups (first child | count << 28: the
// children of the parent's close internal active nodes, expanded on reading)
// rather than node slots; the item is a leaf group (some leaf children of the
// node, at most 64 bodies together, given by a bit mask over its children;
// the field is then the parent's).
enum { kItP = 0, kItAoff = 1, kItAlen = 2, kItEoff = 3, kItElen = 4, kItF = 5, kItFlags = 14, kItLeaves = 15 };
constexpr uint32_t kItemGroups = 1u, kItemLeafGroup = 2u;

__device__ __forceinline__ void write_item(uint32_t *it, int P, uint32_t aoff, uint32_t alen, uint32_t eoff,
                                           uint32_t elen, const LocalF &F, uint32_t flags, uint32_t leaves) {
    uint4 *q = (uint4 *)it;
    q[0] = make_uint4((uint32_t)P, aoff, alen, eoff);
    q[1] = make_uint4(elen, __float_as_uint(F.ax), __float_as_uint(F.ay), __float_as_uint(F.az));
    q[2] = make_uint4(__float_as_uint(F.t0), __float_as_uint(F.t1), __float_as_uint(F.t2), __float_as_uint(F.t3));
    q[3] = make_uint4(__float_as_uint(F.t4), __float_as_uint(F.t5), flags, leaves);
}

// Entry idx of an active list: a node slot, or (groups) child idx % 8 of
// group idx / 8; -1 for an empty group position.
__device__ __forceinline__ int active_entry(const uint32_t *list, uint32_t idx, bool groups) {
    if (!groups) return (int)list[idx];
    const uint32_t g = list[idx >> 3];
    const uint32_t k = idx & 7u;
    return k < (g >> 28) ? (int)(g & 0x0FFFFFFFu) + (int)k : -1;
}

// ---------------------------------------------------------------- leaf groups
// Up to eight leaf children of one passive node, at most 64 bodies together,
// served by one wave: lane t holds body t of the group.  Each leaf takes its
// own decisions against the shared lists (the reference's lockstep step at a
// passive leaf: far active nodes and far bodies into the leaf's local field,
// the other active nodes walked, the other bodies added directly), so the
// interaction set of every body is the reference's; the wave walks each
// active node once for all the lanes whose leaf found it close.
struct GroupLds {
    float4 cen[8];   // leaf centre, side
    int first[8], count[8];
    float F[8][9];   // leaf local fields
};

template <bool SOFT>
__device__ void fmm_leaf_group(const FmmArgs &a, int P, const uint32_t *alist, uint32_t alen, const uint32_t *elist,
                               uint32_t elen, const uint32_t *it, uint32_t lmask, GroupLds &g, int *s_grp,
                               uint64_t *s_msk, uint32_t *ulist, uint64_t &c_pp, uint64_t &c_np, uint64_t &c_pn,
                               uint64_t &c_nn, bool &fail) {
    const int lane = lane_id();
    const OctRec *__restrict__ nodes = a.nodes;
    const float4 *__restrict__ cen = a.centre;
    const QuotientBound theta = {a.theta_mid}, theta2 = {a.theta2_mid};
    const float eps2 = a.eps2;
    const OctRec pr = nodes[P];
    const float4 cP = cen[P];
    const int c0 = pr.topo.z, nch = oct_nchild(pr);
    lmask &= 0xFFu;
    if (c0 < 1 || nch < 1 || nch > 8 || (int64_t)c0 + nch > a.max_nodes || (lmask >> nch) != 0u || lmask == 0u) {
        fail = true;
        return;
    }
    const int k = __popc(lmask);
    // lane j < k: the j-th leaf of the group, its field translated from P's centre
    if (lane < k) {
        uint32_t m = lmask;
        for (int j = 0; j < lane; ++j) m &= m - 1u;
        const int slot = c0 + (__ffs(m) - 1);
        const OctRec lr = nodes[slot];
        const float4 cl = cen[slot];
        g.cen[lane] = cl;
        g.first[lane] = lr.topo.x;
        g.count[lane] = lr.topo.y;
        LocalF F;
        F.ax = __uint_as_float(it[kItF + 0]); F.ay = __uint_as_float(it[kItF + 1]); F.az = __uint_as_float(it[kItF + 2]);
        F.t0 = __uint_as_float(it[kItF + 3]); F.t1 = __uint_as_float(it[kItF + 4]); F.t2 = __uint_as_float(it[kItF + 5]);
        F.t3 = __uint_as_float(it[kItF + 6]); F.t4 = __uint_as_float(it[kItF + 7]); F.t5 = __uint_as_float(it[kItF + 8]);
        float x, y, z;
        fmm_at(F, cl.x - cP.x, cl.y - cP.y, cl.z - cP.z, x, y, z);
        g.F[lane][0] = x; g.F[lane][1] = y; g.F[lane][2] = z;
        g.F[lane][3] = F.t0; g.F[lane][4] = F.t1; g.F[lane][5] = F.t2;
        g.F[lane][6] = F.t3; g.F[lane][7] = F.t4; g.F[lane][8] = F.t5;
    }
    wave_mem_sync();
    // lane -> (leaf, body)
    int total = 0, myleaf = 0;
    for (int j = 0; j < k; ++j) {
        const int f = g.first[j], c = g.count[j];
        if (f < 0 || c < 1 || (int64_t)f + c > a.n) fail = true;
        if (lane >= total) myleaf = j;
        total += c;
    }
    if (fail || total > kWave) {
        fail = true;
        return;
    }
    const bool has = lane < total;
    int pre = 0;
    for (int j = 0; j < myleaf; ++j) pre += g.count[j];
    const int me = g.first[myleaf] + (lane - pre);
    const float4 p = a.bodies[has ? me : g.first[0]];

    // pass A: far active nodes and far bodies into the leaves' fields.  Lane
    // t < kk evaluates entry t / k of a batch against leaf t % k, so a lane
    // always serves one leaf (its partial field is that leaf's) and the k
    // lanes of an entry read the same record.
    // The active nodes some leaf found close are collected (while they fit
    // the wave's buffer) for the walk pass.
    const uint32_t vlen = 8u * alen;
    int nu = 0;
    bool ufull = false;
    {
        const int epb = kWave / k, kk = epb * k, lj = lane % k, li = lane / k;
        const uint64_t kmask = (1ull << k) - 1;
        const float4 cl = g.cen[lj];
        LocalF Fp = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t b = 0; b < vlen; b += (uint32_t)epb) {
            const uint32_t idx = b + (uint32_t)li;
            int an = lane < kk && idx < vlen ? active_entry(alist, idx, true) : -1;
            if (an >= a.max_nodes) { fail = true; an = -1; }
            const bool he = an >= 0;
            const OctRec ar = nodes[he ? an : 0];
            const int r = he ? fmm_recommend(ar.com, cen[he ? an : 0], cl, theta2) : 1;
            const bool far = he && r == 0;
            if (far) fmm_to_local(ar.com.x, ar.com.y, ar.com.z, ar.com.w, cl, eps2, Fp);
            c_nn += __popcll(__ballot(far));
            const uint64_t cb = __ballot(he && r != 0);
            // the first lane of an entry appends it when any of its leaves found it close
            const bool first = lane < kk && lj == 0 && ((cb >> lane) & kmask) != 0ull;
            nu = wave_append(ulist, nu, a.ulist_cap, first, (uint32_t)an, ufull);
        }
        for (uint32_t b = 0; b < elen; b += (uint32_t)epb) {
            const uint32_t idx = b + (uint32_t)li;
            bool he = lane < kk && idx < elen;
            const int e = he ? (int)elist[idx] : 0;
            if (he && (e < 0 || (int64_t)e >= a.n)) { fail = true; he = false; }
            const float4 bd = a.bodies[he ? e : 0];
            const bool far = he && fmm_particle_far(bd.x, bd.y, bd.z, cl, theta);
            if (far) fmm_to_local(bd.x, bd.y, bd.z, bd.w, cl, eps2, Fp);
            c_pn += __popcll(__ballot(far));
        }
        for (int j = 0; j < k; ++j) {
            const bool mine = lane < kk && lj == j;
            const float s0 = wave_fsum(mine ? Fp.ax : 0.0f), s1 = wave_fsum(mine ? Fp.ay : 0.0f);
            const float s2 = wave_fsum(mine ? Fp.az : 0.0f), s3 = wave_fsum(mine ? Fp.t0 : 0.0f);
            const float s4 = wave_fsum(mine ? Fp.t1 : 0.0f), s5 = wave_fsum(mine ? Fp.t2 : 0.0f);
            const float s6 = wave_fsum(mine ? Fp.t3 : 0.0f), s7 = wave_fsum(mine ? Fp.t4 : 0.0f);
            const float s8 = wave_fsum(mine ? Fp.t5 : 0.0f);
            if (lane == 0) {
                g.F[j][0] += s0; g.F[j][1] += s1; g.F[j][2] += s2;
                g.F[j][3] += s3; g.F[j][4] += s4; g.F[j][5] += s5;
                g.F[j][6] += s6; g.F[j][7] += s7; g.F[j][8] += s8;
            }
        }
    }
    wave_mem_sync();

    // pass B: walk each active node once for the lanes whose leaf found it close
    FastWave W;
    fast_init(W, p.x, p.y, p.z);
    W.G = a.G;
    const uint32_t mybit = has ? 1u << myleaf : 0u;
    // the collected entries (all entries again when the buffer overflowed)
    const uint32_t wlen = ufull ? vlen : (uint32_t)nu;
    for (uint32_t b = 0; b < wlen; b += kWave) {
        int an = b + lane < wlen ? (ufull ? active_entry(alist, b + lane, true) : (int)ulist[b + lane]) : -1;
        if (an >= a.max_nodes) an = -1;
        const bool he = an >= 0;
        const OctRec ar = nodes[he ? an : 0];
        const float4 ca = cen[he ? an : 0];
        uint32_t m = 0;
        if (he) {
#pragma unroll 1
            for (int j = 0; j < k; ++j)
                if (fmm_recommend(ar.com, ca, g.cen[j], theta2) != 0) m |= 1u << j;
        }
        uint64_t todo = __ballot(m != 0u);
        while (todo) {
            const int i = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)m, i);
            const uint64_t lanes = __ballot((mm & mybit) != 0u);
            if (lanes) {
                // the node's record is already in lane i
                OctRec st;
                st.com = make_float4(readlane_f32(ar.com.x, i), readlane_f32(ar.com.y, i), readlane_f32(ar.com.z, i),
                                     readlane_f32(ar.com.w, i));
                st.lo = make_float4(readlane_f32(ar.lo.x, i), readlane_f32(ar.lo.y, i), readlane_f32(ar.lo.z, i),
                                    readlane_f32(ar.lo.w, i));
                st.hi = make_float4(readlane_f32(ar.hi.x, i), readlane_f32(ar.hi.y, i), readlane_f32(ar.hi.z, i),
                                    readlane_f32(ar.hi.w, i));
                st.topo = make_int4(__builtin_amdgcn_readlane(ar.topo.x, i), __builtin_amdgcn_readlane(ar.topo.y, i),
                                    __builtin_amdgcn_readlane(ar.topo.z, i), __builtin_amdgcn_readlane(ar.topo.w, i));
                fmm_walk_from<SOFT>(a, st, lanes, W, s_grp, s_msk, fail);
            }
        }
    }
    uint64_t epp = 0;
    for (uint32_t b = 0; b < elen; b += kWave) {
        bool he = b + lane < elen;
        const int e = he ? (int)elist[b + lane] : 0;
        if (he && (e < 0 || (int64_t)e >= a.n)) he = false;
        const float4 bd = a.bodies[he ? e : 0];
        uint32_t m = 0;
        if (he) {
#pragma unroll 1
            for (int j = 0; j < k; ++j)
                if (!fmm_particle_far(bd.x, bd.y, bd.z, g.cen[j], theta)) m |= 1u << j;
        }
        uint64_t todo = __ballot(m != 0u);
        while (todo) {
            const int i = __ffsll((unsigned long long)todo) - 1;
            todo &= todo - 1;
            const uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)m, i);
            const uint64_t lanes = __ballot((mm & mybit) != 0u);
            const float sx = readlane_f32(bd.x, i), sy = readlane_f32(bd.y, i);
            const float sz = readlane_f32(bd.z, i), sm = readlane_f32(bd.w, i);
            if ((mm & mybit) != 0u) interact<SOFT>(sx, sy, sz, sm, eps2, W);
            epp += __popcll(lanes);
        }
    }
    c_pp += W.pp + epp;
    c_np += W.np;
    if (has) {
        const float4 cl = g.cen[myleaf];
        LocalF F;
        F.ax = g.F[myleaf][0]; F.ay = g.F[myleaf][1]; F.az = g.F[myleaf][2];
        F.t0 = g.F[myleaf][3]; F.t1 = g.F[myleaf][4]; F.t2 = g.F[myleaf][5];
        F.t3 = g.F[myleaf][6]; F.t4 = g.F[myleaf][7]; F.t5 = g.F[myleaf][8];
        float lx, ly, lz;
        fmm_at(F, p.x - cl.x, p.y - cl.y, p.z - cl.z, lx, ly, lz);
        const uint32_t o = a.perm[me];
        if ((int64_t)o < a.n) {
            const float3 r = fast_result(W, a.G, lx, ly, lz);
            a.ax[o] = r.x; a.ay[o] = r.y; a.az[o] = r.z;
        }
    }
    wave_mem_sync();  // the next item rewrites the group's LDS
}

// ---------------------------------------------------------------- seeds
// Descent::balancedLockstepDualTreeImplicitField: for each starting node S,
// the breadth-first descent of the active tree while the criterion asks to
// descend the active node only; the nodes where it stops are S's active list.
__global__ __launch_bounds__(64) void k_fmm_seed(FmmArgs a) {
    const int w = blockIdx.x;
    if (w >= (int)*a.nstarts) return;
    const int lane = lane_id();
    const OctRec *__restrict__ nodes = a.nodes;
    const float4 *__restrict__ cen = a.centre;
    const QuotientBound theta2 = {a.theta2_mid};
    uint32_t *base = a.seed_scratch + (int64_t)w * 3 * a.cap_a;
    uint32_t *Q[2] = {base, base + a.cap_a};
    uint32_t *EQ = base + 2 * a.cap_a;
    const int S = uniform_i32(a.starts[w]);
    bool over = false, bad = false;
    if (S < 0 || S >= a.max_nodes) bad = true;
    const float4 cS = cen[bad ? 0 : S];
    int cur = bad ? 0 : 1, eq = 0, qi = 0, levels = 0;
    if (lane == 0) Q[0][0] = 0u;
    wave_mem_sync();
    while (cur > 0) {
        if (++levels > 128) {  // deeper than any octree: an inconsistent tree
            bad = true;
            break;
        }
        int nxt = 0;
        for (int b = 0; b < cur; b += kWave) {
            const bool has = b + lane < cur;
            const int an = has ? (int)Q[qi][b + lane] : 0;
            const OctRec ar = nodes[an];
            const int r = has ? fmm_recommend(ar.com, cen[an], cS, theta2) : 0;
            const bool down = has && r == 1 && ar.topo.z >= 0;
            eq = wave_append(EQ, eq, a.cap_a, has && !down, (uint32_t)an, over);
            nxt = wave_append_children(Q[qi ^ 1], nxt, a.cap_a, ar.topo.z, down ? oct_nchild(ar) : 0, a.max_nodes,
                                       over, bad);
        }
        qi ^= 1;
        cur = nxt;
        wave_mem_sync();
    }
    bad = __ballot(bad) != 0;
    uint32_t *ctr = a.ctr;  // step 0's input set
    const int64_t off = bad || over ? -1 : pool_reserve(a.pool_ctr, eq, a.pool_cap, a.peak);
    if (off < 0) {
        if (lane == 0) atomicOr(a.err, bad ? 2u : 16u);
        return;
    }
    for (int b = 0; b < eq; b += kWave)
        if (b + lane < eq) a.pool[off + b + lane] = EQ[b + lane];
    uint32_t slot = 0;
    if (lane == 0) slot = atomicAdd(ctr + 0, 1u);
    slot = __shfl(slot, 0, kWave);
    if (lane == 0) {
        const LocalF Z = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        write_item(a.items[0] + (int64_t)slot * kFmmItemWords, S, (uint32_t)off, (uint32_t)eq, 0u, 0u, Z, 0u, 0u);
    }
}

// ---------------------------------------------------------------- one traversal step
// Counter sets rotate over three (input, output, next output): step k reads
// set k % 3 ({items, next item, pool words, -}), fills set (k + 1) % 3 and
// clears set (k + 2) % 3, which step k - 1 read.
template <bool SOFT>
__global__ __launch_bounds__(64) void k_fmm_step(FmmArgs a, int k) {
    __shared__ int s_grp[kStackOct];
    __shared__ uint64_t s_msk[kStackOct];
    __shared__ uint32_t s_desc[8];

    const int lane = lane_id();
    const OctRec *__restrict__ nodes = a.nodes;
    const float4 *__restrict__ cen = a.centre;
    const QuotientBound theta = {a.theta_mid}, theta2 = {a.theta2_mid};
    const float eps2 = a.eps2;
    uint32_t *ci = a.ctr + 4 * (k % 3), *co = a.ctr + 4 * ((k + 1) % 3), *cr = a.ctr + 4 * ((k + 2) % 3);
    if (blockIdx.x == 0 && lane < 4) cr[lane] = 0u;
    const uint32_t *__restrict__ itin = a.items[k & 1];
    uint32_t *itout = a.items[(k + 1) & 1];
    const uint32_t *pin = a.pool;
    uint32_t *pout = a.pool;
    uint32_t *ulist = a.ulist + (int64_t)blockIdx.x * a.ulist_cap;
    // a step that overflowed the frontier counted past the capacity (and flagged it)
    const int nin = (int)(ci[0] < (uint64_t)a.item_cap ? ci[0] : (uint32_t)a.item_cap);
    if ((int)blockIdx.x >= nin) return;  // (steps past the deepest leaf launch with nin = 0)
    uint64_t c_pp = 0, c_np = 0, c_pn = 0, c_nn = 0;
    bool fail = false, over = false;

    for (;;) {
        int job = 0;
        if (lane == 0) job = (int)atomicAdd(ci + 1, 1u);
        job = uniform_i32(__shfl(job, 0, kWave));
        if (job >= nin) break;
        const uint32_t *it = itin + (int64_t)job * kFmmItemWords;
        const int P = uniform_i32((int)it[kItP]);
        const uint32_t aoff = (uint32_t)uniform_i32((int)it[kItAoff]), alen = (uint32_t)uniform_i32((int)it[kItAlen]);
        const uint32_t eoff = (uint32_t)uniform_i32((int)it[kItEoff]), elen = (uint32_t)uniform_i32((int)it[kItElen]);
        const uint32_t flags = (uint32_t)uniform_i32((int)it[kItFlags]);
        const bool groups = (flags & kItemGroups) != 0;
        if (P < 0 || P >= a.max_nodes || (int64_t)aoff + alen > a.pool_cap || (int64_t)eoff + elen > a.pool_cap) {
            fail = true;
            break;
        }
        if (flags & kItemLeafGroup) {  // leaf groups never enter the frontier
            fail = true;
            break;
        }
        const OctRec pr = nodes[P];
        const float4 cP = cen[P];
        const bool leaf = pr.topo.z < 0;
        if (pr.topo.x < 0 || pr.topo.y < 0 || (int64_t)pr.topo.x + pr.topo.y > a.n) {
            fail = true;
            break;
        }
        const uint32_t *alist = pin + aoff;
        const uint32_t vlen = groups ? 8u * alen : alen;
        LocalF Fl = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        // pass 1 over the active list: far nodes into the local field; sizes of
        // what stays, whose slots are collected (while they fit) for pass 2
        int nleafc = 0, ninner = 0, nu = 0;
        bool ufull = false;
        uint32_t lbod = 0;
        for (uint32_t b = 0; b < vlen; b += kWave) {
            int an = b + lane < vlen ? active_entry(alist, b + lane, groups) : -1;
            if (an >= a.max_nodes) { fail = true; an = -1; }
            const bool has = an >= 0;
            const OctRec ar = nodes[has ? an : 0];
            const int r = has ? fmm_recommend(ar.com, cen[an], cP, theta2) : 1;
            const bool far = has && r == 0;
            if (far) fmm_to_local(ar.com.x, ar.com.y, ar.com.z, ar.com.w, cP, eps2, Fl);
            c_nn += __popcll(__ballot(far));
            const bool lf = has && !far && ar.topo.z < 0;
            nleafc += __popcll(__ballot(lf));
            ninner += __popcll(__ballot(has && !far && ar.topo.z >= 0));
            if (lf) lbod += (uint32_t)ar.topo.y;
            nu = wave_append(ulist, nu, a.ulist_cap, has && !far, (uint32_t)an, ufull);
        }
        // pool layout: a leaf P keeps its close active nodes (walk starts);
        // an internal P its close internal ones as child groups, then its
        // close leaves; then the close bodies
        const int nclose = nleafc + ninner;
        const int64_t ecap = (int64_t)elen + (leaf ? 0 : (int64_t)wave_sum(lbod));
        const int64_t off = pool_reserve(a.pool_ctr, nclose + ecap, a.pool_cap, a.peak);
        if (off < 0) {
            over = true;
            break;
        }
        uint32_t *C = pout + off, *LA = C + ninner, *E2 = C + nclose;
        // pass 2: write what stays (the collected slots are all close)
        wave_mem_sync();
        int nc = 0, ng = 0, nl = 0;
        const uint32_t wlen = ufull ? vlen : (uint32_t)nu;
        for (uint32_t b = 0; b < wlen; b += kWave) {
            int an = b + lane < wlen ? (ufull ? active_entry(alist, b + lane, groups) : (int)ulist[b + lane]) : -1;
            if (an >= a.max_nodes) an = -1;
            const bool has = an >= 0;
            const OctRec ar = nodes[has ? an : 0];
            const bool close = has && (!ufull || fmm_recommend(ar.com, cen[an], cP, theta2) != 0);
            if (leaf) {
                nc = wave_append(C, nc, nclose, close, (uint32_t)an, over);
            } else {
                const bool inner = close && ar.topo.z >= 0;
                const int nch = inner ? oct_nchild(ar) : 0;
                if (inner && (ar.topo.z < 1 || nch < 1 || (int64_t)ar.topo.z + nch > a.max_nodes)) fail = true;
                const uint32_t g = (uint32_t)ar.topo.z | ((uint32_t)nch << 28);
                ng = wave_append(C, ng, ninner, inner, g, over);
                nl = wave_append(LA, nl, nleafc, close && !inner, (uint32_t)an, over);
            }
        }
        // the close bodies handed down by the passive-tree traversals
        int ne = 0;
        for (uint32_t b = 0; b < elen; b += kWave) {
            bool has = b + lane < elen;
            const int e = has ? (int)pin[eoff + b + lane] : 0;
            if (has && (e < 0 || (int64_t)e >= a.n)) { fail = true; has = false; }
            const float4 bd = a.bodies[has ? e : 0];
            const bool far = has && fmm_particle_far(bd.x, bd.y, bd.z, cP, theta);
            if (far) fmm_to_local(bd.x, bd.y, bd.z, bd.w, cP, eps2, Fl);
            c_pn += __popcll(__ballot(far));
            ne = wave_append(E2, ne, (int)ecap, has && !far, (uint32_t)e, over);
        }
        wave_mem_sync();
        if (!leaf) {
            // bodies of the close active leaves start a passive-tree traversal at P
            for (int j = 0; j < nl; ++j) {
                const OctRec lf = nodes[uniform_i32((int)LA[j])];
                if (lf.topo.x < 0 || lf.topo.y < 0 || (int64_t)lf.topo.x + lf.topo.y > a.n) {
                    fail = true;
                    break;
                }
                for (int b = 0; b < lf.topo.y; b += kWave) {
                    const bool has = b + lane < lf.topo.y;
                    const int e = lf.topo.x + b + lane;
                    const float4 bd = a.bodies[has ? e : lf.topo.x];
                    const bool far = has && fmm_particle_far(bd.x, bd.y, bd.z, cP, theta);
                    if (far) fmm_to_local(bd.x, bd.y, bd.z, bd.w, cP, eps2, Fl);
                    c_pn += __popcll(__ballot(far));
                    ne = wave_append(E2, ne, (int)ecap, has && !far, (uint32_t)e, over);
                }
            }
        }
        LocalF F;
        F.ax = __uint_as_float(it[kItF + 0]) + wave_fsum(Fl.ax);
        F.ay = __uint_as_float(it[kItF + 1]) + wave_fsum(Fl.ay);
        F.az = __uint_as_float(it[kItF + 2]) + wave_fsum(Fl.az);
        F.t0 = __uint_as_float(it[kItF + 3]) + wave_fsum(Fl.t0);
        F.t1 = __uint_as_float(it[kItF + 4]) + wave_fsum(Fl.t1);
        F.t2 = __uint_as_float(it[kItF + 5]) + wave_fsum(Fl.t2);
        F.t3 = __uint_as_float(it[kItF + 6]) + wave_fsum(Fl.t3);
        F.t4 = __uint_as_float(it[kItF + 7]) + wave_fsum(Fl.t4);
        F.t5 = __uint_as_float(it[kItF + 8]) + wave_fsum(Fl.t5);
        if (!leaf && (ng > 0 || ne > 0)) {
            // The children of P share its two lists.  An internal child gets an
            // item with the field translated to its centre; leaf children of at
            // most 64 bodies share items of at most 64 bodies (one wave serves
            // their bodies together), carrying P's field.
            const int nch = oct_nchild(pr);
            if (nch < 1 || nch > 8 || pr.topo.z < 1 || (int64_t)pr.topo.z + nch > a.max_nodes) {
                fail = true;
                break;
            }
            const OctRec chr = nodes[pr.topo.z + (lane < nch ? lane : 0)];
            const int ccount = chr.topo.y;
            const int cleaf = chr.topo.z < 0 ? 1 : 0;
            int nit = 0, gbod = 0;
            uint32_t gmask = 0;
            for (int j = 0; j < nch; ++j) {
                const int cnt = __shfl(ccount, j, kWave);
                if (__shfl(cleaf, j, kWave) && cnt <= kWave) {
                    if (gbod + cnt > kWave) {
                        if (lane == 0) s_desc[nit] = 0x80000000u | gmask;
                        ++nit;
                        gmask = 0;
                        gbod = 0;
                    }
                    gmask |= 1u << j;
                    gbod += cnt;
                } else {
                    if (lane == 0) s_desc[nit] = (uint32_t)(pr.topo.z + j);
                    ++nit;
                }
            }
            if (gmask) {
                if (lane == 0) s_desc[nit] = 0x80000000u | gmask;
                ++nit;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const uint32_t d = lane < nit ? s_desc[lane] : 0u;
            const uint64_t gb = __ballot(lane < nit && (d & 0x80000000u) != 0u);
            const uint64_t nb = __ballot(lane < nit && (d & 0x80000000u) == 0u);
            const int ngrp = __popcll(gb), nnode = __popcll(nb);
            uint32_t base = 0, gbase = 0;
            if (lane == 0) {
                if (nnode) {
                    base = atomicAdd(co + 0, (uint32_t)nnode);
                    atomicMax(a.peak + 1, base + (uint32_t)nnode);
                }
                if (ngrp) gbase = atomicAdd(a.group_ctr, (uint32_t)ngrp);
            }
            base = __shfl(base, 0, kWave);
            gbase = __shfl(gbase, 0, kWave);
            if ((int64_t)base + nnode > a.item_cap || (int64_t)gbase + ngrp > a.item_cap) {
                over = true;
                break;
            }
            if (lane < nit) {
                if (d & 0x80000000u) {
                    uint32_t *dst = a.groups + (int64_t)(gbase + __popcll(gb & lanemask_lt())) * kFmmItemWords;
                    write_item(dst, P, (uint32_t)off, (uint32_t)ng, (uint32_t)(off + nclose), (uint32_t)ne, F,
                               kItemGroups | kItemLeafGroup, d & 0xFFu);
                } else {
                    uint32_t *dst = itout + (int64_t)(base + __popcll(nb & lanemask_lt())) * kFmmItemWords;
                    const int c = (int)d;
                    const float4 cc = cen[c];
                    LocalF Fc = F;
                    fmm_at(F, cc.x - cP.x, cc.y - cP.y, cc.z - cP.z, Fc.ax, Fc.ay, Fc.az);
                    write_item(dst, c, (uint32_t)off, (uint32_t)ng, (uint32_t)(off + nclose), (uint32_t)ne, Fc,
                               kItemGroups, 0u);
                }
            }
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        // P is a leaf, or nothing below P interacts any more: finish P's bodies here
        wave_mem_sync();
        for (int b = 0; b < pr.topo.y; b += kWave) {
            const bool has = b + lane < pr.topo.y;
            const int me = pr.topo.x + b + lane;
            const float4 p = a.bodies[has ? me : pr.topo.x];
            const uint64_t lanes = __ballot(has);
            FastWave W;
            fast_init(W, p.x, p.y, p.z);
            W.G = a.G;
            if (leaf) {
                for (int j = 0; j < nc; ++j)
                    fmm_walk_from<SOFT>(a, nodes[uniform_i32((int)C[j])], lanes, W, s_grp, s_msk, fail);
                for (int j = 0; j < ne; ++j) {
                    const float4 s = a.bodies[uniform_i32((int)E2[j])];
                    if (__builtin_amdgcn_inverse_ballot_w64(lanes)) interact<SOFT>(s.x, s.y, s.z, s.w, eps2, W);
                }
                c_pp += W.pp + (uint64_t)ne * __popcll(lanes);
                c_np += W.np;
            }
            if (has) {
                float lx, ly, lz;
                fmm_at(F, p.x - cP.x, p.y - cP.y, p.z - cP.z, lx, ly, lz);
                const uint32_t o = a.perm[me];
                if ((int64_t)o < a.n) {
                    const float3 r = fast_result(W, a.G, lx, ly, lz);
                    a.ax[o] = r.x; a.ay[o] = r.y; a.az[o] = r.z;
                }
            }
        }
        fail = __ballot(fail) != 0;  // wave-uniform from here on
        over = __ballot(over) != 0;
        if (fail || over) break;
    }
    fail = __ballot(fail) != 0;
    over = __ballot(over) != 0;
    if (lane == 0) {
        if (fail) atomicOr(a.err, 2u);
        if (over) atomicOr(a.err, 16u);
        if (c_pp) atomicAdd(&a.counts[0], (unsigned long long)c_pp);
        if (c_np) atomicAdd(&a.counts[1], (unsigned long long)c_np);
        if (c_pn) atomicAdd(&a.counts[2], (unsigned long long)c_pn);
        if (c_nn) atomicAdd(&a.counts[3], (unsigned long long)c_nn);
    }
}

// The passive leaves, in groups collected by all the steps: one launch, one
// group per wave at a time.  No waves-per-EU cap: held to 64 VGPRs the
// kernel spilled wave-wide values inside divergent branches (see
// tests/test_code_object.py); it needs ~77.
template <bool SOFT>
__global__ __launch_bounds__(64) void k_fmm_groups(FmmArgs a) {
    __shared__ int s_grp[kStackOct];
    __shared__ uint64_t s_msk[kStackOct];
    __shared__ GroupLds s_g;

    const int lane = lane_id();
    const int64_t ng = a.group_ctr[0] < (uint64_t)a.item_cap ? (int64_t)a.group_ctr[0] : a.item_cap;
    if ((int64_t)blockIdx.x >= ng) return;
    uint32_t *ulist = a.ulist + (int64_t)blockIdx.x * a.ulist_cap;
    uint64_t c_pp = 0, c_np = 0, c_pn = 0, c_nn = 0;
    bool fail = false;
    for (;;) {
        int job = 0;
        if (lane == 0) job = (int)atomicAdd(a.group_ctr + 1, 1u);
        job = uniform_i32(__shfl(job, 0, kWave));
        if (job >= ng) break;
        const uint32_t *it = a.groups + (int64_t)job * kFmmItemWords;
        const int P = uniform_i32((int)it[kItP]);
        const uint32_t aoff = (uint32_t)uniform_i32((int)it[kItAoff]), alen = (uint32_t)uniform_i32((int)it[kItAlen]);
        const uint32_t eoff = (uint32_t)uniform_i32((int)it[kItEoff]), elen = (uint32_t)uniform_i32((int)it[kItElen]);
        if (P < 0 || P >= a.max_nodes || (int64_t)aoff + alen > a.pool_cap || (int64_t)eoff + elen > a.pool_cap) {
            fail = true;
            break;
        }
        fmm_leaf_group<SOFT>(a, P, a.pool + aoff, alen, a.pool + eoff, elen, it, (uint32_t)uniform_i32((int)it[kItLeaves]),
                             s_g, s_grp, s_msk, ulist, c_pp, c_np, c_pn, c_nn, fail);
        fail = __ballot(fail) != 0;
        if (fail) break;
    }
    if (lane == 0) {
        if (fail) atomicOr(a.err, 2u);
        if (c_pp) atomicAdd(&a.counts[0], (unsigned long long)c_pp);
        if (c_np) atomicAdd(&a.counts[1], (unsigned long long)c_np);
        if (c_pn) atomicAdd(&a.counts[2], (unsigned long long)c_pn);
        if (c_nn) atomicAdd(&a.counts[3], (unsigned long long)c_nn);
    }
}

// After the last step: a non-empty next frontier means the tree was deeper
// than the steps launched.
__global__ void k_fmm_tail(const uint32_t *ctr_next, uint32_t *peak, int steps, uint32_t *err) {
    if (threadIdx.x == 0) {
        if (ctr_next[0] != 0u) atomicOr(err, 2u);
        peak[2] = (uint32_t)steps;
    }
}

}  // namespace

namespace {
// The midpoint between t and the float below it (exact in FP64).
double quotient_bound(float t) {
    const float p = std::nextafter(t, -std::numeric_limits<float>::infinity());
    return ((double)p + (double)t) / 2.0;
}
}  // namespace

void launch_fmm(const FmmArgs &args, hipStream_t s) {
    if (args.n <= 0) return;
    FmmArgs a = args;
    const float theta2 = args.theta * args.theta;
    a.theta2_mid = quotient_bound(theta2);
    a.theta_mid = quotient_bound(args.theta);
    launch_fill_u32(a.ctr, 0u, 12, s);
    launch_fill_u32(a.pool_ctr, 0u, 1, s);
    launch_fill_u32(a.group_ctr, 0u, 2, s);
    launch_fill_u32(a.peak, 0u, 4, s);
    hipLaunchKernelGGL(k_fmm_starts, dim3(1), dim3(64), 0, s, a.nodes, a.misc, a.starts, a.nstarts, a.err);
    hipLaunchKernelGGL(k_fmm_seed, dim3(kHeapCap), dim3(64), 0, s, a);
    for (int k = 0; k < a.steps; ++k) {
        if (a.eps2 > 0.0f) hipLaunchKernelGGL(k_fmm_step<true>, dim3(a.waves), dim3(64), 0, s, a, k);
        else hipLaunchKernelGGL(k_fmm_step<false>, dim3(a.waves), dim3(64), 0, s, a, k);
    }
    if (a.eps2 > 0.0f) hipLaunchKernelGGL(k_fmm_groups<true>, dim3(kFmmGroupWaves), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_fmm_groups<false>, dim3(kFmmGroupWaves), dim3(64), 0, s, a);
    hipLaunchKernelGGL(k_fmm_tail, dim3(1), dim3(64), 0, s, a.ctr + 4 * (a.steps % 3), a.peak, a.steps, a.err);
}

}  // namespace nbg
