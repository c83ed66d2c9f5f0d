// Implicit fast multipole method (MultipoleImplicitFMMSolver<1>,
// Solvers/ImplicitFMMSolver.h:14-33) on the octree of the Barnes-Hut build.
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
// Here one wavefront serves one passive leaf L: it replays the traversal
// along the path from L's starting node down to L with the lists held in
// per-wave scratch (lanes classify 64 list entries at a time), so the
// interaction set of every body in L is the reference's: the same criteria
// (SideLengthOverDistance, node-node and particle-node forms, in FP32 with
// the reference's operation order) take the same decisions.  Far-field
// contributions are kept as per-lane partial local fields (translation is
// linear, so partials are translated separately and summed once at L).
// Work above L is repeated by every leaf below it; interaction counts count
// each passive node's far-field events once (by the wave of the leaf that
// holds the node's first body).  Empty octants, which the reference keeps as
// body-less nodes, contribute nothing and have no record here.

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

__global__ __launch_bounds__(64) void k_fmm_starts(const OctRec *__restrict__ nodes, const uint32_t *__restrict__ misc,
                                                   uint32_t *flag, int32_t *starts, uint32_t *nstarts, uint32_t *err) {
    __shared__ HeapEnt q[kHeapCap];
    if (threadIdx.x != 0) return;
    const int64_t T = misc[0];
    int len = 1;
    q[0] = HeapEnt{0, nodes[0].topo.y};
    while (len < kFmmStarts) {
        heap_pop(q, len);
        const HeapEnt t = q[len - 1];
        if (t.slot < 0 || nodes[t.slot].topo.z < 0) break;
        --len;
        const OctRec &r = nodes[t.slot];
        const int c0 = r.topo.z, nch = oct_nchild(r);
        if (c0 <= t.slot || (int64_t)c0 + nch > T) { atomicOr(err, 8u); break; }
        int j = 0;
        for (int oct = 0; oct < 8; ++oct) {
            HeapEnt e{-1, 0};
            if (j < nch && oct_octant(nodes[c0 + j]) == oct) {
                e = HeapEnt{c0 + j, nodes[c0 + j].topo.y};
                ++j;
            }
            q[len++] = e;
            heap_push(q, len);
        }
    }
    int k = 0;
    for (int i = 0; i < len; ++i)
        if (q[i].slot >= 0) {
            flag[q[i].slot] = 1u;
            starts[k++] = q[i].slot;
        }
    *nstarts = (uint32_t)k;
}

// Leaf slots, one per leaf (order immaterial).
__global__ __launch_bounds__(256) void k_fmm_leaves(const int32_t *__restrict__ leaf_start, int64_t n, uint32_t *list,
                                                    uint32_t *count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int s = leaf_start[i];
        if (s >= 0) list[atomicAdd(count, 1u)] = (uint32_t)s;
    }
}

// ---------------------------------------------------------------- criteria
// SideLengthOverDistance, node-node (Descent/DescentCriterion/SideLengthOverDistance.h:33-56):
// 0 = Approximate, else DescendActiveNode (1) | DescendPassiveNode (2).
__device__ __forceinline__ int fmm_recommend(float4 acom, float4 ac, float4 pc, float theta2) {
    const float sa2 = ac.w * ac.w, sp2 = pc.w * pc.w;
    const float dx = pc.x - acom.x, dy = pc.y - acom.y, dz = pc.z - acom.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float big = sa2 < sp2 ? sp2 : sa2;
    if (ieee_div(4.0f * big, d2) < theta2) {
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
__device__ __forceinline__ bool fmm_particle_far(float x, float y, float z, float4 pc, float theta) {
    const float dx = x - pc.x, dy = y - pc.y, dz = z - pc.z;
    const float dist = ieee_sqrt(dx * dx + dy * dy + dz * dz);
    if (!(ieee_div(2.0f * pc.w, dist) < theta)) return false;
    const float h = pc.w / 2.0f;
    const bool in = pc.x - h <= x && pc.y - h <= y && pc.z - h <= z && x < pc.x + h && y < pc.y + h && z < pc.z + h;
    return !in;
}

// A per-lane partial local field (MultipoleAcceleration<2>: vector and
// symmetric gradient xx xy xz yy yz zz), without the factor G.
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

// MultipoleAcceleration::translated (MultipoleAcceleration.h:49-70): a += T d.
__device__ __forceinline__ void fmm_translate(LocalF &F, float dx, float dy, float dz) {
    F.ax = __builtin_fmaf(F.t0, dx, __builtin_fmaf(F.t1, dy, __builtin_fmaf(F.t2, dz, F.ax)));
    F.ay = __builtin_fmaf(F.t1, dx, __builtin_fmaf(F.t3, dy, __builtin_fmaf(F.t4, dz, F.ay)));
    F.az = __builtin_fmaf(F.t2, dx, __builtin_fmaf(F.t4, dy, __builtin_fmaf(F.t5, dz, F.az)));
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

// Appends v for every lane with `want`; returns the new (uniform) length.
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

// Appends the non-empty children (contiguous slots c0 .. c0 + k - 1) of every
// lane's node; returns the new length.  A child range outside the node array
// (an inconsistent tree) sets `bad` and appends nothing for that lane.
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

// Descent::activeTree from `start` for the lanes in `lanes` (the octree's
// FAST group walk, counting interactions).
template <bool SOFT>
__device__ void fmm_walk_from(const FmmArgs &a, int start, uint64_t lanes, FastWave &W, int *s_grp, uint64_t *s_msk,
                              bool &fail) {
    const OctRec *__restrict__ nodes = a.nodes;
    const OctRec st = nodes[start];
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

constexpr int kMaxPath = 40;

template <bool SOFT>
__global__ __launch_bounds__(64) void k_fmm_leaf(FmmArgs a) {
    __shared__ int s_path[kMaxPath];
    __shared__ int s_grp[kStackOct];
    __shared__ uint64_t s_msk[kStackOct];

    const int lane = lane_id();
    const OctRec *__restrict__ nodes = a.nodes;
    const float4 *__restrict__ cen = a.centre;
    const float theta = a.theta, theta2 = a.theta * a.theta, eps2 = a.eps2;
    uint32_t *base = a.scratch + (int64_t)blockIdx.x * a.wave_words;
    uint32_t *A[2] = {base, base + a.cap_a};
    uint32_t *LA = base + 2 * a.cap_a, *NA = base + 3 * a.cap_a;
    uint32_t *E[2] = {base + 4 * a.cap_a, base + 4 * a.cap_a + a.cap_e};
    uint64_t c_pp = 0, c_np = 0, c_pn = 0, c_nn = 0;
    bool fail = false, over = false;

    for (;;) {
        int job = 0;
        if (lane == 0) job = (int)atomicAdd(a.next, 1u);
        job = uniform_i32(__shfl(job, 0, kWave));
        if (job >= (int)*a.nleaves) break;
        const int L = uniform_i32((int)a.leaves[job]);
        const OctRec lr = nodes[L];
        // path from the starting node down to L
        int d = 0;
        {
            int v = L;
            while (v >= 0 && v < a.max_nodes && d < kMaxPath) {
                if (lane == 0) s_path[d] = v;
                ++d;
                if (a.start_flag[v]) break;
                v = nodes[v].topo.w;
            }
            if (v < 0 || v >= a.max_nodes || d >= kMaxPath) { fail = true; break; }
            if (lr.topo.x < 0 || lr.topo.y < 0 || (int64_t)lr.topo.x + lr.topo.y > a.n) { fail = true; break; }
        }
        __builtin_amdgcn_wave_barrier();
        const int S = uniform_i32(s_path[d - 1]);
        // breadth-first active list for S (balancedLockstep), kept in A[0]
        int cur = 1, eq = 0;
        uint32_t *Q[2] = {LA, NA};  // breadth-first queues
        if (lane == 0) Q[0][0] = 0u;
        wave_mem_sync();
        const float4 cS = cen[S];
        int qi = 0, levels = 0;
        while (cur > 0) {
            if (++levels > kMaxPath + 32) {  // deeper than any octree: an inconsistent tree
                fail = true;
                break;
            }
            int nxt = 0;
            for (int b = 0; b < cur; b += kWave) {
                const bool has = b + lane < cur;
                const int an = has ? (int)Q[qi][b + lane] : 0;
                const OctRec ar = nodes[an];
                const int r = has ? fmm_recommend(ar.com, cen[an], cS, theta2) : 0;
                const bool down = has && r == 1 && ar.topo.z >= 0;
                eq = wave_append(A[0], eq, a.cap_a, has && !down, (uint32_t)an, over);
                nxt = wave_append_children(Q[qi ^ 1], nxt, a.cap_a, ar.topo.z, down ? oct_nchild(ar) : 0, a.max_nodes,
                                           over, fail);
            }
            qi ^= 1;
            cur = nxt;
            wave_mem_sync();
        }
        if (__ballot(fail)) {
            fail = true;
            break;
        }
        int ai = 0, na = eq, ne = 0, ei = 0;
        LocalF F = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = d - 1; k >= 0; --k) {
            const int P = uniform_i32(s_path[k]);
            const float4 cP = cen[P];
            const bool leaf = k == 0;
            const bool count_far = nodes[P].topo.x == lr.topo.x;  // one wave counts P's far-field events
            // lockstep partition of the active list against P
            int nla = 0, nna = 0;
            for (int b = 0; b < na; b += kWave) {
                const bool has = b + lane < na;
                const int an = has ? (int)A[ai][b + lane] : 0;
                const OctRec ar = nodes[an];
                const int r = has ? fmm_recommend(ar.com, cen[an], cP, theta2) : 1;
                const bool far = has && r == 0;
                if (far) fmm_to_local(ar.com.x, ar.com.y, ar.com.z, ar.com.w, cP, eps2, F);
                if (count_far) c_nn += __popcll(__ballot(far));
                const bool isleaf = ar.topo.z < 0;
                nla = wave_append(LA, nla, a.cap_a, has && !far && isleaf, (uint32_t)an, over);
                nna = wave_append(NA, nna, a.cap_a, has && !far && !isleaf, (uint32_t)an, over);
            }
            // close bodies of the passive-tree traversals against P
            int nne = 0;
            for (int b = 0; b < ne; b += kWave) {
                const bool has = b + lane < ne;
                const int e = has ? (int)E[ei][b + lane] : 0;
                const float4 bd = a.bodies[e];
                const bool far = has && fmm_particle_far(bd.x, bd.y, bd.z, cP, theta);
                if (far) fmm_to_local(bd.x, bd.y, bd.z, bd.w, cP, eps2, F);
                if (count_far) c_pn += __popcll(__ballot(far));
                nne = wave_append(E[ei ^ 1], nne, a.cap_e, has && !far, (uint32_t)e, over);
            }
            ei ^= 1;
            ne = nne;
            wave_mem_sync();
            if (!leaf) {
                // bodies of the active leaves start a passive-tree traversal at P
                for (int j = 0; j < nla; ++j) {
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
                        if (far) fmm_to_local(bd.x, bd.y, bd.z, bd.w, cP, eps2, F);
                        if (count_far) c_pn += __popcll(__ballot(far));
                        ne = wave_append(E[ei], ne, a.cap_e, has && !far, (uint32_t)e, over);
                    }
                }
                // children of the internal active nodes form the next list
                int nxt = 0;
                for (int b = 0; b < nna; b += kWave) {
                    const bool has = b + lane < nna;
                    const OctRec ar = nodes[has ? (int)NA[b + lane] : 0];
                    nxt = wave_append_children(A[ai ^ 1], nxt, a.cap_a, ar.topo.z, has ? oct_nchild(ar) : 0, a.max_nodes,
                                               over, fail);
                }
                ai ^= 1;
                na = nxt;
                const float4 cN = cen[uniform_i32(s_path[k - 1])];
                fmm_translate(F, cN.x - cP.x, cN.y - cP.y, cN.z - cP.z);
                wave_mem_sync();
                continue;
            }
            // passive leaf: near field and the local field for its bodies
            LocalF T;
            T.ax = wave_fsum(F.ax); T.ay = wave_fsum(F.ay); T.az = wave_fsum(F.az);
            T.t0 = wave_fsum(F.t0); T.t1 = wave_fsum(F.t1); T.t2 = wave_fsum(F.t2);
            T.t3 = wave_fsum(F.t3); T.t4 = wave_fsum(F.t4); T.t5 = wave_fsum(F.t5);
            wave_mem_sync();
            for (int b = 0; b < lr.topo.y; b += kWave) {
                const bool has = b + lane < lr.topo.y;
                const int me = lr.topo.x + b + lane;
                const float4 p = a.bodies[has ? me : lr.topo.x];
                const uint64_t lanes = __ballot(has);
                FastWave W;
                W.px = p.x; W.py = p.y; W.pz = p.z;
                W.rx = 0.0f; W.ry = 0.0f; W.rz = 0.0f;
                W.pp = 0; W.np = 0;
                W.visits = 0; W.lanes = 0; W.leaf_visits = 0; W.leaf_lanes = 0;
                W.G = a.G;
                for (int j = 0; j < nla; ++j) fmm_walk_from<SOFT>(a, uniform_i32((int)LA[j]), lanes, W, s_grp, s_msk, fail);
                for (int j = 0; j < nna; ++j) fmm_walk_from<SOFT>(a, uniform_i32((int)NA[j]), lanes, W, s_grp, s_msk, fail);
                for (int j = 0; j < ne; ++j) {
                    const float4 s = a.bodies[uniform_i32((int)E[ei][j])];
                    if (__builtin_amdgcn_inverse_ballot_w64(lanes)) interact<SOFT>(s.x, s.y, s.z, s.w, eps2, W);
                }
                c_pp += W.pp + (uint64_t)ne * __popcll(lanes);
                c_np += W.np;
                if (has) {
                    const float ox = p.x - cP.x, oy = p.y - cP.y, oz = p.z - cP.z;
                    const float lx = __builtin_fmaf(T.t0, ox, __builtin_fmaf(T.t1, oy, __builtin_fmaf(T.t2, oz, T.ax)));
                    const float ly = __builtin_fmaf(T.t1, ox, __builtin_fmaf(T.t3, oy, __builtin_fmaf(T.t4, oz, T.ay)));
                    const float lz = __builtin_fmaf(T.t2, ox, __builtin_fmaf(T.t4, oy, __builtin_fmaf(T.t5, oz, T.az)));
                    const uint32_t o = a.perm[me];
                    if ((int64_t)o < a.n) {
                        a.ax[o] = (W.rx + lx) * a.G;
                        a.ay[o] = (W.ry + ly) * a.G;
                        a.az[o] = (W.rz + lz) * a.G;
                    }
                }
            }
        }
        fail = __ballot(fail) != 0;  // wave-uniform from here on
        over = __ballot(over) != 0;
        if (fail || over) break;
    }
    if (lane == 0) {
        if (fail) atomicOr(a.err, 2u);
        if (over) atomicOr(a.err, 16u);
        atomicAdd(&a.counts[0], (unsigned long long)c_pp);
        atomicAdd(&a.counts[1], (unsigned long long)c_np);
        atomicAdd(&a.counts[2], (unsigned long long)c_pn);
        atomicAdd(&a.counts[3], (unsigned long long)c_nn);
    }
}

}  // namespace

void launch_fmm(const FmmArgs &a, const int32_t *leaf_start, hipStream_t s) {
    if (a.n <= 0) return;
    launch_fill_u32(a.start_flag, 0u, a.max_nodes, s);
    launch_fill_u32(a.nleaves, 0u, 1, s);
    launch_fill_u32(a.next, 0u, 1, s);
    hipLaunchKernelGGL(k_fmm_starts, dim3(1), dim3(64), 0, s, a.nodes, a.misc, a.start_flag, a.starts, a.nstarts, a.err);
    hipLaunchKernelGGL(k_fmm_leaves, dim3(256), dim3(256), 0, s, leaf_start, a.n, a.leaves, a.nleaves);
    if (a.eps2 > 0.0f) hipLaunchKernelGGL(k_fmm_leaf<true>, dim3(a.waves), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_fmm_leaf<false>, dim3(a.waves), dim3(64), 0, s, a);
}

}  // namespace nbg
