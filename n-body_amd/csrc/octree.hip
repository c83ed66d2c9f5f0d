// Barnes-Hut octree (BarnesHutSolver, Solvers/BarnesHutSolver.h:14-32) built
// on the device.
//
// The reference builds top-down (Trees/Octree.h:33-61,100-147,185-205): the
// root cube has centre min + diag/2 and side maxSide * 1.1; a split
// partitions on p[d] < c[d] (lower half) and gives child k = 4x + 2y + z
// the centre c +- (side/2)/2 and side side/2; levels 0 and 1 always split;
// deeper nodes split while they hold more than maxLeafSize bodies and the
// depth budget (maxDepth below level 2) lasts.
//
// Here every body computes its whole octant path with that float centre
// recurrence (the centres along a path depend only on the path, so all
// bodies of a cell see the same centres the reference computes for it).
// Sorting the paths makes every cell a contiguous run of positions; a cell
// at level L starts at position i iff the paths at i-1 and i share fewer
// than L digits.  Per position, the nodes whose ranges start there form a
// chain (a node and its first non-empty child, ...) that ends in a leaf;
// numbering chains in position order gives the pre-order node ids.  Node
// records then go to "slots" in which the non-empty children of a node are
// contiguous in octant order (empty children, which the reference keeps and
// skips, have no record).  Leaf contents are put in ascending original
// index, the order the reference's partitions leave them in when they are
// stable.

#include "device.hpp"
#include "launch.hpp"

namespace nbg {

namespace {

__device__ __forceinline__ void oct_root(const float *bb, float &cx, float &cy, float &cz, float &side) {
    const float dx = bb[3] - bb[0], dy = bb[4] - bb[1], dz = bb[5] - bb[2];
    cx = bb[0] + dx / 2.0f;
    cy = bb[1] + dy / 2.0f;
    cz = bb[2] + dz / 2.0f;
    float s = dx;  // std::max({x, y, z}): the first largest by operator<
    if (s < dy) s = dy;
    if (s < dz) s = dz;
    side = s * 1.1f;
}

__device__ __forceinline__ int oct_digit(uint64_t hi, uint64_t lo, int l) {
    return l < kOctHiLevels ? (int)((hi >> (3 * (kOctHiLevels - 1 - l))) & 7u)
                            : (int)((lo >> (3 * (kOctLevels - 1 - l))) & 7u);
}

// Centre and side of the level-L cell on a path (OctreeNode::split's
// recurrence, Trees/Octree.h:112-131).
__device__ __forceinline__ void oct_cell(const float *bb, uint64_t hi, uint64_t lo, int L, float &cx, float &cy,
                                         float &cz, float &side) {
    oct_root(bb, cx, cy, cz, side);
    for (int l = 0; l < L; ++l) {
        const int d = oct_digit(hi, lo, l);
        const float cs = side / 2.0f, off = cs / 2.0f;
        cx = cx + ((d & 4) ? off : -off);
        cy = cy + ((d & 2) ? off : -off);
        cz = cz + ((d & 1) ? off : -off);
        side = cs;
    }
}

// Number of leading path digits two keys share (capped at levels).
__device__ __forceinline__ int oct_cpd(uint64_t ha, uint64_t la, uint64_t hb, uint64_t lb, int levels) {
    int c;
    const uint64_t x = ha ^ hb;
    if (x) {
        c = (__clzll((long long)x) - 1) / 3;
    } else {
        const uint64_t y = la ^ lb;
        c = y ? kOctHiLevels + (__clzll((long long)y) - 25) / 3 : kOctLevels;
    }
    return c < levels ? c : levels;
}

struct OctKeys {
    const uint64_t *hi;  // sorted
    const uint64_t *lo;  // sorted; null when levels <= kOctHiLevels
    int64_t n;
    int levels;
    // Forest (many systems, sorted by system, then path): the system of each
    // position, each system's first position (nsys + 1 entries) and its
    // levels and leaf size; null for one system.  The system acts as a digit
    // above level 0: positions of different systems share no digit (-1), so
    // every cell, the root included, ends at its system's boundary.
    const uint64_t *sys = nullptr;
    const int64_t *sys_first = nullptr;
    const int32_t *lv_tab = nullptr;
    const int32_t *leaf_tab = nullptr;
    __device__ __forceinline__ uint64_t H(int64_t i) const { return hi[i]; }
    __device__ __forceinline__ uint64_t L(int64_t i) const { return lo ? lo[i] : 0ull; }
    __device__ __forceinline__ int S(int64_t i) const { return sys ? (int)sys[i] : 0; }
    __device__ __forceinline__ int lev(int s) const { return lv_tab ? lv_tab[s] : levels; }
    __device__ __forceinline__ int64_t end_of(int s) const { return sys ? sys_first[s + 1] : n; }
    __device__ __forceinline__ int cpd_to(int64_t k, uint64_t h, uint64_t l, int s) const {
        if (sys && (int)sys[k] != s) return -1;
        return oct_cpd(H(k), L(k), h, l, lev(s));
    }
};

// First position after i outside the level-L cell that contains i, searched
// in (i, lim]: exponential probe, then bisection.
__device__ int64_t cell_end(const OctKeys &K, int64_t i, uint64_t h, uint64_t l, int sy, int L, int64_t lim) {
    int64_t in = i, out = lim;
    for (int64_t step = 1;; step *= 2) {
        const int64_t k = i + step;
        if (k >= lim) break;
        if (K.cpd_to(k, h, l, sy) < L) { out = k; break; }
        in = k;
    }
    while (out - in > 1) {
        const int64_t mid = in + (out - in) / 2;
        if (K.cpd_to(mid, h, l, sy) < L) out = mid;
        else in = mid;
    }
    return out;
}

// First position of the level-L cell that contains i.
__device__ int64_t cell_start(const OctKeys &K, int64_t i, uint64_t h, uint64_t l, int sy, int L) {
    int64_t in = i, out = -1;
    for (int64_t step = 1;; step *= 2) {
        const int64_t k = i - step;
        if (k < 0) break;
        if (K.cpd_to(k, h, l, sy) < L) { out = k; break; }
        in = k;
    }
    while (in - out > 1) {
        const int64_t mid = out + (in - out) / 2;
        if (K.cpd_to(mid, h, l, sy) < L) out = mid;
        else in = mid;
    }
    return in;
}

__device__ __forceinline__ int first_level(const OctKeys &K, int64_t i) {
    return i == 0 ? 0 : K.cpd_to(i - 1, K.H(i), K.L(i), K.S(i)) + 1;
}

// The chain of nodes whose ranges start at position i, shallowest first:
// emit(level, count, leaf).  A node at level L >= 3 exists iff its parent
// holds more than `leaf` bodies (NodeBase::refine, Trees/NodeBase.h:57-75);
// levels 0-2 always exist (Tree::depthSplit(2), Trees/Tree.h:172-194).
template <typename F>
__device__ void oct_chain(const OctKeys &K, int64_t i, int leaf, F &&emit) {
    const uint64_t h = K.H(i), l = K.L(i);
    const int sy = K.S(i), levels = K.lev(sy);
    if (K.leaf_tab) leaf = K.leaf_tab[sy];
    const int64_t send = K.end_of(sy);
    const int lv0 = first_level(K, i);
    if (lv0 > levels) return;
    int64_t lim = send;
    if (lv0 > 2) {
        const int64_t s0 = cell_start(K, i, h, l, sy, lv0 - 1);
        const int64_t e0 = cell_end(K, i, h, l, sy, lv0 - 1, send);
        if (e0 - s0 <= leaf) return;
        lim = e0;
    }
    for (int L = lv0;; ++L) {
        const int64_t end = L == 0 ? send : cell_end(K, i, h, l, sy, L, lim);
        const int64_t cnt = end - i;
        const bool is_leaf = L >= 2 && (cnt <= leaf || L >= levels);
        emit(L, cnt, is_leaf);
        if (is_leaf) break;
        lim = end;
    }
}

// SideLengthOverDistance (Descent/DescentCriterion/SideLengthOverDistance.h:17-22):
// side / sqrt(d2) < theta in FP32 is monotone in d2, so it holds exactly
// when d2 exceeds the largest float d2 for which it fails.
__device__ __forceinline__ bool oct_mac_passes(float side, float d2, float theta) {
    return ieee_div(side, ieee_sqrt(d2)) < theta;
}

__device__ float oct_threshold(float side, float theta) {
    uint32_t lo = 0u, hi = 0x7f800000u;  // 0 .. +inf
    if (!oct_mac_passes(side, __uint_as_float(hi), theta)) return __uint_as_float(hi);
    if (oct_mac_passes(side, 0.0f, theta)) return -1.0f;
    while (hi - lo > 1u) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (oct_mac_passes(side, __uint_as_float(mid), theta)) hi = mid;
        else lo = mid;
    }
    return __uint_as_float(lo);
}

// ------------------------------------------------------------------ kernels

// The octant path of a point over `levels` levels of the cube of bbox bb.
__device__ __forceinline__ void oct_path(const float *bb, int levels, float px, float py, float pz, uint64_t &hi,
                                         uint64_t &lo) {
    float cx, cy, cz, side;
    oct_root(bb, cx, cy, cz, side);
    hi = 0;
    lo = 0;
    for (int l = 0; l < levels; ++l) {
        // partition<D> (Trees/Octree.h:33-61): p[d] < c[d] goes to the lower half
        const int xb = !(px < cx), yb = !(py < cy), zb = !(pz < cz);
        const uint64_t d = (uint64_t)(xb * 4 + yb * 2 + zb);
        if (l < kOctHiLevels) hi |= d << (3 * (kOctHiLevels - 1 - l));
        else lo |= d << (3 * (kOctLevels - 1 - l));
        const float cs = side / 2.0f, off = cs / 2.0f;
        cx = cx + (xb ? off : -off);
        cy = cy + (yb ? off : -off);
        cz = cz + (zb ? off : -off);
        side = cs;
    }
}

template <bool LO>
__global__ __launch_bounds__(256) void k_oct_keys(const float *__restrict__ x, const float *__restrict__ y,
                                                  const float *__restrict__ z, int64_t n,
                                                  const float *__restrict__ bb, int levels, uint64_t *khi,
                                                  uint64_t *klo, uint64_t *klo_copy, uint32_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t hi, lo;
    oct_path(bb, levels, x[i], y[i], z[i], hi, lo);
    khi[i] = hi;
    if (LO) {
        klo[i] = lo;
        klo_copy[i] = lo;
    }
    idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_gather_u64(const uint64_t *__restrict__ src,
                                                    const uint32_t *__restrict__ idx, int64_t n,
                                                    uint64_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    dst[i] = j < n ? src[j] : 0ull;
}

// Forest: each body's path over its own system's cube and depth, kept in
// original order (khi, klo) and copied into the sort buffer (the low word
// when some system has more than kOctHiLevels levels, else the high word).
template <bool LO>
__global__ __launch_bounds__(256) void k_oct_keys_forest(const float *__restrict__ x, const float *__restrict__ y,
                                                         const float *__restrict__ z, int64_t n,
                                                         const uint32_t *__restrict__ bsys,
                                                         const float *__restrict__ bbox6,
                                                         const int32_t *__restrict__ lv_tab, uint64_t *khi,
                                                         uint64_t *klo, uint64_t *ksort, uint32_t *idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sy = bsys[i];
    uint64_t hi, lo;
    oct_path(bbox6 + 6 * (int64_t)sy, lv_tab[sy], x[i], y[i], z[i], hi, lo);
    khi[i] = hi;
    if (LO) klo[i] = lo;
    ksort[i] = LO ? lo : hi;
    idx[i] = (uint32_t)i;
}

// dst[i] = system of the body at sorted position i (a sort key)
__global__ __launch_bounds__(256) void k_gather_sys(const uint32_t *__restrict__ bsys,
                                                    const uint32_t *__restrict__ idx, int64_t n,
                                                    uint64_t *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    dst[i] = j < n ? (uint64_t)bsys[j] : ~0ull;
}

__global__ __launch_bounds__(256) void k_oct_count(OctKeys K, int leaf, uint32_t *nown) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.n) return;
    uint32_t c = 0;
    oct_chain(K, i, leaf, [&](int, int64_t, bool) { ++c; });
    nown[i] = c;
}

__global__ __launch_bounds__(256) void k_oct_nodes(OctKeys K, int leaf, const uint32_t *__restrict__ nbase,
                                                   int64_t cap, int32_t *nfirst, int32_t *ncount,
                                                   int32_t *nparent, uint32_t *nmeta, uint32_t *occ,
                                                   uint32_t *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.n) return;
    const uint64_t h = K.H(i), l = K.L(i);
    const int64_t base = nbase[i];
    int k = 0;
    oct_chain(K, i, leaf, [&](int L, int64_t cnt, bool is_leaf) {
        const int64_t me = base + k;
        ++k;
        int64_t parent;
        if (L == 0) {
            parent = -1;
        } else if (k > 1) {
            parent = me - 1;
        } else {
            // first node of the chain: its parent's range starts further left
            const int64_t j = cell_start(K, i, h, l, K.S(i), L - 1);
            parent = (int64_t)nbase[j] + (L - 1 - first_level(K, j));
        }
        if (me >= cap || parent >= me || (L > 0 && parent < 0)) {
            atomicOr(err, 8u);
            return;
        }
        const int oct = L > 0 ? oct_digit(h, l, L - 1) : 0;
        nfirst[me] = (int32_t)i;
        ncount[me] = (int32_t)cnt;
        nparent[me] = (int32_t)parent;
        nmeta[me] = (uint32_t)L | ((uint32_t)oct << 8) | ((is_leaf ? 1u : 0u) << 12);
        if (parent >= 0) atomicOr(&occ[parent], 1u << oct);
    });
}

__global__ __launch_bounds__(256) void k_oct_nchild(const uint32_t *__restrict__ misc, int64_t cap,
                                                    const uint32_t *__restrict__ nmeta,
                                                    const uint32_t *__restrict__ occ, uint32_t *nch) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= cap) return;
    const int64_t T = misc[0];
    nch[id] = (id < T && !((nmeta[id] >> 12) & 1u)) ? (uint32_t)__popc(occ[id]) : 0u;
}

__global__ void k_oct_levels(const float *__restrict__ bb, int levels, float theta, float *thr) {
    const int L = threadIdx.x;
    if (L > levels) return;
    float cx, cy, cz, side;
    oct_root(bb, cx, cy, cz, side);
    for (int l = 0; l < L; ++l) side = side / 2.0f;
    thr[L] = oct_threshold(side, theta);
}

// Forest: block s computes system s's table (thr + kOctThrStride s) from
// its cube, depth and theta (sys_prm[s].z).
__global__ void k_oct_levels_forest(const float *__restrict__ bbox6, const int32_t *__restrict__ lv_tab,
                                    const float4 *__restrict__ sys_prm, float *thr) {
    const int sy = blockIdx.x, L = threadIdx.x;
    if (L > lv_tab[sy]) return;
    float cx, cy, cz, side;
    oct_root(bbox6 + 6 * sy, cx, cy, cz, side);
    for (int l = 0; l < L; ++l) side = side / 2.0f;
    thr[kOctThrStride * sy + L] = oct_threshold(side, sys_prm[sy].z);
}

// Slot of node id of system sy: a root takes slot sy (the roots of the
// nroots systems come first), any other node its parent's first-child slot
// plus the number of non-empty siblings in lower octants.  -1 if the parent
// link is invalid (parents precede their children in pre-order).
__device__ __forceinline__ int64_t oct_slot(int64_t id, int sy, int nroots, const int32_t *__restrict__ nparent,
                                            const uint32_t *__restrict__ nmeta, const uint32_t *__restrict__ occ,
                                            const uint32_t *__restrict__ cb) {
    const int32_t p = nparent[id];
    if (p < 0) return ((nmeta[id] & 0xFFu) == 0u && sy < nroots) ? sy : -1;
    if (p >= id) return -1;
    const uint32_t oct = (nmeta[id] >> 8) & 7u;
    return nroots + (int64_t)cb[p] + __popc(occ[p] & ((1u << oct) - 1u));
}

__global__ __launch_bounds__(256) void k_oct_emit(const uint32_t *__restrict__ misc, int64_t cap, int nroots,
                                                  const float *__restrict__ bbox6, OctKeys K,
                                                  const float *__restrict__ thr,
                                                  const int32_t *__restrict__ nfirst,
                                                  const int32_t *__restrict__ ncount,
                                                  const int32_t *__restrict__ nparent,
                                                  const uint32_t *__restrict__ nmeta,
                                                  const uint32_t *__restrict__ occ,
                                                  const uint32_t *__restrict__ cb, OctRec *nodes, float4 *centre,
                                                  int32_t *leaf_start, uint32_t *err) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t T = misc[0];
    if (id >= T || id >= cap) return;
    const uint32_t meta = nmeta[id];
    const int L = (int)(meta & 0xFFu);
    const uint32_t oct = (meta >> 8) & 7u;
    const bool is_leaf = (meta >> 12) & 1u;
    const int32_t p = nparent[id];
    const int32_t first = nfirst[id];
    if ((L == 0) != (p < 0) || p >= id || first < 0 || first >= K.n) {
        atomicOr(err, 8u);
        return;
    }
    const int sy = K.S(first);
    if (L > K.lev(sy)) {
        atomicOr(err, 8u);
        return;
    }
    const int64_t slot = oct_slot(id, sy, nroots, nparent, nmeta, occ, cb);
    const int64_t pslot = p < 0 ? -1 : oct_slot(p, sy, nroots, nparent, nmeta, occ, cb);
    const uint32_t nch = is_leaf ? 0u : (uint32_t)__popc(occ[id]);
    const int64_t child = is_leaf ? -1 : nroots + (int64_t)cb[id];
    if (slot < 0 || slot >= T || (p >= 0 && pslot < 0) || child + (int64_t)nch > T) {
        atomicOr(err, 8u);
        return;
    }
    float cx, cy, cz, side;
    oct_cell(bbox6 + 6 * (int64_t)sy, K.H(first), K.L(first), L, cx, cy, cz, side);
    OctRec r;
    r.com = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    // exclusionRegion(OctreeNode) = [c - side, c + side) (Descent/DescentCriterionType.h:81-85)
    r.lo = make_float4(cx - side, cy - side, cz - side, thr[kOctThrStride * (int64_t)sy + L]);
    r.hi = make_float4(cx + side, cy + side, cz + side, __uint_as_float(oct | (nch << 8) | ((uint32_t)L << 16)));
    r.topo = make_int4(first, ncount[id], (int)child, (int)pslot);
    nodes[slot] = r;
    centre[slot] = make_float4(cx, cy, cz, side);
    if (is_leaf) leaf_start[first] = (int32_t)slot;
}

// Leaf contents in ascending original index (insertion sort; a leaf at the
// depth limit holding more than `leaf` bodies has equal paths and is already
// in that order, which insertion sort passes through in linear time).
__global__ __launch_bounds__(256) void k_oct_leafsort(int64_t n, const int32_t *__restrict__ leaf_start,
                                                      const OctRec *__restrict__ nodes, int64_t T_cap,
                                                      const uint32_t *__restrict__ in, uint32_t *out,
                                                      uint32_t *err) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t s = leaf_start[i];
    if (s < 0) return;
    if (s >= T_cap) { atomicOr(err, 8u); return; }
    const int4 t = nodes[s].topo;
    if (t.x != (int)i || t.y < 1 || i + t.y > n) { atomicOr(err, 8u); return; }
    for (int j = 0; j < t.y; ++j) {
        const uint32_t v = in[i + j];
        int k = j;
        while (k > 0 && out[i + k - 1] > v) {
            out[i + k] = out[i + k - 1];
            --k;
        }
        out[i + k] = v;
    }
}

// ------------------------------------------------------------------ summaries
// CenterOfMassSummary (Summaries/CenterOfMassSummary.h:30-55): a leaf sums
// M and C = sum p*m over its bodies in order, an internal node over its
// children in octant order; then C / M.  Empty children contribute +0 to
// both sums in the reference, which leaves them unchanged, so they are
// skipped.  The root's level-1 children are summarised even when empty
// (Tree::summarizeTreeTop, Trees/Tree.h:196-216), giving 0/0 = NaN there and
// so a NaN root COM whenever a level-1 octant is empty.
//
// Scheduling as for the LBVH: workgroups own kSummBodies consecutive bodies,
// climb with workgroup-scope arrival counters while the parent's range stays
// inside, and queue straddling parents for one final workgroup.

__device__ __forceinline__ uint32_t arrive(uint32_t *flag) {
    return __hip_atomic_fetch_add(flag, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void oct_summ_leaf(int s, int4 t, const float4 *__restrict__ bodies, OctRec *nodes,
                                              float4 *quad) {
    float M = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (int i = t.x; i < t.x + t.y; ++i) {
        const float4 b = bodies[i];
        M += b.w;
        cx = cx + b.x * b.w;
        cy = cy + b.y * b.w;
        cz = cz + b.z * b.w;
    }
    nodes[s].com = make_float4(ieee_div(cx, M), ieee_div(cy, M), ieee_div(cz, M), M);
    if (quad) {
        float Q[6];
        quad_bodies(bodies, t.x, t.y, ieee_div(cx, M), ieee_div(cy, M), ieee_div(cz, M), Q);
        quad_store(quad, s, Q);
    }
}

// MultipoleMassSummary::summarize(childNodes) (MultipoleMassSummary.h:54-65)
// over the 8 octants: 0 + ((t0 + t1) + (t2 + t3)), then + ((t4 + t5) +
// (t6 + t7)).  An empty octant's term, outer(0 - c, 0 - c) * 0 + 0, is +0
// for a finite c, and is used as such.
__device__ __forceinline__ void oct_quad_inner(int p, const OctRec &r, const OctRec *nodes, float4 *quad,
                                               float4 com) {
    const int c0 = r.topo.z, nch = oct_nchild(r);
    float t[8][6];
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int e = 0; e < 6; ++e) t[q][e] = 0.0f;
    for (int k = 0; k < nch; ++k) {
        const OctRec &ch = nodes[c0 + k];
        const float4 a = ch.com;
        const float4 q0 = quad[2 * (c0 + k)], q1 = quad[2 * (c0 + k) + 1];
        const float cq[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
        float u[6];
        quad_term(a.x - com.x, a.y - com.y, a.z - com.z, a.w, u);
        const int o = oct_octant(ch) & 7;
        // a dynamic octant index into t would spill it to scratch
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (q == o)
#pragma unroll
                for (int e = 0; e < 6; ++e) t[q][e] = u[e] + cq[e];
    }
    float Q[6];
#pragma unroll
    for (int e = 0; e < 6; ++e) {
        Q[e] = 0.0f + ((t[0][e] + t[1][e]) + (t[2][e] + t[3][e]));
        Q[e] = Q[e] + ((t[4][e] + t[5][e]) + (t[6][e] + t[7][e]));
    }
    quad_store(quad, p, Q);
}

__device__ __forceinline__ void oct_summ_inner(int p, const OctRec &r, OctRec *nodes, float4 *quad) {
    const int c0 = r.topo.z, nch = oct_nchild(r);
    float M = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    for (int k = 0; k < nch; ++k) {
        const float4 c = nodes[c0 + k].com;
        M += c.w;
        cx = cx + c.x * c.w;
        cy = cy + c.y * c.w;
        cz = cz + c.z * c.w;
    }
    float4 com = make_float4(ieee_div(cx, M), ieee_div(cy, M), ieee_div(cz, M), M);
    if (oct_level(r) == 0 && nch < 8) {
        const float q = __builtin_nanf("");
        com.x = q; com.y = q; com.z = q;
    }
    nodes[p].com = com;
    if (quad) oct_quad_inner(p, r, nodes, quad, com);
}

__global__ __launch_bounds__(kSummThreads) void k_oct_summ_local(const float4 *__restrict__ bodies, int64_t n,
                                                        OctRec *nodes, const int32_t *__restrict__ leaf_start,
                                                        const uint32_t *__restrict__ misc, uint32_t *flags,
                                                        uint32_t *glist, int64_t glist_cap, float4 *quad,
                                                        uint32_t *err) {
    const int64_t lo = (int64_t)blockIdx.x * kSummBodies, hi = lo + kSummBodies;
    const int64_t end = hi < n ? hi : n;
    const int64_t T = misc[0];
    for (int64_t i = lo + threadIdx.x; i < end; i += blockDim.x) {
        const int s = leaf_start[i];
        if (s < 0) continue;
        if (s >= T) { atomicOr(err, 4u); continue; }
        const int4 t = nodes[s].topo;
        if (t.x != (int)i || t.y < 1 || (int64_t)t.x + t.y > n || t.z >= 0) { atomicOr(err, 4u); continue; }
        oct_summ_leaf(s, t, bodies, nodes, quad);
        int parent = t.w;
        for (int hops = 0; parent >= 0; ++hops) {
            // parents sit in lower slots than their children; a valid tree
            // is at most kOctLevels deep
            if (parent >= s || hops > kOctLevels) { atomicOr(err, 4u); break; }
            const OctRec r = nodes[parent];
            const int4 tp = r.topo;
            const bool local = tp.x >= lo && (int64_t)tp.x + tp.y <= hi;
            if (!local) {
                const uint32_t k = atomicAdd(&glist[0], 1u);
                if (k < glist_cap) glist[64 + k] = (uint32_t)parent;
                else atomicOr(err, 4u);
                break;
            }
            if (arrive(&flags[parent]) + 1u < (uint32_t)oct_nchild(r)) break;
            if (tp.z <= parent || tp.z + oct_nchild(r) > T || tp.w >= parent) { atomicOr(err, 4u); break; }
            oct_summ_inner(parent, r, nodes, quad);
            parent = tp.w;
        }
    }
}

// Middle pass (as the linear BVH's): workgroup g owns kSummMidBodies bodies
// and the queued parents whose first body lies there, climbs while the node
// stays inside, and forwards a parent that leaves the range in place of the
// arrival.
__global__ __launch_bounds__(1024) void k_oct_summ_mid(OctRec *nodes, const uint32_t *__restrict__ misc,
                                                       uint32_t *flags, const uint32_t *gin, uint32_t *gout,
                                                       int64_t qcap, float4 *quad, uint32_t *err) {
    const int64_t T = misc[0];
    const int64_t lo = (int64_t)blockIdx.x * kSummMidBodies, hi = lo + kSummMidBodies;
    const uint32_t cnt = gin[0] < qcap ? gin[0] : (uint32_t)qcap;
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        int parent = (int)gin[64 + j];
        if (parent < 0 || parent >= T) { atomicOr(err, 4u); continue; }
        const int first = nodes[parent].topo.x;
        if (first < lo || first >= hi) continue;  // another workgroup's entry
        for (int hops = 0; parent >= 0; ++hops) {
            if (parent >= T || hops > kOctLevels) { atomicOr(err, 4u); break; }
            const OctRec r = nodes[parent];
            const int4 tp = r.topo;
            if (!(tp.x >= lo && (int64_t)tp.x + tp.y <= hi)) {
                const uint32_t k = atomicAdd(&gout[0], 1u);
                if (k < qcap) gout[64 + k] = (uint32_t)parent;
                else atomicOr(err, 4u);
                break;
            }
            if (arrive(&flags[parent]) + 1u < (uint32_t)oct_nchild(r)) break;
            if (tp.z <= parent || tp.z + oct_nchild(r) > T || tp.w >= parent) { atomicOr(err, 4u); break; }
            oct_summ_inner(parent, r, nodes, quad);
            parent = tp.w;
        }
    }
}

__global__ __launch_bounds__(1024) void k_oct_summ_global(OctRec *nodes, const uint32_t *__restrict__ misc,
                                                          uint32_t *flags, const uint32_t *glist,
                                                          int64_t glist_cap, float4 *quad, uint32_t *err) {
    const int64_t T = misc[0];
    const uint32_t cnt = glist[0] < glist_cap ? glist[0] : (uint32_t)glist_cap;
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        int parent = (int)glist[64 + j];
        for (int hops = 0; parent >= 0; ++hops) {
            if (parent >= T || hops > kOctLevels) { atomicOr(err, 4u); break; }
            const OctRec r = nodes[parent];
            if (arrive(&flags[parent]) + 1u < (uint32_t)oct_nchild(r)) break;
            if (r.topo.z <= parent || r.topo.z + oct_nchild(r) > T || r.topo.w >= parent) {
                atomicOr(err, 4u);
                break;
            }
            oct_summ_inner(parent, r, nodes, quad);
            parent = r.topo.w;
        }
    }
}

inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

inline OctKeys make_keys(const uint64_t *keys, const OctBuffers &ob, int64_t n, int max_depth) {
    OctKeys K;
    K.hi = keys;
    K.levels = max_depth + 2;
    K.lo = K.levels > kOctHiLevels ? ob.lo_sorted : nullptr;
    K.n = n;
    return K;
}

}  // namespace

void launch_octree_keys(const float *x, const float *y, const float *z, int64_t n, const float *bbox6,
                        int max_depth, OctBuffers &ob, uint64_t *keys, uint32_t *idx, uint32_t *idx_alt,
                        hipStream_t s) {
    if (n <= 0) return;
    const int levels = max_depth + 2;
    const unsigned g = blocks_for(n, 256);
    // With lower-level digits the low word is sorted first from `keys`/`idx`;
    // otherwise the high word is sorted directly from `keys`/`idx_alt`.
    if (levels > kOctHiLevels)
        hipLaunchKernelGGL(k_oct_keys<true>, dim3(g), dim3(256), 0, s, x, y, z, n, bbox6, levels, ob.key_hi, keys,
                           ob.key_lo, idx);
    else
        hipLaunchKernelGGL(k_oct_keys<false>, dim3(g), dim3(256), 0, s, x, y, z, n, bbox6, levels, keys,
                           (uint64_t *)nullptr, (uint64_t *)nullptr, idx_alt);
}

void launch_octree_sort(int64_t n, int max_depth, OctBuffers &ob, uint64_t *keys, uint32_t *idx,
                        uint64_t *keys_alt, uint32_t *idx_alt, uint32_t *sort_status, uint32_t *err,
                        hipStream_t s) {
    if (n <= 0) return;
    const int levels = max_depth + 2;
    const unsigned g = blocks_for(n, 256);
    if (levels > kOctHiLevels) {
        const int lo_bits = 3 * (levels - kOctHiLevels);
        // low word: (keys, idx) -> (keys_alt, idx_alt) after an odd pass count
        int passes = (lo_bits + 7) / 8;
        if (!(passes & 1)) ++passes;
        launch_radix_sort_u64(keys, idx, keys_alt, idx_alt, n, passes, sort_status, err, s);
        if (n == 1) launch_fill_u32(idx_alt, 0u, 1, s);
        hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, ob.key_hi, idx_alt, n, keys);
    }
    // high word (63 bits, 8 passes): (keys, idx_alt) -> (keys, idx_alt)
    launch_radix_sort_u64(keys, idx_alt, keys_alt, idx, n, 8, sort_status, err, s);
    if (levels > kOctHiLevels)
        hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, ob.key_lo, idx_alt, n, ob.lo_sorted);
}

void launch_octree_topology(int64_t n, const float *bbox6, int leaf_size, int max_depth, float theta,
                            OctBuffers &ob, const uint64_t *keys, const uint32_t *idx_alt, uint32_t *idx,
                            int32_t *leaf_start, uint32_t *err, hipStream_t s) {
    if (n <= 0) return;
    const int levels = max_depth + 2;
    const OctKeys K = make_keys(keys, ob, n, max_depth);
    const unsigned g = blocks_for(n, 256), gc = blocks_for(ob.cap, 256);
    hipLaunchKernelGGL(k_oct_count, dim3(g), dim3(256), 0, s, K, leaf_size, ob.nown);
    launch_scan_u32(ob.nown, ob.nbase, n, ob.scan_ws, ob.misc, s);
    launch_fill_u32(ob.occ, 0u, ob.cap, s);
    hipLaunchKernelGGL(k_oct_nodes, dim3(g), dim3(256), 0, s, K, leaf_size, ob.nbase, ob.cap, ob.nfirst, ob.ncount,
                       ob.nparent, ob.nmeta, ob.occ, err);
    hipLaunchKernelGGL(k_oct_nchild, dim3(gc), dim3(256), 0, s, ob.misc, ob.cap, ob.nmeta, ob.occ, ob.cb);
    launch_scan_u32(ob.cb, ob.cb, ob.cap, ob.scan_ws, ob.misc + 1, s);
    hipLaunchKernelGGL(k_oct_levels, dim3(1), dim3(64), 0, s, bbox6, levels, theta, ob.thr);
    launch_fill_u32((uint32_t *)leaf_start, 0xFFFFFFFFu, n, s);
    hipLaunchKernelGGL(k_oct_emit, dim3(gc), dim3(256), 0, s, ob.misc, ob.cap, 1, bbox6, K, ob.thr, ob.nfirst,
                       ob.ncount, ob.nparent, ob.nmeta, ob.occ, ob.cb, ob.nodes, ob.centre, leaf_start, err);
    hipLaunchKernelGGL(k_oct_leafsort, dim3(g), dim3(256), 0, s, n, leaf_start, ob.nodes, ob.cap, idx_alt, idx,
                       err);
}

void launch_octree_forest_build(const float *x, const float *y, const float *z, int64_t n, int nsys,
                                const uint32_t *bsys, const float *bbox6, const int32_t *lv_tab, int max_levels,
                                OctBuffers &ob, uint64_t *keys, uint32_t *idx, uint64_t *keys_alt, uint32_t *idx_alt,
                                uint32_t *sort_status, uint32_t *err, hipStream_t s) {
    if (n <= 0 || nsys <= 0) return;
    const unsigned g = blocks_for(n, 256);
    const bool lo = max_levels > kOctHiLevels;
    // paths in original order (ob.key_hi / ob.key_lo) and the first sort key
    if (lo)
        hipLaunchKernelGGL(k_oct_keys_forest<true>, dim3(g), dim3(256), 0, s, x, y, z, n, bsys, bbox6, lv_tab,
                           ob.key_hi, ob.key_lo, keys, idx);
    else
        hipLaunchKernelGGL(k_oct_keys_forest<false>, dim3(g), dim3(256), 0, s, x, y, z, n, bsys, bbox6, lv_tab,
                           ob.key_hi, (uint64_t *)nullptr, keys, idx_alt);
    if (lo) {
        // low word (odd pass count): (keys, idx) -> (keys_alt, idx_alt)
        int passes = (3 * (max_levels - kOctHiLevels) + 7) / 8;
        if (!(passes & 1)) ++passes;
        launch_radix_sort_u64(keys, idx, keys_alt, idx_alt, n, passes, sort_status, err, s);
        if (n == 1) launch_fill_u32(idx_alt, 0u, 1, s);
        hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, ob.key_hi, idx_alt, n, keys);
    }
    // high word, 8 passes: (keys, idx_alt) -> (keys, idx_alt)
    launch_radix_sort_u64(keys, idx_alt, keys_alt, idx, n, 8, sort_status, err, s);
    // the system above every digit (an even pass count): (keys_alt, idx_alt) -> (keys_alt, idx_alt)
    hipLaunchKernelGGL(k_gather_sys, dim3(g), dim3(256), 0, s, bsys, idx_alt, n, keys_alt);
    launch_radix_sort_u64(keys_alt, idx_alt, keys, idx, n, nsys <= 65536 ? 2 : 4, sort_status, err, s);
    // sorted paths
    hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, ob.key_hi, idx_alt, n, keys);
    if (lo) hipLaunchKernelGGL(k_gather_u64, dim3(g), dim3(256), 0, s, ob.key_lo, idx_alt, n, ob.lo_sorted);
}

void launch_octree_forest_topology(int64_t n, int nsys, const float *bbox6, const float4 *sys_prm,
                                   const int32_t *lv_tab, const int32_t *leaf_tab, const int64_t *sys_first,
                                   int max_levels, OctBuffers &ob, const uint64_t *keys, const uint64_t *sys_sorted,
                                   const uint32_t *idx_alt, uint32_t *idx, int32_t *leaf_start, uint32_t *err,
                                   hipStream_t s) {
    if (n <= 0 || nsys <= 0) return;
    OctKeys K;
    K.hi = keys;
    K.levels = max_levels;
    K.lo = max_levels > kOctHiLevels ? ob.lo_sorted : nullptr;
    K.n = n;
    K.sys = sys_sorted;
    K.sys_first = sys_first;
    K.lv_tab = lv_tab;
    K.leaf_tab = leaf_tab;
    const unsigned g = blocks_for(n, 256), gc = blocks_for(ob.cap, 256);
    hipLaunchKernelGGL(k_oct_count, dim3(g), dim3(256), 0, s, K, 0, ob.nown);
    launch_scan_u32(ob.nown, ob.nbase, n, ob.scan_ws, ob.misc, s);
    launch_fill_u32(ob.occ, 0u, ob.cap, s);
    hipLaunchKernelGGL(k_oct_nodes, dim3(g), dim3(256), 0, s, K, 0, ob.nbase, ob.cap, ob.nfirst, ob.ncount,
                       ob.nparent, ob.nmeta, ob.occ, err);
    hipLaunchKernelGGL(k_oct_nchild, dim3(gc), dim3(256), 0, s, ob.misc, ob.cap, ob.nmeta, ob.occ, ob.cb);
    launch_scan_u32(ob.cb, ob.cb, ob.cap, ob.scan_ws, ob.misc + 1, s);
    hipLaunchKernelGGL(k_oct_levels_forest, dim3((unsigned)nsys), dim3(64), 0, s, bbox6, lv_tab, sys_prm, ob.thr);
    launch_fill_u32((uint32_t *)leaf_start, 0xFFFFFFFFu, n, s);
    hipLaunchKernelGGL(k_oct_emit, dim3(gc), dim3(256), 0, s, ob.misc, ob.cap, nsys, bbox6, K, ob.thr, ob.nfirst,
                       ob.ncount, ob.nparent, ob.nmeta, ob.occ, ob.cb, ob.nodes, ob.centre, leaf_start, err);
    hipLaunchKernelGGL(k_oct_leafsort, dim3(g), dim3(256), 0, s, n, leaf_start, ob.nodes, ob.cap, idx_alt, idx,
                       err);
}

void launch_octree_summaries(const float4 *bodies, int64_t n, OctBuffers &ob, const int32_t *leaf_start,
                             uint32_t *err, hipStream_t s, float4 *quad) {
    if (n <= 0) return;
    // queue layout: local -> middle at glist, middle -> global at glist + cap + 64
    uint32_t *gmid = ob.glist + ob.cap + 64;
    launch_fill_u32(ob.flags, 0u, ob.cap, s);
    launch_fill_u32(ob.glist, 0u, 64, s);
    launch_fill_u32(gmid, 0u, 64, s);
    hipLaunchKernelGGL(k_oct_summ_local, dim3(blocks_for(n, kSummBodies)), dim3(kSummThreads), 0, s, bodies, n, ob.nodes,
                       leaf_start, ob.misc, ob.flags, ob.glist, ob.cap, quad, err);
    const unsigned mblocks = blocks_for(n, kSummMidBodies);
    const bool mid = mblocks >= kSummMidMinBlocks;  // see launch_lbvh_summaries
    if (mid)
        hipLaunchKernelGGL(k_oct_summ_mid, dim3(mblocks), dim3(1024), 0, s, ob.nodes, ob.misc, ob.flags, ob.glist, gmid,
                           ob.cap, quad, err);
    hipLaunchKernelGGL(k_oct_summ_global, dim3(1), dim3(1024), 0, s, ob.nodes, ob.misc, ob.flags,
                       mid ? gmid : ob.glist, ob.cap, quad, err);
}

}  // namespace nbg
