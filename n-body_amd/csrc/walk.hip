// Barnes-Hut traversal for gfx950.
//
// One wavefront walks the tree for 64 spatially adjacent targets (consecutive
// in Morton order).  At every node, each lane in the wave's active mask
// evaluates its own opening criterion; lanes that accept the node take its
// monopole, lanes that reject a leaf interact with its bodies, and lanes that
// reject an internal node (a ballot) form the mask with which its children
// are visited.  Each lane therefore meets exactly the nodes and bodies the
// reference's per-body recursion meets (Descent/activeTree.h:24-79); only the
// schedule is shared.  Node records and leaf bodies are wave-uniform and are
// read with scalar loads.
//
// ORDERED (bit-exact): depth-first in child order with completion markers on
// an LDS stack, IEEE sqrt/division in Gravity::particleParticle's order
// (Physics/Rules/Gravity.h:199-209), leaf sums sequential from zero
// (activeTree.h:44-52) and child results bracketed as libstdc++'s
// std::transform_reduce does for two children, (0 + r0) + r1
// (activeTree.h:66-76).  Each lane's pending partial sums of the open depths
// (the depth is wave-uniform) sit in an LDS ring for the deepest levels and a
// private array for the rest (linear BVH), or in registers for the top frame
// and a private array below it (octree).  A subtree's result does not depend
// on where it is computed, so a small system's walk can be split: frontier
// subtrees walked from zero by several waves, their results used where the
// walk over the top nodes meets them, with unchanged bits.
//
// FAST: the same opening decisions (the MAC is still exact) and so the same
// interaction set, evaluated two siblings (linear BVH) or one child group
// (octree) per step.  All control flow is on wave-uniform ballots; the lanes
// that take an interaction are selected by turning the ballot back into an
// execution mask (inverse ballot), so no per-lane predicate is computed in
// vector registers.  Interactions use rsqrt and FMA, summed in FP32 per leaf
// and in FP64 across leaves (walk_common.hpp); the linear-BVH walks add those
// sums per fixed block of subtrees, which lets a small system spread each
// group of 64 targets over several waves with unchanged bits.

#include "device.hpp"
#include "launch.hpp"
#include "walk_common.hpp"

namespace nbg {

constexpr int kMetaComplete = 0x10000;
constexpr int kStackFast = 96;
// Upper bound on loop iterations per wave; a valid tree needs at most ~3 per node.
constexpr uint32_t kMaxWalkIters = 1u << 28;

// The top of a linear BVH (see the FAST walks below): internal nodes above
// depth kTopDepth holding more than `cut` bodies.  Their other children are
// the frontier, at most 2^kTopDepth nodes.
constexpr int kTopDepth = 7;
constexpr int kMaxFrontier = 1 << kTopDepth;
constexpr int kFrontierBlock = 8;
static_assert(1 + kMaxFrontier / kFrontierBlock == kSplitSlots, "split slots: s_top and one per block");
static_assert(kMaxFrontier == kOrderedSplitSlots, "ORDERED split: one slot per frontier entry");

__device__ __forceinline__ bool lbvh_top(const NodeRec &r, int depth, int cut) {
    return r.topo.z >= 0 && r.topo.y > cut && depth < kTopDepth;
}

// The targets, root and rule of wave w: the whole tree for one system, or
// the system that owns wave w in a forest (all wave-uniform).
struct WaveTask {
    int64_t base, end;
    int root;
    float G, eps2;
};

// Waves are 64 TPL targets wide (TPL targets per lane); a forest's wave
// tables count waves of that width.
template <bool FOREST, int TPL = 1>
__device__ __forceinline__ WaveTask wave_task(const WalkArgs &a, int64_t w) {
    WaveTask t;
    if (!FOREST) {
        t.base = w * (kWave * TPL); t.end = a.nt; t.root = 0; t.G = a.G; t.eps2 = a.eps2;
        return t;
    }
    if (w >= a.nwaves) {
        t.base = 0; t.end = 0; t.root = 0; t.G = a.G; t.eps2 = a.eps2;
        return t;
    }
    int lo = 0, hi = a.nsys;  // wave_first[lo] <= w < wave_first[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.wave_first[mid] <= w) lo = mid;
        else hi = mid;
    }
    t.base = a.sys_first[lo] + (w - a.wave_first[lo]) * (kWave * TPL);
    t.end = a.sys_first[lo + 1];
    t.root = a.sys_root[lo];
    const float4 q = a.sys_prm[lo];
    t.G = q.x;
    t.eps2 = q.y;
    return t;
}


// DiagonalOverCenterDistance (DiagonalOverCenterDistance.h:18-22) for one
// lane: the squared distance to the bbox centre in glm's order, compared
// with the node's threshold (mac_lbvh_threshold, computed with the node's
// summary), which is exactly equivalent to diag2 / d2 < theta^2 in FP32.
__device__ __forceinline__ bool lbvh_accepts(float4 mc, float px, float py, float pz) {
    const float dx = px - mc.x, dy = py - mc.y, dz = pz - mc.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    return d2 > mc.w;
}

// ---------------------------------------------------------------- ORDERED
constexpr int kOrderedWaves = 4;  // waves per workgroup of the ORDERED walks

// Levels of pending partial sums an ORDERED linear-BVH wave keeps in LDS.
constexpr int kSaveRing = 4;

// The ORDERED descent from node `start` for the lanes of `smask`, the
// reference's recursion (activeTree.h:24-79) with its bracketing: each node's
// result into r (from zero), counts into cpp / cnp.  COMBINE: a visited node
// that is not a top node (lbvh_top with the root's cut) is a frontier node:
// its result is not walked but taken from part (slot fi, counted in visit
// order) and added as the walk would add it.
template <int MAXD, bool QUAD, bool COMBINE>
__device__ __forceinline__ void ordered_descent(const WalkArgs &a, int start, uint64_t smask, int cut, float px,
                                                float py, float pz, float G, float eps2, int *s_node, int *s_meta,
                                                uint64_t *s_mask, float *ring, float &rx, float &ry, float &rz,
                                                uint32_t &cpp, uint32_t &cnp, bool &fail, int64_t tgt) {
    constexpr int kStack = 2 * MAXD + 4;
    float s_save[MAXD * 3];
    const float4 *__restrict__ bodies = a.bodies;
    const NodeRec *__restrict__ nodes = a.nodes;
    const int lane = lane_id();
    int fi = 0;
    if (lane == 0) { s_node[0] = start; s_meta[0] = 0; s_mask[0] = smask; }
    __builtin_amdgcn_wave_barrier();
    int sp = 1;
    int ring_base = 0;
    uint32_t iters = 0;
    while (sp > 0) {
        if (++iters > kMaxWalkIters) { fail = true; break; }
        --sp;
        __builtin_amdgcn_wave_barrier();
        const int node = uniform_i32(s_node[sp]);
        const int meta = uniform_i32(s_meta[sp]);
        const uint64_t mask = uniform_u64(s_mask[sp]);
        const bool act = (mask >> lane) & 1ull;
        const int depth = meta & 0xFFFF;
        if (meta & kMetaComplete) {
            // the deepest pending depth comes back: from the ring, or (the
            // ring then holds nothing) from scratch
            if (depth >= ring_base) {
                const float *slot = ring + (depth % kSaveRing) * 3 * kWave + lane;
                if (act) {
                    const float vx = rx, vy = ry, vz = rz;
                    rx = slot[0] + vx;
                    ry = slot[kWave] + vy;
                    rz = slot[2 * kWave] + vz;
                }
            } else {
                ring_base = depth;
                if (act) {
                    const float vx = rx, vy = ry, vz = rz;
                    rx = s_save[depth * 3 + 0] + vx;
                    ry = s_save[depth * 3 + 1] + vy;
                    rz = s_save[depth * 3 + 2] + vz;
                }
            }
            continue;
        }
        if (node < 0 || node >= a.max_nodes) { fail = true; break; }
        const NodeRec nd = nodes[node];
        if (COMBINE && !lbvh_top(nd, depth, cut)) {
            // a frontier node: its subtree's result, walked from zero elsewhere
            if (fi >= kMaxFrontier) { fail = true; break; }
            if (act) {
                const int64_t o = (int64_t)fi * 3 * a.split_stride + tgt;
                const float vx = a.split_part_f[o], vy = a.split_part_f[o + a.split_stride],
                            vz = a.split_part_f[o + 2 * a.split_stride];
                rx = rx + vx; ry = ry + vy; rz = rz + vz;
            }
            ++fi;
            continue;
        }
        const bool pass = act && lbvh_accepts(nd.mac, px, py, pz);
        if (pass) {
            ++cnp;
            float vx, vy, vz;
            if (QUAD)
                np_quad_exact(nd.com.x, nd.com.y, nd.com.z, nd.com.w, a.quad[2 * node], a.quad[2 * node + 1], px, py,
                              pz, G, eps2, vx, vy, vz);
            else
                pp_exact(nd.com.x, nd.com.y, nd.com.z, nd.com.w, px, py, pz, G, eps2, vx, vy, vz);
            rx = rx + vx; ry = ry + vy; rz = rz + vz;
        }
        const uint64_t open = mask & ~__ballot(pass);
        if (open == 0) continue;
        const bool op = (open >> lane) & 1ull;
        if (nd.topo.z < 0) {
            const int first = nd.topo.x, cnt = nd.topo.y;
            if (first < 0 || cnt < 0 || (int64_t)first + cnt > a.n) { fail = true; break; }
            // the body loop runs in wave-uniform flow so its loads are scalar
            // (one per body for the wave); the opened lanes sum
            float lx = 0.0f, ly = 0.0f, lz = 0.0f;
            for (int j = 0; j < cnt; ++j) {
                const float4 b = bodies[first + j];
                if (op) {
                    float vx, vy, vz;
                    pp_exact(b.x, b.y, b.z, b.w, px, py, pz, G, eps2, vx, vy, vz);
                    lx = lx + vx; ly = ly + vy; lz = lz + vz;
                }
            }
            if (op) {
                cpp += (uint32_t)cnt;
                rx = rx + lx; ry = ry + ly; rz = rz + lz;
            }
        } else {
            if (sp + 3 > kStack || depth >= MAXD) { fail = true; break; }
            // pending depths ring_base .. depth - 1 are in the ring; a full
            // ring moves its shallowest level (all lanes) to scratch
            if (depth - ring_base >= kSaveRing) {
                const float *old = ring + (ring_base % kSaveRing) * 3 * kWave + lane;
                s_save[ring_base * 3 + 0] = old[0];
                s_save[ring_base * 3 + 1] = old[kWave];
                s_save[ring_base * 3 + 2] = old[2 * kWave];
                ++ring_base;
            }
            if (op) {
                float *slot = ring + (depth % kSaveRing) * 3 * kWave + lane;
                slot[0] = rx;
                slot[kWave] = ry;
                slot[2 * kWave] = rz;
                rx = 0.0f; ry = 0.0f; rz = 0.0f;
            }
            if (lane == 0) {
                s_node[sp] = node; s_meta[sp] = depth | kMetaComplete; s_mask[sp] = open;
                s_node[sp + 1] = nd.topo.z + 1; s_meta[sp + 1] = depth + 1; s_mask[sp + 1] = open;
                s_node[sp + 2] = nd.topo.z; s_meta[sp + 2] = depth + 1; s_mask[sp + 2] = open;
            }
            sp += 3;
        }
    }
}

// ORDERED linear-BVH walk, one wave per 64 targets from the root.
template <int MAXD, bool FOREST, bool QUAD>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_lbvh_ordered(WalkArgs a) {
    constexpr int kStack = 2 * MAXD + 4;
    __shared__ int s_node_all[kOrderedWaves][kStack];
    __shared__ int s_meta_all[kOrderedWaves][kStack];
    __shared__ uint64_t s_mask_all[kOrderedWaves][kStack];
    // Per lane, the pending partial sum of every open depth.  The pending
    // depths are those of the current path, saved in increasing depth and
    // taken back in decreasing depth, so the deepest kSaveRing of them live
    // in an LDS ring (slot d mod kSaveRing, lane-major: no bank conflicts)
    // and only older ones go to a private array indexed by the wave-uniform
    // depth (scratch memory): depths >= ring_base are in the ring.
    __shared__ float s_ring_all[kOrderedWaves][kSaveRing * 3 * kWave];

    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const WaveTask T = wave_task<FOREST>(a, (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    uint32_t cpp = 0, cnp = 0;
    bool fail = false;
    ordered_descent<MAXD, QUAD, false>(a, T.root, vmask, 0, me.x, me.y, me.z, T.G, T.eps2, s_node_all[wv],
                                       s_meta_all[wv], s_mask_all[wv], s_ring_all[wv], rx, ry, rz, cpp, cnp, fail,
                                       tgt);
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) { a.ax[o] = rx; a.ay[o] = ry; a.az[o] = rz; }
    }
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// ORDERED split (one system): a.split waves per group of 64 targets.  Each
// lists the frontier (the non-top children of opened top nodes, in the
// ORDERED walk's visit order, masks from the exact criterion), and wave k
// walks entries i with i mod a.split == k from zero, storing each subtree's
// result per target in split_part_f for k_walk_lbvh_ordered_combine.
template <int MAXD, bool QUAD>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_lbvh_ordered_split(WalkArgs a) {
    constexpr int kStack = 2 * MAXD + 4;
    __shared__ int s_node_all[kOrderedWaves][kStack];
    __shared__ int s_meta_all[kOrderedWaves][kStack];
    __shared__ uint64_t s_mask_all[kOrderedWaves][kStack];
    __shared__ float s_ring_all[kOrderedWaves][kSaveRing * 3 * kWave];
    __shared__ int f_node_all[kOrderedWaves][kMaxFrontier];
    __shared__ uint64_t f_mask_all[kOrderedWaves][kMaxFrontier];

    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const int K = a.split;
    const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv;
    const int64_t grp = gw / K;
    const int k = (int)(gw - grp * K);
    const WaveTask T = wave_task<false>(a, grp);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const NodeRec *__restrict__ nodes = a.nodes;
    int *s_node = s_node_all[wv];
    int *s_meta = s_meta_all[wv];
    uint64_t *s_mask = s_mask_all[wv];
    int *f_node = f_node_all[wv];
    uint64_t *f_mask = f_mask_all[wv];
    bool fail = false;

    // the frontier, in the order the ORDERED walk visits it
    const NodeRec root = nodes[T.root];
    const int cut = (root.topo.y + 31) / 32;
    int F = 0, sp = 1;
    if (lane == 0) { s_node[0] = T.root; s_meta[0] = 0; s_mask[0] = vmask; }
    __builtin_amdgcn_wave_barrier();
    while (sp > 0) {
        --sp;
        __builtin_amdgcn_wave_barrier();
        const int node = uniform_i32(s_node[sp]);
        const int depth = uniform_i32(s_meta[sp]);
        const uint64_t mask = uniform_u64(s_mask[sp]);
        if (node < 0 || node >= a.max_nodes) { fail = true; break; }
        const NodeRec nd = nodes[node];
        if (!lbvh_top(nd, depth, cut)) {
            if (F >= kMaxFrontier) { fail = true; break; }  // unreachable: depth <= kTopDepth
            if (lane == 0) { f_node[F] = node; f_mask[F] = mask; }
            ++F;
            continue;
        }
        const bool act = (mask >> lane) & 1ull;
        const uint64_t open = mask & ~__ballot(act && lbvh_accepts(nd.mac, me.x, me.y, me.z));
        if (open == 0) continue;
        if (sp + 2 > kStack) { fail = true; break; }
        if (lane == 0) {
            s_node[sp] = nd.topo.z + 1; s_meta[sp] = depth + 1; s_mask[sp] = open;
            s_node[sp + 1] = nd.topo.z; s_meta[sp + 1] = depth + 1; s_mask[sp + 1] = open;
        }
        __builtin_amdgcn_wave_barrier();
        sp += 2;
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t cpp = 0, cnp = 0;
    const int64_t stride = a.split_stride;
    for (int i = k; i < F && !fail; i += K) {
        const int f = uniform_i32(f_node[i]);
        const uint64_t m = uniform_u64(f_mask[i]);
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
        ordered_descent<MAXD, QUAD, false>(a, f, m, cut, me.x, me.y, me.z, T.G, T.eps2, s_node, s_meta, s_mask,
                                           s_ring_all[wv], rx, ry, rz, cpp, cnp, fail, tgt);
        if (valid && ((m >> lane) & 1ull)) {
            const int64_t o = (int64_t)i * 3 * stride + tgt;
            a.split_part_f[o] = rx;
            a.split_part_f[o + stride] = ry;
            a.split_part_f[o + 2 * stride] = rz;
        }
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// The ORDERED split's combine: the walk from the root over the top nodes,
// with each frontier subtree's stored result added where the walk meets it.
template <int MAXD, bool QUAD>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_lbvh_ordered_combine(WalkArgs a) {
    constexpr int kStack = 2 * MAXD + 4;
    __shared__ int s_node_all[kOrderedWaves][kStack];
    __shared__ int s_meta_all[kOrderedWaves][kStack];
    __shared__ uint64_t s_mask_all[kOrderedWaves][kStack];
    __shared__ float s_ring_all[kOrderedWaves][kSaveRing * 3 * kWave];

    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const WaveTask T = wave_task<false>(a, (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int cut = (a.nodes[T.root].topo.y + 31) / 32;
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    uint32_t cpp = 0, cnp = 0;
    bool fail = false;
    ordered_descent<MAXD, QUAD, true>(a, T.root, vmask, cut, me.x, me.y, me.z, T.G, T.eps2, s_node_all[wv],
                                      s_meta_all[wv], s_mask_all[wv], s_ring_all[wv], rx, ry, rz, cpp, cnp, fail,
                                      tgt);
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) { a.ax[o] = rx; a.ay[o] = ry; a.az[o] = rz; }
    }
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// qn: the node's two moment records (QUAD only; uniform).
template <bool SOFT, int STATS, bool QUAD>
__device__ __forceinline__ uint64_t visit(const NodeRec &nd, const float4 *__restrict__ qn, uint64_t mask,
                                          const float4 *__restrict__ bodies, int64_t n, float eps2, FastWave &W,
                                          bool &fail, float4 *sh) {
    const uint64_t acc = __ballot(lbvh_accepts(nd.mac, W.px, W.py, W.pz)) & mask;
    if (STATS >= 2) { W.visits += 1; W.lanes += __popcll(mask); }
    if (STATS >= 1) W.np += __popcll(acc);
    if (QUAD) {
        if (acc) {
            const float4 q0 = qn[0], q1 = qn[1];
            if (__builtin_amdgcn_inverse_ballot_w64(acc))
                interact_quad<SOFT>(nd.com.x, nd.com.y, nd.com.z, nd.com.w, q0, q1, eps2, W);
        }
    } else if (__builtin_amdgcn_inverse_ballot_w64(acc)) {
        interact<SOFT>(nd.com.x, nd.com.y, nd.com.z, nd.com.w, eps2, W);
    }
    const uint64_t open = mask & ~acc;
    if (nd.topo.z >= 0) return open;
    if (open == 0) return 0;
    const int first = nd.topo.x, cnt = nd.topo.y;
    // first in [0, n) and cnt in [0, n - first] (n < 2^29: 32-bit unsigned, two scalar tests)
    if (__builtin_expect((uint32_t)first >= (uint32_t)n, 0)) { fail = true; return 0; }
    if (__builtin_expect((uint32_t)cnt > (uint32_t)n - (uint32_t)first, 0)) { fail = true; return 0; }
    if (STATS >= 2) { W.leaf_visits += 1; W.leaf_lanes += __popcll(open); }
    if (STATS >= 1) W.pp += (uint64_t)__popcll(open) * (uint64_t)cnt;
    leaf_eval<SOFT>(bodies, first, cnt, open, eps2, W, sh);
    return 0;
}

// Visits the sibling pair (r0, r1) for the lanes in `mask`: both opening
// tests first (independent loads and ballots the scheduler can overlap),
// then both node interactions, then the leaves; returns in o0 / o1 the
// lanes that open each internal child.  The same interactions as visit() on
// r0 then r1, summed in another order (FAST).
template <bool SOFT, int STATS, bool QUAD>
__device__ __forceinline__ void visit_pair(const NodeRec &r0, const NodeRec &r1, const float4 *__restrict__ qn,
                                           uint64_t mask, const float4 *__restrict__ bodies, int64_t n, float eps2,
                                           FastWave &W, bool &fail, float4 *sh, uint64_t &o0, uint64_t &o1) {
    const uint64_t a0 = __ballot(lbvh_accepts(r0.mac, W.px, W.py, W.pz)) & mask;
    const uint64_t a1 = __ballot(lbvh_accepts(r1.mac, W.px, W.py, W.pz)) & mask;
    if (STATS >= 2) { W.visits += 2; W.lanes += 2 * __popcll(mask); }
    if (STATS >= 1) W.np += __popcll(a0) + __popcll(a1);
    if (QUAD) {
        if (a0) {
            const float4 q0 = load_uniform(qn), q1 = load_uniform(qn + 1);
            if (__builtin_amdgcn_inverse_ballot_w64(a0))
                interact_quad<SOFT>(r0.com.x, r0.com.y, r0.com.z, r0.com.w, q0, q1, eps2, W);
        }
        if (a1) {
            const float4 q0 = load_uniform(qn + 2), q1 = load_uniform(qn + 3);
            if (__builtin_amdgcn_inverse_ballot_w64(a1))
                interact_quad<SOFT>(r1.com.x, r1.com.y, r1.com.z, r1.com.w, q0, q1, eps2, W);
        }
    } else {
        if (__builtin_amdgcn_inverse_ballot_w64(a0)) interact<SOFT>(r0.com.x, r0.com.y, r0.com.z, r0.com.w, eps2, W);
        if (__builtin_amdgcn_inverse_ballot_w64(a1)) interact<SOFT>(r1.com.x, r1.com.y, r1.com.z, r1.com.w, eps2, W);
    }
    o0 = mask & ~a0;
    o1 = mask & ~a1;
    const NodeRec *rs[2] = {&r0, &r1};
    uint64_t *os[2] = {&o0, &o1};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const NodeRec &nd = *rs[k];
        uint64_t &open = *os[k];
        if (nd.topo.z >= 0 || open == 0) continue;
        const int first = nd.topo.x, cnt = nd.topo.y;
        const uint64_t take = open;
        open = 0;
        // first in [0, n) and cnt in [0, n - first] (n < 2^29: 32-bit unsigned, two scalar tests)
        if (__builtin_expect((uint32_t)first >= (uint32_t)n, 0)) { fail = true; continue; }
        if (__builtin_expect((uint32_t)cnt > (uint32_t)n - (uint32_t)first, 0)) { fail = true; continue; }
        if (STATS >= 2) { W.leaf_visits += 1; W.leaf_lanes += __popcll(take); }
        if (STATS >= 1) W.pp += (uint64_t)__popcll(take) * (uint64_t)cnt;
        leaf_eval<SOFT>(bodies, first, cnt, take, eps2, W, sh);
    }
}

// ---------------------------------------------------------------- FAST LBVH
// Every FAST linear-BVH walk (one or several targets per lane, one system or
// a forest, one wave or several per group of 64 targets) sums a target's
// interactions with the same structure, so its bits do not depend on which
// of them walks it:
//   top nodes: internal nodes above depth kTopDepth holding more than
//     ceil(root count / 32) bodies; the walk over them ("top walk") sums the
//     nodes it accepts into s_top and lists, in the order it meets them, the
//     children of top nodes that are not top nodes ("frontier": at most
//     2^kTopDepth entries);
//   blocks: each run of kFrontierBlock consecutive frontier entries that a
//     target reaches is walked depth-first as before, its interactions summed
//     from zero (FP32 per leaf, FP64 across leaves) into B_b;
//   result: ((s_top + B_0) + B_1) + ... in FP64, times G, rounded once.
// A tree of more than kSplitMaxBodies bodies (one that never splits) has no
// top nodes: its whole walk is one entry, the plain walk.  Only node records
// decide the structure, so a wave walking a group alone and K waves
// splitting its blocks (the sums meeting in k_walk_split_reduce) give the
// same bits.  The opening decisions, and so the interaction set, are those
// of the plain depth-first walk.

// A lane accepts node nd (slot sums W): its interaction (monopole or
// quadrupole, qn the node's two moment records).
template <bool SOFT, bool QUAD>
__device__ __forceinline__ void take_node(const NodeRec &nd, const float4 *__restrict__ qn, uint64_t acc, float eps2,
                                          FastWave &W) {
    if (QUAD) {
        if (acc) {
            const float4 q0 = load_uniform(qn), q1 = load_uniform(qn + 1);
            if (__builtin_amdgcn_inverse_ballot_w64(acc))
                interact_quad<SOFT>(nd.com.x, nd.com.y, nd.com.z, nd.com.w, q0, q1, eps2, W);
        }
    } else if (__builtin_amdgcn_inverse_ballot_w64(acc)) {
        interact<SOFT>(nd.com.x, nd.com.y, nd.com.z, nd.com.w, eps2, W);
    }
}

// The top walk from node `root` for TPL slots of lanes (vmask[t]): the
// accepted top nodes' interactions go into W[t] when INTERACT; the frontier
// (node, per-slot masks) goes to f_node / f_mask[i * TPL + t].  Returns the
// number of frontier entries.  st_* is a stack of kTopDepth entries.
template <bool SOFT, int STATS, bool QUAD, int TPL, bool INTERACT>
__device__ int top_walk(const NodeRec *__restrict__ nodes, const float4 *__restrict__ quad, int root,
                        const uint64_t (&vmask)[TPL], float eps2, FastWave (&W)[TPL], int *f_node, uint64_t *f_mask,
                        int *st_node, uint64_t *st_mask, int *st_dep, uint32_t last_pair, bool &fail) {
    const NodeRec r = load_uniform(nodes + root);
    const int cut = (r.topo.y + 31) / 32;
    // a tree too large to split is one frontier entry: the plain walk
    if (!lbvh_top(r, 0, cut) || r.topo.y > kSplitMaxBodies) {
        if (lane_id() == 0) {
            f_node[0] = root;
#pragma unroll
            for (int t = 0; t < TPL; ++t) f_mask[t] = vmask[t];
        }
        __builtin_amdgcn_wave_barrier();
        return 1;
    }
    uint64_t m[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const uint64_t acc = __ballot(lbvh_accepts(r.mac, W[t].px, W[t].py, W[t].pz)) & vmask[t];
        if (INTERACT) {
            if (STATS >= 2) { W[t].visits += 1; W[t].lanes += __popcll(vmask[t]); }
            if (STATS >= 1) W[t].np += __popcll(acc);
            take_node<SOFT, QUAD>(r, quad + 2 * root, acc, eps2, W[t]);
        }
        m[t] = vmask[t] & ~acc;
    }
    int F = 0, sp = 0, c = r.topo.z, dep = 1;
    while (true) {
        uint64_t any = 0;
#pragma unroll
        for (int t = 0; t < TPL; ++t) any |= m[t];
        if (any == 0) {
            if (sp == 0) break;
            --sp;
            __builtin_amdgcn_wave_barrier();
            c = uniform_i32(st_node[sp]);
            dep = uniform_i32(st_dep[sp]);
#pragma unroll
            for (int t = 0; t < TPL; ++t) m[t] = uniform_u64(st_mask[sp * TPL + t]);
        }
        c = uniform_i32(c);
        if (__builtin_expect((uint32_t)c >= last_pair, 0)) { fail = true; break; }
        const NodeRec r0 = load_uniform(nodes + c), r1 = load_uniform(nodes + c + 1);
        uint64_t o0[TPL], o1[TPL], any0 = 0, any1 = 0;
        const NodeRec *rs[2] = {&r0, &r1};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const NodeRec &nd = *rs[k];
            uint64_t *o = k ? o1 : o0;
            if (lbvh_top(nd, dep, cut)) {
#pragma unroll
                for (int t = 0; t < TPL; ++t) {
                    const uint64_t acc = __ballot(lbvh_accepts(nd.mac, W[t].px, W[t].py, W[t].pz)) & m[t];
                    if (INTERACT) {
                        if (STATS >= 2 && m[t]) { W[t].visits += 1; W[t].lanes += __popcll(m[t]); }
                        if (STATS >= 1) W[t].np += __popcll(acc);
                        take_node<SOFT, QUAD>(nd, quad + 2 * (c + k), acc, eps2, W[t]);
                    }
                    o[t] = m[t] & ~acc;
                }
            } else {
                if (F >= kMaxFrontier) { fail = true; break; }  // unreachable: depth < kTopDepth
                if (lane_id() == 0) {
                    f_node[F] = c + k;
#pragma unroll
                    for (int t = 0; t < TPL; ++t) f_mask[F * TPL + t] = m[t];
                }
                ++F;
#pragma unroll
                for (int t = 0; t < TPL; ++t) o[t] = 0;
            }
        }
        if (fail) break;
#pragma unroll
        for (int t = 0; t < TPL; ++t) { any0 |= o0[t]; any1 |= o1[t]; }
        if (any0 != 0) {
            if (any1 != 0) {
                if (sp >= kTopDepth) { fail = true; break; }
                if (lane_id() == 0) {
                    st_node[sp] = r1.topo.z;
                    st_dep[sp] = dep + 1;
#pragma unroll
                    for (int t = 0; t < TPL; ++t) st_mask[sp * TPL + t] = o1[t];
                }
                __builtin_amdgcn_wave_barrier();
                ++sp;
            }
            c = r0.topo.z;
#pragma unroll
            for (int t = 0; t < TPL; ++t) m[t] = o0[t];
        } else if (any1 != 0) {
            c = r1.topo.z;
#pragma unroll
            for (int t = 0; t < TPL; ++t) m[t] = o1[t];
        } else {
#pragma unroll
            for (int t = 0; t < TPL; ++t) m[t] = 0;
        }
        ++dep;
    }
    __builtin_amdgcn_wave_barrier();
    return F;
}

// Depth-first walk of the subtree of frontier node f (visited on its own,
// then sibling pairs) for the slots' lanes m0[t]; node loads, traversal and
// stack are shared by the TPL slots, each keeping its own masks and sums.
template <bool SOFT, int STATS, bool QUAD, int TPL>
__device__ __forceinline__ void walk_subtree(const NodeRec *__restrict__ nodes, const float4 *__restrict__ quad,
                                             int f, const uint64_t (&m0)[TPL], const float4 *__restrict__ bodies,
                                             int64_t n, float eps2, FastWave (&W)[TPL], bool &fail, int *s_node,
                                             uint64_t *s_mask, float4 *sh, uint32_t last_pair, uint32_t &iters) {
    const NodeRec r = load_uniform(nodes + f);
    uint64_t mask[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t)
        mask[t] = m0[t] ? visit<SOFT, STATS, QUAD>(r, quad + 2 * f, m0[t], bodies, n, eps2, W[t], fail, sh) : 0ull;
    int c = r.topo.z;
    int sp = 0;
    while (true) {
        uint64_t any = 0;
#pragma unroll
        for (int t = 0; t < TPL; ++t) any |= mask[t];
        if (any == 0) {
            if (sp == 0) break;
            --sp;
            __builtin_amdgcn_wave_barrier();
            c = uniform_i32(s_node[sp]);
#pragma unroll
            for (int t = 0; t < TPL; ++t) mask[t] = uniform_u64(s_mask[sp * TPL + t]);
        }
        c = uniform_i32(c);  // keeps the node loads scalar
        if (__builtin_expect((uint32_t)c >= last_pair, 0)) { fail = true; break; }
        if (__builtin_expect(++iters > kMaxWalkIters, 0)) { fail = true; break; }
        const NodeRec r0 = load_uniform(nodes + c), r1 = load_uniform(nodes + c + 1);
        uint64_t o0[TPL], o1[TPL], any0 = 0, any1 = 0;
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            o0[t] = 0;
            o1[t] = 0;
            if (TPL == 1 || mask[t])
                visit_pair<SOFT, STATS, QUAD>(r0, r1, quad + 2 * c, mask[t], bodies, n, eps2, W[t], fail, sh, o0[t],
                                              o1[t]);
            any0 |= o0[t];
            any1 |= o1[t];
        }
        // a failed visit opens nothing, so the walk still ends; fail is
        // reported by the kernel
        if (any0 != 0) {
            if (any1 != 0) {
                if (sp >= kStackFast) { fail = true; break; }
                if (lane_id() == 0) {
                    s_node[sp] = r1.topo.z;
#pragma unroll
                    for (int t = 0; t < TPL; ++t) s_mask[sp * TPL + t] = o1[t];
                }
                __builtin_amdgcn_wave_barrier();
                ++sp;
            }
            c = r0.topo.z;
#pragma unroll
            for (int t = 0; t < TPL; ++t) mask[t] = o0[t];
        } else if (any1 != 0) {
            c = r1.topo.z;
#pragma unroll
            for (int t = 0; t < TPL; ++t) mask[t] = o1[t];
        } else {
#pragma unroll
            for (int t = 0; t < TPL; ++t) mask[t] = 0;
        }
    }
}

// FAST linear-BVH walk.  TPL targets per lane: the wave serves 64 TPL
// targets from its task's base, slot t of lane l being target base + 64 t
// + l.  FOREST: the wave's system from the forest's wave table (wave_task).
// SPLIT (TPL 1, one system or a forest): a.split waves serve each group of
// 64 targets; wave k of a group walks the blocks b with b mod a.split == k
// (wave 0 also the top nodes) and stores their sums for k_walk_split_reduce.
template <int WAVES, bool SOFT, int STATS, bool QUAD, int TPL, bool FOREST, bool SPLIT>
__global__ __launch_bounds__(64 * WAVES) void k_walk_lbvh_fast(WalkArgs a) {
    static_assert(!SPLIT || TPL == 1, "split groups are one-slot walks");
    __shared__ int s_node[WAVES][kStackFast];
    __shared__ uint64_t s_mask[WAVES][kStackFast * TPL];
    __shared__ float4 s_leaf[WAVES][kLeafScratch];
    __shared__ int f_node[WAVES][kMaxFrontier];
    __shared__ uint64_t f_mask[WAVES][kMaxFrontier * TPL];
    __shared__ int st_dep[WAVES][kTopDepth];

    const float4 *__restrict__ bodies = a.bodies;
    const NodeRec *__restrict__ nodes = a.nodes;
    // the wave index is uniform: readfirstlane makes the per-wave task
    // scalar, while LDS addresses use the per-lane copy (one VALU op each)
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int wvu = uniform_i32(wv);
    const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * WAVES + wvu;
    const int K = SPLIT ? a.split : 1;
    const int64_t grp = SPLIT ? gw / K : gw;
    const int k = SPLIT ? (int)(gw - grp * K) : 0;
    const WaveTask T = wave_task<FOREST, TPL>(a, grp);
    const int64_t base = T.base;
    FastWave W[TPL];
    uint64_t vmask[TPL];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const int64_t tgt = base + t * kWave + lane;
        const bool valid = tgt < T.end;
        vmask[t] = __ballot(valid);
        const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        fast_init(W[t], me.x, me.y, me.z);
        W[t].G = T.G;
    }
    if (vmask[0] == 0) return;  // slot 0 holds the wave's first targets
    if (FOREST && (T.root < 0 || T.root >= a.max_nodes)) {
        if (lane == 0) atomicOr(a.err, 2u);
        return;
    }
    const float eps2 = T.eps2;
    const int64_t n = a.n;
    const uint32_t last_pair = (uint32_t)(a.max_nodes - 1);
    bool fail = false;

    int F;
    if (!SPLIT || k == 0)
        F = top_walk<SOFT, STATS, QUAD, TPL, true>(nodes, a.quad, T.root, vmask, eps2, W, f_node[wvu], f_mask[wvu],
                                                   s_node[wvu], s_mask[wvu], st_dep[wvu], last_pair, fail);
    else
        F = top_walk<SOFT, STATS, QUAD, TPL, false>(nodes, a.quad, T.root, vmask, eps2, W, f_node[wvu], f_mask[wvu],
                                                    s_node[wvu], s_mask[wvu], st_dep[wvu], last_pair, fail);
    F = uniform_i32(F);
    // s_top, then each block's sum from zero
    double S[TPL][3];
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        fast_flush(W[t]);
        S[t][0] = W[t].sx; S[t][1] = W[t].sy; S[t][2] = W[t].sz;
        W[t].sx = 0.0; W[t].sy = 0.0; W[t].sz = 0.0;
    }
    const int64_t tgt0 = base + lane;
    const int64_t stride = a.split_stride;
    if (SPLIT && k == 0 && tgt0 < T.end) {
#pragma unroll
        for (int d = 0; d < 3; ++d) a.split_part[d * stride + tgt0] = S[0][d];
        if (lane == 0) a.split_nb[grp] = (F + kFrontierBlock - 1) / kFrontierBlock;
    }
    uint32_t iters = 0;
    if constexpr (TPL == 1) {
        // One loop, shaped as the plain depth-first walk's: when a frontier
        // subtree is done (empty mask, empty stack) the exit path closes the
        // block if it ends there and starts the next owned entry.
        FastWave &V = W[0];
        int i = -1, c = 0, sp = 0;
        uint64_t mask = 0;
        while (true) {
            if (mask == 0) {
              if (sp == 0) {
                // next frontier entry whose visit opens an internal node
                bool done = false;
                do {
                    if (i >= 0 && ((i + 1) % kFrontierBlock == 0 || i + 1 == F)) {
                        fast_flush(V);
                        if (SPLIT) {
                            if (tgt0 < T.end) {
                                const int64_t o = (int64_t)(1 + i / kFrontierBlock) * 3 * stride + tgt0;
                                a.split_part[o] = V.sx;
                                a.split_part[o + stride] = V.sy;
                                a.split_part[o + 2 * stride] = V.sz;
                            }
                        } else {
                            S[0][0] += V.sx; S[0][1] += V.sy; S[0][2] += V.sz;
                        }
                        V.sx = 0.0; V.sy = 0.0; V.sz = 0.0;
                    }
                    ++i;
                    if (SPLIT && i % kFrontierBlock == 0) {
                        // the first entry of the next block this wave owns
                        const int blk = i / kFrontierBlock;
                        i = (blk + (k - blk % K + K) % K) * kFrontierBlock;
                    }
                    if (i >= F) {
                        done = true;
                        break;
                    }
                    const int f = uniform_i32(f_node[wv][i]);
                    const NodeRec r = load_uniform(nodes + f);
                    mask = visit<SOFT, STATS, QUAD>(r, a.quad + 2 * f, uniform_u64(f_mask[wv][i]), bodies, n, eps2, V,
                                                    fail, s_leaf[wv]);
                    c = r.topo.z;
                } while (mask == 0);
                if (done) break;
              } else {
                --sp;
                __builtin_amdgcn_wave_barrier();
                c = uniform_i32(s_node[wv][sp]);
                mask = uniform_u64(s_mask[wv][sp]);
              }
            }
            c = uniform_i32(c);  // keeps the node loads scalar
            if (__builtin_expect((uint32_t)c >= last_pair, 0)) { fail = true; break; }
            if (__builtin_expect(++iters > kMaxWalkIters, 0)) { fail = true; break; }
            const NodeRec r0 = load_uniform(nodes + c), r1 = load_uniform(nodes + c + 1);
            uint64_t o0, o1;
            visit_pair<SOFT, STATS, QUAD>(r0, r1, a.quad + 2 * c, mask, bodies, n, eps2, V, fail, s_leaf[wv], o0, o1);
            // a failed visit opens nothing, so the walk still ends; fail is
            // reported after the loop
            if (o0 != 0) {
                if (o1 != 0) {
                    if (sp >= kStackFast) { fail = true; break; }
                    if (lane == 0) { s_node[wv][sp] = r1.topo.z; s_mask[wv][sp] = o1; }
                    __builtin_amdgcn_wave_barrier();
                    ++sp;
                }
                c = r0.topo.z;
                mask = o0;
            } else if (o1 != 0) {
                c = r1.topo.z;
                mask = o1;
            } else {
                mask = 0;
            }
        }
    } else {
        int cnt[TPL];
#pragma unroll
        for (int t = 0; t < TPL; ++t) cnt[t] = 0;
        for (int i = 0; i < F; ++i) {
            uint64_t m[TPL];
#pragma unroll
            for (int t = 0; t < TPL; ++t) m[t] = uniform_u64(f_mask[wvu][i * TPL + t]);
            walk_subtree<SOFT, STATS, QUAD, TPL>(nodes, a.quad, uniform_i32(f_node[wvu][i]), m, bodies, n, eps2, W,
                                                 fail, s_node[wvu], s_mask[wvu], s_leaf[wvu], last_pair, iters);
            // each slot's blocks are runs of kFrontierBlock of its own entries
#pragma unroll
            for (int t = 0; t < TPL; ++t) {
                if (m[t] == 0 || ++cnt[t] % kFrontierBlock != 0) continue;
                fast_flush(W[t]);
                S[t][0] += W[t].sx; S[t][1] += W[t].sy; S[t][2] += W[t].sz;
                W[t].sx = 0.0; W[t].sy = 0.0; W[t].sz = 0.0;
            }
        }
        // the last, partial block of each slot
#pragma unroll
        for (int t = 0; t < TPL; ++t) {
            if (cnt[t] % kFrontierBlock == 0) continue;
            fast_flush(W[t]);
            S[t][0] += W[t].sx; S[t][1] += W[t].sy; S[t][2] += W[t].sz;
        }
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    uint64_t pp = 0, np = 0, visits = 0, lanes = 0, lvis = 0, llanes = 0;
#pragma unroll
    for (int t = 0; t < TPL; ++t) {
        const int64_t tgt = base + t * kWave + lane;
        if (!SPLIT && tgt < T.end) {
            const uint32_t o = a.perm[tgt];
            if ((int64_t)o < a.nt) {
                const double g = (double)W[t].G;
                a.ax[o] = (float)(S[t][0] * g);
                a.ay[o] = (float)(S[t][1] * g);
                a.az[o] = (float)(S[t][2] * g);
            }
        }
        pp += W[t].pp; np += W[t].np;
        visits += W[t].visits; lanes += W[t].lanes; lvis += W[t].leaf_visits; llanes += W[t].leaf_lanes;
    }
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)pp);
        atomicAdd(&a.counts[1], (unsigned long long)np);
        if (STATS >= 2) {
            atomicAdd(&a.counts[2], (unsigned long long)visits);
            atomicAdd(&a.counts[3], (unsigned long long)lanes);
            atomicAdd(&a.counts[4], (unsigned long long)lvis);
            atomicAdd(&a.counts[5], (unsigned long long)llanes);
        }
    }
}

// A single system above kSplitMaxBodies bodies, one target per lane: the
// plain walk (wave w walks targets 64 w .. 64 w + 63 from the root, one FP32
// partial flushed into the FP64 sum after every leaf).  Such a tree has no
// top nodes, so the block form's result is the same bits: one entry, the
// root, walked the same way, its FP64 sum added to s_top = +0.0 (a sum that
// starts at +0.0 never becomes -0.0, so the addition is exact).
template <int WAVES, bool SOFT, int STATS, bool QUAD>
__global__ __launch_bounds__(64 * WAVES) void k_walk_lbvh_plain(WalkArgs a) {
    __shared__ int s_node[WAVES][kStackFast];
    __shared__ uint64_t s_mask[WAVES][kStackFast];
    __shared__ float4 s_leaf[WAVES][kLeafScratch];

    const float4 *__restrict__ bodies = a.bodies;
    const NodeRec *__restrict__ nodes = a.nodes;
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t tgt = (xcd_block(blockIdx.x, gridDim.x) * WAVES + wv) * kWave + lane;
    const bool valid = tgt < a.nt;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;

    FastWave W;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    fast_init(W, me.x, me.y, me.z);
    // A VGPR copy of G (non-volatile asm: no memory effects, so the node and
    // body loads below stay scalar).
    asm("v_mov_b32 %0, %1" : "=v"(W.G) : "s"(a.G));
    const float eps2 = a.eps2;
    const int64_t n = a.n;
    const uint32_t last_pair = (uint32_t)(a.max_nodes - 1);
    bool fail = false;

    // The root is visited on its own; afterwards every step visits the two
    // children (c, c + 1) of a node that the lanes in `mask` opened.
    const NodeRec root = nodes[0];
    uint64_t mask = visit<SOFT, STATS, QUAD>(root, a.quad, vmask, bodies, n, eps2, W, fail, s_leaf[wv]);
    int c = root.topo.z;
    int sp = 0;
    uint32_t iters = 0;
    while (true) {
        if (mask == 0) {
            if (sp == 0) break;
            --sp;
            __builtin_amdgcn_wave_barrier();
            c = uniform_i32(s_node[wv][sp]);
            mask = uniform_u64(s_mask[wv][sp]);
        }
        c = uniform_i32(c);  // keeps the node loads scalar
        if (__builtin_expect((uint32_t)c >= last_pair, 0)) { fail = true; break; }
        if (__builtin_expect(++iters > kMaxWalkIters, 0)) { fail = true; break; }
        const NodeRec r0 = load_uniform(nodes + c), r1 = load_uniform(nodes + c + 1);
        uint64_t o0, o1;
        visit_pair<SOFT, STATS, QUAD>(r0, r1, a.quad + 2 * c, mask, bodies, n, eps2, W, fail, s_leaf[wv], o0, o1);
        // a failed visit opens nothing, so the walk still ends; fail is
        // reported after the loop
        if (o0 != 0) {
            if (o1 != 0) {
                if (sp >= kStackFast) { fail = true; break; }
                if (lane == 0) { s_node[wv][sp] = r1.topo.z; s_mask[wv][sp] = o1; }
                __builtin_amdgcn_wave_barrier();
                ++sp;
            }
            c = r0.topo.z;
            mask = o0;
        } else if (o1 != 0) {
            c = r1.topo.z;
            mask = o1;
        } else {
            mask = 0;
        }
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) {
            const float3 r = fast_result(W, W.G);
            a.ax[o] = r.x; a.ay[o] = r.y; a.az[o] = r.z;
        }
    }
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)W.pp);
        atomicAdd(&a.counts[1], (unsigned long long)W.np);
        if (STATS >= 2) {
            atomicAdd(&a.counts[2], (unsigned long long)W.visits);
            atomicAdd(&a.counts[3], (unsigned long long)W.lanes);
            atomicAdd(&a.counts[4], (unsigned long long)W.leaf_visits);
            atomicAdd(&a.counts[5], (unsigned long long)W.leaf_lanes);
        }
    }
}

// The split walk's sums, ((s_top + B_0) + B_1) + ... for each target in
// FP64, times G, rounded once and stored in output order.  FOREST: the
// target's group and G from its system (the forest's waves start at each
// system's first body).
template <bool FOREST>
__global__ __launch_bounds__(256) void k_walk_split_reduce(WalkArgs a) {
    const int64_t tgt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tgt >= a.nt) return;
    const int64_t stride = a.split_stride;
    int64_t grp = tgt / kWave;
    double g = (double)a.G;
    if (FOREST) {
        int lo = 0, hi = a.nsys;  // sys_first[lo] <= tgt < sys_first[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (a.sys_first[mid] <= tgt) lo = mid;
            else hi = mid;
        }
        grp = a.wave_first[lo] + (tgt - a.sys_first[lo]) / kWave;
        g = (double)a.sys_prm[lo].x;
    }
    int nb = a.split_nb[grp];
    if (nb < 0 || nb > kMaxFrontier / kFrontierBlock) {
        atomicOr(a.err, 2u);
        nb = 0;
    }
    double S[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) S[d] = a.split_part[d * stride + tgt];
    for (int b = 0; b < nb; ++b) {
#pragma unroll
        for (int d = 0; d < 3; ++d) S[d] += a.split_part[((int64_t)(1 + b) * 3 + d) * stride + tgt];
    }
    const uint32_t o = a.perm[tgt];
    if ((int64_t)o < a.nt) {
        a.ax[o] = (float)(S[0] * g);
        a.ay[o] = (float)(S[1] * g);
        a.az[o] = (float)(S[2] * g);
    }
}

// ================================================================ octree
// ORDERED: libstdc++'s std::transform_reduce over the 8 children of an
// octree node (activeTree.h:66-76) computes init + ((r0 + r1) + (r2 + r3)),
// then + ((r4 + r5) + (r6 + r7)), with init = 0 and r = 0 for an empty
// child.  Per lane and per tree level the walk keeps that bracket's state
// in a private array (R: the running init, A: the open pair sum, B: the
// pending third term) and applies octant q's result with oct_step.
template <int MAXL>
__device__ __forceinline__ void oct_step(float *st, int q, bool act, float vx, float vy, float vz) {
    if (!act) return;
    float *R = st, *A = st + 3, *B = st + 6;
    const float v[3] = {vx, vy, vz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int o = c;
        switch (q & 3) {
        case 0: A[o] = v[c]; break;
        case 1: A[o] = A[o] + v[c]; break;
        case 2: B[o] = v[c]; break;
        default: {
            const float a = A[o] + (B[o] + v[c]);
            A[o] = a;
            R[o] = q == 3 ? 0.0f + a : R[o] + a;
        }
        }
    }
}

// Evaluates node `nd` for the lanes in O: accepting lanes take its monopole
// and leaf lanes their leaf sum as this node's result r; returns the lanes
// that open it when it is internal.
// qn: the node's two moment records (QUAD only).
template <bool QUAD>
__device__ __forceinline__ uint64_t oct_eval_ordered(const OctRec &nd, const float4 *__restrict__ qn, uint64_t O,
                                                     int lane, float px, float py, float pz, float G, float eps2,
                                                     const float4 *__restrict__ bodies, int64_t n, float &rx,
                                                     float &ry, float &rz, uint32_t &cpp, uint32_t &cnp, bool &fail) {
    const bool act = (O >> lane) & 1ull;
    const bool pass = act && oct_accepts(nd, px, py, pz);
    if (pass) {
        ++cnp;
        if (QUAD)
            np_quad_exact(nd.com.x, nd.com.y, nd.com.z, nd.com.w, qn[0], qn[1], px, py, pz, G, eps2, rx, ry, rz);
        else
            pp_exact(nd.com.x, nd.com.y, nd.com.z, nd.com.w, px, py, pz, G, eps2, rx, ry, rz);
    }
    const uint64_t open = O & ~__ballot(pass);
    if (open == 0) return 0;
    if (nd.topo.z >= 0) return open;
    const int first = nd.topo.x, cnt = nd.topo.y;
    if (first < 0 || cnt < 0 || (int64_t)first + cnt > n) { fail = true; return 0; }
    // the body loop runs in wave-uniform flow so its loads are scalar
    const bool op = (open >> lane) & 1ull;
    float lx = 0.0f, ly = 0.0f, lz = 0.0f;
    for (int j = 0; j < cnt; ++j) {
        const float4 b = bodies[first + j];
        if (op) {
            float vx, vy, vz;
            pp_exact(b.x, b.y, b.z, b.w, px, py, pz, G, eps2, vx, vy, vz);
            lx = lx + vx; ly = ly + vy; lz = lz + vz;
        }
    }
    if (op) {
        cpp += (uint32_t)cnt;
        rx = lx; ry = ly; rz = lz;
    }
    return 0;
}

// The top of an octree for the ORDERED split: internal nodes at depth 0 or 1
// holding more than `cut` bodies.  Their other children are the frontier
// (at most 8 + 64 nodes).
__device__ __forceinline__ bool oct_top_ordered(const OctRec &r, int depth, int cut) {
    return r.topo.z >= 0 && r.topo.y > cut && depth < 2;
}

// The ORDERED octree descent from node `start` for the lanes of `smask`: the
// node's result into r (assigned, as the reference's recursion returns it),
// counts into cpp / cnp.  COMBINE: a child that is not a top node (with the
// root's cut) is a frontier node: its result is taken from split_part_f
// (slot fi, counted in visit order) instead of being evaluated.
template <int MAXL, bool QUAD, bool COMBINE>
__device__ __forceinline__ void oct_ordered_descent(const WalkArgs &a, int start, uint64_t smask, int cut, float px,
                                                    float py, float pz, float G, float eps2, int *s_node, int *s_pos,
                                                    uint64_t *s_mask, float &rx, float &ry, float &rz, uint32_t &cpp,
                                                    uint32_t &cnp, bool &fail, int64_t tgt) {
    // Per lane and level, the bracket state (R, A, B; 3 floats each).  The
    // top frame's state lives in registers (cur); the frames below it are
    // parked in a private array indexed by the wave-uniform level (scratch
    // memory), written when a child is opened and read back when it
    // completes, so the child visits in between touch registers only.
    float s_st[MAXL * 9];
    float cur[9];
    const float4 *__restrict__ bodies = a.bodies;
    const OctRec *__restrict__ nodes = a.onodes;
    const int lane = lane_id();
    int fi = 0;
    auto frontier = [&](uint64_t O) {
        if (fi >= kOrderedSplitSlots) { fail = true; return; }
        if ((O >> lane) & 1ull) {
            const int64_t o = (int64_t)fi * 3 * a.split_stride + tgt;
            rx = a.split_part_f[o];
            ry = a.split_part_f[o + a.split_stride];
            rz = a.split_part_f[o + 2 * a.split_stride];
        }
        ++fi;
    };
    int sp = 0;
    {
        const OctRec root = nodes[start];
        if (COMBINE && !oct_top_ordered(root, 0, cut)) {
            frontier(smask);
            return;
        }
        const uint64_t O = oct_eval_ordered<QUAD>(root, a.quad + 2 * start, smask, lane, px, py, pz, G, eps2,
                                                  bodies, a.n, rx, ry, rz, cpp, cnp, fail);
        if (O) {
            if (lane == 0) { s_node[0] = start; s_pos[0] = 0; s_mask[0] = O; }
            sp = 1;
        }
    }
    // each frame is visited nchild + 1 <= 9 times, each node opened once
    const int64_t max_iters = 9 * a.max_nodes + 64;
    int64_t iters = 0;
    while (sp > 0 && !fail) {
        if (++iters > max_iters) { fail = true; break; }
        __builtin_amdgcn_wave_barrier();
        const int top = sp - 1;
        const int X = uniform_i32(s_node[top]);
        const int pos = uniform_i32(s_pos[top]);
        const uint64_t O = uniform_u64(s_mask[top]);
        const bool act = (O >> lane) & 1ull;
        if (X < 0 || X >= a.max_nodes) { fail = true; break; }
        const OctRec xr = nodes[X];
        const int cb = xr.topo.z, nch = oct_nchild(xr);
        if (cb < 0 || nch < 1 || nch > 8 || (int64_t)cb + nch > a.max_nodes) { fail = true; break; }
        int j = pos & 0xFF, q = pos >> 8;
        if (j > 0) {
            const int k = oct_octant(nodes[cb + j - 1]);
            for (; q < k; ++q) oct_step<MAXL>(cur, q, act, 0.0f, 0.0f, 0.0f);
            oct_step<MAXL>(cur, k, act, rx, ry, rz);
            q = k + 1;
        }
        if (j == nch) {
            for (; q < 8; ++q) oct_step<MAXL>(cur, q, act, 0.0f, 0.0f, 0.0f);
            if (act) { rx = cur[0]; ry = cur[1]; rz = cur[2]; }
            sp = top;
            if (top > 0) {  // the parent frame is the top again
#pragma unroll
                for (int i = 0; i < 9; ++i) cur[i] = s_st[(top - 1) * 9 + i];
            }
            continue;
        }
        if (lane == 0) s_pos[top] = (j + 1) | (q << 8);
        const OctRec cr = nodes[cb + j];
        if (COMBINE && !oct_top_ordered(cr, top + 1, cut)) {
            frontier(O);
            continue;
        }
        const uint64_t open = oct_eval_ordered<QUAD>(cr, a.quad + 2 * (cb + j), O, lane, px, py, pz, G, eps2, bodies,
                                                     a.n, rx, ry, rz, cpp, cnp, fail);
        if (open) {
            if (sp >= MAXL) { fail = true; break; }
            if (lane == 0) { s_node[sp] = cb + j; s_pos[sp] = 0; s_mask[sp] = open; }
            // park this frame's state; the child's frame starts in registers
            // (its first octant step writes before anything reads)
#pragma unroll
            for (int i = 0; i < 9; ++i) s_st[top * 9 + i] = cur[i];
            ++sp;
        }
    }
}

// The ORDERED octree walk, one wave per 64 targets from the root.
template <int MAXL, bool QUAD, bool FOREST>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_oct_ordered(WalkArgs a) {
    __shared__ int s_node_all[kOrderedWaves][MAXL];
    __shared__ int s_pos_all[kOrderedWaves][MAXL];  // next child index | next octant << 8
    __shared__ uint64_t s_mask_all[kOrderedWaves][MAXL];
    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const WaveTask T = wave_task<FOREST>(a, (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    if (FOREST && (T.root < 0 || T.root >= a.max_nodes)) {
        if (lane == 0) atomicOr(a.err, 2u);
        return;
    }
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    uint32_t cpp = 0, cnp = 0;
    bool fail = false;
    oct_ordered_descent<MAXL, QUAD, false>(a, T.root, vmask, 0, me.x, me.y, me.z, T.G, T.eps2, s_node_all[wv],
                                           s_pos_all[wv], s_mask_all[wv], rx, ry, rz, cpp, cnp, fail, tgt);
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) { a.ax[o] = rx; a.ay[o] = ry; a.az[o] = rz; }
    }
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// ORDERED octree split (one system): a.split waves per group of 64 targets.
// Each lists the frontier in the ORDERED walk's visit order (masks: the lanes
// that evaluate the node), and wave k computes the results of entries
// i with i mod a.split == k, as the walk would (each node's result is
// assigned, not accumulated), for k_walk_oct_ordered_combine.
template <int MAXL, bool QUAD>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_oct_ordered_split(WalkArgs a) {
    __shared__ int s_node_all[kOrderedWaves][MAXL];
    __shared__ int s_pos_all[kOrderedWaves][MAXL];
    __shared__ uint64_t s_mask_all[kOrderedWaves][MAXL];
    __shared__ int f_node_all[kOrderedWaves][kOrderedSplitSlots];
    __shared__ uint64_t f_mask_all[kOrderedWaves][kOrderedSplitSlots];
    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const int K = a.split;
    const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv;
    const int64_t grp = gw / K;
    const int k = (int)(gw - grp * K);
    const WaveTask T = wave_task<false>(a, grp);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const OctRec *__restrict__ nodes = a.onodes;
    int *s_node = s_node_all[wv];
    int *s_pos = s_pos_all[wv];
    uint64_t *s_mask = s_mask_all[wv];
    int *f_node = f_node_all[wv];
    uint64_t *f_mask = f_mask_all[wv];
    bool fail = false;

    // the frontier, in the order the ORDERED walk meets it: frames over the
    // top nodes, children in index order, a top child's frame walked before
    // its later siblings
    const OctRec root = nodes[T.root];
    const int cut = (root.topo.y + 31) / 32;
    int F = 0, sp = 0;
    auto add = [&](int node, uint64_t m) {
        if (F >= kOrderedSplitSlots) { fail = true; return; }
        if (lane == 0) { f_node[F] = node; f_mask[F] = m; }
        ++F;
    };
    if (!oct_top_ordered(root, 0, cut)) {
        add(T.root, vmask);
    } else {
        const bool act = (vmask >> lane) & 1ull;
        const uint64_t O = vmask & ~__ballot(act && oct_accepts(root, me.x, me.y, me.z));
        if (O) {
            if (lane == 0) { s_node[0] = T.root; s_pos[0] = 0; s_mask[0] = O; }
            sp = 1;
        }
    }
    while (sp > 0 && !fail) {
        __builtin_amdgcn_wave_barrier();
        const int top = sp - 1;
        const int X = uniform_i32(s_node[top]);
        const int j = uniform_i32(s_pos[top]);
        const uint64_t O = uniform_u64(s_mask[top]);
        if (X < 0 || X >= a.max_nodes) { fail = true; break; }
        const OctRec xr = nodes[X];
        const int cb = xr.topo.z, nch = oct_nchild(xr);
        if (cb < 0 || nch < 1 || nch > 8 || (int64_t)cb + nch > a.max_nodes) { fail = true; break; }
        if (j == nch) {
            sp = top;
            continue;
        }
        if (lane == 0) s_pos[top] = j + 1;
        const OctRec cr = nodes[cb + j];
        if (!oct_top_ordered(cr, top + 1, cut)) {
            add(cb + j, O);
            continue;
        }
        const bool act = (O >> lane) & 1ull;
        const uint64_t open = O & ~__ballot(act && oct_accepts(cr, me.x, me.y, me.z));
        if (open) {
            if (sp >= MAXL) { fail = true; break; }
            if (lane == 0) { s_node[sp] = cb + j; s_pos[sp] = 0; s_mask[sp] = open; }
            ++sp;
        }
    }
    __builtin_amdgcn_wave_barrier();
    uint32_t cpp = 0, cnp = 0;
    const int64_t stride = a.split_stride;
    for (int i = k; i < F && !fail; i += K) {
        const int f = uniform_i32(f_node[i]);
        const uint64_t m = uniform_u64(f_mask[i]);
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
        oct_ordered_descent<MAXL, QUAD, false>(a, f, m, cut, me.x, me.y, me.z, T.G, T.eps2, s_node, s_pos, s_mask, rx,
                                               ry, rz, cpp, cnp, fail, tgt);
        if (valid && ((m >> lane) & 1ull)) {
            const int64_t o = (int64_t)i * 3 * stride + tgt;
            a.split_part_f[o] = rx;
            a.split_part_f[o + stride] = ry;
            a.split_part_f[o + 2 * stride] = rz;
        }
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// The ORDERED octree split's combine: the walk over the top nodes, each
// frontier child's stored result applied to its parent's bracket.
template <int MAXL, bool QUAD>
__global__ __launch_bounds__(64 * kOrderedWaves) void k_walk_oct_ordered_combine(WalkArgs a) {
    __shared__ int s_node_all[kOrderedWaves][MAXL];
    __shared__ int s_pos_all[kOrderedWaves][MAXL];
    __shared__ uint64_t s_mask_all[kOrderedWaves][MAXL];
    const int lane = lane_id(), wv = uniform_i32(threadIdx.x / kWave);
    const WaveTask T = wave_task<false>(a, (int64_t)xcd_block(blockIdx.x, gridDim.x) * kOrderedWaves + wv);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int cut = (a.onodes[T.root].topo.y + 31) / 32;
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    uint32_t cpp = 0, cnp = 0;
    bool fail = false;
    oct_ordered_descent<MAXL, QUAD, true>(a, T.root, vmask, cut, me.x, me.y, me.z, T.G, T.eps2, s_node_all[wv],
                                          s_pos_all[wv], s_mask_all[wv], rx, ry, rz, cpp, cnp, fail, tgt);
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) { a.ax[o] = rx; a.ay[o] = ry; a.az[o] = rz; }
    }
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

// The octree's top walk: the root, then (for a tree of at most
// kSplitMaxBodies bodies) its child group when every child in it is internal
// and holds more than ceil(root count / 32) bodies; the nodes
// it accepts sum into s_top.  The frontier is the child group of each root
// child some lane opens (or the root's own group when it is not expanded):
// at most 8 entries, each one block.  Sums per block as in the FAST
// linear-BVH walks; an entry (group, mask) is exactly the group loop's own
// state.
constexpr int kOctMaxFrontier = 8;

__device__ __forceinline__ bool oct_top(const OctRec &r, int cut) { return r.topo.z >= 0 && r.topo.y > cut; }

template <bool SOFT, int STATS, bool QUAD, bool INTERACT>
__device__ int oct_top_walk(const OctRec *__restrict__ nodes, const float4 *__restrict__ quad,
                            const float4 *__restrict__ bodies, int64_t n, int root, uint64_t vmask, float eps2,
                            FastWave &W, int *f_grp, uint64_t *f_mask, int64_t max_nodes, float4 *sh, bool &fail) {
    const OctRec r = load_uniform(nodes + root);
    float ex, ey, ez, d2;
    uint64_t open;
    if (INTERACT) open = visit_oct<SOFT, STATS, QUAD>(r, quad + 2 * root, vmask, bodies, n, eps2, W, fail, sh);
    else open = r.topo.z >= 0 ? vmask & ~oct_accept_mask(r, vmask, W, ex, ey, ez, d2) : 0ull;
    if (open == 0) return 0;  // a leaf root, or every lane accepted it
    const int c = r.topo.z, nch = oct_nchild(r);
    if (c < 0 || nch < 1 || nch > 8 || (int64_t)c + nch > max_nodes) {
        fail = true;
        return 0;
    }
    const int cut = (r.topo.y + 31) / 32;
    // a tree too large to split keeps the root's group as its one entry
    bool expand = r.topo.y <= kSplitMaxBodies;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (k < nch) expand = expand && oct_top(load_uniform(nodes + c + k), cut);
    if (!expand) {
        if (lane_id() == 0) { f_grp[0] = c | (nch << 28); f_mask[0] = open; }
        __builtin_amdgcn_wave_barrier();
        return 1;
    }
    int F = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (k >= nch) continue;
        const OctRec x = load_uniform(nodes + c + k);
        const uint64_t o = INTERACT ? oct_node<SOFT, STATS, QUAD>(x, quad + 2 * (c + k), open, eps2, W)
                                    : open & ~oct_accept_mask(x, open, W, ex, ey, ez, d2);
        if (o == 0) continue;
        const int cc = x.topo.z, nn = oct_nchild(x);
        if (cc < 0 || nn < 1 || nn > 8 || (int64_t)cc + nn > max_nodes) {
            fail = true;
            break;
        }
        if (lane_id() == 0) { f_grp[F] = cc | (nn << 28); f_mask[F] = o; }
        ++F;
    }
    __builtin_amdgcn_wave_barrier();
    return F;
}

// FAST octree walk.  FOREST: the wave's system from the forest's wave table.
// SPLIT: a.split waves per group of 64 targets, wave k walking
// the blocks b with b mod a.split == k (wave 0 also the top nodes), the sums
// meeting in k_walk_split_reduce.
template <int WAVES, bool SOFT, int STATS, bool QUAD, bool FOREST, bool SPLIT>
__global__ __launch_bounds__(64 * WAVES) void k_walk_oct_fast(WalkArgs a) {
    __shared__ int s_grp[WAVES][kStackOct];       // first child | nchild << 28
    __shared__ uint64_t s_mask[WAVES][kStackOct];
    __shared__ float4 s_leaf[WAVES][kLeafScratch];
    __shared__ int4 s_lq[WAVES][8];                // a group's opened leaves: first, cnt, open mask
    __shared__ int f_grp[WAVES][kOctMaxFrontier];  // first child | nchild << 28
    __shared__ uint64_t f_mask[WAVES][kOctMaxFrontier];

    const float4 *__restrict__ bodies = a.bodies;
    const OctRec *__restrict__ nodes = a.onodes;
    // the wave index is uniform: readfirstlane makes the per-wave task
    // scalar, while LDS addresses use the per-lane copy (fewer SGPRs)
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t gw = (int64_t)xcd_block(blockIdx.x, gridDim.x) * WAVES + uniform_i32(wv);
    const int K = SPLIT ? a.split : 1;
    const int64_t grp = SPLIT ? gw / K : gw;
    const int k = SPLIT ? (int)(gw - grp * K) : 0;
    const WaveTask T = wave_task<FOREST>(a, grp);
    const int64_t tgt = T.base + lane;
    const bool valid = tgt < T.end;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;
    if (FOREST && (T.root < 0 || T.root >= a.max_nodes)) {
        if (lane == 0) atomicOr(a.err, 2u);
        return;
    }

    FastWave W;
    const float4 me = valid ? a.targets[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    fast_init(W, me.x, me.y, me.z);
    W.G = T.G;
    const float eps2 = T.eps2;
    const int64_t n = a.n;
    bool fail = false;

    int F;
    if (!SPLIT || k == 0)
        F = oct_top_walk<SOFT, STATS, QUAD, true>(nodes, a.quad, bodies, n, T.root, vmask, eps2, W, f_grp[wv],
                                                  f_mask[wv], a.max_nodes, s_leaf[wv], fail);
    else
        F = oct_top_walk<SOFT, STATS, QUAD, false>(nodes, a.quad, bodies, n, T.root, vmask, eps2, W, f_grp[wv],
                                                   f_mask[wv], a.max_nodes, s_leaf[wv], fail);
    F = uniform_i32(F);
    // s_top, then each block's sum from zero
    fast_flush(W);
    double S0 = W.sx, S1 = W.sy, S2 = W.sz;
    W.sx = 0.0; W.sy = 0.0; W.sz = 0.0;
    const int64_t stride = a.split_stride;
    if (SPLIT && k == 0 && valid) {
        a.split_part[tgt] = S0;
        a.split_part[stride + tgt] = S1;
        a.split_part[2 * stride + tgt] = S2;
        if (lane == 0) a.split_nb[grp] = F;
    }

    // One loop, shaped as the plain walk's: when a frontier subtree is done
    // (empty mask, empty stack) the refill path closes its block and starts
    // the next owned entry.
    int i = -1, c = 0, nch = 0, sp = 0;
    uint64_t mask = 0;
    const int64_t max_iters = a.max_nodes + 64;  // each node's group is visited at most once
    int64_t iters = 0;
    while (!fail) {
        if (mask == 0) {
            if (sp == 0) {
                // the next frontier group this wave owns, after closing the
                // block of the entry just walked (each entry is a block)
                if (i >= 0) {
                    fast_flush(W);
                    if (SPLIT) {
                        if (valid) {
                            const int64_t o = (int64_t)(1 + i) * 3 * stride + tgt;
                            a.split_part[o] = W.sx;
                            a.split_part[o + stride] = W.sy;
                            a.split_part[o + 2 * stride] = W.sz;
                        }
                    } else {
                        S0 += W.sx; S1 += W.sy; S2 += W.sz;
                    }
                    W.sx = 0.0; W.sy = 0.0; W.sz = 0.0;
                }
                ++i;
                if (SPLIT) i += (k - i % K + K) % K;
                if (i >= F) break;
                const int g = uniform_i32(f_grp[wv][i]);
                c = g & 0x0FFFFFFF;
                nch = (int)((uint32_t)g >> 28);
                mask = uniform_u64(f_mask[wv][i]);
            } else {
                --sp;
                __builtin_amdgcn_wave_barrier();
                const int g = uniform_i32(s_grp[wv][sp]);
                c = g & 0x0FFFFFFF;
                nch = (int)((uint32_t)g >> 28);
                mask = uniform_u64(s_mask[wv][sp]);
            }
        }
        c = uniform_i32(c);
        nch = uniform_i32(nch);
        if (++iters > max_iters || c < 0 || nch < 1 || nch > 8 || (int64_t)c + nch > a.max_nodes) {
            fail = true;
            break;
        }
        // The group's children, unrolled (constant record offsets, one
        // uniform guard per child): node interactions now, opened internal
        // children pushed (the last one is walked next), opened leaves queued
        // in LDS and evaluated after the group.
        uint64_t nm = 0;
        int nc = 0, nn = 0, nl = 0;
        const OctRec *grp_nodes = nodes + c;
        const float4 *qg = a.quad + 2 * c;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (q >= nch) break;
            const OctRec r = load_uniform(grp_nodes + q);
            const uint64_t o = oct_node<SOFT, STATS, QUAD>(r, qg + 2 * q, mask, eps2, W);
            if (o == 0) continue;
            if (r.topo.z >= 0) {
                if (nm) {
                    if (sp >= kStackOct) { fail = true; break; }
                    if (lane == 0) { s_grp[wv][sp] = nc | (nn << 28); s_mask[wv][sp] = nm; }
                    __builtin_amdgcn_wave_barrier();
                    ++sp;
                }
                nm = o;
                nc = r.topo.z;
                nn = oct_nchild(r);
            } else {
                if (!oct_leaf_ok<STATS>(r, o, n, W)) { fail = true; break; }
                if (lane == 0) s_lq[wv][nl] = make_int4(r.topo.x, r.topo.y, (int)(uint32_t)o, (int)(uint32_t)(o >> 32));
                ++nl;
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < nl; ++j) {
            const int4 q = s_lq[wv][j];
            const int first = uniform_i32(q.x), cnt = uniform_i32(q.y);
            const uint64_t o = ((uint64_t)(uint32_t)uniform_i32(q.w) << 32) | (uint32_t)uniform_i32(q.z);
            leaf_eval<SOFT>(bodies, first, cnt, o, eps2, W, s_leaf[wv]);
        }
        __builtin_amdgcn_wave_barrier();  // the next group rewrites s_lq
        mask = nm;
        c = nc;
        nch = nn;
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (!SPLIT && valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.nt) {
            const double g = (double)W.G;
            a.ax[o] = (float)(S0 * g);
            a.ay[o] = (float)(S1 * g);
            a.az[o] = (float)(S2 * g);
        }
    }
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)W.pp);
        atomicAdd(&a.counts[1], (unsigned long long)W.np);
        if (STATS >= 2) {
            atomicAdd(&a.counts[2], (unsigned long long)W.visits);
            atomicAdd(&a.counts[3], (unsigned long long)W.lanes);
            atomicAdd(&a.counts[4], (unsigned long long)W.leaf_visits);
            atomicAdd(&a.counts[5], (unsigned long long)W.leaf_lanes);
        }
    }
}

// Waves the walk launches: one per 64 targets, or the forest's aligned waves.
static int64_t walk_waves(const WalkArgs &a) { return a.nsys > 0 ? a.nwaves : (a.nt + kWave - 1) / kWave; }

template <int STATS, bool QUAD>
static void launch_oct_fast_q(const WalkArgs &a, hipStream_t s) {
    constexpr int W = 4;
    const unsigned blocks = (unsigned)((walk_waves(a) + W - 1) / W);
    const dim3 g(blocks), t(64 * W);
    if (a.nsys > 0 && a.split > 1 && a.split_part && a.split_nb) {
        const dim3 gs((unsigned)((a.nwaves * a.split + W - 1) / W));
        if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_oct_fast<W, true, STATS, QUAD, true, true>), gs, t, 0, s, a);
        else hipLaunchKernelGGL((k_walk_oct_fast<W, false, STATS, QUAD, true, true>), gs, t, 0, s, a);
        hipLaunchKernelGGL(k_walk_split_reduce<true>, dim3((unsigned)((a.nt + 255) / 256)), dim3(256), 0, s, a);
    } else if (a.nsys > 0) {
        if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_oct_fast<W, true, STATS, QUAD, true, false>), g, t, 0, s, a);
        else hipLaunchKernelGGL((k_walk_oct_fast<W, false, STATS, QUAD, true, false>), g, t, 0, s, a);
    } else if (a.split > 1 && a.split_part && a.split_nb) {
        const int64_t waves = (a.nt + kWave - 1) / kWave * a.split;
        const dim3 gs((unsigned)((waves + W - 1) / W));
        if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_oct_fast<W, true, STATS, QUAD, false, true>), gs, t, 0, s, a);
        else hipLaunchKernelGGL((k_walk_oct_fast<W, false, STATS, QUAD, false, true>), gs, t, 0, s, a);
        hipLaunchKernelGGL(k_walk_split_reduce<false>, dim3((unsigned)((a.nt + 255) / 256)), dim3(256), 0, s, a);
    } else {
        if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_oct_fast<W, true, STATS, QUAD, false, false>), g, t, 0, s, a);
        else hipLaunchKernelGGL((k_walk_oct_fast<W, false, STATS, QUAD, false, false>), g, t, 0, s, a);
    }
}

template <int STATS>
static void launch_oct_fast(const WalkArgs &a, hipStream_t s) {
    if (a.quad) launch_oct_fast_q<STATS, true>(a, s);
    else launch_oct_fast_q<STATS, false>(a, s);
}

void launch_walk_oct(const WalkArgs &a, int ordered, int stats, hipStream_t s) {
    if (a.n <= 0 || a.nt <= 0) return;
    if (!ordered) {
        if (stats >= 2) launch_oct_fast<2>(a, s);
        else if (stats == 1) launch_oct_fast<1>(a, s);
        else launch_oct_fast<0>(a, s);
    } else if (a.nsys == 0 && a.split > 1 && a.split_part_f) {
        constexpr int W = kOrderedWaves;
        const int64_t groups = (a.nt + kWave - 1) / kWave;
        const dim3 gs((unsigned)((groups * a.split + W - 1) / W)), gc((unsigned)((groups + W - 1) / W)), t(64 * W);
        const bool shallow = a.levels <= 18;
        if (a.quad && shallow) {
            hipLaunchKernelGGL((k_walk_oct_ordered_split<18, true>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_oct_ordered_combine<18, true>), gc, t, 0, s, a);
        } else if (a.quad) {
            hipLaunchKernelGGL((k_walk_oct_ordered_split<kOctLevels, true>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_oct_ordered_combine<kOctLevels, true>), gc, t, 0, s, a);
        } else if (shallow) {
            hipLaunchKernelGGL((k_walk_oct_ordered_split<18, false>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_oct_ordered_combine<18, false>), gc, t, 0, s, a);
        } else {
            hipLaunchKernelGGL((k_walk_oct_ordered_split<kOctLevels, false>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_oct_ordered_combine<kOctLevels, false>), gc, t, 0, s, a);
        }
    } else {
        constexpr int W = kOrderedWaves;
        const unsigned blocks = (unsigned)((walk_waves(a) + W - 1) / W);
        const dim3 g(blocks), t(64 * W);
        const bool shallow = a.levels <= 18;  // a.levels: the deepest member's levels in a forest
        if (a.nsys > 0) {
            if (a.quad) {
                if (shallow) hipLaunchKernelGGL((k_walk_oct_ordered<18, true, true>), g, t, 0, s, a);
                else hipLaunchKernelGGL((k_walk_oct_ordered<kOctLevels, true, true>), g, t, 0, s, a);
            } else {
                if (shallow) hipLaunchKernelGGL((k_walk_oct_ordered<18, false, true>), g, t, 0, s, a);
                else hipLaunchKernelGGL((k_walk_oct_ordered<kOctLevels, false, true>), g, t, 0, s, a);
            }
        } else if (a.quad) {
            if (shallow) hipLaunchKernelGGL((k_walk_oct_ordered<18, true, false>), g, t, 0, s, a);
            else hipLaunchKernelGGL((k_walk_oct_ordered<kOctLevels, true, false>), g, t, 0, s, a);
        } else {
            if (shallow) hipLaunchKernelGGL((k_walk_oct_ordered<18, false, false>), g, t, 0, s, a);
            else hipLaunchKernelGGL((k_walk_oct_ordered<kOctLevels, false, false>), g, t, 0, s, a);
        }
    }
}

template <int STATS, bool QUAD, int TPL, bool FOREST, bool SPLIT>
static void launch_lbvh_fast_k(const WalkArgs &a, int64_t waves, hipStream_t s) {
    constexpr int W = 4;
    const dim3 g((unsigned)((waves + W - 1) / W)), t(64 * W);
    if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_lbvh_fast<W, true, STATS, QUAD, TPL, FOREST, SPLIT>), g, t, 0, s, a);
    else hipLaunchKernelGGL((k_walk_lbvh_fast<W, false, STATS, QUAD, TPL, FOREST, SPLIT>), g, t, 0, s, a);
}

template <int STATS, bool QUAD>
static void launch_fast_q(const WalkArgs &a, hipStream_t s) {
    const int tpl = a.tpl == 2 || a.tpl == 4 ? a.tpl : 1;
    if (a.nsys > 0) {
        // a forest's wave table counts waves of 64 tpl targets
        if (tpl == 1 && a.split > 1 && a.split_part && a.split_nb) {
            launch_lbvh_fast_k<STATS, QUAD, 1, true, true>(a, a.nwaves * a.split, s);
            hipLaunchKernelGGL(k_walk_split_reduce<true>, dim3((unsigned)((a.nt + 255) / 256)), dim3(256), 0, s, a);
            return;
        }
        if (tpl == 2) launch_lbvh_fast_k<STATS, QUAD, 2, true, false>(a, a.nwaves, s);
        else if (tpl == 4) launch_lbvh_fast_k<STATS, QUAD, 4, true, false>(a, a.nwaves, s);
        else launch_lbvh_fast_k<STATS, QUAD, 1, true, false>(a, a.nwaves, s);
        return;
    }
    const int64_t groups = (a.nt + kWave - 1) / kWave;
    if (tpl == 2) return launch_lbvh_fast_k<STATS, QUAD, 2, false, false>(a, (groups + 1) / 2, s);
    if (tpl == 4) return launch_lbvh_fast_k<STATS, QUAD, 4, false, false>(a, (groups + 3) / 4, s);
    if (a.split > 1 && a.split_part && a.split_nb) {
        launch_lbvh_fast_k<STATS, QUAD, 1, false, true>(a, groups * a.split, s);
        hipLaunchKernelGGL(k_walk_split_reduce<false>, dim3((unsigned)((a.nt + 255) / 256)), dim3(256), 0, s, a);
        return;
    }
    if (a.n > kSplitMaxBodies) {
        constexpr int W = 4;
        const dim3 g((unsigned)((groups + W - 1) / W)), t(64 * W);
        if (a.eps2 > 0.0f) hipLaunchKernelGGL((k_walk_lbvh_plain<W, true, STATS, QUAD>), g, t, 0, s, a);
        else hipLaunchKernelGGL((k_walk_lbvh_plain<W, false, STATS, QUAD>), g, t, 0, s, a);
        return;
    }
    launch_lbvh_fast_k<STATS, QUAD, 1, false, false>(a, groups, s);
}

template <int STATS>
static void launch_fast(const WalkArgs &a, hipStream_t s) {
    if (a.quad) launch_fast_q<STATS, true>(a, s);
    else launch_fast_q<STATS, false>(a, s);
}

void launch_walk_lbvh(const WalkArgs &a, int ordered, int max_depth, int stats, hipStream_t s) {
    if (a.n <= 0 || a.nt <= 0) return;
    if (!ordered) {
        if (stats >= 2) launch_fast<2>(a, s);
        else if (stats == 1) launch_fast<1>(a, s);
        else launch_fast<0>(a, s);
    } else if (a.nsys == 0 && a.split > 1 && a.split_part_f) {
        constexpr int W = kOrderedWaves;
        const int64_t groups = (a.nt + kWave - 1) / kWave;
        const dim3 gs((unsigned)((groups * a.split + W - 1) / W)), gc((unsigned)((groups + W - 1) / W)), t(64 * W);
        if (a.quad && max_depth <= 32) {
            hipLaunchKernelGGL((k_walk_lbvh_ordered_split<32, true>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_lbvh_ordered_combine<32, true>), gc, t, 0, s, a);
        } else if (a.quad) {
            hipLaunchKernelGGL((k_walk_lbvh_ordered_split<64, true>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_lbvh_ordered_combine<64, true>), gc, t, 0, s, a);
        } else if (max_depth <= 32) {
            hipLaunchKernelGGL((k_walk_lbvh_ordered_split<32, false>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_lbvh_ordered_combine<32, false>), gc, t, 0, s, a);
        } else {
            hipLaunchKernelGGL((k_walk_lbvh_ordered_split<64, false>), gs, t, 0, s, a);
            hipLaunchKernelGGL((k_walk_lbvh_ordered_combine<64, false>), gc, t, 0, s, a);
        }
    } else {
        constexpr int W = kOrderedWaves;
        const unsigned blocks = (unsigned)((walk_waves(a) + W - 1) / W);
        const dim3 g(blocks), t(64 * W);
        if (a.nsys > 0 && a.quad)
            hipLaunchKernelGGL((k_walk_lbvh_ordered<32, true, true>), g, t, 0, s, a);
        else if (a.nsys > 0)
            hipLaunchKernelGGL((k_walk_lbvh_ordered<32, true, false>), g, t, 0, s, a);
        else if (a.quad && max_depth <= 32)
            hipLaunchKernelGGL((k_walk_lbvh_ordered<32, false, true>), g, t, 0, s, a);
        else if (a.quad)
            hipLaunchKernelGGL((k_walk_lbvh_ordered<64, false, true>), g, t, 0, s, a);
        else if (max_depth <= 32)
            hipLaunchKernelGGL((k_walk_lbvh_ordered<32, false, false>), g, t, 0, s, a);
        else
            hipLaunchKernelGGL((k_walk_lbvh_ordered<64, false, false>), g, t, 0, s, a);
    }
}

}  // namespace nbg
