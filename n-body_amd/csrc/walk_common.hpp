// Building blocks of the FAST walks shared by the Barnes-Hut traversals
// (walk.hip) and the implicit FMM near field (fmm.hip).
#pragma once

#include "device.hpp"
#include "launch.hpp"

namespace nbg {

// A lane's acceleration is summed in two stages: interactions accumulate
// into an FP32 partial (rx, ry, rz) that is added into an FP64 running sum
// (sx, sy, sz) after every leaf, so no FP32 sum runs over more than one
// leaf and the node interactions met since the previous one.  A single FP32
// accumulator over the whole walk (thousands of terms in traversal order)
// drifts from the reference's pairwise tree-ordered sum by more than 1e-5
// of |a| at theta 0.3; the two-stage sum stays within FP32 rounding of the
// exact value of the interaction set.
struct FastWave {
    float px, py, pz;          // this lane's target
    float rx, ry, rz;          // FP32 partial since the last flush
    double sx, sy, sz;         // FP64 running sum
    float G;                   // per lane, so G * (scalar mass) is one VALU op
    uint64_t pp, np;           // wave totals (interaction counts)
    uint64_t visits, lanes, leaf_visits, leaf_lanes;  // wave-uniform diagnostics
};

__device__ __forceinline__ void fast_init(FastWave &W, float px, float py, float pz) {
    W.px = px; W.py = py; W.pz = pz;
    W.rx = 0.0f; W.ry = 0.0f; W.rz = 0.0f;
    W.sx = 0.0; W.sy = 0.0; W.sz = 0.0;
    W.pp = 0; W.np = 0;
    W.visits = 0; W.lanes = 0; W.leaf_visits = 0; W.leaf_lanes = 0;
}

__device__ __forceinline__ void fast_flush(FastWave &W) {
    W.sx += (double)W.rx; W.sy += (double)W.ry; W.sz += (double)W.rz;
    W.rx = 0.0f; W.ry = 0.0f; W.rz = 0.0f;
}

// The lane's acceleration, (running sum + partial + extra) * G, rounded once.
__device__ __forceinline__ float3 fast_result(const FastWave &W, float G, float ex = 0.0f, float ey = 0.0f,
                                              float ez = 0.0f) {
    const double g = (double)G;
    return make_float3((float)((W.sx + ((double)W.rx + (double)ex)) * g),
                       (float)((W.sy + ((double)W.ry + (double)ey)) * g),
                       (float)((W.sz + ((double)W.rz + (double)ez)) * g));
}

// Monopole (or body) interaction with rsqrt + FMA: s = m * inv^3; the walk
// multiplies the sum by G once at the end.  eps2 > 0 (SOFT) makes r2 > 0 for
// every pair, coincident ones included (their d = 0 contributes exactly
// zero); otherwise r2 = 0 is guarded.
template <bool SOFT>
__device__ __forceinline__ void interact(float bx, float by, float bz, float bm, float eps2, FastWave &W) {
    const float dx = bx - W.px, dy = by - W.py, dz = bz - W.pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
    const float s = bm * (inv * (inv * inv));
    W.rx = __builtin_fmaf(dx, s, W.rx);
    W.ry = __builtin_fmaf(dy, s, W.ry);
    W.rz = __builtin_fmaf(dz, s, W.rz);
}

// Quadrupole node interaction (Gravity::nodeParticle<2>, Gravity.h:238-259)
// with rsqrt + FMA.  With d = c - p (the reference's R = -d), the reference's
// R (-M) / r^3 + D3(R) . Q / 2 expands to
//   d (M r^-3 + 7.5 r^-7 d.Q.d - 1.5 r^-5 tr Q) - 3 r^-5 Q d,
// r^2 = d.d + eps2; G is applied once at the end of the walk.  d = 0 (the
// target at the node's COM) contributes exactly zero.
template <bool SOFT>
__device__ __forceinline__ void interact_quad(float bx, float by, float bz, float bm, float4 q0, float4 q1,
                                              float eps2, FastWave &W) {
    const float dx = bx - W.px, dy = by - W.py, dz = bz - W.pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
    const float inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
    // Q d with Q = [[xx xy xz] [xy yy yz] [xz yz zz]] = [[q0.x q0.y q0.z] [q0.y q0.w q1.x] [q0.z q1.x q1.y]]
    const float qx = __builtin_fmaf(q0.x, dx, __builtin_fmaf(q0.y, dy, q0.z * dz));
    const float qy = __builtin_fmaf(q0.y, dx, __builtin_fmaf(q0.w, dy, q1.x * dz));
    const float qz = __builtin_fmaf(q0.z, dx, __builtin_fmaf(q1.x, dy, q1.y * dz));
    const float dqd = __builtin_fmaf(dx, qx, __builtin_fmaf(dy, qy, dz * qz));
    const float s = __builtin_fmaf(bm, inv3, inv5 * __builtin_fmaf(7.5f * inv2, dqd, -1.5f * q1.z));
    const float t = -3.0f * inv5;
    W.rx = __builtin_fmaf(dx, s, __builtin_fmaf(t, qx, W.rx));
    W.ry = __builtin_fmaf(dy, s, __builtin_fmaf(t, qy, W.ry));
    W.rz = __builtin_fmaf(dz, s, __builtin_fmaf(t, qz, W.rz));
}

// Plain four-word records, read field by field out of the constant address
// space (HIP's vector types do not copy out of it).
struct alignas(16) Body4 {
    float x, y, z, w;
};
struct alignas(16) Int4Pod {
    int x, y, z, w;
};
__device__ __forceinline__ Body4 load_body(const_ptr<Body4> p) { return Body4{p->x, p->y, p->z, p->w}; }

// Scalar load of a wave-uniform float4 or octree record that the kernel
// never writes (see as_const).
__device__ __forceinline__ float4 load_uniform(const float4 *p) {
    const const_ptr<Body4> q = as_const(reinterpret_cast<const Body4 *>(p));
    return make_float4(q->x, q->y, q->z, q->w);
}
__device__ __forceinline__ OctRec load_uniform(const OctRec *p) {
    const const_ptr<Body4> f = as_const(reinterpret_cast<const Body4 *>(p));
    const const_ptr<Int4Pod> t = as_const(reinterpret_cast<const Int4Pod *>(p));
    OctRec r;
    r.com = make_float4(f[0].x, f[0].y, f[0].z, f[0].w);
    r.lo = make_float4(f[1].x, f[1].y, f[1].z, f[1].w);
    r.hi = make_float4(f[2].x, f[2].y, f[2].z, f[2].w);
    r.topo = make_int4(t[3].x, t[3].y, t[3].z, t[3].w);
    return r;
}

__device__ __forceinline__ NodeRec load_uniform(const NodeRec *p) {
    const const_ptr<Body4> f = as_const(reinterpret_cast<const Body4 *>(p));
    const const_ptr<Int4Pod> t = as_const(reinterpret_cast<const Int4Pod *>(p));
    NodeRec r;
    r.mac = make_float4(f[0].x, f[0].y, f[0].z, f[0].w);
    r.com = make_float4(f[1].x, f[1].y, f[1].z, f[1].w);
    r.topo = make_int4(t[2].x, t[2].y, t[2].z, t[2].w);
    return r;
}

// The lanes in `take` interact with the `cnt` bodies starting at `first`.
template <bool SOFT>
__device__ __forceinline__ void leaf_bodies(const float4 *__restrict__ bodies, int first, int cnt, uint64_t take,
                                            float eps2, FastWave &W) {
    if (__builtin_amdgcn_inverse_ballot_w64(take)) {
        // uniform: scalar loads (a plain four-float record, which copies out
        // of the constant address space)
        const const_ptr<Body4> lb = as_const(reinterpret_cast<const Body4 *>(bodies + first));
        int j = 0;
        for (; j + 4 <= cnt; j += 4) {
            const Body4 b0 = load_body(lb + j), b1 = load_body(lb + j + 1);
            const Body4 b2 = load_body(lb + j + 2), b3 = load_body(lb + j + 3);
            interact<SOFT>(b0.x, b0.y, b0.z, b0.w, eps2, W);
            interact<SOFT>(b1.x, b1.y, b1.z, b1.w, eps2, W);
            interact<SOFT>(b2.x, b2.y, b2.z, b2.w, eps2, W);
            interact<SOFT>(b3.x, b3.y, b3.z, b3.w, eps2, W);
        }
        for (; j < cnt; ++j) {
            const Body4 b = load_body(lb + j);
            interact<SOFT>(b.x, b.y, b.z, b.w, eps2, W);
        }
    }
    fast_flush(W);
}

// Sum of v over each 16-lane row of the wave (four DPP row shifts); the
// row's last lane holds the row's total.  Every lane of the wave must be
// active.
__device__ __forceinline__ float row_sum16(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, true));  // row_shr:1
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xf, 0xf, true));  // row_shr:2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xf, 0xf, true));  // row_shr:4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xf, 0xf, true));  // row_shr:8
    return v;
}

// A leaf opened by few lanes (k = popcount(open)) and holding at most 16
// bodies.  The per-lane loop costs cnt wave instruction-sets whatever k is;
// here the k x cnt (target, body) pairs are spread over the wave instead:
// lane j of each 16-lane row holds body j of the leaf (one vector load,
// issued before any round) and each row takes one target per round, four
// per round; rows are summed by DPP and the row's last lane returns the sum
// to the target's lane through LDS.  `sh` is the wave's 32-entry LDS
// scratch (targets in slots 0..15 by rank, sums in 16..31).  Rows past the
// last target repeat it and are not written back; lanes past the last body
// hold a massless body at the origin (an exact zero term).  Requires
// 1 <= k <= 16, cnt <= 16 and every lane of the wave active (the walks call
// it in wave-uniform flow).
template <bool SOFT>
__device__ __forceinline__ void leaf_rows(const float4 *__restrict__ bodies, int first, int cnt, uint64_t open,
                                          float eps2, FastWave &W, float4 *sh) {
    const int lane = lane_id();
    const int row = lane >> 4, j = lane & 15;
    const float4 b = j < cnt ? bodies[first + j] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int k = __popcll(open);
    const bool mine = (open >> lane) & 1ull;
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(open >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)open, 0u));
    if (mine) sh[rank] = make_float4(W.px, W.py, W.pz, 0.0f);
    __builtin_amdgcn_wave_barrier();
    for (int q = 0; q < k; q += 4) {
        const int tr = q + row;
        const float4 t = sh[tr < k ? tr : k - 1];
        const float dx = b.x - t.x, dy = b.y - t.y, dz = b.z - t.z;
        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
        float inv = __builtin_amdgcn_rsqf(r2);
        if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
        const float sc = b.w * (inv * (inv * inv));
        const float fx = row_sum16(dx * sc), fy = row_sum16(dy * sc), fz = row_sum16(dz * sc);
        if (j == 15 && tr < k) sh[16 + tr] = make_float4(fx, fy, fz, 0.0f);
    }
    __builtin_amdgcn_wave_barrier();
    if (mine) {
        const float4 r = sh[16 + rank];
        W.rx += r.x;
        W.ry += r.y;
        W.rz += r.z;
    }
    __builtin_amdgcn_wave_barrier();  // the next leaf rewrites sh
    fast_flush(W);
}

// The opened lanes of a leaf interact with its bodies: spread over rows
// (leaf_rows) when that takes fewer wave instructions than the per-lane loop
// (leaf_bodies), judged per leaf from k and cnt (wave-uniform); `sh` null
// always takes the per-lane loop.
template <bool SOFT>
__device__ __forceinline__ void leaf_eval(const float4 *__restrict__ bodies, int first, int cnt, uint64_t open,
                                          float eps2, FastWave &W, float4 *sh) {
    const int k = __popcll(open);
    // rows: ~16 instructions of set-up and ~30 per round of four targets;
    // per-lane loop: ~12 per body
    if (sh && k <= 16 && cnt <= 16 && 16 + 30 * ((k + 3) >> 2) < 12 * cnt)
        leaf_rows<SOFT>(bodies, first, cnt, open, eps2, W, sh);
    else
        leaf_bodies<SOFT>(bodies, first, cnt, open, eps2, W);
}

// Visits one node for the lanes in `mask` (wave-uniform); returns the lanes
// that open it when it is internal, 0 when it is a leaf or all accept.

// SideLengthOverDistance with the octree exclusion region
// (Descent/DescentCriterion/SideLengthOverDistance.h:17-22,
// Descent/DescentCriterionType.h:81-85): the distance to the centre of mass
// in glm's order against the level's threshold (side / sqrt(d2) < theta is
// monotone in d2), and the target outside [c - side, c + side).
__device__ __forceinline__ bool oct_accepts(const OctRec &r, float px, float py, float pz) {
    const float dx = px - r.com.x, dy = py - r.com.y, dz = pz - r.com.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    const bool inside = r.lo.x <= px && r.lo.y <= py && r.lo.z <= pz && px < r.hi.x && py < r.hi.y && pz < r.hi.z;
    return d2 > r.lo.w && !inside;
}

// Monopole node interaction from the criterion's offset e = p - com and its
// squared length d2 (computed in glm's order for the criterion):
// r^2 = d2 + eps2, a = -e m r^-3 (interact() with d = -e; r^2 rounds once
// more than interact's FMA chain, well inside FP32 accuracy).
template <bool SOFT>
__device__ __forceinline__ void interact_e(float ex, float ey, float ez, float d2, float bm, float eps2,
                                           FastWave &W) {
    const float r2 = d2 + eps2;
    float inv = __builtin_amdgcn_rsqf(r2);
    if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
    const float s = bm * (inv * (inv * inv));
    W.rx = __builtin_fmaf(-ex, s, W.rx);
    W.ry = __builtin_fmaf(-ey, s, W.ry);
    W.rz = __builtin_fmaf(-ez, s, W.rz);
}

// Quadrupole node interaction (interact_quad) from e = p - com and d2.
// With d = -e: Q d = -(Q e), d.Q.d = e.Q.e, so
//   a = -e (M r^-3 + 7.5 r^-7 e.Q.e - 1.5 r^-5 tr Q) + 3 r^-5 Q e.
template <bool SOFT>
__device__ __forceinline__ void interact_quad_e(float ex, float ey, float ez, float d2, float bm, float4 q0,
                                                float4 q1, float eps2, FastWave &W) {
    const float r2 = d2 + eps2;
    float inv = __builtin_amdgcn_rsqf(r2);
    if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
    const float inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
    const float qx = __builtin_fmaf(q0.x, ex, __builtin_fmaf(q0.y, ey, q0.z * ez));
    const float qy = __builtin_fmaf(q0.y, ex, __builtin_fmaf(q0.w, ey, q1.x * ez));
    const float qz = __builtin_fmaf(q0.z, ex, __builtin_fmaf(q1.x, ey, q1.y * ez));
    const float eqe = __builtin_fmaf(ex, qx, __builtin_fmaf(ey, qy, ez * qz));
    const float s = __builtin_fmaf(bm, inv3, inv5 * __builtin_fmaf(7.5f * inv2, eqe, -1.5f * q1.z));
    const float t = 3.0f * inv5;
    W.rx = __builtin_fmaf(-ex, s, __builtin_fmaf(t, qx, W.rx));
    W.ry = __builtin_fmaf(-ey, s, __builtin_fmaf(t, qy, W.ry));
    W.rz = __builtin_fmaf(-ez, s, __builtin_fmaf(t, qz, W.rz));
}

// The lanes of `mask` that accept octree node nd (oct_accepts_e as one
// ballot per comparison: the lane masks combine in scalar registers without
// turning the predicate into a vector value); e = p - com and d2 = |e|^2.
__device__ __forceinline__ uint64_t oct_accept_mask(const OctRec &nd, uint64_t mask, const FastWave &W, float &ex,
                                                    float &ey, float &ez, float &d2) {
    ex = W.px - nd.com.x;
    ey = W.py - nd.com.y;
    ez = W.pz - nd.com.z;
    d2 = ex * ex + ey * ey + ez * ez;
    const uint64_t inside = __ballot(nd.lo.x <= W.px) & __ballot(nd.lo.y <= W.py) & __ballot(nd.lo.z <= W.pz) &
                            __ballot(W.px < nd.hi.x) & __ballot(W.py < nd.hi.y) & __ballot(W.pz < nd.hi.z);
    return __ballot(d2 > nd.lo.w) & ~inside & mask;
}

// The node half of an octree visit: the criterion for the lanes in `mask`
// (wave-uniform), the node interaction for those that accept; returns the
// lanes that open the node.
template <bool SOFT, int STATS, bool QUAD>
__device__ __forceinline__ uint64_t oct_node(const OctRec &nd, const float4 *__restrict__ qn, uint64_t mask,
                                             float eps2, FastWave &W) {
    float ex, ey, ez, d2;
    const uint64_t acc = oct_accept_mask(nd, mask, W, ex, ey, ez, d2);
    if (STATS >= 2) { W.visits += 1; W.lanes += __popcll(mask); }
    if (STATS >= 1) W.np += __popcll(acc);
    if (QUAD) {
        if (acc) {
            const float4 q0 = load_uniform(qn), q1 = load_uniform(qn + 1);
            if (__builtin_amdgcn_inverse_ballot_w64(acc))
                interact_quad_e<SOFT>(ex, ey, ez, d2, nd.com.w, q0, q1, eps2, W);
        }
    } else if (__builtin_amdgcn_inverse_ballot_w64(acc)) {
        interact_e<SOFT>(ex, ey, ez, d2, nd.com.w, eps2, W);
    }
    return mask & ~acc;
}

// A leaf's body range checked against the n bodies: first in [0, n) and cnt
// in [0, n - first] (n < 2^31); counts its interactions.
template <int STATS>
__device__ __forceinline__ bool oct_leaf_ok(const OctRec &nd, uint64_t open, int64_t n, FastWave &W) {
    const int first = nd.topo.x, cnt = nd.topo.y;
    if (__builtin_expect((uint64_t)(uint32_t)first >= (uint64_t)n, 0)) return false;
    if (__builtin_expect((uint64_t)(uint32_t)cnt > (uint64_t)n - (uint32_t)first, 0)) return false;
    if (STATS >= 2) { W.leaf_visits += 1; W.leaf_lanes += __popcll(open); }
    if (STATS >= 1) W.pp += (uint64_t)__popcll(open) * (uint64_t)cnt;
    return true;
}

// FAST over the octree: the children of an opened node are contiguous and
// visited in one group; opened internal children are pushed as groups.
template <bool SOFT, int STATS, bool QUAD>
__device__ __forceinline__ uint64_t visit_oct(const OctRec &nd, const float4 *__restrict__ qn, uint64_t mask,
                                              const float4 *__restrict__ bodies, int64_t n, float eps2, FastWave &W,
                                              bool &fail, float4 *sh = nullptr) {
    const uint64_t open = oct_node<SOFT, STATS, QUAD>(nd, qn, mask, eps2, W);
    if (nd.topo.z >= 0) return open;
    if (open == 0) return 0;
    if (!oct_leaf_ok<STATS>(nd, open, n, W)) { fail = true; return 0; }
    leaf_eval<SOFT>(bodies, nd.topo.x, nd.topo.y, open, eps2, W, sh);
    return 0;
}

constexpr int kStackOct = 256;
constexpr int kLeafScratch = 32;  // float4 per wave for leaf_rows

}  // namespace nbg
