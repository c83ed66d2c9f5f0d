// Device-side helpers shared by the gfx950 kernels (wave64, CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbg {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Lanes strictly below this one.
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Correctly rounded binary32 square root and division.  hipcc lowers these
// to v_sqrt_f32 + a +-1 ulp FMA correction and to the v_div_scale /
// v_div_fmas / v_div_fixup sequence (HIP's default is correctly rounded fp32
// divide/sqrt).  HIP's __fsqrt_rn is NOT used: without
// OCML_BASIC_ROUNDED_OPERATIONS it maps to the approximate native sqrt.
__device__ __forceinline__ float ieee_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ float ieee_div(float a, float b) { return a / b; }

// Three correctly rounded binary32 quotients with one denominator:
// (float)((double)n * r) with r = 1 / (double)d from v_rcp_f64 and two
// Newton steps (within ~2^-52 of 1/d).  The exact quotient of two binary32
// numbers is either a binary32 number or at least ~2^-49 (relative) away
// from every binary32 rounding midpoint (it can never lie on one), so an FP64
// value within ~2^-51 of it rounds to the IEEE quotient.  d = 0, inf or NaN
// take the IEEE division itself.  The identity is checked on 10^8 random
// pairs with the reciprocal perturbed by up to 2 FP64 ulp (CPU test).
__device__ __forceinline__ void ieee_div3(float nx, float ny, float nz, float d, float &qx, float &qy, float &qz) {
    if (__builtin_expect(d > 0.0f && d < __builtin_huge_valf(), 1)) {
        const double dd = (double)d;
        double r = __builtin_amdgcn_rcp(dd);
        r = __builtin_fma(r, __builtin_fma(-dd, r, 1.0), r);
        r = __builtin_fma(r, __builtin_fma(-dd, r, 1.0), r);
        qx = (float)((double)nx * r);
        qy = (float)((double)ny * r);
        qz = (float)((double)nz * r);
    } else {
        qx = ieee_div(nx, d);
        qy = ieee_div(ny, d);
        qz = ieee_div(nz, d);
    }
}

// Gravity::particleParticle (Physics/Rules/Gravity.h:199-209) with the
// reference's FP32 operation order: R = p - a, r = sqrt(R.R + eps2),
// -R m G / (r (r r)); zero when the positions are equal.  Compile without
// FMA contraction.
__device__ __forceinline__ void pp_exact(float bx, float by, float bz, float bm, float px, float py, float pz,
                                         float G, float eps2, float &vx, float &vy, float &vz) {
    if (bx == px && by == py && bz == pz) {
        vx = 0.0f; vy = 0.0f; vz = 0.0f;
        return;
    }
    const float Rx = px - bx, Ry = py - by, Rz = pz - bz;
    const float r2 = Rx * Rx + Ry * Ry + Rz * Rz;
    const float r = ieee_sqrt(r2 + eps2);
    const float r3 = r * (r * r);
    ieee_div3((-Rx * bm) * G, (-Ry * bm) * G, (-Rz * bm) * G, r3, vx, vy, vz);
}

// ---------------------------------------------------------------- quadrupole
// Second moments are stored per node slot as two float4:
//   q[2 s] = (xx, xy, xz, yy),  q[2 s + 1] = (yz, zz, xx + yy + zz, 0).
// The entries are MultipoleMoment<2>'s raw sum m (x - c)(x - c)^T about the
// node's centre of mass c (Physics/MultipoleMoment.h:20-55), in its storage
// order; the trace is for the fast evaluation only.

// One term outerProduct(o, o) * m (SymmetricTensor3.h:531-536, Multipole.h:96-101).
__device__ __forceinline__ void quad_term(float ox, float oy, float oz, float m, float t[6]) {
    t[0] = (ox * ox) * m; t[1] = (ox * oy) * m; t[2] = (ox * oz) * m;
    t[3] = (oy * oy) * m; t[4] = (oy * oz) * m; t[5] = (oz * oz) * m;
}

__device__ __forceinline__ void quad_store(float4 *q, int64_t s, const float Q[6]) {
    q[2 * s] = make_float4(Q[0], Q[1], Q[2], Q[3]);
    q[2 * s + 1] = make_float4(Q[4], Q[5], (Q[0] + Q[3]) + Q[5], 0.0f);
}

// MultipoleMassSummary::summarize(entities) (MultipoleMassSummary.h:39-51):
// libstdc++'s transform_reduce over the bodies, 0 + ((t0 + t1) + (t2 + t3))
// per block of four, then the rest one at a time; t = outer(x - c, x - c) m.
__device__ __forceinline__ void quad_bodies(const float4 *__restrict__ bodies, int first, int cnt, float cx,
                                            float cy, float cz, float Q[6]) {
#pragma unroll
    for (int e = 0; e < 6; ++e) Q[e] = 0.0f;
    int i = 0;
    for (; cnt - i >= 4; i += 4) {
        float t[4][6];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 b = bodies[first + i + k];
            quad_term(b.x - cx, b.y - cy, b.z - cz, b.w, t[k]);
        }
#pragma unroll
        for (int e = 0; e < 6; ++e) Q[e] = Q[e] + ((t[0][e] + t[1][e]) + (t[2][e] + t[3][e]));
    }
    for (; i < cnt; ++i) {
        const float4 b = bodies[first + i];
        float t[6];
        quad_term(b.x - cx, b.y - cy, b.z - cz, b.w, t);
#pragma unroll
        for (int e = 0; e < 6; ++e) Q[e] = Q[e] + t[e];
    }
}

// Gravity::nodeParticle<2> (Physics/Rules/Gravity.h:238-259) in the
// reference's FP32 operation order: R = p - c, r = sqrt(R.R + eps2),
// dPhi = (R * -M) / (r (r r)) + (D<3>(R, r) . Q) / 2, result G * dPhi.
// D<3> (Gravity.h:43-93): entry (a,b,c) = R_a (R_b R_c) * g3 + B * g2 with
// B = 3 R_a on (a,a,a), the odd index's component on (a,a,b) / (a,b,b) and
// 0 on (x,y,z); f2 = 3 / r^3, f3 = -15 / r^4, g2 = f2 / r^2, g3 = f3 / r^3.
// The contraction (SymmetricTensor3.h:266-276) adds D_ijk Q_jk into
// component i over (j, k) in lexicographic order from zero.  Compile
// without FMA contraction.
__device__ __forceinline__ void np_quad_exact(float cx, float cy, float cz, float M, float4 q0, float4 q1, float px,
                                              float py, float pz, float G, float eps2, float &vx, float &vy,
                                              float &vz) {
    if (cx == px && cy == py && cz == pz) {
        vx = 0.0f; vy = 0.0f; vz = 0.0f;
        return;
    }
    const float R[3] = {px - cx, py - cy, pz - cz};
    const float r = ieee_sqrt((R[0] * R[0] + R[1] * R[1] + R[2] * R[2]) + eps2);
    const float r3 = r * (r * r);
    const float f2 = ieee_div(3.0f, r3), f3 = ieee_div(-15.0f, r * r3);
    const float g2 = ieee_div(f2, r * r), g3 = ieee_div(f3, r3);
    // unique entries: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    constexpr int kA[10] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 2};
    constexpr int kB[10] = {0, 0, 0, 1, 1, 2, 1, 1, 2, 2};
    constexpr int kC[10] = {0, 1, 2, 1, 2, 2, 1, 2, 2, 2};
    float D[10];
#pragma unroll
    for (int u = 0; u < 10; ++u) {
        const int a = kA[u], b = kB[u], c = kC[u];
        const float cp = R[a] * (R[b] * R[c]);
        float B;
        if (a == b && b == c) B = 0.0f + R[a] * 3.0f;
        else if (a != b && b != c) B = 0.0f + 0.0f * 3.0f;
        else B = (a == b ? R[c] : R[a]) + 0.0f * 3.0f;
        D[u] = cp * g3 + B * g2;
    }
    // sorted (i, j, k) -> unique index; (j, k) -> moment index
    constexpr int kU3[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}},
                                  {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                                  {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    constexpr int kU2[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const float Q[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
    float h[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) acc = acc + D[kU3[i][j][k]] * Q[kU2[j][k]];
        h[i] = acc;
    }
    const float dx = ieee_div(R[0] * -M, r3) + h[0] / 2.0f;
    const float dy = ieee_div(R[1] * -M, r3) + h[1] / 2.0f;
    const float dz = ieee_div(R[2] * -M, r3) + h[2] / 2.0f;
    vx = G * dx; vy = G * dy; vz = G * dz;
}

// Exact form of the reference's LBVH opening test
//   RN(diag2 / d2) < T,   T = RN(theta * theta)   (DiagonalOverCenterDistance.h:18-22)
// without a division.  RN is monotone, so RN(q) < T iff q < mid, where mid is
// the midpoint between T and its predecessor, or q == mid and round-to-even
// picks the predecessor (T has an odd significand).  With q = diag2 / d2 and
// d2 > 0 this is diag2 < mid * d2 (or ==), and mid * d2 (25 x 24 significant
// bits) is exact in FP64, so the comparison is exact.  d2 == 0 gives q = inf
// or NaN, never accepted.  mid and tie come from mac_threshold() on the host.
__device__ __forceinline__ bool mac_lbvh_exact(float diag2, float d2, double mid, bool tie) {
    const double l = (double)diag2, r = mid * (double)d2;
    return tie ? ((l < r) || (l == r && d2 > 0.0f && l < __builtin_huge_val())) : (l < r);
}

// Per-node form of the same test: the reference accepts iff d2 > thr, where
// thr is the largest float f >= 0 with f * mid <= diag2 (f * mid < diag2 when
// a tie rounds down).  For float d2, "d2 > tau = diag2 / mid" is equivalent
// to "d2 > largest float <= tau"; every product f * mid is exact in FP64.
// mid <= 0 (theta = 0) never accepts; a NaN diag2 never accepts.
__device__ __forceinline__ float mac_lbvh_threshold(float diag2, double mid, bool tie) {
    if (!(mid > 0.0) || diag2 == __builtin_huge_valf()) return __builtin_huge_valf();
    if (diag2 != diag2) return diag2;
    const double D = (double)diag2;
    float f = (float)(D / mid);  // within an ulp of diag2 / mid, >= 0
    auto ok = [&](float v) { const double p = (double)v * mid; return tie ? (p < D) : (p <= D); };
    // neighbours of a non-negative finite float by its bit pattern
    auto up = [](float v) { return __uint_as_float(__float_as_uint(v) + 1u); };
    auto down = [](float v) { return __uint_as_float(__float_as_uint(v) - 1u); };
    while (f > 0.0f && !ok(f)) f = down(f);
    while (f < __FLT_MAX__ && ok(up(f))) f = up(f);
    return f > 0.0f ? f : 0.0f;
}

// Workgroup -> target-range mapping that keeps each XCD on one contiguous
// eighth of the Morton order (workgroups are dispatched round-robin over the
// 8 XCDs, and each XCD has its own L2).  A bijection on [0, grid).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ int64_t xcd_block(unsigned b, unsigned grid) {
    const unsigned per = grid / kXcds, rem = grid % kXcds;
    const unsigned x = b % kXcds, k = b / kXcds;
    return x < rem ? (int64_t)x * (per + 1) + k : (int64_t)rem * (per + 1) + (int64_t)(x - rem) * per + k;
}

__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }

// A read-only device array seen through the constant address space.  A
// uniform load through it is always a scalar load: the compiler need not
// prove that no store in the kernel reaches it (its proof gives up in large
// unrolled loops that write LDS, and then emits per-lane vector loads).
// Only for data the kernel never writes.
template <typename T>
using const_ptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ const_ptr<T> as_const(const T *p) {
    return (const_ptr<T>)p;
}

__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ float readlane_f32(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Relaxed agent-scope atomics on 32-bit status words (look-back protocols).
__device__ __forceinline__ uint32_t ld_acquire_agent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed_agent(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive scan across one wave64.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        uint32_t t = __shfl_up(v, o, kWave);
        if (l >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Exclusive scan over a block of NT threads (one value each); `tmp` holds
// NT/64 words of LDS.  Every thread of the block must call it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *tmp, uint32_t *total) {
    constexpr int W = NT / kWave;
    const uint32_t inc = wave_incl_scan(v);
    const int w = threadIdx.x / kWave;
    if (lane_id() == kWave - 1) tmp[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        uint32_t t = tmp[i];
        off += (i < w) ? t : 0u;
        tot += t;
    }
    __syncthreads();
    if (total) *total = tot;
    return off + inc - v;
}

// Bit spread of the reference's expandBits10 / expandBits21 magic-constant
// forms (MortonSort.h:27-44).
__device__ __forceinline__ uint32_t expand_bits10(uint32_t i) {
    i = (i * 0x00010001u) & 0xFF0000FFu;
    i = (i * 0x00000101u) & 0x0F00F00Fu;
    i = (i * 0x00000011u) & 0xC30C30C3u;
    i = (i * 0x00000005u) & 0x49249249u;
    return i;
}

__device__ __forceinline__ uint64_t expand_bits21(uint64_t i) {
    i &= 0x1fffffull;
    i = (i | i << 32) & 0x1f00000000ffffull;
    i = (i | i << 16) & 0x1f0000ff0000ffull;
    i = (i | i << 8) & 0x100f00f00f00f00full;
    i = (i | i << 4) & 0x10c30c30c30c30c3ull;
    i = (i | i << 2) & 0x1249249249249249ull;
    return i;
}

}  // namespace nbg
