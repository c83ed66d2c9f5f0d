// Barnes-Hut traversal for gfx950.
//
// One wavefront walks the tree for 64 spatially adjacent targets (consecutive
// in Morton order).  The wave keeps one DFS stack in LDS whose entries carry
// a 64-bit lane mask: at each node every lane in the mask evaluates its own
// opening criterion, lanes that accept the node take the monopole
// interaction, and the lanes that reject it (a ballot) are the mask pushed
// with the children.  Each lane therefore visits exactly the nodes the
// reference's per-body recursion visits, in the same order
// (Descent/activeTree.h:24-79).  Leaf bodies are loaded once per leaf, one
// per lane (coalesced), and broadcast to the lanes with v_readlane.
//
// ORDERED arithmetic reproduces the reference bit for bit: IEEE sqrt and
// division in Gravity::particleParticle's order (Physics/Rules/Gravity.h:199-209),
// leaf sums sequential from zero (activeTree.h:44-52) and child results
// bracketed as std::transform_reduce does for two children, (0 + r0) + r1
// (activeTree.h:66-76).  Each lane's pending partial sums live in LDS, one
// slot per tree depth (the stack depth is wave-uniform).  FAST arithmetic
// keeps the same opening decisions (the MAC is still evaluated exactly) and
// order, but accumulates with rsqrt and FMA into one register triple.

#include "device.hpp"
#include "launch.hpp"

namespace nbg {

constexpr int kMetaComplete = 0x10000;
// Upper bound on stack pops per wave; a valid tree needs at most ~3 per node.
constexpr uint32_t kMaxWalkIters = 1u << 28;

// Gravity::particleParticle, exact reference arithmetic.
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
    vx = ieee_div((-Rx * bm) * G, r3);
    vy = ieee_div((-Ry * bm) * G, r3);
    vz = ieee_div((-Rz * bm) * G, r3);
}

// Same interaction, rsqrt + FMA, accumulated in place.  Coincident bodies
// contribute exactly zero when eps2 > 0 (d = 0); the guard covers eps2 = 0.
__device__ __forceinline__ void pp_fast(float bx, float by, float bz, float bmG, float px, float py, float pz,
                                        float eps2, float &rx, float &ry, float &rz) {
    const float dx = bx - px, dy = by - py, dz = bz - pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    float inv = __builtin_amdgcn_rsqf(r2);
    inv = r2 > 0.0f ? inv : 0.0f;
    const float s = bmG * (inv * inv * inv);
    rx = __builtin_fmaf(dx, s, rx);
    ry = __builtin_fmaf(dy, s, ry);
    rz = __builtin_fmaf(dz, s, rz);
}

template <bool ORDERED, int MAXD, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_walk_lbvh(WalkArgs a) {
    __shared__ int s_node[WAVES][kStackLBVH];
    __shared__ int s_meta[WAVES][kStackLBVH];
    __shared__ uint64_t s_mask[WAVES][kStackLBVH];
    __shared__ float s_save[ORDERED ? WAVES * MAXD * 3 * kWave : 1];

    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t tgt = ((int64_t)blockIdx.x * WAVES + wv) * kWave + lane;
    const bool valid = tgt < a.n;
    const uint64_t vmask = __ballot(valid);
    if (vmask == 0) return;

    const float4 me = valid ? a.bodies[tgt] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float px = me.x, py = me.y, pz = me.z;
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    uint32_t cpp = 0, cnp = 0;
    int *stn = s_node[wv];
    int *stm = s_meta[wv];
    uint64_t *stk = s_mask[wv];
    float *sv = ORDERED ? &s_save[wv * MAXD * 3 * kWave] : nullptr;
    const float G = a.G, eps2 = a.eps2, theta2 = a.theta2;

    if (lane == 0) { stn[0] = 0; stm[0] = 0; stk[0] = vmask; }
    int sp = 1;
    bool fail = false;
    uint32_t iters = 0;
    while (sp > 0) {
        if (++iters > kMaxWalkIters) { fail = true; break; }
        --sp;
        __builtin_amdgcn_wave_barrier();
        const int node = uniform_i32(stn[sp]);
        const int meta = uniform_i32(stm[sp]);
        const uint64_t mask = uniform_u64(stk[sp]);
        const bool act = (mask >> lane) & 1ull;
        const int depth = meta & 0xFFFF;
        if (ORDERED && (meta & kMetaComplete)) {
            if (act) {
                const float vx = rx, vy = ry, vz = rz;
                rx = sv[(depth * 3 + 0) * kWave + lane] + vx;
                ry = sv[(depth * 3 + 1) * kWave + lane] + vy;
                rz = sv[(depth * 3 + 2) * kWave + lane] + vz;
            }
            continue;
        }
        if (node < 0 || node >= a.max_nodes) { fail = true; break; }
        const float4 mc = a.mac[node];
        const float4 cm = a.com[node];
        const int4 tp = a.topo[node];
        // DiagonalOverCenterDistance (DiagonalOverCenterDistance.h:18-22).
        bool pass = false;
        if (act) {
            const float dx = px - mc.x, dy = py - mc.y, dz = pz - mc.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            pass = ieee_div(mc.w, d2) < theta2;
        }
        if (pass) {
            ++cnp;
            if (ORDERED) {
                float vx, vy, vz;
                pp_exact(cm.x, cm.y, cm.z, cm.w, px, py, pz, G, eps2, vx, vy, vz);
                rx = rx + vx; ry = ry + vy; rz = rz + vz;
            } else {
                pp_fast(cm.x, cm.y, cm.z, cm.w * G, px, py, pz, eps2, rx, ry, rz);
            }
        }
        const uint64_t open = mask & ~__ballot(pass);
        if (open == 0) continue;
        const bool op = (open >> lane) & 1ull;
        if (tp.z < 0) {
            // Leaf: every rejecting lane interacts with every body in order.
            const int first = tp.x, cnt = tp.y;
            if (first < 0 || cnt < 0 || (int64_t)first + cnt > a.n) { fail = true; break; }
            float lx = 0.0f, ly = 0.0f, lz = 0.0f;
            for (int base = 0; base < cnt; base += kWave) {
                const int k = (cnt - base) < kWave ? (cnt - base) : kWave;
                float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (lane < k) b = a.bodies[first + base + lane];
                for (int j = 0; j < k; ++j) {
                    const float bx = readlane_f32(b.x, j), by = readlane_f32(b.y, j);
                    const float bz = readlane_f32(b.z, j), bm = readlane_f32(b.w, j);
                    if (op) {
                        if (ORDERED) {
                            float vx, vy, vz;
                            pp_exact(bx, by, bz, bm, px, py, pz, G, eps2, vx, vy, vz);
                            lx = lx + vx; ly = ly + vy; lz = lz + vz;
                        } else {
                            pp_fast(bx, by, bz, bm * G, px, py, pz, eps2, rx, ry, rz);
                        }
                    }
                }
            }
            if (op) {
                cpp += (uint32_t)cnt;
                if (ORDERED) { rx = rx + lx; ry = ry + ly; rz = rz + lz; }
            }
        } else {
            if (sp + 3 > kStackLBVH || (ORDERED && depth >= MAXD)) { fail = true; break; }
            if (ORDERED) {
                if (op) {
                    sv[(depth * 3 + 0) * kWave + lane] = rx;
                    sv[(depth * 3 + 1) * kWave + lane] = ry;
                    sv[(depth * 3 + 2) * kWave + lane] = rz;
                    rx = 0.0f; ry = 0.0f; rz = 0.0f;
                }
                if (lane == 0) { stn[sp] = node; stm[sp] = depth | kMetaComplete; stk[sp] = open; }
                ++sp;
            }
            if (lane == 0) {
                stn[sp] = tp.z + 1; stm[sp] = depth + 1; stk[sp] = open;
                stn[sp + 1] = tp.z; stm[sp + 1] = depth + 1; stk[sp + 1] = open;
            }
            sp += 2;
        }
    }
    if (fail && lane == 0) atomicOr(a.err, 2u);
    if (valid) {
        const uint32_t o = a.perm[tgt];
        if ((int64_t)o < a.n) { a.ax[o] = rx; a.ay[o] = ry; a.az[o] = rz; }
    }
    const uint32_t wpp = wave_sum(valid ? cpp : 0u), wnp = wave_sum(valid ? cnp : 0u);
    if (lane == 0) {
        atomicAdd(&a.counts[0], (unsigned long long)wpp);
        atomicAdd(&a.counts[1], (unsigned long long)wnp);
    }
}

void launch_walk_lbvh(const WalkArgs &a, int ordered, int max_depth, hipStream_t s) {
    if (a.n <= 0) return;
    if (!ordered) {
        constexpr int W = 4;
        const unsigned blocks = (unsigned)((a.n + 64 * W - 1) / (64 * W));
        hipLaunchKernelGGL((k_walk_lbvh<false, 1, W>), dim3(blocks), dim3(64 * W), 0, s, a);
    } else {
        const unsigned blocks = (unsigned)((a.n + 63) / 64);
        if (max_depth <= 32)
            hipLaunchKernelGGL((k_walk_lbvh<true, 32, 1>), dim3(blocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((k_walk_lbvh<true, 64, 1>), dim3(blocks), dim3(64), 0, s, a);
    }
}

}  // namespace nbg
