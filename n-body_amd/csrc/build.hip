// Tree build kernels for gfx950: bounding box, Morton keys, onesweep radix
// sort, gather, Karras LBVH topology and atomic-flag bottom-up summaries.
//
// Arithmetic that feeds a tree decision is IEEE binary32 in the reference's
// operation order (this file is compiled with -ffp-contract=off), so keys,
// topology and summaries are bit-exact with the reference's CPU path.

#include <cfloat>

#include "device.hpp"
#include "launch.hpp"

namespace nbg {

// ------------------------------------------------------------------ fill
__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}

// ------------------------------------------------------------------ bbox
// outerBoundingBox (Physics/Summaries/SummaryType.h:73-81).  Min/max are
// exact, so a tree-shaped reduction equals the reference's serial loop for
// finite inputs (up to which of +0/-0 survives a tie).  The last block to
// finish reduces the per-block partials.
// glm::min(x, y) = (y < x) ? y : x and glm::max(x, y) = (x < y) ? y : x; the
// argument order matters only for ties between +0 and -0.
__device__ __forceinline__ float fmin_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float fmax_ref(float a, float b) { return (a < b) ? b : a; }

__global__ __launch_bounds__(256) void k_bbox(const float *__restrict__ x, const float *__restrict__ y,
                                              const float *__restrict__ z, int64_t n, float *partials) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        float v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d) { mn[d] = fmin_ref(mn[d], v[d]); mx[d] = fmax_ref(mx[d], v[d]); }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin_ref(mn[d], __shfl_xor(mn[d], o, kWave));
            mx[d] = fmax_ref(mx[d], __shfl_xor(mx[d], o, kWave));
        }
    __shared__ float s_part[4][6];
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int d = 0; d < 3; ++d) { s_part[w][d] = mn[d]; s_part[w][3 + d] = mx[d]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[6];
        for (int d = 0; d < 6; ++d) r[d] = s_part[0][d];
        for (int ww = 1; ww < 4; ++ww)
            for (int d = 0; d < 3; ++d) { r[d] = fmin_ref(r[d], s_part[ww][d]); r[3 + d] = fmax_ref(r[3 + d], s_part[ww][3 + d]); }
        for (int d = 0; d < 6; ++d) partials[blockIdx.x * 6 + d] = r[d];
    }
}

// Second stage: one block folds the per-block partials (the kernel boundary
// makes them visible; no agent-scope fence, which on gfx950 is an L2
// write-back + invalidate).
__global__ __launch_bounds__(256) void k_bbox_final(const float *__restrict__ partials, int blocks, float *out6) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int b = threadIdx.x; b < blocks; b += blockDim.x)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin_ref(mn[d], partials[b * 6 + d]);
            mx[d] = fmax_ref(mx[d], partials[b * 6 + 3 + d]);
        }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin_ref(mn[d], __shfl_xor(mn[d], o, kWave));
            mx[d] = fmax_ref(mx[d], __shfl_xor(mx[d], o, kWave));
        }
    __shared__ float s_part[4][6];
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int d = 0; d < 3; ++d) { s_part[w][d] = mn[d]; s_part[w][3 + d] = mx[d]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[6];
        for (int d = 0; d < 6; ++d) r[d] = s_part[0][d];
        for (int ww = 1; ww < 4; ++ww)
            for (int d = 0; d < 3; ++d) { r[d] = fmin_ref(r[d], s_part[ww][d]); r[3 + d] = fmax_ref(r[3 + d], s_part[ww][3 + d]); }
        for (int d = 0; d < 6; ++d) out6[d] = r[d];
    }
}

void launch_bbox(const float *x, const float *y, const float *z, int64_t n, float *partials, float *out6,
                 hipStream_t s) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > kBBoxBlocks) blocks = kBBoxBlocks;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_bbox, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, n, partials);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, partials, (int)blocks, out6);
}

// ------------------------------------------------------------------ Morton
// toMortonCode / setMortonCodes (Solvers/MortonSort.h:46-73):
// (p - min) / maxSide, * 2^k, glm::clamp to [0, 2^k - 1], truncate,
// interleave as 4x + 2y + z.  NaN (zero-extent box) maps to 0.
__device__ __forceinline__ uint32_t quantize(float v, float mul, float hi) {
    float c = v * mul;
    float t = (c < 0.0f) ? 0.0f : c;      // glm::max(c, 0)
    t = (hi < t) ? hi : t;                // glm::min(t, hi)
    return (t != t) ? 0u : (uint32_t)t;
}

// Morton key of one point over the bounding cube of bb (min xyz, max xyz).
template <typename K>
__device__ __forceinline__ K morton_key(float px, float py, float pz, const float *__restrict__ bb) {
    const float mnx = bb[0], mny = bb[1], mnz = bb[2];
    const float dx = bb[3] - mnx, dy = bb[4] - mny, dz = bb[5] - mnz;
    const float m1 = (dx < dy) ? dy : dx;   // std::max(std::max(dx, dy), dz)
    const float scale = (m1 < dz) ? dz : m1;
    const float vx = ieee_div(px - mnx, scale);
    const float vy = ieee_div(py - mny, scale);
    const float vz = ieee_div(pz - mnz, scale);
    if constexpr (sizeof(K) == 4) {
        uint32_t qx = quantize(vx, 1024.0f, 1023.0f), qy = quantize(vy, 1024.0f, 1023.0f),
                 qz = quantize(vz, 1024.0f, 1023.0f);
        return expand_bits10(qx) * 4u + expand_bits10(qy) * 2u + expand_bits10(qz);
    } else {
        uint64_t qx = quantize(vx, 2097152.0f, 2097151.0f), qy = quantize(vy, 2097152.0f, 2097151.0f),
                 qz = quantize(vz, 2097152.0f, 2097151.0f);
        return expand_bits21(qx) * 4u + expand_bits21(qy) * 2u + expand_bits21(qz);
    }
}

template <typename K>
__global__ __launch_bounds__(256) void k_morton(const float *__restrict__ x, const float *__restrict__ y,
                                                const float *__restrict__ z, int64_t n,
                                                const float *__restrict__ bb, K *__restrict__ keys,
                                                uint32_t *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = morton_key<K>(x[i], y[i], z[i], bb);
        idx[i] = (uint32_t)i;
    }
}

void launch_morton(const float *x, const float *y, const float *z, int64_t n, const float *bbox6, int bits,
                   void *keys, uint32_t *idx, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (bits == 30)
        hipLaunchKernelGGL(k_morton<uint32_t>, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, n, bbox6,
                           (uint32_t *)keys, idx);
    else
        hipLaunchKernelGGL(k_morton<uint64_t>, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, n, bbox6,
                           (uint64_t *)keys, idx);
}

// ------------------------------------------------------------------ sort
// Stable LSD radix sort (8-bit digits), one "onesweep" kernel per pass:
// tiles take their index from an atomic counter, rank their keys in LDS with
// ballot-based digit matching (stable: wave-striped items in input order),
// publish per-digit tile counts, and resolve the global offset by decoupled
// look-back over earlier tiles.  Starting from an iota payload, the result
// is the order by (key, original index).
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPrefix = 2u << 30;
constexpr uint32_t kFlagMask = 3u << 30;
constexpr uint32_t kValueMask = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const K *__restrict__ keys, int64_t n, int passes,
                                                           uint32_t *__restrict__ ghist) {
    __shared__ uint32_t h[8 * kRadix];
    for (int i = threadIdx.x; i < passes * kRadix; i += blockDim.x) h[i] = 0;
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        K k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p * kRadix + (uint32_t)((k >> (8 * p)) & 0xFF)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * kRadix; i += blockDim.x) {
        uint32_t v = h[i];
        if (v) atomicAdd(&ghist[i], v);
    }
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_onesweep(const K *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                         K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                         int64_t n, int shift, const uint32_t *__restrict__ ghist,
                                                         uint32_t *status, uint32_t *tile_counter, uint32_t *err) {
    constexpr int W = kSortBlock / kWave;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wh[W][kRadix];
    __shared__ uint32_t s_local_off[kRadix];
    __shared__ uint32_t s_gadj[kRadix];
    __shared__ uint32_t s_tmp[W];
    __shared__ K s_k[kSortTile];
    __shared__ uint32_t s_v[kSortTile];

    const int tid = threadIdx.x, lane = lane_id(), w = tid / kWave;
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < W * kRadix; i += kSortBlock) (&s_wh[0][0])[i] = 0;
    __syncthreads();
    const int64_t tile = s_tile;
    const int64_t tile_base = tile * kSortTile;
    const int64_t wave_base = tile_base + (int64_t)w * (kWave * kSortItems);

    K key[kSortItems];
    uint32_t val[kSortItems], rnk[kSortItems];
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = wave_base + j * kWave + lane;
        if (i < n) { key[j] = kin[i]; val[j] = vin[i]; } else { key[j] = 0; val[j] = 0; }
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const bool valid = (wave_base + j * kWave + lane) < n;
        const uint32_t d = (uint32_t)(key[j] >> shift) & 0xFFu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        uint32_t base = 0;
        if (valid) base = s_wh[w][d];
        const uint32_t below = __popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) s_wh[w][d] = base + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rnk[j] = base + below;
    }
    __syncthreads();

    // One thread per digit: exclusive prefix across waves, tile total.
    const uint32_t d = tid;
    uint32_t tot = 0;
#pragma unroll
    for (int ww = 0; ww < W; ++ww) {
        uint32_t c = s_wh[ww][d];
        s_wh[ww][d] = tot;
        tot += c;
    }
    uint32_t *st = status + tile * kRadix;
    st_relaxed_agent(st + d, (tile == 0 ? kFlagPrefix : kFlagAgg) | tot);
    const uint32_t local_off = block_excl_scan<kSortBlock>(tot, s_tmp, nullptr);
    const uint32_t gbase = block_excl_scan<kSortBlock>(ghist[d], s_tmp, nullptr);

    uint32_t excl = 0;
    if (tile > 0) {
        int64_t p = tile - 1;
        uint32_t spins = 0;
        while (p >= 0) {
            const uint32_t sv = ld_acquire_agent(status + p * kRadix + d);
            const uint32_t f = sv & kFlagMask;
            if (f == 0) {
                if (++spins > kSpinLimit) { atomicOr(err, 1u); break; }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += sv & kValueMask;
            if (f == kFlagPrefix) break;
            --p;
        }
        st_relaxed_agent(st + d, kFlagPrefix | (excl + tot));
    }
    s_local_off[d] = local_off;
    s_gadj[d] = gbase + excl - local_off;
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        const bool valid = (wave_base + j * kWave + lane) < n;
        if (valid) {
            const uint32_t dj = (uint32_t)(key[j] >> shift) & 0xFFu;
            const uint32_t pos = s_local_off[dj] + s_wh[w][dj] + rnk[j];
            s_k[pos] = key[j];
            s_v[pos] = val[j];
        }
    }
    __syncthreads();
    const int64_t rem = n - tile_base;
    const int ntile = rem < kSortTile ? (int)rem : kSortTile;
    for (int i = tid; i < ntile; i += kSortBlock) {
        const K k = s_k[i];
        const uint32_t dj = (uint32_t)(k >> shift) & 0xFFu;
        const uint32_t g = s_gadj[dj] + (uint32_t)i;
        if ((int64_t)g < n) {  // always true unless a look-back timed out
            kout[g] = k;
            vout[g] = s_v[i];
        } else {
            atomicOr(err, 1u);
        }
    }
}

template <typename K>
static void radix_sort_impl(K *keys, uint32_t *idx, K *keys_alt, uint32_t *idx_alt, int64_t n, int passes,
                            uint32_t *status, uint32_t *err, hipStream_t s) {
    const int64_t tiles = sort_tiles(n);
    uint32_t *ghist = status;                       // passes * 256
    uint32_t *counters = status + 8 * kRadix;       // passes
    uint32_t *tstat = status + 8 * kRadix + 64;     // passes * tiles * 256
    launch_fill_u32(status, 0u, 8 * kRadix + 64 + passes * tiles * kRadix, s);
    int64_t hb = tiles;
    if (hb > 2048) hb = 2048;
    hipLaunchKernelGGL(k_radix_hist<K>, dim3((unsigned)hb), dim3(kSortBlock), 0, s, keys, n, passes, ghist);
    K *ka = keys, *kb = keys_alt;
    uint32_t *va = idx, *vb = idx_alt;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_onesweep<K>, dim3((unsigned)tiles), dim3(kSortBlock), 0, s, ka, va, kb, vb, n, 8 * p,
                           ghist + p * kRadix, tstat + (int64_t)p * tiles * kRadix, counters + p, err);
        K *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
}

void launch_radix_sort(void *keys, uint32_t *idx, void *keys_alt, uint32_t *idx_alt, int64_t n, int bits,
                       uint32_t *status, uint32_t *err, hipStream_t s) {
    if (n <= 1) return;
    if (bits == 30)
        radix_sort_impl<uint32_t>((uint32_t *)keys, idx, (uint32_t *)keys_alt, idx_alt, n, 4, status, err, s);
    else
        radix_sort_impl<uint64_t>((uint64_t *)keys, idx, (uint64_t *)keys_alt, idx_alt, n, 8, status, err, s);
}

bool launch_radix_sort_u64(uint64_t *keys, uint32_t *idx, uint64_t *keys_alt, uint32_t *idx_alt, int64_t n,
                           int passes, uint32_t *status, uint32_t *err, hipStream_t s) {
    if (n <= 1 || passes <= 0) return false;
    radix_sort_impl<uint64_t>(keys, idx, keys_alt, idx_alt, n, passes, status, err, s);
    return (passes & 1) != 0;
}

// ------------------------------------------------------------------ gather
__global__ __launch_bounds__(256) void k_gather(const float *__restrict__ x, const float *__restrict__ y,
                                                const float *__restrict__ z, const float *__restrict__ m,
                                                const uint32_t *__restrict__ perm, int64_t n,
                                                float4 *__restrict__ bodies) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t j = perm[i];
        bodies[i] = (int64_t)j < n ? make_float4(x[j], y[j], z[j], m[j]) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

void launch_gather(const float *x, const float *y, const float *z, const float *m, const uint32_t *perm,
                   int64_t n, float4 *bodies, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, m, perm, n, bodies);
}

// ------------------------------------------------------------------ LBVH topology
// Karras (2012) radix tree over the sorted keys with index-augmented key
// comparison.  For a range whose end keys differ, the Karras split is the
// first index whose key has the highest differing bit set, i.e. exactly
// LinearBVHNode::split (Trees/LinearBVH.h:47-76).  The reference stops
// splitting at size <= leaf or all-equal keys (LinearBVH.h:114-119); since a
// splittable node's ancestors are splittable too, the reference tree is the
// root plus the children of the splittable Karras nodes.
template <typename K>
__device__ __forceinline__ int kdelta(const K *__restrict__ keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const K a = keys[i], b = keys[j];
    if (a == b) return (int)(8 * sizeof(K)) + __clz((uint32_t)(i ^ j));
    if constexpr (sizeof(K) == 8) return __clzll((unsigned long long)(a ^ b));
    else return __clz((uint32_t)(a ^ b));
}

// Forest (sys_shift > 0): keys carry the system index above bit sys_shift; a
// node whose range spans systems always splits, and a node inside system s
// uses that system's leaf size leaf_tab[s].
template <typename K>
__global__ __launch_bounds__(256) void k_karras(const K *__restrict__ keys, int n, int leaf, int32_t *kfirst,
                                                int32_t *klast, int32_t *kgamma, uint32_t *ksplit,
                                                int32_t *kparent, int sys_shift, const int32_t *__restrict__ leaf_tab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = kdelta(keys, n, i, i - d);
    int lmax = 2;
    while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = kdelta(keys, n, i, j);
    int sp = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (kdelta(keys, n, i, i + (sp + t) * d) > dnode) sp += t;
    } while (t > 1);
    const int gamma = i + sp * d + (d < 0 ? -1 : 0);
    const int first = d > 0 ? i : j, last = d > 0 ? j : i;
    kfirst[i] = first;
    klast[i] = last;
    kgamma[i] = gamma;
    int L = leaf;
    bool span = false;
    if (sys_shift > 0) {
        const uint64_t s0 = (uint64_t)keys[first] >> sys_shift, s1 = (uint64_t)keys[last] >> sys_shift;
        span = s0 != s1;
        L = leaf_tab[s0];
    }
    ksplit[i] = ((span || last - first + 1 > L) && keys[first] != keys[last]) ? 1u : 0u;
    if (gamma != first) kparent[gamma] = 2 * i;
    if (gamma + 1 != last) kparent[gamma + 1] = 2 * i + 1;
}

// Three-phase exclusive scan of u32 flags (tile sums, scan of sums, tiles).
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(const uint32_t *__restrict__ in, int64_t n,
                                                          uint32_t *__restrict__ sums) {
    __shared__ uint32_t s_tmp[kScanBlock / kWave];
    const int64_t base = blockIdx.x * (int64_t)kScanTile;
    uint32_t acc = 0;
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + (int64_t)j * kScanBlock + threadIdx.x;
        if (i < n) acc += in[i];
    }
    uint32_t tot;
    block_excl_scan<kScanBlock>(acc, s_tmp, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_top(uint32_t *sums, int64_t nt, uint32_t *total) {
    __shared__ uint32_t s_tmp[kScanBlock / kWave];
    uint32_t carry = 0;
    for (int64_t b = 0; b < nt; b += kScanBlock) {
        int64_t i = b + threadIdx.x;
        uint32_t v = i < nt ? sums[i] : 0u;
        uint32_t tot;
        uint32_t e = block_excl_scan<kScanBlock>(v, s_tmp, &tot);
        if (i < nt) sums[i] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

// in may equal out: every thread reads its items before it writes them.
__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t *in, int64_t n,
                                                           const uint32_t *__restrict__ sums, uint32_t *out) {
    __shared__ uint32_t s_tmp[kScanBlock / kWave];
    const int64_t base = blockIdx.x * (int64_t)kScanTile;
    // Blocked arrangement: thread t owns items [t*I, t*I + I).
    uint32_t v[kScanItems], acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + (int64_t)threadIdx.x * kScanItems + j;
        v[j] = i < n ? in[i] : 0u;
        acc += v[j];
    }
    uint32_t run = sums[blockIdx.x] + block_excl_scan<kScanBlock>(acc, s_tmp, nullptr);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        int64_t i = base + (int64_t)threadIdx.x * kScanItems + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
}

void launch_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *ws, uint32_t *total, hipStream_t s) {
    uint32_t *tot = total ? total : ws;
    if (n <= 0) {
        launch_fill_u32(tot, 0u, 1, s);
        return;
    }
    uint32_t *sums = ws + 64;
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nt), dim3(kScanBlock), 0, s, in, n, sums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, s, sums, nt, tot);
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanBlock), 0, s, in, n, sums, out);
}

template <typename K>
__global__ __launch_bounds__(256) void k_emit(const K *__restrict__ keys, int n, const int32_t *__restrict__ kfirst,
                                              const int32_t *__restrict__ klast, const int32_t *__restrict__ kgamma,
                                              const uint32_t *__restrict__ ksplit, const uint32_t *__restrict__ krank,
                                              const int32_t *__restrict__ kparent, NodeRec *__restrict__ nodes,
                                              int32_t *__restrict__ leaf_start) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1 || !ksplit[i]) return;
    const int r = (int)krank[i];
    const int first = kfirst[i], last = klast[i], gamma = kgamma[i];
    int my_slot = 0;
    if (i != 0) {
        const int pk = kparent[i];
        my_slot = 1 + 2 * (int)krank[pk >> 1] + (pk & 1);
    } else {
        nodes[0].topo = make_int4(first, last - first + 1, 1 + 2 * r, -1);
    }
    const int kl = gamma, kr = gamma + 1;
    const bool l_int = (gamma != first), r_int = (gamma + 1 != last);
    const int lptr = (l_int && ksplit[kl]) ? 1 + 2 * (int)krank[kl] : -1;
    const int rptr = (r_int && ksplit[kr]) ? 1 + 2 * (int)krank[kr] : -1;
    nodes[1 + 2 * r].topo = make_int4(first, gamma - first + 1, lptr, my_slot);
    nodes[2 + 2 * r].topo = make_int4(gamma + 1, last - gamma, rptr, my_slot);
    if (lptr < 0) leaf_start[first] = 1 + 2 * r;
    if (rptr < 0) leaf_start[gamma + 1] = 2 + 2 * r;
    (void)keys;
}

template <typename K>
__global__ void k_root_fix(int n, const uint32_t *__restrict__ ksplit, const uint32_t *__restrict__ total,
                           NodeRec *__restrict__ nodes, int32_t *__restrict__ leaf_start,
                           uint32_t *__restrict__ node_count) {
    if (n <= 1 || !ksplit[0]) {
        nodes[0].topo = make_int4(0, n, -1, -1);
        if (n > 0) leaf_start[0] = 0;
        *node_count = n > 0 ? 1u : 0u;
    } else {
        *node_count = 1u + 2u * *total;
    }
}

void launch_lbvh_topology(const void *keys, int bits, int64_t n64, int leaf_size, int32_t *kfirst, int32_t *klast,
                          int32_t *kgamma, uint32_t *ksplit, uint32_t *krank, int32_t *kparent, uint32_t *scan_ws,
                          NodeRec *nodes, int32_t *leaf_start, uint32_t *node_count, hipStream_t s, int sys_shift,
                          const int32_t *leaf_tab) {
    const int n = (int)n64;
    uint32_t *total = scan_ws;
    uint32_t *sums = scan_ws + 64;
    launch_fill_u32((uint32_t *)leaf_start, 0xFFFFFFFFu, n, s);
    if (n >= 2) {
        const int m = n - 1;
        const unsigned blocks = (unsigned)((m + 255) / 256);
        if (bits == 30)
            hipLaunchKernelGGL(k_karras<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t *)keys, n, leaf_size,
                               kfirst, klast, kgamma, ksplit, kparent, 0, (const int32_t *)nullptr);
        else
            hipLaunchKernelGGL(k_karras<uint64_t>, dim3(blocks), dim3(256), 0, s, (const uint64_t *)keys, n, leaf_size,
                               kfirst, klast, kgamma, ksplit, kparent, sys_shift, leaf_tab);
        const int64_t nt = (m + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)nt), dim3(kScanBlock), 0, s, ksplit, (int64_t)m, sums);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, s, sums, nt, total);
        hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(kScanBlock), 0, s, ksplit, (int64_t)m, sums, krank);
        if (bits == 30)
            hipLaunchKernelGGL(k_emit<uint32_t>, dim3(blocks), dim3(256), 0, s, (const uint32_t *)keys, n, kfirst, klast,
                               kgamma, ksplit, krank, kparent, nodes, leaf_start);
        else
            hipLaunchKernelGGL(k_emit<uint64_t>, dim3(blocks), dim3(256), 0, s, (const uint64_t *)keys, n, kfirst, klast,
                               kgamma, ksplit, krank, kparent, nodes, leaf_start);
    }
    hipLaunchKernelGGL(k_root_fix<uint32_t>, dim3(1), dim3(1), 0, s, n, ksplit, total, nodes, leaf_start, node_count);
}

// ------------------------------------------------------------------ summaries
// NodeBase::summarize (Trees/NodeBase.h:43-55) bottom-up.  A leaf takes
// CenterOfMassSummary's sequential sums over its bodies in sort order
// (Summaries/CenterOfMassSummary.h:30-42) and its bounding box
// (Summaries/BoundingBoxSummary.h:30-57); an internal node combines its two
// children (CenterOfMassSummary.h:44-55, BoundingBoxSummary.h:59-68).  The
// combine reads only finished children, so the result does not depend on
// which child finishes first.  The MAC record is (bbox centre (max+min)/2,
// |max-min|^2) (Physics/BoundingBox.h:47,55).
//
// Scheduling: each workgroup owns kSummBodies consecutive bodies.  Its
// threads summarise the leaves starting in that range and climb with
// per-node arrival counters (the second arrival combines) for as long as
// the parent's range stays inside the workgroup's range, synchronising with
// workgroup-scope acquire/release only.  Parents that straddle two ranges
// are queued and finished by a single workgroup in a second launch.  No
// agent-scope fence is used: on gfx950 that is an L2 write-back and
// invalidate across the XCDs.
// The MAC record: bbox centre and the squared-distance threshold above which
// a target accepts the node (mac_lbvh_threshold of the squared diagonal).
// Per-node choice of the MAC parameters: one set, or per system of a forest
// (the system index is the key prefix of the node's first body).
struct MacSel {
    MacParam one;
    const MacParam *tab;
    const uint64_t *keys;
    int shift;
    __device__ __forceinline__ MacParam at(int first) const { return tab ? tab[keys[first] >> shift] : one; }
    // Forest: a node whose range spans systems needs no summary (every walk
    // starts at a system's root), so climbs stop below it.
    __device__ __forceinline__ bool spans(int4 t) const {
        return tab && (keys[t.x] >> shift) != (keys[t.x + t.y - 1] >> shift);
    }
};

__device__ __forceinline__ float4 mac_record(float4 mn, float4 mx, MacParam mp) {
    const float cx = (mx.x + mn.x) / 2.0f, cy = (mx.y + mn.y) / 2.0f, cz = (mx.z + mn.z) / 2.0f;
    const float dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    const float diag2 = dx * dx + dy * dy + dz * dz;
    return make_float4(cx, cy, cz, mac_lbvh_threshold(diag2, mp.mid, mp.tie != 0));
}

__device__ __forceinline__ uint32_t arrive_wg(uint32_t *flag) {
    return __hip_atomic_fetch_add(flag, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// MultipoleMassSummary::summarize(childNodes) (MultipoleMassSummary.h:54-65)
// for two children: (0 + t0) + t1 with t = outer(c_child - c, c_child) m_child
// + Q_child.
__device__ __forceinline__ void quad_pair(int p, int c0, const NodeRec *nodes, float4 *quad) {
    const float4 c = nodes[p].com;
    float Q[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float4 a = nodes[c0 + k].com;
        const float4 q0 = quad[2 * (c0 + k)], q1 = quad[2 * (c0 + k) + 1];
        const float cq[6] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y};
        float t[6];
        quad_term(a.x - c.x, a.y - c.y, a.z - c.z, a.w, t);
#pragma unroll
        for (int e = 0; e < 6; ++e) Q[e] = Q[e] + (t[e] + cq[e]);
    }
    quad_store(quad, p, Q);
}

__device__ void summarize_pair(int p, int c0, NodeRec *nodes, float4 *bmin, float4 *bmax, MacParam mp) {
    const float4 a = nodes[c0].com;
    const float4 b = nodes[c0 + 1].com;
    const float4 an = bmin[c0], bn = bmin[c0 + 1];
    const float4 ax = bmax[c0], bx = bmax[c0 + 1];
    float Mp = 0.0f;
    Mp += a.w;
    float px = 0.0f + a.x * a.w, py = 0.0f + a.y * a.w, pz = 0.0f + a.z * a.w;
    Mp += b.w;
    px = px + b.x * b.w;
    py = py + b.y * b.w;
    pz = pz + b.z * b.w;
    nodes[p].com = make_float4(ieee_div(px, Mp), ieee_div(py, Mp), ieee_div(pz, Mp), Mp);
    float4 pn = make_float4(fmin_ref(an.x, FLT_MAX), fmin_ref(an.y, FLT_MAX), fmin_ref(an.z, FLT_MAX), 0.0f);
    float4 px4 = make_float4(fmax_ref(ax.x, -FLT_MAX), fmax_ref(ax.y, -FLT_MAX), fmax_ref(ax.z, -FLT_MAX), 0.0f);
    pn = make_float4(fmin_ref(bn.x, pn.x), fmin_ref(bn.y, pn.y), fmin_ref(bn.z, pn.z), 0.0f);
    px4 = make_float4(fmax_ref(bx.x, px4.x), fmax_ref(bx.y, px4.y), fmax_ref(bx.z, px4.z), 0.0f);
    bmin[p] = pn;
    bmax[p] = px4;
    nodes[p].mac = mac_record(pn, px4, mp);
}

__device__ void summarize_leaf(int s, int4 t, const float4 *__restrict__ bodies, NodeRec *nodes, float4 *bmin,
                               float4 *bmax, MacParam mp) {
    float M = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    float nx = FLT_MAX, ny = FLT_MAX, nz = FLT_MAX, xx = -FLT_MAX, xy = -FLT_MAX, xz = -FLT_MAX;
    for (int i = t.x; i < t.x + t.y; ++i) {
        const float4 b = bodies[i];
        M += b.w;
        cx = cx + b.x * b.w;
        cy = cy + b.y * b.w;
        cz = cz + b.z * b.w;
    }
    for (int i = t.x; i < t.x + t.y; ++i) {
        const float4 b = bodies[i];
        nx = fmin_ref(b.x, nx); ny = fmin_ref(b.y, ny); nz = fmin_ref(b.z, nz);
        xx = fmax_ref(b.x, xx); xy = fmax_ref(b.y, xy); xz = fmax_ref(b.z, xz);
    }
    float4 mn = make_float4(nx, ny, nz, 0.0f), mx = make_float4(xx, xy, xz, 0.0f);
    nodes[s].com = make_float4(ieee_div(cx, M), ieee_div(cy, M), ieee_div(cz, M), M);
    bmin[s] = mn;
    bmax[s] = mx;
    nodes[s].mac = mac_record(mn, mx, mp);
}

// Leaf and pair summaries plus, for a quadrupole tree (quad != null), the
// second moments.
__device__ __forceinline__ void summarize_leaf_q(int s, int4 t, const float4 *__restrict__ bodies, NodeRec *nodes,
                                                 float4 *bmin, float4 *bmax, MacParam mp, float4 *quad) {
    summarize_leaf(s, t, bodies, nodes, bmin, bmax, mp);
    if (quad) {
        const float4 c = nodes[s].com;
        float Q[6];
        quad_bodies(bodies, t.x, t.y, c.x, c.y, c.z, Q);
        quad_store(quad, s, Q);
    }
}

__device__ __forceinline__ void summarize_pair_q(int p, int c0, NodeRec *nodes, float4 *bmin, float4 *bmax,
                                                 MacParam mp, float4 *quad) {
    summarize_pair(p, c0, nodes, bmin, bmax, mp);
    if (quad) quad_pair(p, c0, nodes, quad);
}

__global__ __launch_bounds__(kSummThreads) void k_summ_local(const float4 *__restrict__ bodies, int64_t n,
                                                    NodeRec *nodes,
                                                    const int32_t *__restrict__ leaf_start, int64_t max_nodes,
                                                    uint32_t *flags, uint32_t *glist, float4 *bmin,
                                                    float4 *bmax, MacSel ms, float4 *quad, uint32_t *err) {
    const int64_t lo = (int64_t)blockIdx.x * kSummBodies, hi = lo + kSummBodies;
    const int64_t end = hi < n ? hi : n;
    for (int64_t i = lo + threadIdx.x; i < end; i += blockDim.x) {
        const int s = leaf_start[i];
        if (s < 0) continue;
        if (s >= max_nodes) { atomicOr(err, 4u); continue; }
        const int4 t = nodes[s].topo;
        if (t.x != (int)i || t.y < 1 || (int64_t)t.x + t.y > n || t.z >= 0) { atomicOr(err, 4u); continue; }
        summarize_leaf_q(s, t, bodies, nodes, bmin, bmax, ms.at(t.x), quad);
        int parent = t.w;
        while (parent >= 0) {
            if (parent >= max_nodes) { atomicOr(err, 4u); break; }
            const int4 tp = nodes[parent].topo;
            if (tp.x < 0 || tp.y < 1 || (int64_t)tp.x + tp.y > n) { atomicOr(err, 4u); break; }
            if (ms.spans(tp)) break;
            const bool local = tp.x >= lo && (int64_t)tp.x + tp.y <= hi;
            if (!local) {
                const uint32_t k = atomicAdd(&glist[0], 1u);
                if (k < 2 * n) glist[64 + k] = (uint32_t)parent;
                else atomicOr(err, 4u);
                break;
            }
            if (arrive_wg(&flags[parent]) == 0) break;
            if (tp.z < 0 || tp.z + 1 >= max_nodes) { atomicOr(err, 4u); break; }
            summarize_pair_q(parent, tp.z, nodes, bmin, bmax, ms.at(tp.x), quad);
            parent = tp.w;
        }
    }
}

__global__ __launch_bounds__(1024) void k_summ_global(NodeRec *nodes, int64_t n, int64_t max_nodes,
                                                      uint32_t *flags, const uint32_t *glist, int64_t qcap,
                                                      float4 *bmin, float4 *bmax, MacSel ms, float4 *quad,
                                                      uint32_t *err) {
    const uint32_t cnt = glist[0] < qcap ? glist[0] : (uint32_t)qcap;
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        int parent = (int)glist[64 + j];
        while (parent >= 0) {
            if (parent >= max_nodes) { atomicOr(err, 4u); break; }
            const int4 tp = nodes[parent].topo;
            if (tp.x < 0 || tp.y < 1 || (int64_t)tp.x + tp.y > n) { atomicOr(err, 4u); break; }
            if (ms.spans(tp)) break;
            if (arrive_wg(&flags[parent]) == 0) break;
            if (tp.z < 0 || tp.z + 1 >= max_nodes) { atomicOr(err, 4u); break; }
            summarize_pair_q(parent, tp.z, nodes, bmin, bmax, ms.at(tp.x), quad);
            parent = tp.w;
        }
    }
}

// Middle pass: workgroup g owns the kSummMidBodies bodies from g *
// kSummMidBodies and the queued parents whose first body lies there.  It
// climbs from them while the node stays inside its range (all arrivals at
// such a node come from this workgroup: its children lie in the same range),
// and forwards a parent that leaves the range to the output queue in place of
// the arrival, exactly as the local pass does.  Each input entry forwards at
// most once, so the output queue never holds more entries than the input.
__global__ __launch_bounds__(1024) void k_summ_mid(NodeRec *nodes, int64_t n, int64_t max_nodes, uint32_t *flags,
                                                   const uint32_t *gin, uint32_t *gout, int64_t qcap, float4 *bmin,
                                                   float4 *bmax, MacSel ms, float4 *quad, uint32_t *err) {
    const int64_t lo = (int64_t)blockIdx.x * kSummMidBodies, hi = lo + kSummMidBodies;
    const uint32_t cnt = gin[0] < qcap ? gin[0] : (uint32_t)qcap;
    for (uint32_t j = threadIdx.x; j < cnt; j += blockDim.x) {
        int parent = (int)gin[64 + j];
        if (parent < 0 || parent >= max_nodes) { atomicOr(err, 4u); continue; }
        const int first = nodes[parent].topo.x;
        if (first < lo || first >= hi) continue;  // another workgroup's entry
        while (parent >= 0) {
            if (parent >= max_nodes) { atomicOr(err, 4u); break; }
            const int4 tp = nodes[parent].topo;
            if (tp.x < 0 || tp.y < 1 || (int64_t)tp.x + tp.y > n) { atomicOr(err, 4u); break; }
            if (ms.spans(tp)) break;
            if (!(tp.x >= lo && (int64_t)tp.x + tp.y <= hi)) {
                const uint32_t k = atomicAdd(&gout[0], 1u);
                if (k < qcap) gout[64 + k] = (uint32_t)parent;
                else atomicOr(err, 4u);
                break;
            }
            if (arrive_wg(&flags[parent]) == 0) break;
            if (tp.z < 0 || tp.z + 1 >= max_nodes) { atomicOr(err, 4u); break; }
            summarize_pair_q(parent, tp.z, nodes, bmin, bmax, ms.at(tp.x), quad);
            parent = tp.w;
        }
    }
}

void launch_lbvh_summaries(const float4 *bodies, int64_t n, NodeRec *nodes, const int32_t *leaf_start,
                           int64_t max_nodes, uint32_t *flags, uint32_t *glist, float4 *bmin, float4 *bmax,
                           float theta, uint32_t *err, hipStream_t s, const MacParam *mp_tab,
                           const uint64_t *keys, int sys_shift, float4 *quad) {
    if (n <= 0) return;
    MacSel ms;
    mac_threshold(theta, &ms.one.mid, &ms.one.tie);
    ms.tab = mp_tab;
    ms.keys = keys;
    ms.shift = sys_shift;
    // queue layout: local -> middle at glist, middle -> global at glist + qcap + 64
    const int64_t qcap = 2 * n;
    uint32_t *gmid = glist + qcap + 64;
    launch_fill_u32(flags, 0u, max_nodes, s);
    launch_fill_u32(glist, 0u, 64, s);
    launch_fill_u32(gmid, 0u, 64, s);
    const unsigned blocks = (unsigned)((n + kSummBodies - 1) / kSummBodies);
    hipLaunchKernelGGL(k_summ_local, dim3(blocks), dim3(kSummThreads), 0, s, bodies, n, nodes, leaf_start, max_nodes, flags,
                       glist, bmin, bmax, ms, quad, err);
    // the middle pass pays only when it has enough workgroups (a 48M-body
    // forest: -6 % step time; a 1M-body system: +25 us), so small inputs go
    // from the local pass to the single-workgroup pass directly
    const unsigned mblocks = (unsigned)((n + kSummMidBodies - 1) / kSummMidBodies);
    const bool mid = mblocks >= kSummMidMinBlocks;
    if (mid)
        hipLaunchKernelGGL(k_summ_mid, dim3(mblocks), dim3(1024), 0, s, nodes, n, max_nodes, flags, glist, gmid, qcap,
                           bmin, bmax, ms, quad, err);
    hipLaunchKernelGGL(k_summ_global, dim3(1), dim3(1024), 0, s, nodes, n, max_nodes, flags, mid ? gmid : glist, qcap,
                       bmin, bmax, ms, quad, err);
}

// ------------------------------------------------------------------ forest
// Independent systems s = 0..S-1 own the ranges [sys_first[s], sys_first[s+1])
// of the concatenated state.  Each gets its own bounding box and 30-bit Morton
// codes over its own bounding cube; the system index sits above the code
// (key = s << 30 | code), so one sort orders every system and, with nodes
// that span systems always split, the radix tree over the concatenation holds
// each system's reference tree as the subtree whose range is that system.

__global__ __launch_bounds__(256) void k_seg_bbox(const float *__restrict__ x, const float *__restrict__ y,
                                                  const float *__restrict__ z, const int64_t *__restrict__ sys_first,
                                                  float *bbox6) {
    const int sys = blockIdx.x;
    const int64_t lo = sys_first[sys], hi = sys_first[sys + 1];
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        float v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d) { mn[d] = fmin_ref(mn[d], v[d]); mx[d] = fmax_ref(mx[d], v[d]); }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin_ref(mn[d], __shfl_xor(mn[d], o, kWave));
            mx[d] = fmax_ref(mx[d], __shfl_xor(mx[d], o, kWave));
        }
    __shared__ float s_part[4][6];
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0)
        for (int d = 0; d < 3; ++d) { s_part[w][d] = mn[d]; s_part[w][3 + d] = mx[d]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[6];
        for (int d = 0; d < 6; ++d) r[d] = s_part[0][d];
        for (int ww = 1; ww < 4; ++ww)
            for (int d = 0; d < 3; ++d) { r[d] = fmin_ref(r[d], s_part[ww][d]); r[3 + d] = fmax_ref(r[3 + d], s_part[ww][3 + d]); }
        for (int d = 0; d < 6; ++d) bbox6[sys * 6 + d] = r[d];
    }
}

__global__ __launch_bounds__(256) void k_forest_keys(const float *__restrict__ x, const float *__restrict__ y,
                                                     const float *__restrict__ z, int64_t n,
                                                     const uint32_t *__restrict__ bsys,
                                                     const float *__restrict__ bbox6, uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ idx) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t sys = bsys[i];
        keys[i] = ((uint64_t)sys << 30) | (uint64_t)morton_key<uint32_t>(x[i], y[i], z[i], bbox6 + 6 * sys);
        idx[i] = (uint32_t)i;
    }
}

// A node is the root of system s iff its range lies inside s and its
// parent's range does not.
__global__ __launch_bounds__(256) void k_forest_roots(const NodeRec *__restrict__ nodes,
                                                      const uint32_t *__restrict__ node_count, int64_t max_nodes,
                                                      const uint64_t *__restrict__ keys, int64_t n, int nsys,
                                                      int32_t *sys_root, uint32_t *err) {
    const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (slot >= *node_count || slot >= max_nodes) return;
    const int4 t = nodes[slot].topo;
    if (t.x < 0 || t.y < 1 || (int64_t)t.x + t.y > n) { atomicOr(err, 4u); return; }
    const uint64_t s0 = keys[t.x] >> 30, s1 = keys[t.x + t.y - 1] >> 30;
    if (s0 != s1) return;
    if (t.w >= 0) {
        if ((int64_t)t.w >= max_nodes) { atomicOr(err, 4u); return; }
        const int4 tp = nodes[t.w].topo;
        if (tp.x < 0 || tp.y < 1 || (int64_t)tp.x + tp.y > n) { atomicOr(err, 4u); return; }
        if ((keys[tp.x] >> 30) == (keys[tp.x + tp.y - 1] >> 30)) return;
    }
    if (s0 < (uint64_t)nsys) sys_root[s0] = (int32_t)slot;
}

void launch_forest_build(const float *x, const float *y, const float *z, int64_t n, int nsys,
                         const int64_t *sys_first, const uint32_t *bsys, float *bbox6, uint64_t *keys, uint32_t *idx,
                         uint64_t *keys_alt, uint32_t *idx_alt, uint32_t *sort_status, uint32_t *err, hipStream_t s) {
    if (n <= 0 || nsys <= 0) return;
    hipLaunchKernelGGL(k_seg_bbox, dim3((unsigned)nsys), dim3(256), 0, s, x, y, z, sys_first, bbox6);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_forest_keys, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, n, bsys, bbox6, keys, idx);
    int sbits = 0;
    while ((1ll << sbits) < nsys) ++sbits;
    int passes = (30 + sbits + 7) / 8;
    if (passes & 1) ++passes;  // an even pass count leaves the result in keys / idx
    launch_radix_sort_u64(keys, idx, keys_alt, idx_alt, n, passes, sort_status, err, s);
}

void launch_seg_bbox(const float *x, const float *y, const float *z, int nsys, const int64_t *sys_first,
                     float *bbox6, hipStream_t s) {
    if (nsys <= 0) return;
    hipLaunchKernelGGL(k_seg_bbox, dim3((unsigned)nsys), dim3(256), 0, s, x, y, z, sys_first, bbox6);
}

void launch_forest_roots(const NodeRec *nodes, const uint32_t *node_count, int64_t max_nodes, const uint64_t *keys,
                         int64_t n, int nsys, int32_t *sys_root, uint32_t *err, hipStream_t s) {
    if (n <= 0 || nsys <= 0) return;
    launch_fill_u32((uint32_t *)sys_root, 0xFFFFFFFFu, nsys, s);
    hipLaunchKernelGGL(k_forest_roots, dim3((unsigned)((max_nodes + 255) / 256)), dim3(256), 0, s, nodes, node_count,
                       max_nodes, keys, n, nsys, sys_root, err);
}

}  // namespace nbg
