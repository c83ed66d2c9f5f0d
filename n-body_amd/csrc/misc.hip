// Integrator and direct-summation kernels for gfx950.

#include "device.hpp"
#include "launch.hpp"

namespace nbg {

// Solver::updateVelocities + updatePositions (Simulation/Solver.h:117-139),
// fused: v = v + a*dt, then p = p + v*dt, two IEEE ops each (no FMA; this
// file is compiled with -ffp-contract=off).
__global__ __launch_bounds__(256) void k_update(float *__restrict__ x, float *__restrict__ y, float *__restrict__ z,
                                                float *__restrict__ vx, float *__restrict__ vy,
                                                float *__restrict__ vz, const float *__restrict__ ax,
                                                const float *__restrict__ ay, const float *__restrict__ az,
                                                float dt, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float ux = vx[i] + ax[i] * dt, uy = vy[i] + ay[i] * dt, uz = vz[i] + az[i] * dt;
        vx[i] = ux; vy[i] = uy; vz[i] = uz;
        x[i] = x[i] + ux * dt;
        y[i] = y[i] + uy * dt;
        z[i] = z[i] + uz * dt;
    }
}

void launch_update(float *x, float *y, float *z, float *vx, float *vy, float *vz, const float *ax, const float *ay,
                   const float *az, float dt, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_update, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, vx, vy, vz, ax, ay, az, dt, n);
}

// The same update for a forest: each body takes its system's time step.
__global__ __launch_bounds__(256) void k_update_forest(float *__restrict__ x, float *__restrict__ y,
                                                       float *__restrict__ z, float *__restrict__ vx,
                                                       float *__restrict__ vy, float *__restrict__ vz,
                                                       const float *__restrict__ ax, const float *__restrict__ ay,
                                                       const float *__restrict__ az,
                                                       const uint32_t *__restrict__ bsys,
                                                       const float4 *__restrict__ sys_prm, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float dt = sys_prm[bsys[i]].w;
        const float ux = vx[i] + ax[i] * dt, uy = vy[i] + ay[i] * dt, uz = vz[i] + az[i] * dt;
        vx[i] = ux; vy[i] = uy; vz[i] = uz;
        x[i] = x[i] + ux * dt;
        y[i] = y[i] + uy * dt;
        z[i] = z[i] + uz * dt;
    }
}

void launch_update_forest(float *x, float *y, float *z, float *vx, float *vy, float *vz, const float *ax,
                          const float *ay, const float *az, const uint32_t *bsys, const float4 *sys_prm, int64_t n,
                          hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_update_forest, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, vx, vy, vz, ax, ay, az, bsys,
                       sys_prm, n);
}

// Non-finite guard of a forest: bad[s] becomes nonzero when a body of
// system s holds a NaN or infinite position or velocity.  bad holds one word
// per system and is cleared by the caller.
__global__ __launch_bounds__(256) void k_forest_nonfinite(const float *__restrict__ x, const float *__restrict__ y,
                                                          const float *__restrict__ z, const float *__restrict__ vx,
                                                          const float *__restrict__ vy, const float *__restrict__ vz,
                                                          const uint32_t *__restrict__ bsys, int64_t n,
                                                          uint32_t *__restrict__ bad) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        // x - x is 0 for a finite x and NaN for NaN or +-inf
        const float t = (x[i] - x[i]) + (y[i] - y[i]) + (z[i] - z[i]) + (vx[i] - vx[i]) + (vy[i] - vy[i]) +
                        (vz[i] - vz[i]);
        if (t != 0.0f) bad[bsys[i]] = 1u;
    }
}

void launch_forest_nonfinite(const float *x, const float *y, const float *z, const float *vx, const float *vy,
                             const float *vz, const uint32_t *bsys, int64_t n, uint32_t *bad, hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_forest_nonfinite, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, vx, vy, vz, bsys, n, bad);
}

// Forest compaction after a system leaves: new position i of system s
// takes old position old_first[s] + (i - new_first[s]), for up to 8 arrays.
struct ArrayPtrs {
    float *p[8];
};

__global__ __launch_bounds__(256) void k_forest_compact(ArrayPtrs dst, ArrayPtrs src, int arrays, int64_t n, int nsys,
                                                        const int64_t *__restrict__ new_first,
                                                        const int64_t *__restrict__ old_first) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nsys;  // new_first[lo] <= i < new_first[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (new_first[mid] <= i) lo = mid;
            else hi = mid;
        }
        const int64_t j = old_first[lo] + (i - new_first[lo]);
        for (int k = 0; k < arrays; ++k) dst.p[k][i] = src.p[k][j];
    }
}

void launch_forest_compact(float *const *dst, const float *const *src, int arrays, int64_t n, int nsys,
                           const int64_t *new_first, const int64_t *old_first, hipStream_t s) {
    if (n <= 0 || nsys <= 0 || arrays <= 0) return;
    ArrayPtrs d{}, r{};
    for (int k = 0; k < arrays && k < 8; ++k) {
        d.p[k] = dst[k];
        r.p[k] = const_cast<float *>(src[k]);
    }
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_forest_compact, dim3((unsigned)blocks), dim3(256), 0, s, d, r, arrays < 8 ? arrays : 8, n,
                       nsys, new_first, old_first);
}

// NaiveSolver::updateAccelerations (Solvers/NaiveSolver.h:22-52) in FP32:
// each target sums every source in source order, starting from zero.
// ORDERED uses Gravity::particleParticle's exact arithmetic (bit-exact with
// the reference's sequential sum); FAST uses rsqrt + FMA with G applied
// once at the end.  Sources are staged through LDS 256 at a time.
template <bool ORDERED, bool SOFT>
__global__ __launch_bounds__(256) void k_direct_f32(const float4 *__restrict__ src, int64_t ns,
                                                    const float4 *__restrict__ tgt, int64_t nt,
                                                    const uint32_t *__restrict__ out_index, float G, float eps2,
                                                    float *__restrict__ ax, float *__restrict__ ay,
                                                    float *__restrict__ az) {
    __shared__ float4 s[256];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const float4 p = i < nt ? tgt[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float rx = 0.0f, ry = 0.0f, rz = 0.0f;
    for (int64_t base = 0; base < ns; base += 256) {
        __syncthreads();
        if (base + threadIdx.x < ns) s[threadIdx.x] = src[base + threadIdx.x];
        __syncthreads();
        const int lim = (ns - base) < 256 ? (int)(ns - base) : 256;
        for (int j = 0; j < lim; ++j) {
            const float4 b = s[j];
            if (ORDERED) {
                float vx, vy, vz;
                pp_exact(b.x, b.y, b.z, b.w, p.x, p.y, p.z, G, eps2, vx, vy, vz);
                rx = rx + vx; ry = ry + vy; rz = rz + vz;
            } else {
                const float dx = b.x - p.x, dy = b.y - p.y, dz = b.z - p.z;
                const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (!SOFT) inv = r2 > 0.0f ? inv : 0.0f;
                const float sc = b.w * (inv * (inv * inv));
                rx = __builtin_fmaf(dx, sc, rx);
                ry = __builtin_fmaf(dy, sc, ry);
                rz = __builtin_fmaf(dz, sc, rz);
            }
        }
    }
    if (!ORDERED) { rx *= G; ry *= G; rz *= G; }
    if (i < nt) {
        const int64_t o = out_index ? (int64_t)out_index[i] : i;
        if (o < nt) { ax[o] = rx; ay[o] = ry; az[o] = rz; }
    }
}

void launch_direct_f32(const float4 *src, int64_t ns, const float4 *tgt, int64_t nt, const uint32_t *out_index,
                       float G, float eps2, int ordered, float *ax, float *ay, float *az, hipStream_t s) {
    if (nt <= 0) return;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
    if (ordered)
        hipLaunchKernelGGL((k_direct_f32<true, true>), dim3(blocks), dim3(256), 0, s, src, ns, tgt, nt, out_index, G,
                           eps2, ax, ay, az);
    else if (eps2 > 0.0f)
        hipLaunchKernelGGL((k_direct_f32<false, true>), dim3(blocks), dim3(256), 0, s, src, ns, tgt, nt, out_index,
                           G, eps2, ax, ay, az);
    else
        hipLaunchKernelGGL((k_direct_f32<false, false>), dim3(blocks), dim3(256), 0, s, src, ns, tgt, nt, out_index,
                           G, eps2, ax, ay, az);
}

// FP64 direct sum (NaiveSolver generalised, Solvers/NaiveSolver.h:22-52):
// one thread per target, sources staged through LDS.  Sources at exactly the
// target position are skipped, as Gravity::particleParticle does
// (Physics/Rules/Gravity.h:203).
__global__ __launch_bounds__(256) void k_direct_f64(const float4 *__restrict__ src, int64_t ns,
                                                    const float4 *__restrict__ tgt, int64_t nt, double G,
                                                    double eps2, double *__restrict__ ax, double *__restrict__ ay,
                                                    double *__restrict__ az) {
    __shared__ float4 s[256];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const float4 p = i < nt ? tgt[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int64_t base = 0; base < ns; base += 256) {
        __syncthreads();
        if (base + threadIdx.x < ns) s[threadIdx.x] = src[base + threadIdx.x];
        __syncthreads();
        const int lim = (ns - base) < 256 ? (int)(ns - base) : 256;
        for (int j = 0; j < lim; ++j) {
            const float4 b = s[j];
            if (b.x == p.x && b.y == p.y && b.z == p.z) continue;
            const double rx = (double)p.x - (double)b.x, ry = (double)p.y - (double)b.y,
                         rz = (double)p.z - (double)b.z;
            const double r2 = rx * rx + ry * ry + rz * rz + eps2;
            const double sc = G * (double)b.w / (sqrt(r2) * r2);
            a0 -= rx * sc; a1 -= ry * sc; a2 -= rz * sc;
        }
    }
    if (i < nt) { ax[i] = a0; ay[i] = a1; az[i] = a2; }
}

void launch_direct_f64(const float4 *src, int64_t ns, const float4 *tgt, int64_t nt, double G, double eps2,
                       double *ax, double *ay, double *az, hipStream_t s) {
    if (nt <= 0) return;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
    hipLaunchKernelGGL(k_direct_f64, dim3(blocks), dim3(256), 0, s, src, ns, tgt, nt, G, eps2, ax, ay, az);
}

// FP64 softened potential W = -sum_{i<j} G m_i m_j / sqrt(r^2 + eps2),
// each pair once.  Wave w takes targets i = 64 w + lane and the sources
// j > i in ascending order (wave-uniform, read with scalar loads from an
// FP64 copy of the bodies); 1 / sqrt is the hardware estimate refined by two
// Newton steps in FP64 (relative error near FP64 rounding).  Waves
// [w0, w1) of a chunk write one partial each (the lanes' sums, fixed order),
// so the host's in-order sum is deterministic.
__device__ __forceinline__ double rsqrt_f64(double r2) {
    double y = __builtin_amdgcn_rsq(r2);
    const double h = 0.5 * r2;
    y = y * __builtin_fma(-h, y * y, 1.5);
    y = y * __builtin_fma(-h, y * y, 1.5);
    return y;
}

__global__ __launch_bounds__(256) void k_potential_f64(const double4 *__restrict__ b, int64_t n, double eps2,
                                                       int64_t w0, int64_t w1, double *__restrict__ partial) {
    const int lane = lane_id();
    const int64_t w = w0 + (int64_t)blockIdx.x * 4 + uniform_i32(threadIdx.x / kWave);
    if (w >= w1) return;
    const int64_t base = w * kWave, i = base + lane;
    const bool valid = i < n;
    const double4 p = valid ? b[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    double acc = 0.0;
    // the wave's own tile: j in (i, min(base + 64, n))
    const int64_t tile_end = base + kWave < n ? base + kWave : n;
    for (int64_t j = base; j < tile_end; ++j) {
        const double4 q = b[j];
        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
        const double t = q.w * rsqrt_f64(r2);
        acc += j > i ? t : 0.0;
    }
    int64_t j = tile_end;
    for (; j + 4 <= n; j += 4) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const double4 q = b[j + k];
            const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
            acc = __builtin_fma(q.w, rsqrt_f64(r2), acc);
        }
    }
    for (; j < n; ++j) {
        const double4 q = b[j];
        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
        acc = __builtin_fma(q.w, rsqrt_f64(r2), acc);
    }
    double v = valid ? p.w * acc : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) partial[w] = v;
}

__global__ __launch_bounds__(256) void k_pack_f64(const float *__restrict__ x, const float *__restrict__ y,
                                                  const float *__restrict__ z, const float *__restrict__ m, int64_t n,
                                                  double4 *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_double4(x[i], y[i], z[i], m[i]);
}

__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ x, const float *__restrict__ y,
                                              const float *__restrict__ z, const float *__restrict__ m, int64_t n,
                                              float4 *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = make_float4(x[i], y[i], z[i], m[i]);
}

void launch_pack(const float *x, const float *y, const float *z, const float *m, int64_t n, float4 *out,
                 hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_pack, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, m, n, out);
}

int64_t potential_waves(int64_t n) { return (n + kWave - 1) / kWave; }

void launch_pack_f64(const float *x, const float *y, const float *z, const float *m, int64_t n, double4 *out,
                     hipStream_t s) {
    if (n <= 0) return;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_pack_f64, dim3((unsigned)blocks), dim3(256), 0, s, x, y, z, m, n, out);
}

void launch_potential_f64(const double4 *bodies, int64_t n, double eps2, int64_t w0, int64_t w1, double *partial,
                          hipStream_t s) {
    if (n <= 0 || w1 <= w0) return;
    const unsigned blocks = (unsigned)((w1 - w0 + 3) / 4);
    hipLaunchKernelGGL(k_potential_f64, dim3(blocks), dim3(256), 0, s, bodies, n, eps2, w0, w1, partial);
}

}  // namespace nbg
