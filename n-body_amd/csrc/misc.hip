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

// FP64 softened potential W = -1/2 sum_i sum_{j != i} G m_i m_j / sqrt(r^2 + eps2).
__global__ __launch_bounds__(256) void k_potential_f64(const float4 *__restrict__ b, int64_t n, double G,
                                                       double eps2, double *__restrict__ partial) {
    __shared__ float4 s[256];
    __shared__ double s_red[256];
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const float4 p = i < n ? b[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    double w = 0.0;
    for (int64_t base = 0; base < n; base += 256) {
        __syncthreads();
        if (base + threadIdx.x < n) s[threadIdx.x] = b[base + threadIdx.x];
        __syncthreads();
        const int lim = (n - base) < 256 ? (int)(n - base) : 256;
        for (int j = 0; j < lim; ++j) {
            if (base + j == i) continue;
            const float4 q = s[j];
            const double rx = (double)p.x - q.x, ry = (double)p.y - q.y, rz = (double)p.z - q.z;
            w += (double)q.w / sqrt(rx * rx + ry * ry + rz * rz + eps2);
        }
    }
    s_red[threadIdx.x] = (i < n) ? -0.5 * G * (double)p.w * w : 0.0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s_red[threadIdx.x] += s_red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s_red[0];
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

int64_t potential_blocks(int64_t n) { return (n + 255) / 256; }

void launch_potential_f64(const float4 *bodies, int64_t n, double G, double eps2, double *partial, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_potential_f64, dim3((unsigned)potential_blocks(n)), dim3(256), 0, s, bodies, n, G, eps2,
                       partial);
}

}  // namespace nbg
