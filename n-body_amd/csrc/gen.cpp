// Synthetic inputs (host CPU): Plummer spheres and uniform cubes, generated
// in FP64 from std::mt19937_64 with 53-bit uniform doubles and rounded to
// FP32.  The reference has no Plummer generator (its Benchmark generators are
// random/Perlin/galaxy, src/Benchmark/Generator.h:18-186); these fix the
// workloads the force pass is measured on.

#include <cmath>
#include <cstdint>
#include <random>

#include "../../include/nbg.h"

namespace {

struct Uniform53 {
    std::mt19937_64 eng;
    explicit Uniform53(uint64_t seed) : eng(seed) {}
    double operator()() { return (double)(eng() >> 11) * 0x1.0p-53; }
};

// Isotropic unit vector: cos(polar) uniform in [-1, 1], azimuth uniform.
void direction(Uniform53 &u, double &x, double &y, double &z) {
    const double c = 2.0 * u() - 1.0;
    const double ph = 2.0 * M_PI * u();
    const double s = std::sqrt(std::fmax(0.0, 1.0 - c * c));
    x = s * std::cos(ph);
    y = s * std::sin(ph);
    z = c;
}

}  // namespace

extern "C" {

// Plummer (1911) sphere sampled as Aarseth, Henon & Wielen (1974):
// r = a / sqrt(u^(-2/3) - 1), resampled while r >= rmax*a; speed q * v_esc
// with q drawn by rejection of 0.1*U against q^2 (1 - q^2)^(7/2) and
// v_esc = sqrt(2 G M / a) (1 + r^2/a^2)^(-1/4).
int nbg_generate_plummer(uint64_t seed, int64_t n, double a, double M, double G, double rmax, float *x, float *y,
                         float *z, float *m, float *vx, float *vy, float *vz) {
    if (n < 0 || !(a > 0.0) || !(M >= 0.0) || !(rmax > 0.0)) return NBG_ERR_ARG;
    if (n > 0 && (!x || !y || !z || !m || !vx || !vy || !vz)) return NBG_ERR_ARG;
    Uniform53 u(seed);
    const double mi = n > 0 ? M / (double)n : 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double r;
        do {
            const double v = u();
            r = a / std::sqrt(std::pow(v, -2.0 / 3.0) - 1.0);
        } while (!(r < rmax * a));
        double dx, dy, dz;
        direction(u, dx, dy, dz);
        double q, g;
        do {
            q = u();
            g = 0.1 * u();
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vesc = std::sqrt(2.0 * G * M / a) * std::pow(1.0 + (r * r) / (a * a), -0.25);
        const double sp = q * vesc;
        double ex, ey, ez;
        direction(u, ex, ey, ez);
        x[i] = (float)(r * dx);
        y[i] = (float)(r * dy);
        z[i] = (float)(r * dz);
        m[i] = (float)mi;
        vx[i] = (float)(sp * ex);
        vy[i] = (float)(sp * ey);
        vz[i] = (float)(sp * ez);
    }
    return NBG_OK;
}

int nbg_generate_cube(uint64_t seed, int64_t n, double side, double M, float *x, float *y, float *z, float *m,
                      float *vx, float *vy, float *vz) {
    if (n < 0 || !(side > 0.0)) return NBG_ERR_ARG;
    if (n > 0 && (!x || !y || !z || !m || !vx || !vy || !vz)) return NBG_ERR_ARG;
    Uniform53 u(seed);
    const double mi = n > 0 ? M / (double)n : 0.0;
    for (int64_t i = 0; i < n; ++i) {
        x[i] = (float)((u() - 0.5) * side);
        y[i] = (float)((u() - 0.5) * side);
        z[i] = (float)((u() - 0.5) * side);
        m[i] = (float)mi;
        vx[i] = 0.0f;
        vy[i] = 0.0f;
        vz[i] = 0.0f;
    }
    return NBG_OK;
}

}  // extern "C"
