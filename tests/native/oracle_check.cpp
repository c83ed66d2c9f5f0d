// CPU check of the oracle's C++ (oracle/oracle.cpp) under AddressSanitizer
// and UndefinedBehaviorSanitizer, built by tests/test_oracle_sanitizers.py
// together with oracle.cpp.  Every tree kind (octree, LBVH, their quadrupole
// variants, the implicit FMM) is built and walked over random, clustered,
// coincident, collinear, single-body and empty inputs at several leaf sizes;
// the topology exports are checked for consistency and, at theta 0, the walk
// against the FP64 direct sum.  Exits nonzero on the first failed expectation
// (a sanitizer report aborts the run by itself).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
void *orc_tree_new(const int *ip, const float *fp, int64_t n, const float *x, const float *y, const float *z,
                   const float *m);
void orc_tree_free(void *h);
int64_t orc_tree_num_nodes(void *h);
int orc_tree_max_depth(void *h);
int orc_tree_nodes(void *h, int64_t *first, int64_t *count, int32_t *leaf, int32_t *depth, float *com4,
                   float *geo6);
int orc_tree_quads(void *h, float *q6);
int orc_tree_parents(void *h, int32_t *parent);
int orc_tree_order(void *h, uint32_t *perm, uint64_t *keys_or_null);
int orc_tree_bbox(void *h, float *out6);
int orc_tree_accel(void *h, int64_t nt, const float *tx, const float *ty, const float *tz, float *ax, float *ay,
                   float *az, uint64_t *pp, uint64_t *np, int threads);
int orc_tree_accel_f64(void *h, int64_t nt, const float *tx, const float *ty, const float *tz, double *ax,
                       double *ay, double *az, uint64_t *pp, uint64_t *np, int threads);
int orc_fmm_accel(void *h, float *ax, float *ay, float *az, uint64_t *counts4, double *mass_or_null, int threads);
int64_t orc_fmm_starts(void *h, int32_t *out, int64_t cap);
int64_t orc_fmm_sample(void *h, int32_t k, int32_t threads);
int orc_direct_f64(int64_t ns, const float *sx, const float *sy, const float *sz, const float *sm, int64_t nt,
                   const float *tx, const float *ty, const float *tz, double G, double eps2, double *ax,
                   double *ay, double *az);
int orc_step(const int *ip, const float *fp, int64_t n, float *x, float *y, float *z, const float *m, float *vx,
             float *vy, float *vz, float *ax, float *ay, float *az, int nsteps, int threads);
}

static int failures = 0;
#define EXPECT(c, ...)                                                           \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "%s:%d: expected %s: ", __FILE__, __LINE__, #c); \
            std::fprintf(stderr, __VA_ARGS__);                                   \
            std::fprintf(stderr, "\n");                                          \
            ++failures;                                                          \
        }                                                                        \
    } while (0)

struct Bodies {
    std::vector<float> x, y, z, m;
    int64_t n() const { return (int64_t)m.size(); }
};

// kind: 0 uniform cube, 1 Plummer-like cluster, 2 tight clusters with
// coincident pairs, 3 all coincident, 4 collinear.
static Bodies make(int kind, int64_t n, uint32_t seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f), um(0.5f, 1.5f);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    Bodies b;
    for (int64_t i = 0; i < n; ++i) {
        float x = 0, y = 0, z = 0;
        if (kind == 0) { x = u(g); y = u(g); z = u(g); }
        if (kind == 1) {
            const float r = 1.0f / std::sqrt(std::pow(0.05f + 0.9f * (u(g) * 0.5f + 0.5f), -2.0f / 3.0f) - 1.0f);
            const float a = nd(g), c = nd(g), d = nd(g), s = std::sqrt(a * a + c * c + d * d) + 1e-12f;
            x = r * a / s; y = r * c / s; z = r * d / s;
        }
        if (kind == 2) {
            const int64_t c = i / 7;
            x = 0.25f * (float)(c % 5); y = 0.5f * (float)(c % 3); z = -0.125f * (float)(c % 4);
            if (i % 7 > 1) { x += 1e-4f * u(g); y += 1e-4f * u(g); z += 1e-4f * u(g); }
        }
        if (kind == 3) { x = 0.375f; y = -1.25f; z = 2.0f; }
        if (kind == 4) { x = u(g); y = 0.5f; z = -0.5f; }
        b.x.push_back(x); b.y.push_back(y); b.z.push_back(z); b.m.push_back(um(g));
    }
    return b;
}

static bool finite3(float a, float b, float c) { return std::isfinite(a) && std::isfinite(b) && std::isfinite(c); }

static void check_tree(int solver, int leaf, const Bodies &b, float theta, float eps2, bool expect_finite) {
    const int ip[4] = {solver, leaf, 16, 63};
    const float fp[4] = {theta, 1.0f, eps2, 1e-3f};
    const int64_t n = b.n();
    void *h = orc_tree_new(ip, fp, n, b.x.data(), b.y.data(), b.z.data(), b.m.data());
    EXPECT(h != nullptr, "solver %d leaf %d n %lld", solver, leaf, (long long)n);
    if (!h) return;
    const int64_t nn = orc_tree_num_nodes(h);
    EXPECT(nn >= 0, "solver %d n %lld nodes %lld", solver, (long long)n, (long long)nn);
    (void)orc_tree_max_depth(h);
    std::vector<int64_t> first(nn), count(nn);
    std::vector<int32_t> isleaf(nn), depth(nn), parent(nn);
    std::vector<float> com4(4 * nn), geo6(6 * nn), q6(6 * nn), bb(6);
    orc_tree_nodes(h, first.data(), count.data(), isleaf.data(), depth.data(), com4.data(), geo6.data());
    orc_tree_quads(h, q6.data());
    orc_tree_parents(h, parent.data());
    orc_tree_bbox(h, bb.data());
    std::vector<uint32_t> perm(n);
    std::vector<uint64_t> keys(n);
    orc_tree_order(h, perm.data(), solver == 1 || solver == 4 ? keys.data() : nullptr);
    // the body order is a permutation; every node's bodies lie inside the
    // array and inside its parent's range; leaves cover every body once
    std::vector<int> seen(n, 0);
    for (int64_t i = 0; i < n; ++i) {
        EXPECT(perm[i] < (uint64_t)n, "perm[%lld] = %u", (long long)i, perm[i]);
        if (perm[i] < (uint64_t)n) ++seen[perm[i]];
    }
    for (int64_t i = 0; i < n; ++i) EXPECT(seen[i] == 1, "body %lld seen %d times", (long long)i, seen[i]);
    std::vector<int> covered(n, 0);
    for (int64_t i = 0; i < nn; ++i) {
        EXPECT(first[i] >= 0 && count[i] >= 0 && first[i] + count[i] <= n, "node %lld [%lld, +%lld)", (long long)i,
               (long long)first[i], (long long)count[i]);
        EXPECT((parent[i] < 0) == (i == 0), "node %lld parent %d", (long long)i, parent[i]);
        if (parent[i] >= 0 && parent[i] < nn) {
            const int64_t p = parent[i];
            EXPECT(first[i] >= first[p] && first[i] + count[i] <= first[p] + count[p], "node %lld outside parent",
                   (long long)i);
            EXPECT(depth[i] == depth[p] + 1, "node %lld depth %d parent depth %d", (long long)i, depth[i], depth[p]);
        }
        if (isleaf[i] && first[i] >= 0 && first[i] + count[i] <= n)
            for (int64_t j = first[i]; j < first[i] + count[i]; ++j) ++covered[j];
    }
    for (int64_t i = 0; i < n; ++i) EXPECT(covered[i] == 1, "sorted body %lld in %d leaves", (long long)i, covered[i]);

    // walk every body and a few targets off the bodies
    std::vector<float> tx(b.x), ty(b.y), tz(b.z);
    for (int k = 0; k < 5; ++k) { tx.push_back(0.1f * k - 0.2f); ty.push_back(3.0f - k); tz.push_back(0.07f * k); }
    const int64_t nt = (int64_t)tx.size();
    std::vector<float> ax(nt, 0.0f), ay(nt, 0.0f), az(nt, 0.0f);
    if (solver == 5) {
        uint64_t c4[4];
        std::vector<double> mass(n);
        EXPECT(orc_fmm_accel(h, ax.data(), ay.data(), az.data(), c4, nullptr, 1) == 0, "fmm n %lld", (long long)n);
        EXPECT(orc_fmm_accel(h, nullptr, nullptr, nullptr, nullptr, mass.data(), 1) == 0, "fmm mass");
        double total = 0;
        for (float mi : b.m) total += mi;
        for (int64_t i = 0; i < n; ++i)
            EXPECT(std::fabs(mass[i] - total) <= 1e-9 * total, "fmm body %lld receives mass %.17g of %.17g",
                   (long long)i, mass[i], total);
        std::vector<int32_t> starts(4096);
        const int64_t ns = orc_fmm_starts(h, starts.data(), (int64_t)starts.size());
        EXPECT(ns >= 0, "fmm starts %lld", (long long)ns);
        (void)orc_fmm_sample(h, 3, 1);
        if (expect_finite)
            for (int64_t i = 0; i < n; ++i) EXPECT(finite3(ax[i], ay[i], az[i]), "fmm body %lld", (long long)i);
    } else {
        std::vector<uint64_t> pp(nt), np(nt);
        orc_tree_accel(h, nt, tx.data(), ty.data(), tz.data(), ax.data(), ay.data(), az.data(), pp.data(),
                       np.data(), 1);
        if (expect_finite)
            for (int64_t i = 0; i < nt; ++i) EXPECT(finite3(ax[i], ay[i], az[i]), "solver %d target %lld", solver,
                                                    (long long)i);
        if (solver == 0 || solver == 1) {
            std::vector<double> dx(nt), dy(nt), dz(nt);
            std::vector<uint64_t> pp2(nt), np2(nt);
            EXPECT(orc_tree_accel_f64(h, nt, tx.data(), ty.data(), tz.data(), dx.data(), dy.data(), dz.data(),
                                      pp2.data(), np2.data(), 1) == 0, "walk64 solver %d", solver);
            for (int64_t i = 0; i < nt; ++i)
                EXPECT(pp[i] == pp2[i] && np[i] == np2[i], "walk64 counts differ at %lld", (long long)i);
            // theta 0 opens every node: the FP64 same-set walk is the direct sum
            if (theta == 0.0f && expect_finite) {
                std::vector<double> ex(nt), ey(nt), ez(nt);
                orc_direct_f64(n, b.x.data(), b.y.data(), b.z.data(), b.m.data(), nt, tx.data(), ty.data(),
                               tz.data(), 1.0, (double)eps2, ex.data(), ey.data(), ez.data());
                for (int64_t i = 0; i < nt; ++i) {
                    const double e = std::sqrt((dx[i] - ex[i]) * (dx[i] - ex[i]) + (dy[i] - ey[i]) * (dy[i] - ey[i]) +
                                               (dz[i] - ez[i]) * (dz[i] - ez[i]));
                    const double s = std::sqrt(ex[i] * ex[i] + ey[i] * ey[i] + ez[i] * ez[i]);
                    EXPECT(e <= 1e-9 * s + 1e-12, "solver %d theta 0 target %lld: |walk64 - direct| %.3g of %.3g",
                           solver, (long long)i, e, s);
                }
            }
        } else {
            double dx, dy, dz;
            EXPECT(orc_tree_accel_f64(h, 1, tx.data(), ty.data(), tz.data(), &dx, &dy, &dz, nullptr, nullptr, 1) ==
                       -1, "walk64 refuses quadrupole trees");
        }
    }
    orc_tree_free(h);
}

int main() {
    const int solvers[] = {0, 1, 3, 4, 5};
    const int leaves[] = {1, 4, 16};
    const int64_t sizes[] = {1, 2, 3, 17, 300, 3000};
    int cases = 0;
    for (int s : solvers)
        for (int leaf : leaves)
            for (int kind = 0; kind < 5; ++kind)
                for (int64_t n : sizes) {
                    if (kind >= 3 && n > 300) continue;  // coincident / collinear: small sizes suffice
                    const Bodies b = make(kind, n, (uint32_t)(1000 * s + 100 * leaf + 10 * kind + n));
                    const bool coincident = kind == 2 || kind == 3;
                    check_tree(s, leaf, b, 0.5f, 1e-4f, true);
                    if (coincident && s != 5) check_tree(s, leaf, b, 0.5f, 0.0f, false);  // r = 0 pairs: inf allowed
                    if (n <= 300 && kind != 3 && s != 5) check_tree(s, leaf, b, 0.0f, 1e-4f, true);
                    cases += 1;
                }
    {  // empty input: the trees hold no node or a root only, and walks return zero
        const Bodies b;
        for (int s : {0, 1, 3, 4}) {
            const int ip[4] = {s, 16, 16, 63};
            const float fp[4] = {0.5f, 1.0f, 1e-4f, 1e-3f};
            void *h = orc_tree_new(ip, fp, 0, b.x.data(), b.y.data(), b.z.data(), b.m.data());
            EXPECT(h != nullptr, "empty solver %d", s);
            if (!h) continue;
            const float t[3] = {0.5f, 0.25f, 0.125f};
            float a[3] = {1, 1, 1};
            orc_tree_accel(h, 1, &t[0], &t[1], &t[2], &a[0], &a[1], &a[2], nullptr, nullptr, 1);
            EXPECT(a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f, "empty solver %d: %g %g %g", s, a[0], a[1], a[2]);
            orc_tree_free(h);
        }
    }
    {  // refused parameters
        const Bodies b = make(0, 10, 7);
        const int bad_bits[4] = {1, 16, 16, 31};
        const int bad_solver[4] = {9, 16, 16, 63};
        const float fp[4] = {0.5f, 1.0f, 1e-4f, 1e-3f};
        EXPECT(orc_tree_new(bad_bits, fp, 10, b.x.data(), b.y.data(), b.z.data(), b.m.data()) == nullptr, "bits 31");
        EXPECT(orc_tree_new(bad_solver, fp, 10, b.x.data(), b.y.data(), b.z.data(), b.m.data()) == nullptr,
               "solver 9");
    }
    {  // the integrator over a few steps of every solver
        for (int s : {0, 1, 3, 4, 5}) {
            Bodies b = make(1, 500, 11 + s);
            std::vector<float> vx(500, 0.01f), vy(500, 0.0f), vz(500, -0.01f), ax(500), ay(500), az(500);
            const int ip[4] = {s, 16, 16, 63};
            const float fp[4] = {0.5f, 1.0f, 1e-4f, 1e-3f};
            EXPECT(orc_step(ip, fp, 500, b.x.data(), b.y.data(), b.z.data(), b.m.data(), vx.data(), vy.data(),
                            vz.data(), ax.data(), ay.data(), az.data(), 3, 1) == 0, "step solver %d", s);
            for (int i = 0; i < 500; ++i) EXPECT(finite3(b.x[i], b.y[i], b.z[i]), "step solver %d body %d", s, i);
        }
    }
    if (failures) {
        std::printf("%d failed expectations over %d cases\n", failures, cases);
        return 1;
    }
    std::printf("all checks passed (%d input cases)\n", cases);
    return 0;
}
