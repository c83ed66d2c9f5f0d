// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement, written from scratch, of the reference simulator's
// tree-code force pass (JacksonCampolattaro/n-body).  It exists to check the
// HIP implementation: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// Every function cites the reference file:line whose semantics it restates
// (paths relative to the reference's src/NBody/).  Arithmetic follows the
// reference's FP32 operation order exactly (glm scalar paths, no FMA: build
// with -ffp-contract=off, no -ffast-math), so results are bit-comparable.
//
// Where the reference is not deterministic the oracle fixes a rule:
//   * sort ties (MortonSort.h:75-88,141-183 use std::partition + spreadsort,
//     both unstable): ties are broken by original body index (stable sort);
//   * octree partition (Octree.h:33-61 uses std::partition): stable, so the
//     bodies of a leaf appear in ascending original index.
// These two orders are the only freedoms; everything else is the reference's.
// Parity unpinned: these two orders (only the reference binary could pin them,
// and it is itself nondeterministic there); they move summation order inside
// COM and leaf sums, never the set of interactions.
//
// Parity pins (see tests/test_oracle.py): reference KATs gravity.cpp:13-26,
// centerOfMassSummary.cpp:12-47, boundingBox.cpp:12-32; closed forms for
// Morton interleave, bbox, COM, integrator; brute force for sort/topology;
// the octree child-centre recurrence and cell containment; an independent
// float32 restatement of the COM summation order (both trees); truth tables
// of the implicit FMM's two criteria; theta->0 walk == direct sum over leaf
// order; FP64 direct-sum force error; the FP64 same-set walk (walk64) equals
// the FP64 direct sum at theta 0 and has the FP32 walk's counts.  tests/test_oracle_mutations.py plants
// a mistake in the octree, COM and criterion functions and checks that the
// pins catch it.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Test switch: a build with -DORC_MUTATE=k (k = 1..7) plants one deliberate
// mistake, beside the line it changes, so that the tests can show that
// their pins catch it (tests/test_oracle_mutations.py); the default build
// plants none.
#ifndef ORC_MUTATE
#define ORC_MUTATE 0
#endif

namespace {

struct V3 {
    float x, y, z;
};

inline V3 v_add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 v_sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 v_mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 v_div(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline bool v_eq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
// glm::dot on vec3: tmp = a*b; tmp.x + tmp.y + tmp.z (left to right).
inline float v_len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
const V3 kZero{0.0f, 0.0f, 0.0f};

struct Params {
    int solver;       // 0 = Barnes-Hut octree, 1 = linear BVH
    float theta;
    int leaf_size;
    int max_depth;    // octree refine depth below the 2 forced levels
    int morton_bits;  // 30 or 63
    float G;
    float eps2;       // the reference's Gravity "epsilon" (added to r^2)
    float dt;
    int order = 1;    // 1: monopole; 2: quadrupole (MultipoleMassSummary<2>)
};

struct Counts {
    uint64_t pp = 0, np = 0;
};

// Gravity::particleParticle (Physics/Rules/Gravity.h:199-209) with
// pow<3>(r) = r*(r*(r*1)) (Gravity.h:35-41).
inline V3 particle_particle(V3 a, float m, V3 p, float G, float eps2) {
    if (v_eq(a, p)) return kZero;
    V3 R = v_sub(p, a);
    float r = std::sqrt(v_len2(R) + eps2);
    float r3 = r * (r * r);
    V3 nR{-R.x, -R.y, -R.z};
    return v_div(v_mul(v_mul(nR, m), G), r3);
}

// FP64 evaluation of one monopole interaction (the softened Newtonian
// acceleration of Gravity.h:199-209 without FP32 rounding): the exact value
// of a term of the reference's interaction set, for measuring the FP32
// rounding of a walk (walk64 below).
struct D3 {
    double x, y, z;
};

inline void pp64(V3 a, float m, V3 p, double G, double eps2, D3 &acc) {
    if (v_eq(a, p)) return;
    const double Rx = (double)p.x - a.x, Ry = (double)p.y - a.y, Rz = (double)p.z - a.z;
    const double r = std::sqrt(Rx * Rx + Ry * Ry + Rz * Rz + eps2);
    const double s = -(double)m * G / (r * r * r);
    acc.x += Rx * s; acc.y += Ry * s; acc.z += Rz * s;
}

// ---------------------------------------------------------------- quadrupole
// MultipoleMoment<2> (Physics/MultipoleMoment.h:20-55): the raw second
// moment sum m (x - c)(x - c)^T about the centre of mass c, stored as the
// 6 entries xx, xy, xz, yy, yz, zz (not made traceless: enforceTraceless is
// commented out in Summaries/MultipoleMassSummary.h:50,89).
struct Quad {
    float q[6] = {0, 0, 0, 0, 0, 0};
};

inline Quad quad_of(V3 o, float m) {  // outerProduct(o, o) * mass
    Quad r;
    r.q[0] = (o.x * o.x) * m; r.q[1] = (o.x * o.y) * m; r.q[2] = (o.x * o.z) * m;
    r.q[3] = (o.y * o.y) * m; r.q[4] = (o.y * o.z) * m; r.q[5] = (o.z * o.z) * m;
    return r;
}

inline Quad quad_add(const Quad &a, const Quad &b) {
    Quad r;
    for (int k = 0; k < 6; ++k) r.q[k] = a.q[k] + b.q[k];
    return r;
}

// libstdc++ std::transform_reduce grouping over a sequence of terms:
// init + ((t0 + t1) + (t2 + t3)) per block of four, then the rest one by one.
template <typename F>
Quad quad_reduce(int64_t count, F term) {
    Quad init;
    int64_t i = 0;
    for (; count - i >= 4; i += 4)
        init = quad_add(init, quad_add(quad_add(term(i), term(i + 1)), quad_add(term(i + 2), term(i + 3))));
    for (; i < count; ++i) init = quad_add(init, term(i));
    return init;
}

// Gravity::nodeParticle<2> (Physics/Rules/Gravity.h:238-259) in the
// reference's FP32 operation order:
//   R = p - c, r = sqrt(R.R + eps2), dPhi = (R * -M) / (r (r r)),
//   dPhi += (D<3>(R, r) . Q) / 2, result G * dPhi.
// D<3> (Gravity.h:43-93) holds the 10 unique entries of the third derivative
// of 1/r, A + B with A = g3 * R(x)R(x)R (each entry R_a (R_b R_c), a right
// fold, SymmetricTensor3.h:106-113,471-476) and B = g2 * (R_a on the entries
// (a,b,b) and (b,b,a) with one odd index a, 3 R_a on (a,a,a), 0 on (x,y,z));
// f2 = 3 / r^3, f3 = -15 / r^4, g2 = f2 / r^2, g3 = f3 / r^3 with
// pow<k>(r) = r (r (... r)).  The contraction (SymmetricTensor3.h:266-276)
// runs over all 27 index triples (i, j, k) in lexicographic order and adds
// D_ijk Q_jk into component i, starting from zero.
inline V3 node_particle_quad(V3 c, float M, const Quad &Q, V3 p, float G, float eps2) {
    if (v_eq(c, p)) return kZero;
    const V3 R = v_sub(p, c);
    const float r = std::sqrt(v_len2(R) + eps2);
    V3 dphi = v_div(v_mul(R, -M), r * (r * r));
    const float f2 = 3.0f / (r * (r * r)), f3 = -15.0f / (r * (r * (r * r)));
    const float g2 = f2 / (r * r), g3 = f3 / (r * (r * r));
    const float Rv[3] = {R.x, R.y, R.z};
    // unique entries in the reference's storage order: xxx xxy xxz xyy xyz
    // xzz yyy yyz yzz zzz
    static const int kDims[10][3] = {{0, 0, 0}, {0, 0, 1}, {0, 0, 2}, {0, 1, 1}, {0, 1, 2},
                                     {0, 2, 2}, {1, 1, 1}, {1, 1, 2}, {1, 2, 2}, {2, 2, 2}};
    float D[10];
    for (int u = 0; u < 10; ++u) {
        const int a = kDims[u][0], b = kDims[u][1], d = kDims[u][2];
        const float cp = Rv[a] * (Rv[b] * Rv[d]);
        float B;
        if (a == b && b == d) B = 0.0f + Rv[a] * 3.0f;
        else if (a != b && b != d) B = 0.0f + 0.0f * 3.0f;
        else B = (a == b ? Rv[d] : Rv[a]) + 0.0f * 3.0f;
        D[u] = cp * g3 + B * g2;
    }
    // storage index of D_ijk (the sorted triple's place in kDims) and of Q_jk
    static constexpr int kU3[3][3][3] = {{{0, 1, 2}, {1, 3, 4}, {2, 4, 5}},
                                         {{1, 3, 4}, {3, 6, 7}, {4, 7, 8}},
                                         {{2, 4, 5}, {4, 7, 8}, {5, 8, 9}}};
    static constexpr int kU2[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    float h[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) h[i] += D[kU3[i][j][k]] * Q.q[kU2[j][k]];
    dphi = v_add(dphi, V3{h[0] / 2.0f, h[1] / 2.0f, h[2] / 2.0f});
    return V3{G * dphi.x, G * dphi.y, G * dphi.z};
}

// ---------------------------------------------------------------- bbox
// outerBoundingBox (Physics/Summaries/SummaryType.h:73-81), BoundingBox
// defaults +-FLT_MAX (Physics/BoundingBox.h:27-37).
void bbox_of(int64_t n, const float *x, const float *y, const float *z, float out[6]) {
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int64_t i = 0; i < n; ++i) {
        // glm::min(cur, p) = (p < cur) ? p : cur; glm::max(cur, p) = (cur < p) ? p : cur
        mn[0] = std::min(mn[0], x[i]); mx[0] = std::max(mx[0], x[i]);
        mn[1] = std::min(mn[1], y[i]); mx[1] = std::max(mx[1], y[i]);
        mn[2] = std::min(mn[2], z[i]); mx[2] = std::max(mx[2], z[i]);
    }
    for (int d = 0; d < 3; ++d) { out[d] = mn[d]; out[3 + d] = mx[d]; }
}

// ---------------------------------------------------------------- Morton
// expandBits10 / expandBits21 / toMortonCode / setMortonCodes
// (Simulation/Solvers/MortonSort.h:27-73).  Written as a per-bit spread,
// which the tests check against the reference's magic-constant form.
uint64_t spread_bits(uint64_t v, int nbits) {
    uint64_t out = 0;
    for (int b = 0; b < nbits; ++b) out |= ((v >> b) & 1ull) << (3 * b);
    return out;
}

uint64_t morton_of(V3 p, const float bb[6], float scale, int bits) {
    // (position - min) / scale, then * 2^k, clamp to [0, 2^k - 1], truncate.
    V3 v = v_div(v_sub(p, V3{bb[0], bb[1], bb[2]}), scale);
    const int per = bits == 30 ? 10 : 21;
    const float mul = bits == 30 ? 1024.0f : 2097152.0f;
    const float hi = bits == 30 ? 1023.0f : 2097151.0f;
    float c[3] = {v.x * mul, v.y * mul, v.z * mul};
    uint64_t q[3];
    for (int d = 0; d < 3; ++d) {
        // glm::clamp(x, lo, hi) = min(max(x, lo), hi)
        float t = std::min(std::max(c[d], 0.0f), hi);
        // NaN (a zero-extent box) converts to 0 here; the reference's cast of
        // NaN is undefined behaviour, so this is the oracle's defined choice.
        q[d] = (t != t) ? 0 : (bits == 30 ? (uint64_t)(uint32_t)t : (uint64_t)t);
    }
    return spread_bits(q[0], per) * 4u + spread_bits(q[1], per) * 2u + spread_bits(q[2], per);
}

float morton_scale(const float bb[6]) {
    float dx = bb[3] - bb[0], dy = bb[4] - bb[1], dz = bb[5] - bb[2];
    return std::max(std::max(dx, dy), dz);
}

// ---------------------------------------------------------------- LBVH
struct LNode {
    int64_t first, count;
    int32_t child0 = -1, child1 = -1;
    int32_t depth = 0;
    bool leaf = true;
    V3 com{0, 0, 0};
    float mass = 0;
    V3 bmin{FLT_MAX, FLT_MAX, FLT_MAX}, bmax{-FLT_MAX, -FLT_MAX, -FLT_MAX};
    Quad quad;
};

struct LBVH {
    Params prm;
    int64_t n = 0;
    std::vector<uint64_t> keys;      // sorted keys
    std::vector<uint32_t> perm;      // sorted position -> original index
    std::vector<V3> pos;             // sorted positions
    std::vector<float> mass;         // sorted masses
    std::vector<LNode> nodes;        // pre-order
    float bbox[6];
    int max_depth = 0;

    // CenterOfMassSummary::summarize(entities) (CenterOfMassSummary.h:30-42)
    // + BoundingBoxSummary leaf loop (BoundingBoxSummary.h:30-57).
    void summarize_leaf(LNode &nd) {
        float M = 0.0f;
        V3 C = kZero;
        V3 mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
#if ORC_MUTATE == 2
        for (int64_t i = nd.first + nd.count - 1; i >= nd.first; --i) {
#else
        for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
#endif
            float m = mass[i];
            M += m;
            C = v_add(C, v_mul(pos[i], m));
        }
        C = v_div(C, M);
        for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
            V3 p = pos[i];
            mn = {std::min(p.x, mn.x), std::min(p.y, mn.y), std::min(p.z, mn.z)};
            mx = {std::max(p.x, mx.x), std::max(p.y, mx.y), std::max(p.z, mx.z)};
        }
        nd.com = C; nd.mass = M; nd.bmin = mn; nd.bmax = mx;
        // MultipoleMassSummary::summarize(entities) (MultipoleMassSummary.h:39-51)
        if (prm.order == 2)
            nd.quad = quad_reduce(nd.count, [&](int64_t k) {
                return quad_of(v_sub(pos[nd.first + k], C), mass[nd.first + k]);
            });
    }

    // CenterOfMassSummary::summarize(childNodes) (CenterOfMassSummary.h:44-55)
    // + BoundingBoxSummary child loop (BoundingBoxSummary.h:59-68).
    void summarize_internal(LNode &nd) {
        const LNode &a = nodes[nd.child0], &b = nodes[nd.child1];
        float M = 0.0f;
        V3 C = kZero;
        M += a.mass; C = v_add(C, v_mul(a.com, a.mass));
        M += b.mass; C = v_add(C, v_mul(b.com, b.mass));
        nd.com = v_div(C, M); nd.mass = M;
        V3 mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (const LNode *c : {&a, &b}) {
            mn = {std::min(c->bmin.x, mn.x), std::min(c->bmin.y, mn.y), std::min(c->bmin.z, mn.z)};
            mx = {std::max(c->bmax.x, mx.x), std::max(c->bmax.y, mx.y), std::max(c->bmax.z, mx.z)};
        }
        nd.bmin = mn; nd.bmax = mx;
        // MultipoleMassSummary::summarize(childNodes) (MultipoleMassSummary.h:54-65)
        if (prm.order == 2) {
            const LNode *ch[2] = {&a, &b};
            nd.quad = quad_reduce(2, [&](int64_t k) {
                return quad_add(quad_of(v_sub(ch[k]->com, nd.com), ch[k]->mass), ch[k]->quad);
            });
        }
    }

    // LinearBVH::build split criterion (Trees/LinearBVH.h:114-119),
    // LinearBVHNode::split (LinearBVH.h:47-76), NodeBase::refine
    // (Trees/NodeBase.h:57-75).  Nodes are appended in pre-order.
    int32_t build(int64_t first, int64_t count, int depth) {
        int32_t id = (int32_t)nodes.size();
        nodes.push_back(LNode{});
        nodes[id].first = first; nodes[id].count = count; nodes[id].depth = depth;
        max_depth = std::max(max_depth, depth);
        const int64_t last = first + count - 1;
        bool split = count > prm.leaf_size && keys[first] != keys[last];
        if (!split) {
            nodes[id].leaf = true;
            summarize_leaf(nodes[id]);
            return id;
        }
        uint64_t x = keys[first] ^ keys[last];
        int msb = 63 - __builtin_clzll(x);
        uint64_t boundary = 1ull << msb;
        // std::upper_bound with comp(v, e) = (v & key(e)) != 0: first index
        // whose key has the boundary bit set.
        int64_t lo = first, hi = last + 1;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if ((boundary & keys[mid]) != 0) hi = mid; else lo = mid + 1;
        }
        int64_t s = lo;
        nodes[id].leaf = false;
        int32_t c0 = build(first, s - first, depth + 1);
        int32_t c1 = build(s, last + 1 - s, depth + 1);
        nodes[id].child0 = c0; nodes[id].child1 = c1;
        summarize_internal(nodes[id]);
        return id;
    }

    void construct(const Params &p, int64_t n_, const float *x, const float *y, const float *z,
                   const float *m) {
        prm = p; n = n_;
        bbox_of(n, x, y, z, bbox);
        float scale = morton_scale(bbox);
        std::vector<uint64_t> k(n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) k[i] = morton_of(V3{x[i], y[i], z[i]}, bbox, scale, prm.morton_bits);
        perm.resize(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
        keys.resize(n); pos.resize(n); mass.resize(n);
        for (int64_t i = 0; i < n; ++i) {
            keys[i] = k[perm[i]];
            pos[i] = V3{x[perm[i]], y[perm[i]], z[perm[i]]};
            mass[i] = m[perm[i]];
        }
        nodes.clear();
        if (n > 0) build(0, n, 0);
    }

    // DiagonalOverCenterDistance (Descent/DescentCriterion/DiagonalOverCenterDistance.h:18-22):
    // diagonalLength2 / distance2(center, p) < theta^2, centre = (max+min)/2
    // (Physics/BoundingBox.h:47,55).
    bool mac(const LNode &nd, V3 p, float theta2) const {
        V3 c = v_div(v_add(nd.bmax, nd.bmin), 2.0f);
        float diag2 = v_len2(v_sub(nd.bmax, nd.bmin));
        float d2 = v_len2(v_sub(p, c));
        return (diag2 / d2) < theta2;
    }

    // Descent::activeTree (Descent/activeTree.h:24-79); children are combined
    // by std::transform_reduce, which for 2 children is (0 + r0) + r1.
    V3 walk(int32_t id, V3 p, float theta2, Counts &cnt) const {
        const LNode &nd = nodes[id];
        if (nd.count == 0) return kZero;
        if (mac(nd, p, theta2)) {
            cnt.np++;
            if (prm.order == 2) return node_particle_quad(nd.com, nd.mass, nd.quad, p, prm.G, prm.eps2);
            return particle_particle(nd.com, nd.mass, p, prm.G, prm.eps2);
        }
        if (nd.leaf) {
            V3 acc = kZero;
            for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
                cnt.pp++;
                acc = v_add(acc, particle_particle(pos[i], mass[i], p, prm.G, prm.eps2));
            }
            return acc;
        }
        V3 r0 = walk(nd.child0, p, theta2, cnt);
        V3 r1 = walk(nd.child1, p, theta2, cnt);
        return v_add(v_add(kZero, r0), r1);
    }

    // The same opening decisions as walk() (FP32 MAC), each monopole
    // interaction evaluated and summed in FP64.
    void walk64(int32_t id, V3 p, float theta2, D3 &acc, Counts &cnt) const {
        const LNode &nd = nodes[id];
        if (nd.count == 0) return;
        if (mac(nd, p, theta2)) {
            cnt.np++;
            pp64(nd.com, nd.mass, p, prm.G, prm.eps2, acc);
            return;
        }
        if (nd.leaf) {
            for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
                cnt.pp++;
                pp64(pos[i], mass[i], p, prm.G, prm.eps2, acc);
            }
            return;
        }
        walk64(nd.child0, p, theta2, acc, cnt);
        walk64(nd.child1, p, theta2, acc, cnt);
    }
};

// ---------------------------------------------------------------- Octree
struct ONode {
    int64_t first = 0, count = 0;
    int32_t child[8];
    int32_t depth = 0;
    bool leaf = true;
    bool summarized = false;
    V3 center{0, 0, 0};
    float side = 1.0f;
    V3 com{0, 0, 0};
    float mass = 0;
    Quad quad;
};

struct Octree {
    Params prm;
    int64_t n = 0;
    std::vector<uint32_t> idx;   // node contents, permuted in place
    std::vector<V3> P;           // original-order positions
    std::vector<float> Mv;       // original-order masses
    std::vector<ONode> nodes;
    float bbox[6];
    int max_depth = 0;

    // OctreeNode::split + partition<2> (Trees/Octree.h:33-61,100-147):
    // partition on z, then y, then x with p[d] < c[d]; the children array is
    // x-major (index 4x + 2y + z); child centres c +- (side/2)/2.
    void split(int32_t id) {
        ONode nd = nodes[id];
        auto beg = idx.begin() + nd.first, end = beg + nd.count;
        V3 c = nd.center;
        auto less = [&](int d) {
            return [&, d](uint32_t e) {
                float v = d == 0 ? P[e].x : (d == 1 ? P[e].y : P[e].z);
                float cv = d == 0 ? c.x : (d == 1 ? c.y : c.z);
                return v < cv;
            };
        };
        auto mz = std::stable_partition(beg, end, less(2));
        auto my0 = std::stable_partition(beg, mz, less(1));
        auto my1 = std::stable_partition(mz, end, less(1));
        auto mx00 = std::stable_partition(beg, my0, less(0));
        auto mx01 = std::stable_partition(my0, mz, less(0));
        auto mx10 = std::stable_partition(mz, my1, less(0));
        auto mx11 = std::stable_partition(my1, end, less(0));
        // spans in memory order (z,y,x): zyx=000,001(x),010(y),011,100(z),...
        struct Sp { std::vector<uint32_t>::iterator b, e; };
        Sp s_zyx[8] = {{beg, mx00}, {mx00, my0}, {my0, mx01}, {mx01, mz},
                       {mz, mx10}, {mx10, my1}, {my1, mx11}, {mx11, end}};
        float cs = nd.side / 2.0f;
#if ORC_MUTATE == 1
        float off = cs;
#else
        float off = cs / 2.0f;
#endif
        int32_t base = (int32_t)nodes.size();
        for (int k = 0; k < 8; ++k) {
            int xb = (k >> 2) & 1, yb = (k >> 1) & 1, zb = k & 1;   // child index 4x+2y+z
            Sp sp = s_zyx[zb * 4 + yb * 2 + xb];
            ONode ch;
            ch.first = sp.b - idx.begin();
            ch.count = sp.e - sp.b;
            ch.depth = nd.depth + 1;
            ch.center = v_add(c, V3{xb ? off : -off, yb ? off : -off, zb ? off : -off});
            ch.side = cs;
            nodes.push_back(ch);
        }
        nodes[id].leaf = false;
        for (int k = 0; k < 8; ++k) nodes[id].child[k] = base + k;
        max_depth = std::max(max_depth, nd.depth + 1);
    }

    // CenterOfMassSummary leaf / child loops (CenterOfMassSummary.h:30-55).
    void summarize(int32_t id) {
        ONode &nd = nodes[id];
        float M = 0.0f;
        V3 C = kZero;
        if (nd.leaf) {
            for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
                uint32_t e = idx[i];
                M += Mv[e];
                C = v_add(C, v_mul(P[e], Mv[e]));
            }
        } else {
#if ORC_MUTATE == 7
            for (int k = 0; k < 7; ++k) {
#else
            for (int k = 0; k < 8; ++k) {
#endif
                const ONode &ch = nodes[nd.child[k]];
                M += ch.mass;
                C = v_add(C, v_mul(ch.com, ch.mass));
            }
        }
        nodes[id].com = v_div(C, M);
        nodes[id].mass = M;
        nodes[id].summarized = true;
        // MultipoleMassSummary (MultipoleMassSummary.h:39-65); an empty child
        // keeps the default summary (zero COM, mass and moment).
        if (prm.order == 2) {
            const V3 Cn = nodes[id].com;
            const ONode &me = nodes[id];
            if (me.leaf)
                nodes[id].quad = quad_reduce(me.count, [&](int64_t k) {
                    const uint32_t e = idx[me.first + k];
                    return quad_of(v_sub(P[e], Cn), Mv[e]);
                });
            else
                nodes[id].quad = quad_reduce(8, [&](int64_t k) {
                    const ONode &ch = nodes[me.child[k]];
                    return quad_add(quad_of(v_sub(ch.com, Cn), ch.mass), ch.quad);
                });
        }
    }

    // NodeBase::refine (Trees/NodeBase.h:57-75): empty nodes return without
    // a summary (they keep the default zero COM / zero mass).
    void refine(int32_t id, int depth_left) {
        if (nodes[id].count == 0) return;
        if (depth_left > 0 && nodes[id].count > prm.leaf_size) {
            split(id);
            for (int k = 0; k < 8; ++k) refine(nodes[id].child[k], depth_left - 1);
        } else {
            nodes[id].leaf = true;
        }
        summarize(id);
    }

    // Octree::build (Trees/Octree.h:185-205): root centre min + diag/2,
    // side = maxSide * 1.1; depthSplit(2) (Trees/Tree.h:172-194) splits the
    // root and all 8 level-1 nodes unconditionally; level-2 nodes refine with
    // maxDepth; summarizeTreeTop (Tree.h:196-216) summarises levels 1 and 0.
    void construct(const Params &p, int64_t n_, const float *x, const float *y, const float *z,
                   const float *m) {
        prm = p; n = n_;
        P.resize(n); Mv.assign(m, m + n);
        for (int64_t i = 0; i < n; ++i) P[i] = V3{x[i], y[i], z[i]};
        bbox_of(n, x, y, z, bbox);
        idx.resize(n);
        std::iota(idx.begin(), idx.end(), 0u);
        nodes.clear();
        ONode root;
        root.first = 0; root.count = n;
        V3 mn{bbox[0], bbox[1], bbox[2]}, mx{bbox[3], bbox[4], bbox[5]};
        V3 diag = v_sub(mx, mn);
        root.center = v_add(mn, v_div(diag, 2.0f));
        root.side = std::max({diag.x, diag.y, diag.z}) * 1.1f;
        nodes.push_back(root);
        max_depth = 0;
        split(0);
        for (int k = 0; k < 8; ++k) split(nodes[0].child[k]);
        for (int k = 0; k < 8; ++k) {
            int32_t l1 = nodes[0].child[k];
            for (int j = 0; j < 8; ++j) refine(nodes[l1].child[j], prm.max_depth);
        }
        for (int k = 0; k < 8; ++k) summarize(nodes[0].child[k]);
        summarize(0);
    }

    // SideLengthOverDistance (Descent/DescentCriterion/SideLengthOverDistance.h:17-22)
    // with exclusionRegion(OctreeNode) (Descent/DescentCriterionType.h:81-85)
    // and half-open BoundingBox::contains (Physics/BoundingBox.h:61-64).
    bool mac(const ONode &nd, V3 p) const {
        float dist = std::sqrt(v_len2(v_sub(p, nd.com)));
        if (!((nd.side / dist) < prm.theta)) return false;
        V3 lo{nd.center.x - nd.side, nd.center.y - nd.side, nd.center.z - nd.side};
        V3 hi{nd.center.x + nd.side, nd.center.y + nd.side, nd.center.z + nd.side};
        bool inside = lo.x <= p.x && lo.y <= p.y && lo.z <= p.z && p.x < hi.x && p.y < hi.y && p.z < hi.z;
        return !inside;
    }

    // Descent::activeTree (Descent/activeTree.h:24-79); 8 children combined by
    // libstdc++ transform_reduce: init + ((r0+r1)+(r2+r3)), then
    // + ((r4+r5)+(r6+r7)).
    V3 walk(int32_t id, V3 p, Counts &cnt) const {
        const ONode &nd = nodes[id];
        if (nd.count == 0) return kZero;
        if (mac(nd, p)) {
            cnt.np++;
            if (prm.order == 2) return node_particle_quad(nd.com, nd.mass, nd.quad, p, prm.G, prm.eps2);
            return particle_particle(nd.com, nd.mass, p, prm.G, prm.eps2);
        }
        if (nd.leaf) {
            V3 acc = kZero;
            for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
                uint32_t e = idx[i];
                cnt.pp++;
                acc = v_add(acc, particle_particle(P[e], Mv[e], p, prm.G, prm.eps2));
            }
            return acc;
        }
        V3 r[8];
        for (int k = 0; k < 8; ++k) r[k] = walk(nd.child[k], p, cnt);
        V3 init = kZero;
        init = v_add(init, v_add(v_add(r[0], r[1]), v_add(r[2], r[3])));
        init = v_add(init, v_add(v_add(r[4], r[5]), v_add(r[6], r[7])));
        return init;
    }

    // The same opening decisions as walk() (FP32 MAC), each monopole
    // interaction evaluated and summed in FP64.
    void walk64(int32_t id, V3 p, D3 &acc, Counts &cnt) const {
        const ONode &nd = nodes[id];
        if (nd.count == 0) return;
        if (mac(nd, p)) {
            cnt.np++;
            pp64(nd.com, nd.mass, p, prm.G, prm.eps2, acc);
            return;
        }
        if (nd.leaf) {
            for (int64_t i = nd.first; i < nd.first + nd.count; ++i) {
                uint32_t e = idx[i];
                cnt.pp++;
                pp64(P[e], Mv[e], p, prm.G, prm.eps2, acc);
            }
            return;
        }
        for (int k = 0; k < 8; ++k) walk64(nd.child[k], p, acc, cnt);
    }
};

// ---------------------------------------------------------------- implicit FMM
// MultipoleImplicitFMMSolver<1> (Solvers/ImplicitFMMSolver.h:14-33): an
// ImplicitDualTraversalSolver (Solvers/ImplicitDualTraversalSolver.h:17-70)
// over one octree whose nodes carry a centre of mass (active side) and imply
// a MultipoleAccelerationSummary<2> local field (passive side:
// Summaries/EmptySummary.h:37-60), with SideLengthOverDistance at theta
// (default 0.4).

// MultipoleAcceleration<2> (Physics/MultipoleAcceleration.h:14-90): a vector
// and the symmetric 3x3 tensor of its gradient (xx xy xz yy yz zz).  `mass`
// is used only by the accounting mode (mass received per target).
struct Local {
    V3 a = kZero;
    float T[6] = {0, 0, 0, 0, 0, 0};
    double mass = 0.0;
};

// SymmetricTensor3<2> * vec3 (Physics/SymmetricTensor3.h:235-247).
inline V3 t_dot(const float *T, V3 o) {
    return V3{(T[0] * o.x + T[1] * o.y) + T[2] * o.z, (T[1] * o.x + T[3] * o.y) + T[4] * o.z,
              (T[2] * o.x + T[4] * o.y) + T[5] * o.z};
}
// MultipoleAcceleration::at and ::translated (MultipoleAcceleration.h:33-70).
inline V3 local_at(const Local &L, V3 o) { return v_add(L.a, t_dot(L.T, o)); }
inline Local local_translated(const Local &L, V3 o) {
    Local t = L;
    t.a = v_add(t.a, t_dot(L.T, o));
    return t;
}

struct FMMCounts {
    uint64_t pp = 0, np = 0, pn = 0, nn = 0;  // SimpleTrackingRule's four counters (empty sources excluded)
};

// Results shared by the traversals of all starting nodes: each body lies under
// exactly one starting node, so the traversals (one per thread at a time, as the
// reference's tbb::parallel_for_each runs them) write disjoint entries.
struct FMMShared {
    std::vector<V3> acc;       // by original body index
    std::vector<double> recv;  // accounting mode: mass received by each body
    std::vector<double> nmass; // accounting mode: FP64 mass of every node's bodies
    FMMShared(const Octree &T, bool account) {
        acc.assign(T.n, kZero);
        recv.assign(T.n, 0.0);
        if (account) {
            nmass.assign(T.nodes.size(), 0.0);
            for (size_t i = 0; i < T.nodes.size(); ++i)
                for (int64_t j = T.nodes[i].first; j < T.nodes[i].first + T.nodes[i].count; ++j)
                    nmass[i] += (double)T.Mv[T.idx[j]];
        }
    }
};

// SideLengthOverDistance, node-node (SideLengthOverDistance.h:33-56):
// 0 = Approximate, else DescendActiveNode (1) | DescendPassiveNode (2)
// (DescentCriterionType.h:21-27).  The distance runs from the active node's
// centre of mass to the passive node's centre; exclusion regions are
// [c - side, c + side] (DescentCriterionType.h:81-85) with closed-interval
// overlap (BoundingBox.h:77-82).
int fmm_recommend(const ONode &A, const ONode &Pn, float theta) {
    const float sa2 = A.side * A.side, sp2 = Pn.side * Pn.side;
#if ORC_MUTATE == 6
    const float d2 = v_len2(v_sub(Pn.center, A.center));
#else
    const float d2 = v_len2(v_sub(Pn.center, A.com));
#endif
    const float big = sa2 < sp2 ? sp2 : sa2;
    if (((4.0f * big) / d2) < (theta * theta)) {
#if ORC_MUTATE == 3
        const float ps = Pn.side / 2.0f, as = A.side / 2.0f;
#else
        const float ps = Pn.side, as = A.side;
#endif
        const V3 pl{Pn.center.x - ps, Pn.center.y - ps, Pn.center.z - ps};
        const V3 ph{Pn.center.x + ps, Pn.center.y + ps, Pn.center.z + ps};
        const V3 al{A.center.x - as, A.center.y - as, A.center.z - as};
        const V3 ah{A.center.x + as, A.center.y + as, A.center.z + as};
        const bool meet = std::min(ph.x, ah.x) >= std::max(pl.x, al.x) && std::min(ph.y, ah.y) >= std::max(pl.y, al.y) &&
                          std::min(ph.z, ah.z) >= std::max(pl.z, al.z);
        if (!meet) return 0;
    }
#if ORC_MUTATE == 4
    int r = (sp2 <= sa2 * 4.0f) ? 2 : 0;
    r |= (sa2 <= sp2 * 4.0f) ? 1 : 0;
#else
    int r = (sa2 <= sp2 * 4.0f) ? 2 : 0;
    r |= (sp2 <= sa2 * 4.0f) ? 1 : 0;
#endif
    return r;
}

// SideLengthOverDistance, particle-node (SideLengthOverDistance.h:25-29):
// 2 side / |x - c| < theta and x outside the cell [c - side/2, c + side/2)
// (OctreeNode::boundingBox, Trees/Octree.h:92-98; half-open contains,
// Physics/BoundingBox.h:61-64).
bool fmm_particle_far(V3 x, const ONode &Pn, float theta) {
    const float dist = std::sqrt(v_len2(v_sub(x, Pn.center)));
    if (!(((2.0f * Pn.side) / dist) < theta)) return false;
    const float h = Pn.side / 2.0f;
#if ORC_MUTATE == 5
    const bool in = Pn.center.x - h <= x.x && Pn.center.y - h <= x.y && Pn.center.z - h <= x.z &&
                    x.x <= Pn.center.x + h && x.y <= Pn.center.y + h && x.z <= Pn.center.z + h;
#else
    const bool in = Pn.center.x - h <= x.x && Pn.center.y - h <= x.y && Pn.center.z - h <= x.z &&
                    x.x < Pn.center.x + h && x.y < Pn.center.y + h && x.z < Pn.center.z + h;
#endif
    return !in;
}

struct FMM {
    const Octree &T;
    float theta;
    bool account;
    FMMShared &sh;
    std::vector<V3> &acc;
    std::vector<double> &recv, &nmass;
    FMMCounts cnt;

    FMM(const Octree &t, bool acc_mode, FMMShared &shared)
        : T(t), theta(t.prm.theta), account(acc_mode), sh(shared), acc(shared.acc), recv(shared.recv),
          nmass(shared.nmass) {}

    const ONode &nd(int32_t i) const { return T.nodes[i]; }
    V3 pos(uint32_t e) const { return T.P[e]; }

    int recommend(int32_t a, int32_t p) const { return fmm_recommend(nd(a), nd(p), theta); }
    bool particle_far(V3 x, int32_t p) const { return fmm_particle_far(x, nd(p), theta); }

    // Gravity::particleNode into a MultipoleAccelerationSummary<2>
    // (Physics/Rules/Gravity.h:281-304): R = c - s, r = sqrt(R.R + eps2),
    // a -= (R / r^3) m G, T += D<2>(R, r) m G with D<2> = g2 R(x)R + g1 I,
    // g1 = (-1 / r^2) / r, g2 = (3 / r^3) / r^2 (Gravity.h:43-58).
    void to_local(V3 s, float m, V3 c, Local &L, double exact_mass) const {
        if (account) { L.mass += exact_mass; return; }
        if (v_eq(s, c)) return;
        const float G = T.prm.G;
        const V3 R = v_sub(c, s);
        const float r = std::sqrt(v_len2(R) + T.prm.eps2);
        const float r3 = r * (r * r);
        L.a = v_sub(L.a, v_mul(v_mul(v_div(R, r3), m), G));
        const float g1 = (-1.0f / (r * r)) / r, g2 = (3.0f / (r * (r * r))) / (r * r);
        const float Rv[3] = {R.x, R.y, R.z};
        static const int kI[6] = {0, 0, 0, 1, 1, 2}, kJ[6] = {0, 1, 2, 1, 2, 2};
        for (int k = 0; k < 6; ++k) {
            const float D = Rv[kI[k]] * Rv[kJ[k]] * g2 + (kI[k] == kJ[k] ? 1.0f : 0.0f) * g1;
            L.T[k] = L.T[k] + (D * m) * G;
        }
    }

    // Descent::activeTree from node a for body q (counts as SimpleTrackingRule).
    void walk_into(int32_t a, uint32_t q) {
        if (account) { recv[q] += walk_mass(a, pos(q)); return; }
        Counts c;
        acc[q] = v_add(acc[q], T.walk(a, pos(q), c));
        cnt.pp += c.pp;
        cnt.np += c.np;
    }
    double walk_mass(int32_t id, V3 p) {
        const ONode &n = nd(id);
        if (n.count == 0) return 0.0;
        if (T.mac(n, p)) { cnt.np++; return nmass[id]; }
        if (n.leaf) {
            double s = 0.0;
            for (int64_t i = n.first; i < n.first + n.count; ++i) { cnt.pp++; s += T.Mv[T.idx[i]]; }
            return s;
        }
        double s = 0.0;
        for (int k = 0; k < 8; ++k) s += walk_mass(n.child[k], p);
        return s;
    }
    void p2p(uint32_t e, uint32_t q) {
        cnt.pp++;
        if (account) { recv[q] += T.Mv[e]; return; }
        acc[q] = v_add(acc[q], particle_particle(pos(e), T.Mv[e], pos(q), T.prm.G, T.prm.eps2));
    }
    void apply_local(const Local &L, uint32_t q, V3 c) {
        if (account) { recv[q] += L.mass; return; }
        acc[q] = v_add(acc[q], local_at(L, v_sub(pos(q), c)));
    }

    // Descent::passiveTreeImplicitField (Descent/passiveTreeImplicitField.h:13-78)
    // over the entities E[b, e) (reordered in place by std::partition, as the
    // reference's span is).
    void passive_field(std::vector<uint32_t> &E, int64_t b, int64_t e, int32_t p, Local L) {
        const ONode &Pn = nd(p);
        if (Pn.count == 0) return;
        auto mid = std::partition(E.begin() + b, E.begin() + e, [&](uint32_t x) { return particle_far(pos(x), p); });
        const int64_t m = mid - E.begin();
        for (int64_t i = b; i < m; ++i) {
            cnt.pn++;
            to_local(pos(E[i]), T.Mv[E[i]], Pn.center, L, (double)T.Mv[E[i]]);
        }
        if (Pn.leaf) {
            for (int64_t i = Pn.first; i < Pn.first + Pn.count; ++i) {
                const uint32_t q = T.idx[i];
                for (int64_t j = m; j < e; ++j) p2p(E[j], q);
                apply_local(L, q, Pn.center);
            }
        } else {
            for (int k = 0; k < 8; ++k) {
                const int32_t c = Pn.child[k];
                passive_field(E, m, e, c, local_translated(L, v_sub(nd(c).center, Pn.center)));
            }
        }
    }

    // Descent::lockstepDualTreeImplicitField (Descent/lockstepDualTreeImplicitField.h:18-170)
    // over the active nodes A[b, e) (reordered in place, as the reference's span is).
    void lockstep(std::vector<int32_t> &A, int64_t b, int64_t e, int32_t p, Local L) {
        const ONode &Pn = nd(p);
        if (Pn.count == 0) return;
        if (b == e) {
            for (int64_t i = Pn.first; i < Pn.first + Pn.count; ++i) apply_local(L, T.idx[i], Pn.center);
            return;
        }
        // partitionByDescentCriterionLockstep (:22-58): far | leaf | non-leaf
        int64_t end_far = b, begin_nonleaf = e;
        for (int64_t i = b; i < begin_nonleaf; ++i) {
            const int r = recommend(A[i], p);
            if (r == 0) {
                std::swap(A[i], A[end_far]);
                ++end_far;
            } else if (!nd(A[i]).leaf) {
                --begin_nonleaf;
                std::swap(A[i], A[begin_nonleaf]);
                --i;
            }
        }
        for (int64_t i = b; i < end_far; ++i) {
            const ONode &a = nd(A[i]);
            if (a.count > 0) cnt.nn++;
            to_local(a.com, a.mass, Pn.center, L, account ? nmass[A[i]] : 0.0);
        }
        if (Pn.leaf) {
            for (int64_t i = end_far; i < begin_nonleaf; ++i)
                for (int64_t j = Pn.first; j < Pn.first + Pn.count; ++j) walk_into(A[i], T.idx[j]);
            for (int64_t j = Pn.first; j < Pn.first + Pn.count; ++j) {
                const uint32_t q = T.idx[j];
                for (int64_t i = begin_nonleaf; i < e; ++i) walk_into(A[i], q);
                apply_local(L, q, Pn.center);
            }
            return;
        }
        std::vector<uint32_t> ents;
        for (int64_t i = end_far; i < begin_nonleaf; ++i) {
            const ONode &a = nd(A[i]);
            ents.insert(ents.end(), T.idx.begin() + a.first, T.idx.begin() + a.first + a.count);
        }
        passive_field(ents, 0, (int64_t)ents.size(), p, Local{});
        std::vector<int32_t> deeper;
        for (int64_t i = begin_nonleaf; i < e; ++i)
            for (int k = 0; k < 8; ++k) deeper.push_back(nd(A[i]).child[k]);
        for (int k = 0; k < 8; ++k) {
            const int32_t c = Pn.child[k];
            lockstep(deeper, 0, (int64_t)deeper.size(), c, local_translated(L, v_sub(nd(c).center, Pn.center)));
        }
    }

    // Descent::balancedLockstepDualTreeImplicitField (:14-45): breadth-first
    // descent of the active tree while the criterion asks to descend the
    // active node only.
    void balanced(int32_t s) {
        std::deque<int32_t> q{0};
        std::vector<int32_t> eq;
        while (!q.empty()) {
            const int32_t a = q.front();
            const int r = recommend(a, s);
            q.pop_front();
            if (r != 1 || nd(a).leaf) eq.push_back(a);
            else for (int k = 0; k < 8; ++k) q.push_back(nd(a).child[k]);
        }
        lockstep(eq, 0, (int64_t)eq.size(), s, Local{});
    }

    // Tree::loadBalancedBreak (Trees/Tree.h:82-106): a max-heap by body count;
    // the largest node is replaced by its 8 children until n nodes are
    // queued or the largest is a leaf.
    std::vector<int32_t> starting_nodes(size_t n) const {
        std::vector<int32_t> q{0};
        auto less = [&](int32_t a, int32_t b) { return nd(a).count < nd(b).count; };
        while (q.size() < n) {
            std::pop_heap(q.begin(), q.end(), less);
            const int32_t t = q.back();
            if (nd(t).leaf) break;
            q.pop_back();
            for (int k = 0; k < 8; ++k) {
                q.push_back(nd(t).child[k]);
                std::push_heap(q.begin(), q.end(), less);
            }
        }
        return q;
    }

};

// ImplicitDualTraversalSolver::updateAccelerations (:46-68): the traversals of
// the first `k` starting nodes (all of them for k < 0), on `threads` threads
// (0: OpenMP's default); returns the bodies those starting nodes hold.
int64_t fmm_run(const Octree &T, FMMShared &sh, bool account, int k, int threads, FMMCounts &total) {
    if (T.n == 0) return 0;
    FMM f0(T, account, sh);
    std::vector<int32_t> st = f0.starting_nodes(256);
    if (k >= 0 && (size_t)k < st.size()) st.resize((size_t)k);
    int64_t covered = 0;
    for (int32_t s : st) covered += T.nodes[s].count;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
    {
        FMM f(T, account, sh);
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)st.size(); ++i) f.balanced(st[(size_t)i]);
#pragma omp critical
        {
            total.pp += f.cnt.pp;
            total.np += f.cnt.np;
            total.pn += f.cnt.pn;
            total.nn += f.cnt.nn;
        }
    }
    return covered;
}

struct Handle {
    int kind;  // 0 octree, 1 lbvh
    bool fmm = false;
    LBVH *l = nullptr;
    Octree *o = nullptr;
    ~Handle() { delete l; delete o; }
};

Params to_params(const float *fp, const int *ip) {
    Params p;
    p.solver = ip[0]; p.leaf_size = ip[1]; p.max_depth = ip[2]; p.morton_bits = ip[3];
    // 3 / 4: the quadrupole variants of the octree / LBVH solvers
    // (QuadrupoleBarnesHutSolver, QuadrupoleLinearBVHSolver); 5: the implicit
    // FMM (an octree)
    if (p.solver == 3 || p.solver == 4) { p.solver -= 3; p.order = 2; }
    p.theta = fp[0]; p.G = fp[1]; p.eps2 = fp[2]; p.dt = fp[3];
    return p;
}

}  // namespace

extern "C" {

// Params are passed as two small arrays so the ctypes side stays trivial:
//   ip = {solver, leaf_size, max_depth, morton_bits}, fp = {theta, G, eps2, dt}.

int orc_bbox(int64_t n, const float *x, const float *y, const float *z, float *out6) {
    bbox_of(n, x, y, z, out6);
    return 0;
}

int orc_morton(int64_t n, const float *x, const float *y, const float *z, int bits, const float *bb,
               uint64_t *keys) {
    if (bits != 30 && bits != 63) return -1;
    float scale = morton_scale(bb);
    for (int64_t i = 0; i < n; ++i) keys[i] = morton_of(V3{x[i], y[i], z[i]}, bb, scale, bits);
    return 0;
}

// Reference form of the interleave (MortonSort.h:27-44), used only by the
// tests to pin spread_bits.
uint64_t orc_expand_bits_magic(uint64_t i, int bits) {
    if (bits == 30) {
        uint32_t v = (uint32_t)i;
        v = (v * 0x00010001u) & 0xFF0000FFu;
        v = (v * 0x00000101u) & 0x0F00F00Fu;
        v = (v * 0x00000011u) & 0xC30C30C3u;
        v = (v * 0x00000005u) & 0x49249249u;
        return v;
    }
    i &= 0x1fffff;
    i = (i | i << 32) & 0x1f00000000ffffull;
    i = (i | i << 16) & 0x1f0000ff0000ffull;
    i = (i | i << 8) & 0x100f00f00f00f00full;
    i = (i | i << 4) & 0x10c30c30c30c30c3ull;
    i = (i | i << 2) & 0x1249249249249249ull;
    return i;
}

// The implicit FMM's two criterion forms on explicit cells, for the truth
// tables of tests/test_oracle.py: cell = centre xyz, side; com = the active
// node's centre of mass.
int orc_fmm_recommend(const float *a_cell4, const float *a_com3, const float *p_cell4, float theta) {
    ONode A, P;
    A.center = V3{a_cell4[0], a_cell4[1], a_cell4[2]};
    A.side = a_cell4[3];
    A.com = V3{a_com3[0], a_com3[1], a_com3[2]};
    P.center = V3{p_cell4[0], p_cell4[1], p_cell4[2]};
    P.side = p_cell4[3];
    return fmm_recommend(A, P, theta);
}

int orc_fmm_particle_far(const float *x3, const float *p_cell4, float theta) {
    ONode P;
    P.center = V3{p_cell4[0], p_cell4[1], p_cell4[2]};
    P.side = p_cell4[3];
    return fmm_particle_far(V3{x3[0], x3[1], x3[2]}, P, theta) ? 1 : 0;
}

uint64_t orc_expand_bits(uint64_t v, int bits) { return spread_bits(v, bits == 30 ? 10 : 21); }

int orc_sort(int64_t n, const uint64_t *keys, uint32_t *perm) {
    std::iota(perm, perm + n, 0u);
    std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    return 0;
}

void *orc_tree_new(const int *ip, const float *fp, int64_t n, const float *x, const float *y,
                   const float *z, const float *m) {
    Params p = to_params(fp, ip);
    Handle *h = new Handle;
    if (p.solver == 5) { p.solver = 0; h->fmm = true; }
    h->kind = p.solver;
    if (p.solver != 0 && p.solver != 1) { delete h; return nullptr; }
    if (p.solver == 1) {
        if (p.morton_bits != 30 && p.morton_bits != 63) { delete h; return nullptr; }
        h->l = new LBVH;
        h->l->construct(p, n, x, y, z, m);
    } else {
        h->o = new Octree;
        h->o->construct(p, n, x, y, z, m);
    }
    return h;
}

void orc_tree_free(void *hp) { delete (Handle *)hp; }

int64_t orc_tree_num_nodes(void *hp) {
    Handle *h = (Handle *)hp;
    return h->kind == 1 ? (int64_t)h->l->nodes.size() : (int64_t)h->o->nodes.size();
}

int orc_tree_max_depth(void *hp) {
    Handle *h = (Handle *)hp;
    return h->kind == 1 ? h->l->max_depth : h->o->max_depth;
}

// Nodes in pre-order: first, count, leaf flag, depth, com(x,y,z,mass) and,
// per kind, bbox (min,max) for LBVH or (centre, side, 0, 0) for the octree.
int orc_tree_nodes(void *hp, int64_t *first, int64_t *count, int32_t *leaf, int32_t *depth,
                   float *com4, float *geo6) {
    Handle *h = (Handle *)hp;
    if (h->kind == 1) {
        const auto &N = h->l->nodes;
        for (size_t i = 0; i < N.size(); ++i) {
            first[i] = N[i].first; count[i] = N[i].count; leaf[i] = N[i].leaf; depth[i] = N[i].depth;
            com4[4 * i] = N[i].com.x; com4[4 * i + 1] = N[i].com.y; com4[4 * i + 2] = N[i].com.z;
            com4[4 * i + 3] = N[i].mass;
            float g[6] = {N[i].bmin.x, N[i].bmin.y, N[i].bmin.z, N[i].bmax.x, N[i].bmax.y, N[i].bmax.z};
            std::memcpy(geo6 + 6 * i, g, sizeof g);
        }
    } else {
        const auto &N = h->o->nodes;
        for (size_t i = 0; i < N.size(); ++i) {
            first[i] = N[i].first; count[i] = N[i].count; leaf[i] = N[i].leaf; depth[i] = N[i].depth;
            com4[4 * i] = N[i].com.x; com4[4 * i + 1] = N[i].com.y; com4[4 * i + 2] = N[i].com.z;
            com4[4 * i + 3] = N[i].mass;
            float g[6] = {N[i].center.x, N[i].center.y, N[i].center.z, N[i].side, 0.0f, 0.0f};
            std::memcpy(geo6 + 6 * i, g, sizeof g);
        }
    }
    return 0;
}

// Second moments of every node in pre-order (xx xy xz yy yz zz; zeros for
// a monopole tree and for empty octree nodes).
int orc_tree_quads(void *hp, float *q6) {
    Handle *h = (Handle *)hp;
    if (h->kind == 1)
        for (size_t i = 0; i < h->l->nodes.size(); ++i) std::memcpy(q6 + 6 * i, h->l->nodes[i].quad.q, 6 * sizeof(float));
    else
        for (size_t i = 0; i < h->o->nodes.size(); ++i) std::memcpy(q6 + 6 * i, h->o->nodes[i].quad.q, 6 * sizeof(float));
    return 0;
}

// Gravity::nodeParticle<2> for one node and one target (test pin).
int orc_node_particle_quad(const float *c3, float M, const float *q6, const float *p3, float G, float eps2,
                           float *out3) {
    Quad Q;
    std::memcpy(Q.q, q6, sizeof Q.q);
    const V3 r = node_particle_quad(V3{c3[0], c3[1], c3[2]}, M, Q, V3{p3[0], p3[1], p3[2]}, G, eps2);
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
    return 0;
}

// Parent index of every node (-1 for the root), from the child links.
int orc_tree_parents(void *hp, int32_t *parent) {
    Handle *h = (Handle *)hp;
    if (h->kind == 1) {
        const auto &N = h->l->nodes;
        for (size_t i = 0; i < N.size(); ++i) parent[i] = -1;
        for (size_t i = 0; i < N.size(); ++i)
            if (!N[i].leaf) { parent[N[i].child0] = (int32_t)i; parent[N[i].child1] = (int32_t)i; }
    } else {
        const auto &N = h->o->nodes;
        for (size_t i = 0; i < N.size(); ++i) parent[i] = -1;
        for (size_t i = 0; i < N.size(); ++i)
            if (!N[i].leaf)
                for (int k = 0; k < 8; ++k) parent[N[i].child[k]] = (int32_t)i;
    }
    return 0;
}

// Body order of the tree (sorted position -> original index) and, for the
// LBVH, the sorted keys.
int orc_tree_order(void *hp, uint32_t *perm, uint64_t *keys_or_null) {
    Handle *h = (Handle *)hp;
    if (h->kind == 1) {
        std::copy(h->l->perm.begin(), h->l->perm.end(), perm);
        if (keys_or_null) std::copy(h->l->keys.begin(), h->l->keys.end(), keys_or_null);
    } else {
        std::copy(h->o->idx.begin(), h->o->idx.end(), perm);
    }
    return 0;
}

int orc_tree_bbox(void *hp, float *out6) {
    Handle *h = (Handle *)hp;
    std::memcpy(out6, h->kind == 1 ? h->l->bbox : h->o->bbox, 6 * sizeof(float));
    return 0;
}

// ActiveTreeSolver::updateAccelerations walk (Solvers/ActiveTreeSolver.h:61-70)
// for arbitrary target positions; per-target interaction counts (pp, np) as
// SimpleTrackingRule counts them (Physics/Rules/SimpleTrackingRule.h:29-45).
int orc_tree_accel(void *hp, int64_t nt, const float *tx, const float *ty, const float *tz, float *ax,
                   float *ay, float *az, uint64_t *pp_or_null, uint64_t *np_or_null, int threads) {
    Handle *h = (Handle *)hp;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nt; ++i) {
        Counts c;
        V3 p{tx[i], ty[i], tz[i]}, a;
        if (h->kind == 1) {
            float t2 = h->l->prm.theta * h->l->prm.theta;
            a = h->l->nodes.empty() ? kZero : h->l->walk(0, p, t2, c);
        } else {
            a = h->o->nodes.empty() ? kZero : h->o->walk(0, p, c);
        }
        ax[i] = a.x; ay[i] = a.y; az[i] = a.z;
        if (pp_or_null) pp_or_null[i] = c.pp;
        if (np_or_null) np_or_null[i] = c.np;
    }
    return 0;
}

// The monopole walk of orc_tree_accel with every interaction evaluated and
// summed in FP64 (the FP32 opening decisions are unchanged): the exact
// value of the reference's interaction set for each target, against which
// the FP32 rounding of the reference's walk and of a device walk are
// measured.  Returns -1 for quadrupole and FMM trees.
int orc_tree_accel_f64(void *hp, int64_t nt, const float *tx, const float *ty, const float *tz, double *ax,
                       double *ay, double *az, uint64_t *pp_or_null, uint64_t *np_or_null, int threads) {
    Handle *h = (Handle *)hp;
    if (h->fmm || (h->kind == 1 ? h->l->prm.order : h->o->prm.order) != 1) return -1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nt; ++i) {
        Counts c;
        D3 a{0.0, 0.0, 0.0};
        V3 p{tx[i], ty[i], tz[i]};
        if (h->kind == 1) {
            float t2 = h->l->prm.theta * h->l->prm.theta;
            if (!h->l->nodes.empty()) h->l->walk64(0, p, t2, a, c);
        } else if (!h->o->nodes.empty()) {
            h->o->walk64(0, p, a, c);
        }
        ax[i] = a.x; ay[i] = a.y; az[i] = a.z;
        if (pp_or_null) pp_or_null[i] = c.pp;
        if (np_or_null) np_or_null[i] = c.np;
    }
    return 0;
}

// The implicit FMM's accelerations for every body of an FMM tree (solver 5),
// original order, and its four interaction counts (pp, np, pn, nn).  With
// mass_or_null the same traversal instead adds each source's mass to every
// target it reaches (FP64; the local fields carry mass too).
int orc_fmm_accel(void *hp, float *ax, float *ay, float *az, uint64_t *counts4, double *mass_or_null,
                  int32_t threads) {
    Handle *h = (Handle *)hp;
    if (!h->fmm) return -1;
    const bool account = mass_or_null != nullptr;
    FMMShared sh(*h->o, account);
    FMMCounts c;
    fmm_run(*h->o, sh, account, -1, threads, c);
    for (int64_t i = 0; i < h->o->n; ++i) {
        if (mass_or_null) mass_or_null[i] = sh.recv[i];
        else { ax[i] = sh.acc[i].x; ay[i] = sh.acc[i].y; az[i] = sh.acc[i].z; }
    }
    if (counts4) { counts4[0] = c.pp; counts4[1] = c.np; counts4[2] = c.pn; counts4[3] = c.nn; }
    return 0;
}

// One source's local field at c (Gravity::particleNode into a
// MultipoleAccelerationSummary<2>) evaluated at p (MultipoleAcceleration::at):
// test pin for the expansion.
int orc_fmm_p2l(const float *s3, float m, const float *c3, const float *p3, float G, float eps2, float *out3) {
    Octree t;
    t.prm.G = G;
    t.prm.eps2 = eps2;
    t.prm.theta = 0.4f;
    t.n = 0;
    FMMShared sh(t, false);
    FMM f(t, false, sh);
    Local L;
    f.to_local(V3{s3[0], s3[1], s3[2]}, m, V3{c3[0], c3[1], c3[2]}, L, 0.0);
    const V3 a = local_at(L, V3{p3[0] - c3[0], p3[1] - c3[1], p3[2] - c3[2]});
    out3[0] = a.x; out3[1] = a.y; out3[2] = a.z;
    return 0;
}

// The implicit FMM's traversals from its first k starting nodes only (a CPU
// timing sample of the whole pass); returns the number of bodies those
// starting nodes hold.
int64_t orc_fmm_sample(void *hp, int32_t k, int32_t threads) {
    Handle *h = (Handle *)hp;
    if (!h->fmm) return -1;
    FMMShared sh(*h->o, false);
    FMMCounts c;
    return fmm_run(*h->o, sh, false, k < 0 ? 0 : k, threads, c);
}

// The starting passive nodes of the implicit FMM (pre-order node ids).
int64_t orc_fmm_starts(void *hp, int32_t *out, int64_t cap) {
    Handle *h = (Handle *)hp;
    if (!h->fmm) return -1;
    FMMShared sh(*h->o, false);
    FMM f(*h->o, false, sh);
    const std::vector<int32_t> s = f.starting_nodes(256);
    for (size_t i = 0; i < s.size() && (int64_t)i < cap; ++i) out[i] = s[i];
    return (int64_t)s.size();
}

// Solver::step (Simulation/Solver.h:103-139): accelerations from the tree,
// then v = v + a*dt, then p = p + v*dt, each a separate pass, no FMA.
int orc_step(const int *ip, const float *fp, int64_t n, float *x, float *y, float *z, const float *m,
             float *vx, float *vy, float *vz, float *ax, float *ay, float *az, int nsteps, int threads) {
    Params p = to_params(fp, ip);
    for (int s = 0; s < nsteps; ++s) {
        void *h = orc_tree_new(ip, fp, n, x, y, z, m);
        if (!h) return -1;
        if (((Handle *)h)->fmm) orc_fmm_accel(h, ax, ay, az, nullptr, nullptr, threads);
        else orc_tree_accel(h, n, x, y, z, ax, ay, az, nullptr, nullptr, threads);
        orc_tree_free(h);
        for (int64_t i = 0; i < n; ++i) {
            vx[i] = vx[i] + ax[i] * p.dt; vy[i] = vy[i] + ay[i] * p.dt; vz[i] = vz[i] + az[i] * p.dt;
        }
        for (int64_t i = 0; i < n; ++i) {
            x[i] = x[i] + vx[i] * p.dt; y[i] = y[i] + vy[i] * p.dt; z[i] = z[i] + vz[i] * p.dt;
        }
    }
    return 0;
}

// NaiveSolver::updateAccelerations (Simulation/Solvers/NaiveSolver.h:22-52):
// FP32, sources summed in index order into a zeroed accumulator.
int orc_direct_f32(int64_t ns, const float *sx, const float *sy, const float *sz, const float *sm,
                   int64_t nt, const float *tx, const float *ty, const float *tz, float G, float eps2,
                   float *ax, float *ay, float *az) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; ++i) {
        V3 p{tx[i], ty[i], tz[i]}, acc = kZero;
        for (int64_t j = 0; j < ns; ++j)
            acc = v_add(acc, particle_particle(V3{sx[j], sy[j], sz[j]}, sm[j], p, G, eps2));
        ax[i] = acc.x; ay[i] = acc.y; az[i] = acc.z;
    }
    return 0;
}

// FP64 direct sum with the reference's softening (eps2 added to r^2) and
// self-skip by exact position equality (Gravity.h:203).
int orc_direct_f64(int64_t ns, const float *sx, const float *sy, const float *sz, const float *sm,
                   int64_t nt, const float *tx, const float *ty, const float *tz, double G, double eps2,
                   double *ax, double *ay, double *az) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < nt; ++i) {
        double px = tx[i], py = ty[i], pz = tz[i], a0 = 0, a1 = 0, a2 = 0;
        for (int64_t j = 0; j < ns; ++j) {
            if (sx[j] == tx[i] && sy[j] == ty[i] && sz[j] == tz[i]) continue;
            double rx = px - sx[j], ry = py - sy[j], rz = pz - sz[j];
            double r2 = rx * rx + ry * ry + rz * rz + eps2;
            double inv = 1.0 / (std::sqrt(r2) * r2);
            double s = G * (double)sm[j] * inv;
            a0 -= rx * s; a1 -= ry * s; a2 -= rz * s;
        }
        ax[i] = a0; ay[i] = a1; az[i] = a2;
    }
    return 0;
}

// Softened potential energy W = -sum_{i<j} G m_i m_j / sqrt(r^2 + eps2), FP64.
double orc_potential_f64(int64_t n, const float *x, const float *y, const float *z, const float *m,
                         double G, double eps2) {
    double W = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : W)
    for (int64_t i = 0; i < n; ++i) {
        double wi = 0;
        for (int64_t j = i + 1; j < n; ++j) {
            double rx = (double)x[i] - x[j], ry = (double)y[i] - y[j], rz = (double)z[i] - z[j];
            wi += (double)m[j] / std::sqrt(rx * rx + ry * ry + rz * rz + eps2);
        }
        W -= G * (double)m[i] * wi;
    }
    return W;
}

// The reference's P2P rule on one pair (Gravity.h:199-209), for the KATs.
int orc_particle_particle(const float *a3, float m, const float *p3, float G, float eps2, float *out3) {
    V3 r = particle_particle(V3{a3[0], a3[1], a3[2]}, m, V3{p3[0], p3[1], p3[2]}, G, eps2);
    out3[0] = r.x; out3[1] = r.y; out3[2] = r.z;
    return 0;
}

// BoundingBox box-box intersection (Physics/BoundingBox.h:77-82), for the KAT.
int orc_boxes_intersect(const float *a6, const float *b6) {
    for (int d = 0; d < 3; ++d) {
        float lo = std::max(a6[d], b6[d]);
        float hi = std::min(a6[3 + d], b6[3 + d]);
        if (!(hi >= lo)) return 0;
    }
    return 1;
}

// CenterOfMassSummary leaf summary of an explicit body list, for the KAT.
int orc_com(int64_t n, const float *x, const float *y, const float *z, const float *m, float *out4) {
    float M = 0.0f;
    V3 C = kZero;
    for (int64_t i = 0; i < n; ++i) {
        M += m[i];
        C = v_add(C, v_mul(V3{x[i], y[i], z[i]}, m[i]));
    }
    C = v_div(C, M);
    out4[0] = C.x; out4[1] = C.y; out4[2] = C.z; out4[3] = M;
    return 0;
}

}  // extern "C"
