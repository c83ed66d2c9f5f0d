// Host-side launch wrappers for the gfx950 kernels.  Each takes plain device
// pointers and the stream to launch on; none allocates or synchronises.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbg {

// Radix-sort geometry (onesweep): 256 threads x 16 keys per tile.
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kRadix = 256;

constexpr int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }
constexpr int sort_passes(int key_bits) { return (key_bits + 7) / 8; }
// Words of status memory one sort needs (per pass: tiles x 256 + counters).
constexpr int64_t sort_status_words(int64_t n, int key_bits) {
    return 8 * kRadix + 64 + (int64_t)sort_passes(key_bits) * (sort_tiles(n) * kRadix);
}

// Scan geometry.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// Body count above which the tree arrays' 32-bit indices would overflow.
constexpr int64_t kMaxBodies = (int64_t)1 << 29;




// One tree node as the walk reads it (48 bytes; the two children of a node
// are adjacent, so a sibling pair is 96 contiguous bytes):
//   mac  = LBVH: bbox centre xyz, squared bbox diagonal;
//   com  = centre of mass xyz, total mass;
//   topo = first body (sort order), body count, first child (-1 for a
//          leaf; the second child is first child + 1), parent (-1 root).
struct NodeRec {
    float4 mac;
    float4 com;
    int4 topo;
};

// One octree node as the walk reads it (64 bytes).  The non-empty children
// of a node are contiguous, in the reference's child order (octant
// 4x + 2y + z; empty octants have no record):
//   com  = centre of mass xyz, total mass;
//   lo   = cell centre - side (xyz), w = squared-distance opening threshold;
//   hi   = cell centre + side (xyz), w = bits of (octant | nchild << 8 |
//          level << 16);
//   topo = first body, body count, first child (-1 for a leaf), parent.
struct OctRec {
    float4 com;
    float4 lo;
    float4 hi;
    int4 topo;
};

__host__ __device__ inline int oct_nchild(const OctRec &r) { return (int)((__builtin_bit_cast(uint32_t, r.hi.w) >> 8) & 0xFFu); }
__host__ __device__ inline int oct_octant(const OctRec &r) { return (int)(__builtin_bit_cast(uint32_t, r.hi.w) & 0xFFu); }
__host__ __device__ inline int oct_level(const OctRec &r) { return (int)(__builtin_bit_cast(uint32_t, r.hi.w) >> 16); }

// Octree key geometry: the octant of every level is a 3-bit digit; levels
// 0..20 live in key_hi bits 62..0, levels 21..33 in key_lo bits 38..0.
constexpr int kOctHiLevels = 21;
constexpr int kOctLevels = 34;   // max_depth 32 below the two forced levels
constexpr int kOctThrStride = 64;  // opening thresholds per system (one per level)

// Upper bound on the number of non-empty octree nodes: every body is in one
// leaf, and on each level 2..levels-1 an internal node holds > leaf bodies.
inline int64_t octree_capacity(int64_t n, int leaf, int max_depth) {
    const int64_t levels = (int64_t)max_depth + 2;
    return 9 + n + (levels - 2) * (n / ((int64_t)leaf + 1) + 1);
}

// Work buffers of the octree build (capacity `cap` nodes, n bodies).
struct OctBuffers {
    int64_t cap = 0, n = 0;
    uint64_t *key_hi = nullptr, *key_lo = nullptr, *lo_sorted = nullptr;  // n each
    uint32_t *nown = nullptr;      // n: nodes whose range starts at a position
    uint32_t *nbase = nullptr;     // n: exclusive scan of nown (pre-order node ids)
    int32_t *nfirst = nullptr, *ncount = nullptr, *nparent = nullptr;  // cap, by node id
    uint32_t *nmeta = nullptr;     // cap: level | octant << 8 | leaf << 12
    uint32_t *occ = nullptr;       // cap: occupied-child octant mask
    uint32_t *cb = nullptr;        // cap: child counts, then their exclusive scan
    uint32_t *flags = nullptr;     // cap: summary arrival counters
    uint32_t *glist = nullptr;     // 2 * (cap + 64): two queues of straddling parents (summaries)
    uint32_t *scan_ws = nullptr;   // 128 + cap / kScanTile
    OctRec *nodes = nullptr;       // cap, by slot
    float4 *centre = nullptr;      // cap, by slot: cell centre xyz, side
    float *thr = nullptr;          // 64: opening threshold per level
    uint32_t *misc = nullptr;      // 64: [0] node total, [1] child total
};

// Octree build (Octree::build, Trees/Octree.h:185-205) in three stages, each
// enqueue-only.  keys: per-body octant path on the reference's float centre
// recurrence.  sort: stable LSD radix sort by path (ties by original index);
// idx_alt receives the key order.  topology: cells from the sorted keys,
// node records, leaf contents reordered by original index into idx (the
// final body order), leaf_start.  Summaries run after the body gather.
void launch_octree_keys(const float *x, const float *y, const float *z, int64_t n, const float *bbox6,
                        int max_depth, OctBuffers &ob, uint64_t *keys, uint32_t *idx, uint32_t *idx_alt,
                        hipStream_t s);
void launch_octree_sort(int64_t n, int max_depth, OctBuffers &ob, uint64_t *keys, uint32_t *idx,
                        uint64_t *keys_alt, uint32_t *idx_alt, uint32_t *sort_status, uint32_t *err,
                        hipStream_t s);
void launch_octree_topology(int64_t n, const float *bbox6, int leaf_size, int max_depth, float theta,
                            OctBuffers &ob, const uint64_t *keys, const uint32_t *idx_alt, uint32_t *idx,
                            int32_t *leaf_start, uint32_t *err, hipStream_t s);
// Forest of octrees: bodies of system s (bsys, original order) get paths
// over their own system's cube (bbox6 + 6 s) and depth (lv_tab[s] levels);
// the sort orders positions by system, then path.  Leaves the sorted high
// words in keys (low words in ob.lo_sorted when max_levels > 21), the
// sorted system index in keys_alt and the permutation in idx_alt.
void launch_octree_forest_build(const float *x, const float *y, const float *z, int64_t n, int nsys,
                                const uint32_t *bsys, const float *bbox6, const int32_t *lv_tab, int max_levels,
                                OctBuffers &ob, uint64_t *keys, uint32_t *idx, uint64_t *keys_alt, uint32_t *idx_alt,
                                uint32_t *sort_status, uint32_t *err, hipStream_t s);

// Forest topology: every system gets its own octree (root slot s, its own
// leaf size, depth, and opening thresholds from its cube and theta =
// sys_prm[s].z; ob.thr holds kOctThrStride floats per system).
void launch_octree_forest_topology(int64_t n, int nsys, const float *bbox6, const float4 *sys_prm,
                                   const int32_t *lv_tab, const int32_t *leaf_tab, const int64_t *sys_first,
                                   int max_levels, OctBuffers &ob, const uint64_t *keys, const uint64_t *sys_sorted,
                                   const uint32_t *idx_alt, uint32_t *idx, int32_t *leaf_start, uint32_t *err,
                                   hipStream_t s);

// Per-system bounding boxes (one workgroup per system).
void launch_seg_bbox(const float *x, const float *y, const float *z, int nsys, const int64_t *sys_first,
                     float *bbox6, hipStream_t s);

void launch_octree_summaries(const float4 *bodies, int64_t n, OctBuffers &ob, const int32_t *leaf_start,
                             uint32_t *err, hipStream_t s, float4 *quad = nullptr);

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);

// Exclusive scan of n u32 values into out (in may equal out); total (if not
// null) receives the sum.  ws holds 64 + ceil(n / kScanTile) words.
void launch_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *ws, uint32_t *total, hipStream_t s);

// Stable radix sort of u64 keys + u32 payload over the low 8 * passes bits.
// Returns true when the result is in the _alt buffers (odd passes).
bool launch_radix_sort_u64(uint64_t *keys, uint32_t *idx, uint64_t *keys_alt, uint32_t *idx_alt, int64_t n,
                           int passes, uint32_t *status, uint32_t *err, hipStream_t s);

// outerBoundingBox: writes min xyz, max xyz to out6.  partials holds
// 6 * kBBoxBlocks floats.
constexpr int kBBoxBlocks = 512;
void launch_bbox(const float *x, const float *y, const float *z, int64_t n, float *partials, float *out6,
                 hipStream_t s);

// Morton keys (bits = 30 -> u32 keys, 63 -> u64 keys) and iota payload.
void launch_morton(const float *x, const float *y, const float *z, int64_t n, const float *bbox6, int bits,
                   void *keys, uint32_t *idx, hipStream_t s);

// Stable LSD radix sort of (key, idx) with onesweep passes.  Result lands
// back in keys/idx (an even number of passes).  status must hold
// sort_status_words(n, bits) words; err is a device flag word.
void launch_radix_sort(void *keys, uint32_t *idx, void *keys_alt, uint32_t *idx_alt, int64_t n, int bits,
                       uint32_t *status, uint32_t *err, hipStream_t s);

// Sorted SoA -> float4 (x, y, z, m) in sort order.
void launch_gather(const float *x, const float *y, const float *z, const float *m, const uint32_t *perm,
                   int64_t n, float4 *bodies, hipStream_t s);

// LBVH topology (Karras radix tree truncated at the reference's leaf rule)
// into the compact node arrays.  scratch ints: first,last,gamma,split,rank,
// parent (6 * n words) plus scan_ws (sort_tiles-sized block sums + 64).
// node_count receives the number of nodes.
void launch_lbvh_topology(const void *keys, int bits, int64_t n, int leaf_size, int32_t *kfirst,
                          int32_t *klast, int32_t *kgamma, uint32_t *ksplit, uint32_t *krank,
                          int32_t *kparent, uint32_t *scan_ws, NodeRec *nodes, int32_t *leaf_start,
                          uint32_t *node_count, hipStream_t s, int sys_shift = 0,
                          const int32_t *leaf_tab = nullptr);

// Bottom-up LBVH summaries with atomic arrival flags: COM, mass, bbox and the
// MAC record (bbox centre, squared diagonal).  leaf_start[i] holds the leaf
// slot whose first body is i (else -1; written by the topology pass);
// flags holds max_nodes words; glist holds 2 * (2n + 64) words (two queues).
constexpr int kSummBodies = 4096;   // bodies per workgroup in the local pass
constexpr int kSummMidBodies = kSummBodies * 64;  // bodies per workgroup in the middle pass
constexpr unsigned kSummMidMinBlocks = 8;         // fewer middle workgroups: no middle pass
constexpr int kSummThreads = 1024;  // threads per workgroup in the local pass (latency hiding)
struct MacParam {       // mac_threshold's output for one theta
    double mid;
    int tie;
};
// mp_tab (forest): per-system MAC parameters, indexed by keys[first] >> sys_shift.
// quad (quadrupole trees): 2 * max_nodes float4 of second moments (layout in
// device.hpp), or null.
void launch_lbvh_summaries(const float4 *bodies, int64_t n, NodeRec *nodes, const int32_t *leaf_start,
                           int64_t max_nodes, uint32_t *flags, uint32_t *glist, float4 *bmin, float4 *bmax,
                           float theta, uint32_t *err, hipStream_t s, const MacParam *mp_tab = nullptr,
                           const uint64_t *keys = nullptr, int sys_shift = 0, float4 *quad = nullptr);

// Forest of independent systems (30-bit codes, u64 keys = system << 30 | code):
// per-system bounding boxes (bbox6: 6 floats per system), keys, and one sort.
void launch_forest_build(const float *x, const float *y, const float *z, int64_t n, int nsys,
                         const int64_t *sys_first, const uint32_t *bsys, float *bbox6, uint64_t *keys, uint32_t *idx,
                         uint64_t *keys_alt, uint32_t *idx_alt, uint32_t *sort_status, uint32_t *err, hipStream_t s);
// sys_root[s] = slot of the node whose range is exactly system s.
void launch_forest_roots(const NodeRec *nodes, const uint32_t *node_count, int64_t max_nodes, const uint64_t *keys,
                         int64_t n, int nsys, int32_t *sys_root, uint32_t *err, hipStream_t s);
// Forest update: dt of body i is sys_prm[bsys[i]].w.
void launch_update_forest(float *x, float *y, float *z, float *vx, float *vy, float *vz, const float *ax,
                          const float *ay, const float *az, const uint32_t *bsys, const float4 *sys_prm, int64_t n,
                          hipStream_t s);

// Forest non-finite guard: bad[bsys[i]] = 1 for every body i whose position
// or velocity is NaN or infinite (bad is cleared by the caller).
void launch_forest_nonfinite(const float *x, const float *y, const float *z, const float *vx, const float *vy,
                             const float *vz, const uint32_t *bsys, int64_t n, uint32_t *bad, hipStream_t s);

// Host side of mac_lbvh_exact: the FP64 midpoint below T = RN(theta*theta)
// and whether a tie rounds down (T has an odd significand).
inline void mac_threshold(float theta, double *mid, int *tie) {
    const float T = theta * theta;
    if (!(T > 0.0f)) { *mid = -1.0; *tie = 0; return; }
    uint32_t bits;
    __builtin_memcpy(&bits, &T, 4);
    const uint32_t pbits = bits - 1;  // predecessor of a positive finite float
    float P;
    __builtin_memcpy(&P, &pbits, 4);
    *mid = ((double)P + (double)T) * 0.5;
    *tie = (int)(bits & 1u);
}

struct WalkArgs {
    int64_t n;              // bodies in the tree
    const float4 *bodies;   // sorted (x, y, z, m)
    int64_t nt;             // targets walked (the bodies themselves, or passive points)
    const float4 *targets;  // targets in walk order (xyz)
    const uint32_t *perm;   // walk position -> output index (< nt)
    const NodeRec *nodes;   // LBVH nodes
    const OctRec *onodes;   // octree nodes
    int levels;             // octree: max_depth + 2
    float theta;            // octree MAC
    float theta2;           // LBVH MAC: theta * theta (float)
    double mac_mid;         // LBVH MAC threshold (mac_threshold)
    int mac_tie;
    float G, eps2;
    int64_t max_nodes;      // node array capacity (index guard)
    float *ax, *ay, *az;    // output order (nt each)
    unsigned long long *counts;  // [pp, np]
    uint32_t *err;
    // Forest walk (nsys > 0): waves are aligned to systems.  Wave w serves
    // system s with wave_first[s] <= w < wave_first[s+1], targets
    // sys_first[s] + 64 (w - wave_first[s]) + lane below sys_first[s+1],
    // starts at node sys_root[s] and uses G = sys_prm[s].x, eps2 = sys_prm[s].y.
    int nsys = 0;
    int64_t nwaves = 0;
    const int64_t *wave_first = nullptr;
    const int64_t *sys_first = nullptr;
    const int32_t *sys_root = nullptr;
    const float4 *sys_prm = nullptr;
    // Quadrupole trees: second moments per node slot (2 float4, layout in
    // device.hpp); null for monopole trees.
    const float4 *quad = nullptr;
    // FAST single-system walks: targets per lane (1, 2 or 4); a wavefront
    // serves 64 tpl consecutive targets, lane l of slot t target 64 t + l.
    int tpl = 1;
    // FAST single-system linear-BVH walks with one target per lane: split
    // waves per group of 64 targets (1 = none).  With split > 1 the walk
    // stores per-target sums in split_part (kSplitSlots x 3 doubles per
    // target, component-major, split_stride doubles apart) and per-group
    // block counts in split_nb, and a second kernel adds them up.
    int split = 1;
    double *split_part = nullptr;
    float *split_part_f = nullptr;  // ORDERED split: kMaxFrontier x 3 floats per target
    int32_t *split_nb = nullptr;
    int64_t split_stride = 0;
};

// Sums a split walk stores per target: the top nodes' and one per block of
// frontier subtrees.
constexpr int kSplitSlots = 17;
// Largest system whose FAST linear-BVH walk splits its groups over waves.
constexpr int64_t kSplitMaxBodies = 65536;
// Waves per group the FAST linear-BVH walk takes when chosen by size.
constexpr int kSplitAutoLbvh = 4;
// ORDERED split: one result per frontier subtree (at most kOrderedSplitSlots)
// per target, for systems of at most kOrderedSplitMaxBodies bodies.
constexpr int kOrderedSplitSlots = 128;
constexpr int64_t kOrderedSplitMaxBodies = 65536;
// Split waves per group for a walk of `groups` groups of 64 targets, and
// for a system of n bodies (1: the chip is full): the most waves, up to 16
// per group, that keep the launch within 4,096 waves (16 per CU; measured:
// wider launches of the split walks run slower from 30k bodies up).
inline int split_waves_groups(int64_t groups) {
    int k = 1;
    while (k < 16 && groups * k * 2 <= 4096) k *= 2;
    return k;
}
inline int split_waves(int64_t n) {
    if (n > kSplitMaxBodies || n < 1) return 1;
    return split_waves_groups((n + 63) / 64);
}

// Barnes-Hut traversal over the LBVH.  ordered != 0 selects the
// bit-exact reference arithmetic; diag != 0 makes the FAST walk also count
// its schedule statistics (counts[2..5]).
void launch_walk_lbvh(const WalkArgs &a, int ordered, int max_depth, int diag, hipStream_t s);

// Barnes-Hut traversal over the octree (same modes and statistics).
void launch_walk_oct(const WalkArgs &a, int ordered, int diag, hipStream_t s);

// Implicit FMM (fmm.hip): the lockstep dual traversal breadth-first, one
// launch per traversal step, then one launch for the passive leaves.
// Frontier items (kFmmItemWords each) alternate between two halves by step
// parity; ctr holds three rotating sets of {items, next item, -, -}; leaf
// groups collect in `groups` for the last launch; lists live in one pool for
// the whole pass.  counts: [pp, np, pn, nn]; err bit 16: a list, the frontier
// or the leaf groups outgrew their capacity.
constexpr int kFmmItemWords = 16;
constexpr int kFmmGroupWaves = 8192;  // the leaf-group launch (8 waves per SIMD)
constexpr int kFmmUlistCap = 4096;    // close active nodes kept per leaf-group wave
constexpr int kFmmStartCap = 512;
constexpr int kFmmSeedWaves = 272;  // one per starting node (at most 256 + 7)
struct FmmArgs {
    int64_t n;
    const float4 *bodies;     // tree order (x, y, z, m)
    const OctRec *nodes;
    const float4 *centre;     // cell centre xyz, side, by slot
    const uint32_t *misc;     // octree misc ([0] node count)
    int64_t max_nodes;
    const uint32_t *perm;     // tree position -> original index
    float *ax, *ay, *az;
    float theta, G, eps2;
    // the criteria's bounds on FP32 quotients (fmm.hip's quotient_below): the
    // midpoints below theta * theta and below theta (set by launch_fmm)
    double theta2_mid, theta_mid;
    int32_t *starts;          // kFmmStartCap
    uint32_t *nstarts;
    uint32_t *ctr;            // 12 words
    uint32_t *pool_ctr;       // pool words used
    uint32_t *group_ctr;      // [0] leaf groups, [1] next leaf group
    uint32_t *peak;           // 4 words: pool words, items (largest step), steps run
    uint32_t *items[2];       // item_cap items each
    uint32_t *groups;         // item_cap items
    uint32_t *pool;           // pool_cap words
    int64_t item_cap, pool_cap;
    uint32_t *seed_scratch;   // kFmmSeedWaves x 3 x cap_a words
    uint32_t *ulist;          // kFmmGroupWaves x ulist_cap node slots
    int ulist_cap;
    int cap_a, waves, steps;
    unsigned long long *counts;
    uint32_t *err;
};
void launch_fmm(const FmmArgs &a, hipStream_t s);

// v = v + a*dt; p = p + v*dt (original order).
void launch_update(float *x, float *y, float *z, float *vx, float *vy, float *vz, const float *ax,
                   const float *ay, const float *az, float dt, int64_t n, hipStream_t s);

// NaiveSolver in FP32: every target sums every source in order (ORDERED:
// the reference's exact arithmetic).  Result of target i goes to
// out_index[i] (or i when out_index is null).
void launch_direct_f32(const float4 *src, int64_t ns, const float4 *tgt, int64_t nt, const uint32_t *out_index,
                       float G, float eps2, int ordered, float *ax, float *ay, float *az, hipStream_t s);

// FP64 direct sum at targets (float4 sources x,y,z,m; targets float4 x,y,z).
void launch_direct_f64(const float4 *src, int64_t ns, const float4 *tgt, int64_t nt, double G, double eps2,
                       double *ax, double *ay, double *az, hipStream_t s);

// SoA (x, y, z, m) -> float4, original order.
void launch_pack(const float *x, const float *y, const float *z, const float *m, int64_t n, float4 *out,
                 hipStream_t s);

// FP64 softened potential, each pair i < j once: wave w (64 targets) of
// [w0, w1) writes sum_{i in wave} m_i sum_{j > i} m_j / sqrt(r^2 + eps2)
// into partial[w]; -G times the sum of all potential_waves(n) partials is
// the potential.  bodies: (x, y, z, m) in FP64 (launch_pack_f64).
int64_t potential_waves(int64_t n);
void launch_pack_f64(const float *x, const float *y, const float *z, const float *m, int64_t n, double4 *out,
                     hipStream_t s);
void launch_potential_f64(const double4 *bodies, int64_t n, double eps2, int64_t w0, int64_t w1, double *partial,
                          hipStream_t s);

}  // namespace nbg
