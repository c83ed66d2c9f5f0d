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

inline int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }
inline int sort_passes(int key_bits) { return (key_bits + 7) / 8; }
// Words of status memory one sort needs (per pass: tiles x 256 + counters).
inline int64_t sort_status_words(int64_t n, int key_bits) {
    return 8 * kRadix + 64 + (int64_t)sort_passes(key_bits) * (sort_tiles(n) * kRadix);
}

// Scan geometry.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

// Body count above which the tree arrays' 32-bit indices would overflow.
constexpr int64_t kMaxBodies = (int64_t)1 << 29;

// Per-wave traversal stack entries.
constexpr int kStackLBVH = 192;



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

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);

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
                          uint32_t *node_count, hipStream_t s);

// Bottom-up LBVH summaries with atomic arrival flags: COM, mass, bbox and the
// MAC record (bbox centre, squared diagonal).  leaf_start[i] holds the leaf
// slot whose first body is i (else -1; written by the topology pass);
// flags holds max_nodes words; glist holds 2n + 64 words.
constexpr int kSummBodies = 4096;   // bodies per workgroup in the local pass
void launch_lbvh_summaries(const float4 *bodies, int64_t n, NodeRec *nodes, const int32_t *leaf_start,
                           int64_t max_nodes, uint32_t *flags, uint32_t *glist, float4 *bmin, float4 *bmax,
                           float theta, uint32_t *err, hipStream_t s);

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
    int64_t n;
    const float4 *bodies;   // sorted (x, y, z, m)
    const uint32_t *perm;   // sorted position -> original index
    const NodeRec *nodes;   // tree nodes
    float theta;            // octree MAC
    float theta2;           // LBVH MAC: theta * theta (float)
    double mac_mid;         // LBVH MAC threshold (mac_threshold)
    int mac_tie;
    float G, eps2;
    int64_t max_nodes;      // node array capacity (index guard)
    float *ax, *ay, *az;    // original order
    unsigned long long *counts;  // [pp, np]
    uint32_t *err;
};

// Barnes-Hut traversal over the LBVH.  ordered != 0 selects the
// bit-exact reference arithmetic; diag != 0 makes the FAST walk also count
// its schedule statistics (counts[2..5]).
void launch_walk_lbvh(const WalkArgs &a, int ordered, int max_depth, int diag, hipStream_t s);

// v = v + a*dt; p = p + v*dt (original order).
void launch_update(float *x, float *y, float *z, float *vx, float *vy, float *vz, const float *ax,
                   const float *ay, const float *az, float dt, int64_t n, hipStream_t s);

// FP64 direct sum at targets (float4 sources x,y,z,m; targets float4 x,y,z).
void launch_direct_f64(const float4 *src, int64_t ns, const float4 *tgt, int64_t nt, double G, double eps2,
                       double *ax, double *ay, double *az, hipStream_t s);

// SoA (x, y, z, m) -> float4, original order.
void launch_pack(const float *x, const float *y, const float *z, const float *m, int64_t n, float4 *out,
                 hipStream_t s);

// FP64 softened potential: per-block partial sums into partial[blocks].
int64_t potential_blocks(int64_t n);
void launch_potential_f64(const float4 *bodies, int64_t n, double G, double eps2, double *partial,
                          hipStream_t s);

}  // namespace nbg
