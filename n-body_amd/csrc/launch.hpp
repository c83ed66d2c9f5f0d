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

void launch_fill_u32(uint32_t *p, uint32_t v, int64_t n, hipStream_t s);

// outerBoundingBox: writes min xyz, max xyz to out6.  partials holds
// 6 * kBBoxBlocks floats; counter is one zeroed word (left zeroed on exit).
constexpr int kBBoxBlocks = 512;
void launch_bbox(const float *x, const float *y, const float *z, int64_t n, float *partials,
                 uint32_t *counter, float *out6, hipStream_t s);

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
                          int32_t *kparent, uint32_t *scan_ws, int4 *topo, uint32_t *node_count,
                          hipStream_t s);

// Bottom-up LBVH summaries with atomic arrival flags: COM, mass, bbox and the
// MAC record (bbox centre, squared diagonal).  max_nodes bounds the launch.
void launch_lbvh_summaries(const float4 *bodies, int64_t n, const int4 *topo, const uint32_t *node_count,
                           int64_t max_nodes, uint32_t *flags, float4 *com, float4 *bmin, float4 *bmax,
                           float4 *mac, uint32_t *err, hipStream_t s);

struct WalkArgs {
    int64_t n;
    const float4 *bodies;   // sorted (x, y, z, m)
    const uint32_t *perm;   // sorted position -> original index
    const float4 *mac;      // per node MAC record
    const float4 *com;      // per node (com, mass)
    const int4 *topo;       // per node (first, count, child, parent)
    float theta;            // octree MAC
    float theta2;           // LBVH MAC: theta * theta (float)
    float G, eps2;
    int64_t max_nodes;      // node array capacity (index guard)
    float *ax, *ay, *az;    // original order
    unsigned long long *counts;  // [pp, np]
    uint32_t *err;
};

// Barnes-Hut traversal over the LBVH.  ordered != 0 selects the
// bit-exact reference arithmetic.
void launch_walk_lbvh(const WalkArgs &a, int ordered, int max_depth, hipStream_t s);

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
