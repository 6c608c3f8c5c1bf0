// gcs_common.hpp -- layouts shared by the host engine and the gfx950 kernels.
//
// Device state (structure of arrays, one entry per particle, all instances
// back to back; each instance's particles stay contiguous and in Morton order):
//   posm   float4  x, y, z (FP32 "hi" position), mass
//   vel    float4  vx, vy, vz, bit pattern of (type | r<<8 | g<<16 | b<<24)
//   plo    float4  x_lo, y_lo, z_lo, 0   (compensated precision only)
//   acc    float4  ax, ay, az, potential
// Tree of instance s with n_s particles starting at offset o_s occupies node
// indices [2*o_s - s, 2*o_s - s + 2*n_s - 1): internal nodes first, then
// leaves, i.e. the reference's numbering (karras_bvh.cuh:114-124) shifted.
//   nodeA  float4  com x, y, z, mass
//   nodeB  int4    opening value (float bits), first, last, split
//                  (particle indices, global; split = -1 for a leaf)
//   nodeQ  2 x float4  Qxx, Qyy, Qzz, Qxy | Qxz, Qyz, width, 0
//   boxLo/boxHi  float4  AABB of internal nodes
//   parent int     parent node (-1 for a root)
//   flag   int     arrival counter of the bottom-up aggregation
#pragma once
#include <cstdint>

namespace gcs {

constexpr int kWave = 64;
constexpr int kTile = 1024;          // particles per sort tile
constexpr int kTileThreads = 256;    // threads per sort block
constexpr int kTileItems = kTile / kTileThreads;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortPasses = 8;       // 64 bits (bit 63 of a Morton key is 0)
constexpr int kGroup = 64;           // default largest traversal group (one wave)
constexpr int kGroupMax = 192;       // largest group a wave takes (3 particles per lane)
constexpr int kGroupChunk = 128;    // quarter tiles per group-numbering chunk
constexpr int kChunkTiles = 32;      // tiles per sort-histogram and aggregation chunk
constexpr int kTileShift = 10;       // kTile == 1 << kTileShift
// Radix-sort blocks: runs of at most kSortItems x (threads) keys of one
// instance (aligned to that size within it): kSortTileBig keys per block of
// 1,024 threads for sorts of at least kSortBigMin keys, else kTile keys per
// block of 256 threads.
constexpr int kSortItems = 4;
constexpr int kSortTileBig = 4096;
constexpr int64_t kSortBigMin = (int64_t)3 << 18;
static_assert(kSortTileBig == kSortItems * 1024 && kTile == kSortItems * kTileThreads, "tiles");
static_assert(kTile == 1 << kTileShift, "shift");
// Aggregation subtrees: the maximal subtrees of at most kATile particles are
// finished one per block, the nodes above them by the instance pass.
constexpr int kATile = kTile;
// Instance pass in two levels (for instances of at least kMidSize
// particles): the nodes above the aggregation subtrees inside each maximal
// subtree of fewer than kMidSize particles by one block each, then the nodes
// of at least kMidSize particles by one block per instance.
constexpr int kMidShift = 15;
constexpr int kMidSize = 1 << kMidShift;
constexpr int kStackCap = 8192;      // per-wave traversal stack entries
constexpr int kOrderClasses = 8;     // traversal dispatch classes by group size
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, blocks dealt round-robin
// Most particles an engine holds: node records are addressed as 2 * v with a
// 32-bit node index v < 2 n, so 4 n must stay below 2^31.
constexpr int64_t kMaxParticles = (int64_t)1 << 29;

// Sort tile: a run of at most kTile particles of one instance.
struct TileInfo {
    int start;      // first particle (global index)
    int count;      // particles in the tile
    int seg;        // instance slot
    int tile_in_seg;
    int pad0, pad1, pad2, pad3;
};

// Instance slot.
struct SegInfo {
    int start;      // first particle
    int n;          // particles
    int int_base;   // internal node of particle p: p + int_base (= start - slot)
    int leaf_base;  // leaf node of particle p: p + leaf_base (= start + n - slot - 1)
    int root;       // 2*start - slot
    int tile0;      // global index of the first tile
    int ntiles;
    int chunk0;     // first sort-histogram chunk (32 tiles each)
};

// Aggregation chunk: kChunkTiles consecutive tiles of one instance (also the
// sort's histogram chunk).
// Aggregation subtree: a maximal subtree of at most kATile particles.
struct ATileInfo {
    int first;  // first particle (global index)
    int count;  // particles
    int seg;    // instance slot
    int root;   // its root node (a leaf when count == 1)
};

// Traversal group: up to group_size (<= kGroupMax) consecutive particles of one
// instance.
struct GroupInfo {
    int first;
    int count;
    int seg;
    int pad;
};

// Error bits of the device error word.
enum : unsigned {
    kErrStack = 1u,
    kErrNonFinite = 2u,
};

}  // namespace gcs
