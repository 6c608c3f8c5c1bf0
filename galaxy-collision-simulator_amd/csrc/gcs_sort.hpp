// gcs_sort.hpp -- stable LSD radix sort of (u64 key, u32 payload), wave64/LDS.
//
// Replaces cub::DeviceRadixSort::SortPairs (cuda_engine.cu:397): ascending,
// stable, all key bits.  Eight passes of 8-bit digits, two kernels per pass,
// each block one sort tile of at most 4 x THREADS keys of one instance
// (THREADS = 1,024 for large sorts, 256 when 4,096-key tiles would leave the
// GPU's 256 CUs short of blocks):
//   k_sort_count    per-tile digit histogram (LDS atomics into four copies),
//                   stored per tile and added into the histogram of the
//                   tile's chunk (32,768 consecutive keys of one instance)
//   k_sort_scatter  each digit's output offset for the tile -- instance start
//                   + keys of smaller digits in the instance (from the chunk
//                   histograms) + keys of this digit in earlier chunks and in
//                   earlier tiles of the tile's chunk -- then the stable rank
//                   within the tile by 64-lane digit matching (8 ballots),
//                   staged through LDS in sorted order, and written out in
//                   runs that are contiguous in memory (about 16 keys per
//                   digit and 4,096-key tile for uniform digits).
// No global scan: a tile reads at most (chunks of its instance) + 31 rows of
// 256 counts, which stay in L2.  Segmented: tiles never span instances and
// every offset is relative to the instance's first particle, so every
// instance is sorted within its own range by one launch sequence.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_sort_count(const TileInfo* __restrict__ tiles,
                                                        const SegInfo* __restrict__ segs,
                                                        const uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ thist,
                                                        uint32_t* __restrict__ chist, int shift)
{
    constexpr int kTileKeys = kSortItems * THREADS;
    constexpr int kChunkST = kChunkTiles * kTile / kTileKeys;
    static_assert(THREADS >= kRadix && kChunkTiles * kTile % kTileKeys == 0, "tiles");
    __shared__ uint32_t h[4][kRadix];
    const int tile = blockIdx.x;
    const TileInfo t = tiles[tile];
    const int tid = (int)threadIdx.x;
    for (int k = tid; k < 4 * kRadix; k += THREADS) (&h[0][0])[k] = 0u;
    __syncthreads();
    uint32_t* const hw = h[(tid >> 6) & 3];
#pragma unroll
    for (int it = 0; it < kSortItems; ++it) {
        const int idx = it * THREADS + tid;
        if (idx < t.count) {
            const uint32_t d = (uint32_t)(keys[t.start + idx] >> shift) & (kRadix - 1);
            atomicAdd(&hw[d], 1u);
        }
    }
    __syncthreads();
    if (tid < kRadix) {
        const uint32_t c = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
        thist[(size_t)tile * kRadix + tid] = c;
        if (c != 0u) {
            const int chunk = segs[t.seg].chunk0 + t.tile_in_seg / kChunkST;
            atomicAdd(&chist[(size_t)chunk * kRadix + tid], c);
        }
    }
}

// Exclusive scan over the first kRadix threads of the block (one digit
// each); every thread of the block calls it.
__device__ __forceinline__ uint32_t digit_exclusive_scan(uint32_t v, uint32_t* total,
                                                         uint32_t* wsum)
{
    static_assert(kRadix == 4 * kWave, "four waves of digits");
    const int l = lane_id(), w = threadIdx.x >> 6;
    const uint32_t incl = (uint32_t)wave_scan_incl((int)v);
    if (w < 4 && l == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t s = wsum[k];
        if (k < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

// Stable scatter of one pass.  Wave w owns tile items [w*256, w*256+256) in
// four rounds of 64; within a round, lanes holding the same digit are found
// with eight ballots and ranked by lane order, and a per-wave LDS counter per
// digit carries the rank across rounds.  Tile order = wave order, round order,
// lane order = input order, so equal digits keep their input order.
template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_sort_scatter(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, const uint32_t* __restrict__ thist,
    const uint32_t* __restrict__ chist, int shift)
{
    constexpr int kTileKeys = kSortItems * THREADS;
    constexpr int kWaves = THREADS / kWave;
    constexpr int kChunkST = kChunkTiles * kTile / kTileKeys;
    __shared__ uint32_t wcnt[kWaves][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint64_t skey[kTileKeys];
    __shared__ uint32_t sval[kTileKeys];
    __shared__ uint32_t ws[4];

    const int tile = blockIdx.x;
    const TileInfo t = tiles[tile];
    const int tid = (int)threadIdx.x;
    const int l = lane_id(), w = tid >> 6;
    for (int d = tid; d < kWaves * kRadix; d += THREADS) (&wcnt[0][0])[d] = 0u;

    // The tile's keys, loaded first so they are in flight with the offsets.
    uint64_t key[kSortItems];
    uint32_t val[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = w * (kWave * kSortItems) + r * kWave + l;
        const bool valid = idx < t.count;
        key[r] = valid ? kin[t.start + idx] : 0ull;
        val[r] = valid ? vin[t.start + idx] : 0u;
    }

    // Output offset of digit d (= this thread, d < kRadix) for the tile.
    const int d = tid & (kRadix - 1);
    const bool dthread = tid < kRadix;
    const SegInfo sg = segs[t.seg];
    uint32_t seg_tot = 0u, before = 0u;
    if (dthread) {
        const int nch = (sg.ntiles + kChunkTiles - 1) / kChunkTiles;
        const int myc = t.tile_in_seg / kChunkST;
        // All loads of a batch are issued before any is used (one L2 round
        // trip per 16 chunks, one for the tile rows).
        const uint32_t* ch = chist + (size_t)sg.chunk0 * kRadix + d;
        for (int c0 = 0; c0 < nch; c0 += 16) {
            uint32_t v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = c0 + u < nch ? ch[(size_t)(c0 + u) * kRadix] : 0u;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                seg_tot += v[u];
                before += c0 + u < myc ? v[u] : 0u;
            }
        }
        const int tfirst = tile - t.tile_in_seg % kChunkST;
        uint32_t v[kChunkST - 1];
#pragma unroll
        for (int j = 0; j < kChunkST - 1; ++j)
            v[j] = tfirst + j < tile ? thist[(size_t)(tfirst + j) * kRadix + d] : 0u;
#pragma unroll
        for (int j = 0; j < kChunkST - 1; ++j) before += v[j];
    }
    __syncthreads();

    uint32_t dig[kSortItems];
    uint32_t rank[kSortItems];
    const uint64_t lt = (1ull << l) - 1ull;
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = w * (kWave * kSortItems) + r * kWave + l;
        const bool valid = idx < t.count;
        const uint32_t dg = (uint32_t)(key[r] >> shift) & (kRadix - 1);
        dig[r] = dg;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t bb = ballot((dg >> b) & 1u);
            m &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t old = wcnt[w][dg];
        rank[r] = old + (uint32_t)__popcll(m & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && (m & lt) == 0ull) wcnt[w][dg] = old + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (!valid) rank[r] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // Per digit: offsets of the waves within the tile and the tile total.
    uint32_t tot_d = 0;
    if (dthread) {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < kWaves; ++k) {
            const uint32_t c = wcnt[k][d];
            wcnt[k][d] = run;
            run += c;
        }
        tot_d = run;
    }
    uint32_t total;
    const uint32_t ex = digit_exclusive_scan(tot_d, &total, ws);
    const uint32_t dbase = digit_exclusive_scan(seg_tot, &total, ws);
    if (dthread) {
        dstart[d] = ex;
        gbase[d] = (uint32_t)sg.start + dbase + before;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        if (rank[r] != 0xFFFFFFFFu) {
            const uint32_t pos = dstart[dig[r]] + wcnt[w][dig[r]] + rank[r];
            skey[pos] = key[r];
            sval[pos] = val[r];
        }
    }
    __syncthreads();
    for (int i = tid; i < t.count; i += THREADS) {
        const uint64_t k = skey[i];
        const uint32_t dg = (uint32_t)(k >> shift) & (kRadix - 1);
        const uint32_t g = gbase[dg] + ((uint32_t)i - dstart[dg]);
        kout[g] = k;
        vout[g] = sval[i];
    }
}

}  // namespace gcs
