// gcs_sort.hpp -- stable LSD radix sort of (u64 key, u32 payload), wave64/LDS.
//
// Replaces cub::DeviceRadixSort::SortPairs (cuda_engine.cu:397): ascending,
// stable, all key bits.  Eight passes of 8-bit digits, two kernels per pass:
//   k_sort_count    per-tile digit histogram (LDS atomics), stored per tile
//                   and added into the histogram of the tile's chunk (32
//                   consecutive tiles of one instance)
//   k_sort_scatter  each digit's output offset for the tile -- instance start
//                   + keys of smaller digits in the instance (from the chunk
//                   histograms) + keys of this digit in earlier chunks and in
//                   earlier tiles of the tile's chunk -- then the stable rank
//                   within the tile by 64-lane digit matching (8 ballots),
//                   staged through LDS in sorted order, and written out in
//                   runs that are contiguous in memory.
// No global scan: a tile reads at most (chunks of its instance) + 31 rows of
// 256 counts, which stay in L2.  Segmented: tiles never span instances and
// every offset is relative to the instance's first particle, so every
// instance is sorted within its own range by one launch sequence.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

__global__ __launch_bounds__(kTileThreads) void k_sort_count(const TileInfo* __restrict__ tiles,
                                                             const SegInfo* __restrict__ segs,
                                                             const uint64_t* __restrict__ keys,
                                                             uint32_t* __restrict__ thist,
                                                             uint32_t* __restrict__ chist,
                                                             int shift)
{
    static_assert(kRadix == kTileThreads, "one digit per thread");
    __shared__ uint32_t h[kRadix];
    const int tile = blockIdx.x;
    const TileInfo t = tiles[tile];
    h[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int it = 0; it < kTileItems; ++it) {
        const int idx = it * kTileThreads + (int)threadIdx.x;
        if (idx < t.count) {
            const uint32_t d = (uint32_t)(keys[t.start + idx] >> shift) & (kRadix - 1);
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    const uint32_t c = h[threadIdx.x];
    thist[(size_t)tile * kRadix + threadIdx.x] = c;
    if (c != 0u) {
        const int chunk = segs[t.seg].chunk0 + t.tile_in_seg / kChunkTiles;
        atomicAdd(&chist[(size_t)chunk * kRadix + threadIdx.x], c);
    }
}

// Stable scatter of one pass.  Wave w owns tile items [w*256, w*256+256) in
// four rounds of 64; within a round, lanes holding the same digit are found
// with eight ballots and ranked by lane order, and a per-wave LDS counter per
// digit carries the rank across rounds.  Tile order = wave order, round order,
// lane order = input order, so equal digits keep their input order.
__global__ __launch_bounds__(kTileThreads) void k_sort_scatter(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, const uint32_t* __restrict__ thist,
    const uint32_t* __restrict__ chist, int shift)
{
    __shared__ uint32_t wcnt[4][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint64_t skey[kTile];
    __shared__ uint32_t sval[kTile];
    __shared__ uint32_t ws[4];

    const int tile = blockIdx.x;
    const TileInfo t = tiles[tile];
    const int l = lane_id(), w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 4 * kRadix; d += kTileThreads) (&wcnt[0][0])[d] = 0u;

    // The tile's keys, loaded first so they are in flight with the offsets.
    uint64_t key[kTileItems];
    uint32_t val[kTileItems];
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        const int idx = w * (kWave * kTileItems) + r * kWave + l;
        const bool valid = idx < t.count;
        key[r] = valid ? kin[t.start + idx] : 0ull;
        val[r] = valid ? vin[t.start + idx] : 0u;
    }

    // Output offset of digit d (= this thread) for the tile.
    const int d = threadIdx.x;
    const SegInfo sg = segs[t.seg];
    const int nch = (sg.ntiles + kChunkTiles - 1) / kChunkTiles;
    const int myc = t.tile_in_seg / kChunkTiles;
    // All loads of a batch are issued before any is used (one L2 round trip
    // per 16 chunks, one for the tile rows).
    const uint32_t* ch = chist + (size_t)sg.chunk0 * kRadix + d;
    uint32_t seg_tot = 0u, before = 0u;
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
    const int tfirst = tile - t.tile_in_seg % kChunkTiles;
    {
        uint32_t v[kChunkTiles - 1];
#pragma unroll
        for (int j = 0; j < kChunkTiles - 1; ++j)
            v[j] = tfirst + j < tile ? thist[(size_t)(tfirst + j) * kRadix + d] : 0u;
#pragma unroll
        for (int j = 0; j < kChunkTiles - 1; ++j) before += v[j];
    }
    __syncthreads();

    uint32_t dig[kTileItems];
    uint32_t rank[kTileItems];
    const uint64_t lt = (1ull << l) - 1ull;
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        const int idx = w * (kWave * kTileItems) + r * kWave + l;
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
    {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t c = wcnt[k][d];
            wcnt[k][d] = run;
            run += c;
        }
        tot_d = run;
    }
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(tot_d, &total, ws);
    const uint32_t dbase = block_exclusive_scan(seg_tot, &total, ws);
    dstart[d] = ex;
    gbase[d] = (uint32_t)sg.start + dbase + before;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        if (rank[r] != 0xFFFFFFFFu) {
            const uint32_t pos = dstart[dig[r]] + wcnt[w][dig[r]] + rank[r];
            skey[pos] = key[r];
            sval[pos] = val[r];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < t.count; i += kTileThreads) {
        const uint64_t k = skey[i];
        const uint32_t dg = (uint32_t)(k >> shift) & (kRadix - 1);
        const uint32_t g = gbase[dg] + ((uint32_t)i - dstart[dg]);
        kout[g] = k;
        vout[g] = sval[i];
    }
}

}  // namespace gcs
