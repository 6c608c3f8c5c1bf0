// gcs_sort.hpp -- stable LSD radix sort of (u64 key, u32 payload), wave64/LDS.
//
// Replaces cub::DeviceRadixSort::SortPairs (cuda_engine.cu:397): ascending,
// stable, all key bits.  Eight passes of 8-bit digits; each pass is
//   k_sort_count    per-tile digit histogram (LDS atomics)
//   k_scan_*        exclusive scan of all tile histograms (3 kernels)
//   k_sort_scatter  stable rank within the tile by 64-lane digit matching
//                   (8 ballots), staged through LDS in sorted order, then
//                   written out in runs that are contiguous in memory.
// Segmented: tiles never span instances, and the histogram of instance s is
// laid out digit-major at hist_base (digit d, tile j -> hist_base + d*T_s + j),
// so one exclusive scan over the whole array yields, for every instance, its
// own offsets starting at the instance's first particle.  Every instance is
// therefore sorted within its own range by one launch sequence.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

__global__ __launch_bounds__(kTileThreads) void k_sort_count(const TileInfo* __restrict__ tiles,
                                                             const uint64_t* __restrict__ keys,
                                                             uint32_t* __restrict__ hist, int shift)
{
    __shared__ uint32_t h[kRadix];
    const TileInfo t = tiles[blockIdx.x];
    for (int d = threadIdx.x; d < kRadix; d += kTileThreads) h[d] = 0u;
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
    for (int d = threadIdx.x; d < kRadix; d += kTileThreads)
        hist[t.hist_base + d * t.seg_tiles + t.tile_in_seg] = h[d];
}

// Block-wide exclusive scan of one value per thread (256 threads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total,
                                                         uint32_t* wsum /* [4] shared */)
{
    const int l = lane_id(), w = threadIdx.x >> 6;
    uint32_t incl = (uint32_t)wave_scan_incl((int)v);
    if (l == 63) wsum[w] = incl;
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

// Sum of each kScanChunk-entry chunk.
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                     uint32_t* __restrict__ partial)
{
    __shared__ uint32_t ws[4];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk;
    uint32_t s = 0;
    for (int k = threadIdx.x; k < kScanChunk; k += 256) {
        const int64_t i = base + k;
        if (i < n) s += in[i];
    }
    s = (uint32_t)wave_sum((int)s);
    if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Exclusive scan of the chunk sums in one block, 4096 at a time with a
// running carry (any number of chunks).
__global__ __launch_bounds__(256) void k_scan_partials(uint32_t* partial, int nb)
{
    __shared__ uint32_t ws[4];
    constexpr int kPer = 16;
    uint32_t carry = 0;
    for (int base = 0; base < nb; base += 256 * kPer) {
        uint32_t v[kPer];
        uint32_t s = 0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = base + (int)threadIdx.x * kPer + k;
            v[k] = i < nb ? partial[i] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t run = block_exclusive_scan(s, &tot, ws) + carry;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int i = base + (int)threadIdx.x * kPer + k;
            if (i < nb) partial[i] = run;
            run += v[k];
        }
        carry += tot;
    }
}

// Exclusive scan of each chunk plus its chunk offset, in place.
__global__ __launch_bounds__(256) void k_scan_apply(uint32_t* data, int64_t n,
                                                    const uint32_t* __restrict__ partial)
{
    __shared__ uint32_t ws[4];
    constexpr int kPer = kScanChunk / 256;
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kPer;
    uint32_t v[kPer];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t i = base + k;
        v[k] = i < n ? data[i] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t run = block_exclusive_scan(s, &tot, ws) + partial[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int64_t i = base + k;
        if (i < n) data[i] = run;
        run += v[k];
    }
}

// Stable scatter of one pass.  Wave w owns tile items [w*256, w*256+256) in
// four rounds of 64; within a round, lanes holding the same digit are found
// with eight ballots and ranked by lane order, and a per-wave LDS counter per
// digit carries the rank across rounds.  Tile order = wave order, round order,
// lane order = input order, so equal digits keep their input order.
__global__ __launch_bounds__(kTileThreads) void k_sort_scatter(
    const TileInfo* __restrict__ tiles, const uint64_t* __restrict__ kin,
    const uint32_t* __restrict__ vin, uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
    const uint32_t* __restrict__ offs, int shift)
{
    __shared__ uint32_t wcnt[4][kRadix];
    __shared__ uint32_t dstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint64_t skey[kTile];
    __shared__ uint32_t sval[kTile];
    __shared__ uint32_t ws[4];

    const TileInfo t = tiles[blockIdx.x];
    const int l = lane_id(), w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < 4 * kRadix; d += kTileThreads) (&wcnt[0][0])[d] = 0u;
    __syncthreads();

    uint64_t key[kTileItems];
    uint32_t val[kTileItems];
    uint32_t dig[kTileItems];
    uint32_t rank[kTileItems];
    const uint64_t lt = (1ull << l) - 1ull;
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        const int idx = w * (kWave * kTileItems) + r * kWave + l;
        const bool valid = idx < t.count;
        key[r] = valid ? kin[t.start + idx] : 0ull;
        val[r] = valid ? vin[t.start + idx] : 0u;
        const uint32_t d = (uint32_t)(key[r] >> shift) & (kRadix - 1);
        dig[r] = d;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t bb = ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const uint32_t old = wcnt[w][d];
        rank[r] = old + (uint32_t)__popcll(m & lt);
        __builtin_amdgcn_wave_barrier();
        if (valid && (m & lt) == 0ull) wcnt[w][d] = old + (uint32_t)__popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (!valid) rank[r] = 0xFFFFFFFFu;
    }
    __syncthreads();
    // Per digit: offsets of the waves within the tile, tile total, and the
    // tile-local start of the digit.
    uint32_t tot_d = 0;
    for (int d = threadIdx.x; d < kRadix; d += kTileThreads) {
        uint32_t run = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t c = wcnt[k][d];
            wcnt[k][d] = run;
            run += c;
        }
        tot_d = run;
        gbase[d] = offs[t.hist_base + d * t.seg_tiles + t.tile_in_seg];
    }
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(tot_d, &total, ws);
    dstart[threadIdx.x] = ex;
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
        const uint32_t d = (uint32_t)(k >> shift) & (kRadix - 1);
        const uint32_t g = gbase[d] + ((uint32_t)i - dstart[d]);
        kout[g] = k;
        vout[g] = sval[i];
    }
}

}  // namespace gcs
