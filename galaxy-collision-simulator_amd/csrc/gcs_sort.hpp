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

// ---- Bucket sort for keys in nearly sorted order --------------------------
//
// After a step the particles are stored in that step's sorted order, and at
// the next step most keys are within a few hundred places of their sorted
// rank (2 x 1,048,576 particles, dt 0.005: the median moves 88 places, 0.1 %
// move across the key space).  Splitters taken from the keys at every
// kBucketKeys-th place of the current order then cut each instance into
// buckets of about kBucketKeys keys (at most 1,762 there over 20 steps), and
// the sort is
//   k_bucket_cand     per bucket: a splitter candidate from the keys around
//                     every kBucketKeys-th place; bucket counters cleared
//   k_bucket_split    per instance: candidates filtered and made
//                     non-decreasing
//   k_bucket_count    per tile: each key's bucket (a search among the
//                     tile's window of kBucketWin splitters in LDS, or among
//                     all of the instance's for a key outside it), counted
//   k_bucket_scan     per instance: bucket offsets
//   k_bucket_scatter  per tile: keys and their indices moved into their
//                     buckets, in any order within a bucket
//   k_bucket_sort     per bucket: sorted by (key, index) in LDS (a sorting
//                     network of four per thread, then merge-path rounds).
// (key, index) is a total order, so every bucket's order -- and the result
// -- is the stable sort's, whatever order the scatter left.  A bucket of more
// than kBucketCap keys (splitters that fit the keys badly) is sorted in
// kBucketCap-key LDS chunks merged in global memory: slow, still exact.
// Bucket b of an instance holds the keys k with split[b] <= k < split[b+1]
// (split[0] = 0, split[nb] = +inf): its index is the number of splitters
// 1 .. nb-1 at most k.
constexpr int kBucketKeys = 1536;  // splitter spacing
constexpr int kBucketCap = 2048;   // keys a block sorts in LDS
constexpr int kBucketWin = 64;     // splitters per tile in LDS
constexpr int kBucketThreads = 512;
constexpr int kBucketPer = kBucketCap / kBucketThreads;
static_assert(kBucketPer == 4, "the register network sorts four");

__host__ __device__ __forceinline__ uint64_t min_u64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return a < b ? b : a; }

__host__ __device__ __forceinline__ void swap_u64(uint64_t& a, uint64_t& b)
{
    const uint64_t t = a;
    a = b;
    b = t;
}

// Median of nine: odd-even transposition sort (nine rounds of fixed
// compare-exchanges, unrolled into registers).
__host__ __device__ __forceinline__ uint64_t median9(uint64_t* a)
{
#pragma unroll
    for (int r = 0; r < 9; ++r)
#pragma unroll
        for (int j = r & 1; j + 1 < 9; j += 2) {
            const uint64_t lo = min_u64(a[j], a[j + 1]), hi = max_u64(a[j], a[j + 1]);
            a[j] = lo;
            a[j + 1] = hi;
        }
    return a[4];
}

// Splitter candidate b of an instance's keys k[0, n): the median of the
// nine keys at places b * kBucketKeys + 64 m, m = -4 .. 4 (clamped); 0 for b
// = 0 and +inf for b = nb.  Keys that crossed a coarse octree boundary
// together move as runs of neighbours in the current order; nine points
// 256 places either side keep the median off such a run.
// The nine keys are read whatever b is (places clamped into the instance,
// n >= 1) and the result chosen after, so no branch holds the loads back.
__host__ __device__ __forceinline__ uint64_t split_candidate(const uint64_t* k, int n, int b,
                                                             int nb)
{
    uint64_t a[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) {
        const int q = b * kBucketKeys + 64 * (m - 4);
        a[m] = k[q < 0 ? 0 : (q >= n ? n - 1 : q)];
    }
    const uint64_t med = median9(a);
    return b <= 0 ? 0ull : (b >= nb ? ~0ull : med);
}

// Candidate of every bucket b < nbkt (one thread each, over all instances;
// bseg[b] = its instance), and the bucket counters cleared.
__global__ __launch_bounds__(256) void k_bucket_cand(const SegInfo* __restrict__ segs,
                                                     const int* __restrict__ bkt0,
                                                     const int* __restrict__ bseg, int nbkt,
                                                     const uint64_t* __restrict__ keys,
                                                     uint64_t* __restrict__ cand,
                                                     uint32_t* __restrict__ bcnt,
                                                     uint32_t* __restrict__ bfill)
{
    const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (b >= nbkt) return;
    const int s = bseg[b];
    const SegInfo sg = segs[s];
    const int lo = bkt0[s], hi = bkt0[s + 1];
    cand[b] = split_candidate(keys + sg.start, sg.n, b - lo, hi - lo);
    bcnt[b] = 0u;
    bfill[b] = 0u;
}

// Running maximum over the block's threads, exclusive: the maximum of v over
// the threads before this one (0 for the first), and in *total over all of
// them.  Within each wave by shuffles, then over the waves' totals in LDS;
// every thread calls it.
__device__ __forceinline__ uint64_t block_max_before(uint64_t v, uint64_t* wtot, uint64_t* total)
{
    const int l = lane_id(), w = (int)threadIdx.x >> 6, nw = (int)blockDim.x >> 6;
    uint64_t incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const uint64_t u = __shfl_up(incl, o);
        incl = l >= o ? max_u64(incl, u) : incl;
    }
    const uint64_t prev = __shfl_up(incl, 1);
    if (l == kWave - 1) wtot[w] = incl;
    __syncthreads();
    uint64_t before = l > 0 ? prev : 0ull, all = 0ull;
    for (int q = 0; q < nw; ++q) {
        const uint64_t t = wtot[q];
        before = q < w ? max_u64(before, t) : before;
        all = max_u64(all, t);
    }
    __syncthreads();  // wtot is rewritten by the next call
    *total = all;
    return before;
}

// Splitter b: the median of candidates b - 1, b, b + 1 (one bad candidate
// never reaches the running maximum), made non-decreasing by a running
// maximum over the instance.  Each thread takes four consecutive buckets.
__global__ __launch_bounds__(1024) void k_bucket_split(const int* __restrict__ bkt0,
                                                       const uint64_t* __restrict__ cand,
                                                       uint64_t* __restrict__ split)
{
    __shared__ uint64_t wtot[16];
    const int s = blockIdx.x;
    const int b0 = bkt0[s], nb = bkt0[s + 1] - b0;
    const uint64_t* c = cand + b0;
    uint64_t carry = 0ull;
    for (int c0 = 0; c0 < nb; c0 += 4 * 1024) {
        const int bb = c0 + 4 * (int)threadIdx.x;
        // The six candidates around the thread's four buckets in one round
        // trip (places clamped into the instance; the clamped ones are not used).
        uint64_t w[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) w[e] = c[min(max(bb - 1 + e, 0), nb - 1)];
        uint64_t v[4];
        uint64_t run = 0ull;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int b = bb + e;
            uint64_t x = 0ull;
            if (b > 0 && b < nb) {
                x = w[e];
                uint64_t y = w[e + 1];
                const uint64_t z = b + 1 < nb ? w[e + 2] : ~0ull;
                if (x > y) swap_u64(x, y);
                x = max_u64(x, min_u64(y, z));  // median of three
            }
            run = max_u64(run, x);
            v[e] = run;
        }
        uint64_t total;
        const uint64_t before = max_u64(carry, block_max_before(run, wtot, &total));
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (bb + e < nb) split[b0 + bb + e] = max_u64(before, v[e]);
        carry = max_u64(carry, total);
    }
}

// The window of splitters of the tile starting at place j0 of an instance
// of nb buckets: buckets [wbase, wbase + wn).
__host__ __device__ __forceinline__ void bucket_window(int j0, int nb, int& wbase, int& wn)
{
    wn = nb < kBucketWin ? nb : kBucketWin;
    const int g = j0 / kBucketKeys - kBucketWin / 4;
    wbase = g < 0 ? 0 : (g > nb - wn ? nb - wn : g);
}

// Local bucket of key k: the number of splitters 1 .. nb-1 at most k.
__host__ __device__ __forceinline__ int bucket_of(uint64_t k, const uint64_t* ls, int wbase, int wn,
                                         const uint64_t* __restrict__ split, int nb, bool& inwin)
{
    int lo, hi;
    inwin = k >= ls[0] && k < ls[wn];
    if (inwin) {
        lo = 1;
        hi = wn;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (ls[mid] <= k) lo = mid + 1;
            else hi = mid;
        }
        return wbase + lo - 1;
    }
    lo = 1;
    hi = nb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (split[mid] <= k) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

// The window's splitters in LDS: ls[w] = split[wbase + w] for w <= wn
// (+inf past the last bucket; split[0] = 0).
__device__ __forceinline__ void load_window(uint64_t* ls, const uint64_t* __restrict__ split,
                                            int wbase, int wn, int nb)
{
    for (int w = (int)threadIdx.x; w <= wn; w += blockDim.x)
        ls[w] = wbase + w < nb ? split[wbase + w] : ~0ull;
}

__global__ __launch_bounds__(kTileThreads) void k_bucket_count(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const int* __restrict__ bkt0, const uint64_t* __restrict__ keys,
    const uint64_t* __restrict__ split, uint32_t* __restrict__ bid, uint32_t* __restrict__ bcnt)
{
    __shared__ uint64_t ls[kBucketWin + 1];
    __shared__ uint32_t lc[kBucketWin];
    const TileInfo t = tiles[blockIdx.x];
    const int tid = (int)threadIdx.x;
    // The tile's keys in flight while the window is staged.
    uint64_t key[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = r * kTileThreads + tid;
        key[r] = idx < t.count ? keys[t.start + idx] : 0ull;
    }
    const SegInfo sg = segs[t.seg];
    const int b0 = bkt0[t.seg], nb = bkt0[t.seg + 1] - b0;
    int wbase, wn;
    bucket_window(t.start - sg.start, nb, wbase, wn);
    load_window(ls, split + b0, wbase, wn, nb);
    for (int w = tid; w < wn; w += kTileThreads) lc[w] = 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = r * kTileThreads + tid;
        if (idx < t.count) {
            const uint64_t k = key[r];
            bool inwin;
            const int b = bucket_of(k, ls, wbase, wn, split + b0, nb, inwin);
            if (inwin) atomicAdd(&lc[b - wbase], 1u);
            else atomicAdd(&bcnt[b0 + b], 1u);
            bid[t.start + idx] = (uint32_t)(b0 + b);
        }
    }
    __syncthreads();
    for (int w = tid; w < wn; w += kTileThreads)
        if (lc[w] != 0u) atomicAdd(&bcnt[b0 + wbase + w], lc[w]);
}

// Per instance: boff[b] = the instance's first particle + the keys of its
// earlier buckets.
__global__ __launch_bounds__(1024) void k_bucket_scan(const SegInfo* __restrict__ segs,
                                                      const int* __restrict__ bkt0,
                                                      const uint32_t* __restrict__ bcnt,
                                                      uint32_t* __restrict__ boff)
{
    __shared__ uint32_t wsum[16];
    const int s = blockIdx.x;
    const int tid = (int)threadIdx.x;
    const int l = lane_id(), w = tid >> 6;
    const int b0 = bkt0[s], nb = bkt0[s + 1] - b0;
    uint32_t carry = (uint32_t)segs[s].start;
    for (int c0 = 0; c0 < nb; c0 += 1024) {
        const int b = c0 + tid;
        const uint32_t v = b < nb ? bcnt[b0 + b] : 0u;
        const uint32_t incl = (uint32_t)wave_scan_incl((int)v);
        if (l == 63) wsum[w] = incl;
        __syncthreads();
        uint32_t base = carry, tot = 0u;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const uint32_t x = wsum[k];
            if (k < w) base += x;
            tot += x;
        }
        if (b < nb) boff[b0 + b] = base + incl - v;
        carry += tot;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kTileThreads) void k_bucket_scatter(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const int* __restrict__ bkt0, const uint64_t* __restrict__ keys,
    const uint32_t* __restrict__ bid, const uint32_t* __restrict__ boff,
    uint32_t* __restrict__ bfill, uint64_t* __restrict__ kout, uint32_t* __restrict__ vout)
{
    __shared__ uint32_t lc[kBucketWin];
    __shared__ uint32_t lb[kBucketWin];
    const TileInfo t = tiles[blockIdx.x];
    const int tid = (int)threadIdx.x;
    // The tile's keys and buckets first, then the buckets' offsets, all in
    // flight while the window is set up and the ranks are taken.
    uint64_t key[kSortItems];
    uint32_t b[kSortItems], rank[kSortItems], base[kSortItems];
    bool win[kSortItems];
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = r * kTileThreads + tid;
        win[r] = false;
        rank[r] = 0u;
        key[r] = idx < t.count ? keys[t.start + idx] : 0ull;
        b[r] = idx < t.count ? bid[t.start + idx] : 0u;
    }
    // Unconditional, so the four go out together (b is 0 past the tile's end).
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) base[r] = boff[b[r]];
    const SegInfo sg = segs[t.seg];
    const int b0 = bkt0[t.seg], nb = bkt0[t.seg + 1] - b0;
    int wbase, wn;
    bucket_window(t.start - sg.start, nb, wbase, wn);
    for (int w = tid; w < wn; w += kTileThreads) lc[w] = 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = r * kTileThreads + tid;
        if (idx < t.count) {
            const int w = (int)b[r] - b0 - wbase;
            win[r] = w >= 0 && w < wn;
            rank[r] = win[r] ? atomicAdd(&lc[w], 1u) : atomicAdd(&bfill[b[r]], 1u);
        }
    }
    __syncthreads();
    for (int w = tid; w < wn; w += kTileThreads)
        lb[w] = lc[w] != 0u ? atomicAdd(&bfill[b0 + wbase + w], lc[w]) : 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortItems; ++r) {
        const int idx = r * kTileThreads + tid;
        if (idx < t.count) {
            const uint32_t pos =
                base[r] + rank[r] + (win[r] ? lb[(int)b[r] - b0 - wbase] : 0u);
            kout[pos] = key[r];
            vout[pos] = (uint32_t)(t.start + idx);
        }
    }
}

__host__ __device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t va, uint64_t kb, uint32_t vb)
{
    return (ka < kb) | ((ka == kb) & (va < vb));  // bitwise: no branches
}

__host__ __device__ __forceinline__ void cmp_swap(uint64_t& ka, uint32_t& va, uint64_t& kb,
                                                  uint32_t& vb)
{
    const bool sw = pair_less(kb, vb, ka, va);
    const uint64_t lk = sw ? kb : ka, hk = sw ? ka : kb;
    const uint32_t lv = sw ? vb : va, hv = sw ? va : vb;
    ka = lk;
    kb = hk;
    va = lv;
    vb = hv;
}

// Four pairs sorted by a 5-comparator network.
__host__ __device__ __forceinline__ void sort4(uint64_t* k, uint32_t* v)
{
    cmp_swap(k[0], v[0], k[1], v[1]);
    cmp_swap(k[2], v[2], k[3], v[3]);
    cmp_swap(k[0], v[0], k[2], v[2]);
    cmp_swap(k[1], v[1], k[3], v[3]);
    cmp_swap(k[1], v[1], k[2], v[2]);
}

// Outputs [base, base + kBucketPer) of a merge round over sk/sv, where sorted runs of
// w pairs are merged pairwise: the split of the merge path before the first
// output (a binary search), then kBucketPer merge steps, each loading the next
// pair of the run it took from (one LDS load, no branch; an exhausted run's
// head -- at most place kBucketCap, so the arrays hold kBucketCap + 1 -- is
// loaded but never taken).
// Ties (only the padding is equal) go to the second run in both the search
// and the merge, so every output is taken once.
__host__ __device__ __forceinline__ void merge_round(const uint64_t* sk, const uint32_t* sv,
                                                      int w, int base, uint64_t* k, uint32_t* v)
{
    const int a0 = base & ~(2 * w - 1), k0 = base & (2 * w - 1);
    int lo = k0 > w ? k0 - w : 0, hi = k0 < w ? k0 : w;
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        const int pa = (a0 + i), pb = (a0 + w + k0 - i - 1);
        const uint64_t xa = sk[pa], xb = sk[pb];
        bool right = xa < xb;
        if (xa == xb) right = sv[pa] < sv[pb];  // equal keys are rare: indices only then
        lo = right ? i + 1 : lo;
        hi = right ? hi : i;
    }
    int i = lo, j = k0 - lo;
    uint64_t ka = sk[a0 + i], kb = sk[a0 + w + j];
    uint32_t va = sv[a0 + i], vb = sv[a0 + w + j];
#pragma unroll
    for (int e = 0; e < kBucketPer; ++e) {
        const bool take_a = (j >= w) | ((i < w) & pair_less(ka, va, kb, vb));
        k[e] = take_a ? ka : kb;
        v[e] = take_a ? va : vb;
        i += take_a ? 1 : 0;
        j += take_a ? 0 : 1;
        const int q = take_a ? a0 + i : a0 + w + j;
        const uint64_t nk = sk[q];
        const uint32_t nv = sv[q];
        ka = take_a ? nk : ka;
        va = take_a ? nv : va;
        kb = take_a ? kb : nk;
        vb = take_a ? vb : nv;
    }
}

// Sorts kBucketCap (key, index) pairs in LDS, ascending; every thread of the
// block calls it (sk/sv written and synchronised on entry).  Thread t sorts
// places [4t, 4t + 4) in registers, then rounds merge sorted runs of w = 4,
// 8, ..., kBucketCap / 2 pairs: thread t computes outputs [4t, 4t + 4) into
// registers and writes them back after the block has read the round's input.
__device__ __forceinline__ void lds_merge_sort(uint64_t* sk, uint32_t* sv)
{
    const int base = kBucketPer * (int)threadIdx.x;
    uint64_t k[kBucketPer];
    uint32_t v[kBucketPer];
#pragma unroll
    for (int e = 0; e < kBucketPer; ++e) {
        k[e] = sk[base + e];
        v[e] = sv[base + e];
    }
    sort4(k, v);
#pragma unroll
    for (int e = 0; e < kBucketPer; ++e) {
        sk[base + e] = k[e];
        sv[base + e] = v[e];
    }
    __syncthreads();
    for (int w = kBucketPer; w < kBucketCap; w <<= 1) {
        merge_round(sk, sv, w, base, k, v);
        __syncthreads();  // the round's input has been read
#pragma unroll
        for (int e = 0; e < kBucketPer; ++e) {
            sk[base + e] = k[e];
            sv[base + e] = v[e];
        }
        __syncthreads();
    }
}

// One chunk of at most kBucketCap pairs from (kin, vin) sorted into (kout,
// vout) at the same places.
__device__ __forceinline__ void sort_chunk(uint64_t* sk, uint32_t* sv,
                                           const uint64_t* __restrict__ kin,
                                           const uint32_t* __restrict__ vin, uint64_t* kout,
                                           uint32_t* vout, int m)
{
    const int tid = (int)threadIdx.x;
    // All of a thread's loads are issued before any is used.
    uint64_t kr[kBucketPer];
    uint32_t vr[kBucketPer];
#pragma unroll
    for (int r = 0; r < kBucketPer; ++r) {
        const int i = tid + r * kBucketThreads;
        kr[r] = i < m ? kin[i] : ~0ull;  // keys are below 2^63: the pads sort last
        vr[r] = i < m ? vin[i] : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int r = 0; r < kBucketPer; ++r) {
        const int i = tid + r * kBucketThreads;
        sk[i] = kr[r];
        sv[i] = vr[r];
    }
    __syncthreads();
    lds_merge_sort(sk, sv);
#pragma unroll
    for (int r = 0; r < kBucketPer; ++r) {
        const int i = tid + r * kBucketThreads;
        kr[r] = sk[i];
        vr[r] = sv[i];
    }
#pragma unroll
    for (int r = 0; r < kBucketPer; ++r) {
        const int i = tid + r * kBucketThreads;
        if (i < m) {
            kout[i] = kr[r];
            vout[i] = vr[r];
        }
    }
    __syncthreads();
}

// The block merges sorted runs a[0, la) and b[0, lb) into out[0, la + lb):
// thread t writes outputs [t E, t E + E), starting from the split of the
// first t E outputs between the runs (a binary search on the merge path).
__host__ __device__ __forceinline__ void block_merge(int t, const uint64_t* ka,
                                                     const uint32_t* va, int la,
                                                     const uint64_t* kb, const uint32_t* vb,
                                                     int lb, uint64_t* ko, uint32_t* vo)
{
    const int total = la + lb;
    const int per = (total + kBucketThreads - 1) / kBucketThreads;
    const int k0 = t * per;
    if (k0 >= total) return;
    const int k1 = k0 + per < total ? k0 + per : total;
    int lo = k0 - lb > 0 ? k0 - lb : 0, hi = k0 < la ? k0 : la;
    while (lo < hi) {
        const int i = (lo + hi) >> 1;
        // Outputs before k0 take i from a and k0 - i from b: too few from a
        // while a[i] < b[k0 - i - 1].
        if (pair_less(ka[i], va[i], kb[k0 - i - 1], vb[k0 - i - 1])) lo = i + 1;
        else hi = i;
    }
    int i = lo, j = k0 - lo;
    for (int k = k0; k < k1; ++k) {
        const bool take_a = j >= lb || (i < la && pair_less(ka[i], va[i], kb[j], vb[j]));
        if (take_a) {
            ko[k] = ka[i];
            vo[k] = va[i];
            ++i;
        } else {
            ko[k] = kb[j];
            vo[k] = vb[j];
            ++j;
        }
    }
}

// Bucket b from (kin, vin) into (kout, vout) at [boff[b], boff[b] + bcnt[b]).
// More than kBucketCap keys: sorted chunks in (kout, vout), merged pairwise
// between the two buffers (both free over the bucket's places), the result
// copied into (kout, vout) if the last merge left it in (kin, vin).
__global__ __launch_bounds__(kBucketThreads, 3) void k_bucket_sort(
    const uint32_t* __restrict__ bcnt, const uint32_t* __restrict__ boff, uint64_t* kin,
    uint32_t* vin, uint64_t* kout, uint32_t* vout)
{
    __shared__ uint64_t sk[kBucketCap + 1];
    __shared__ uint32_t sv[kBucketCap + 1];
    const int b = blockIdx.x;
    const int m = (int)bcnt[b];
    if (m == 0) return;
    const size_t off = boff[b];
    if (m <= kBucketCap) {
        sort_chunk(sk, sv, kin + off, vin + off, kout + off, vout + off, m);
        return;
    }
    for (int c = 0; c < m; c += kBucketCap)
        sort_chunk(sk, sv, kin + off + c, vin + off + c, kout + off + c, vout + off + c,
                   min(kBucketCap, m - c));
    uint64_t* ks = kout + off;
    uint32_t* vs = vout + off;
    uint64_t* kd = kin + off;
    uint32_t* vd = vin + off;
    for (int w = kBucketCap; w < m; w <<= 1) {
        for (int s0 = 0; s0 < m; s0 += 2 * w) {
            const int la = min(w, m - s0), lb = min(w, m - s0 - la);
            block_merge((int)threadIdx.x, ks + s0, vs + s0, la, ks + s0 + la, vs + s0 + la, lb,
                        kd + s0, vd + s0);
        }
        __syncthreads();  // this round's output is visible to the block
        uint64_t* tk = ks; ks = kd; kd = tk;
        uint32_t* tv = vs; vs = vd; vd = tv;
    }
    if (ks != kout + off)
        for (int i = (int)threadIdx.x; i < m; i += kBucketThreads) {
            kout[off + i] = ks[i];
            vout[off + i] = vs[i];
        }
}

}  // namespace gcs
