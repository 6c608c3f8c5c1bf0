// gcs_kernels.hpp -- per-particle and per-node kernels of the step.
//
// Every per-particle kernel runs one thread per particle over "quarter tiles":
// block b covers tile b/4, particles [q*256, q*256+256) of it with q = b%4, so
// a block never spans two instances and the instance of every thread is the
// tile's.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

struct QuarterTile {
    int i;      // global particle index
    int local;  // index within the tile's instance
    bool valid;
    int seg;
};

__device__ __forceinline__ QuarterTile quarter_tile(const TileInfo* __restrict__ tiles,
                                                    const SegInfo* __restrict__ segs)
{
    const TileInfo t = tiles[blockIdx.x >> 2];
    const int off = (int)(blockIdx.x & 3) * kTileThreads + (int)threadIdx.x;
    QuarterTile q;
    q.valid = off < t.count;
    q.i = t.start + off;
    q.seg = t.seg;
    q.local = q.i - segs[t.seg].start;
    return q;
}

// main.cpp:152-166 -- per-instance cube [min over x,y,z, max over x,y,z], in
// two kernels: k_bounds writes each quarter tile's (min, max) as
// order-preserving integer encodings, k_bounds_reduce folds an instance's
// quarter tiles (no atomics, so no contention on one address).  NaN
// coordinates are ignored, as the reference's `<` comparisons ignore them.
// A step's k_integrate writes the same partials from the positions it stores,
// so the next step needs only the fold; k_bounds runs after the positions
// were set in any other way.
//
// The quarter tile's (min, max) over every coordinate of its positions, as
// order-preserving encodings, into bpart[2 * blockIdx.x ..]; NaN coordinates
// are left out.  Every thread of the 256-thread block calls it.
__device__ __forceinline__ void bounds_partial(const float4 p, bool valid,
                                               uint32_t* __restrict__ bpart)
{
    float lo = __builtin_inff(), hi = -__builtin_inff();
    if (valid) {
        lo = fminf(p.x, fminf(p.y, p.z));
        hi = fmaxf(p.x, fmaxf(p.y, p.z));
        if (lo != lo) lo = __builtin_inff();
        if (hi != hi) hi = -__builtin_inff();
    }
    uint32_t elo = wave_min(fenc(lo)), ehi = wave_max(fenc(hi));
    __shared__ uint32_t slo[4], shi[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        slo[w] = elo;
        shi[w] = ehi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        bpart[2 * blockIdx.x] = min(min(slo[0], slo[1]), min(slo[2], slo[3]));
        bpart[2 * blockIdx.x + 1] = max(max(shi[0], shi[1]), max(shi[2], shi[3]));
    }
}

__global__ __launch_bounds__(256) void k_bounds(const TileInfo* __restrict__ tiles,
                                                const SegInfo* __restrict__ segs,
                                                const float4* __restrict__ posm,
                                                uint32_t* __restrict__ bpart)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    bounds_partial(q.valid ? posm[q.i] : make_float4(0.f, 0.f, 0.f, 0.f), q.valid, bpart);
}

__global__ __launch_bounds__(1024) void k_bounds_reduce(const SegInfo* __restrict__ segs,
                                                        const uint32_t* __restrict__ bpart,
                                                        uint32_t* __restrict__ seg_bounds)
{
    const SegInfo sg = segs[blockIdx.x];
    // (min, max) pairs of the instance's quarter tiles, eight loads in flight
    // per thread.
    const uint2* __restrict__ bp = reinterpret_cast<const uint2*>(bpart);
    const int k0 = 4 * sg.tile0, k1 = 4 * (sg.tile0 + sg.ntiles);
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 8 * 1024) {
        uint2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int kk = k + u * 1024;
            v[u] = kk < k1 ? bp[kk] : make_uint2(0xFFFFFFFFu, 0u);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            lo = min(lo, v[u].x);
            hi = max(hi, v[u].y);
        }
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    __shared__ uint32_t slo[16], shi[16];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 16; ++q) {
            slo[0] = min(slo[0], slo[q]);
            shi[0] = max(shi[0], shi[q]);
        }
        seg_bounds[2 * blockIdx.x] = slo[0];
        seg_bounds[2 * blockIdx.x + 1] = shi[0];
    }
}

// Subtree marks of one kind (traversal groups, aggregation subtrees): per
// particle a flag at a subtree's first particle, its size and root node, and
// the number of subtrees starting in each quarter tile.
struct Marks {
    uint32_t* mark;
    uint32_t* perq;
    int* cnt;
    int* root;
};

// Start-of-step clearing, done by the Morton kernel (no launch of its own):
// aggregation-frontier, work-queue and interaction counters, the sort's
// chunk histograms (all passes) and the subtree marks of the coming build.
// A null chist means no reset (a Morton-only call).
// Kernels rather than hipMemsetAsync, so the step can be captured into a
// hipGraph.
struct StepReset {
    int nseg;
    int* nfront;
    int* ntop;
    int* nmid;
    int* mfail;
    unsigned* qctr;
    unsigned long long* counters;
    uint32_t* chist;
    int64_t chist_words;
    Marks gm, am;
};

// karras_bvh.cuh:34-41 -- keys of the current order, payload = global index.
__global__ __launch_bounds__(256) void k_morton(const TileInfo* __restrict__ tiles,
                                                const SegInfo* __restrict__ segs,
                                                const float4* __restrict__ posm,
                                                uint32_t* __restrict__ seg_bounds,
                                                int use_given, float gmin, float gmax,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals, int vals_seg0,
                                                StepReset rs)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (rs.chist) {
        const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
        const int64_t stride = (int64_t)gridDim.x * blockDim.x;
        if (tid < rs.nseg) {
            rs.nfront[tid] = 0;
            rs.ntop[tid] = 0;
            if (rs.mfail) rs.mfail[tid] = 0;
            // Given bounds are recorded as the ones used (read back per
            // instance); computed ones were written by the bounds kernels.
            if (use_given) {
                seg_bounds[2 * tid] = fenc(gmin);
                seg_bounds[2 * tid + 1] = fenc(gmax);
            }
        }
        if (tid <= kNumXcd) rs.qctr[tid] = 0u;
        if (tid == 0 && rs.nmid) *rs.nmid = 0;
        if (tid < 3) rs.counters[tid] = 0ull;
        for (int64_t k = tid; k < rs.chist_words; k += stride) rs.chist[k] = 0u;
        if (threadIdx.x == 0) {
            rs.gm.perq[blockIdx.x] = 0u;
            rs.am.perq[blockIdx.x] = 0u;
        }
        if (q.valid) {
            rs.gm.mark[q.i] = 0u;
            rs.am.mark[q.i] = 0u;
        }
    }
    if (!q.valid) return;
    // The instance's bounds and the position in one round trip (with given
    // bounds the loaded pair is not used).
    const uint2 sb = reinterpret_cast<const uint2*>(seg_bounds)[q.seg];
    const float4 p = posm[q.i];
    const float mn = use_given ? gmin : fdec(sb.x), mx = use_given ? gmax : fdec(sb.y);
    keys[q.i] = morton_key(p.x, p.y, p.z, mn, mx);
    // The radix sort's payload; the bucket sort writes its own.
    if (q.seg >= vals_seg0) vals[q.i] = (uint32_t)q.i;
}

// karras_bvh.cuh:44-70 -- gather the state into sorted order (and, for the
// reference traversal, karras_bvh.cuh:134-164's leaf records).
struct PermuteArgs {
    const uint32_t* perm;
    const float4* posm_in;
    const float4* vel_in;
    const float4* plo_in;
    float4* posm_out;
    float4* vel_out;
    float4* plo_out;
    int compensated, write_leaves;
};

__device__ __forceinline__ void permute_one(const QuarterTile& q, int src,
                                            const SegInfo* __restrict__ segs,
                                            const PermuteArgs& pa, float4* __restrict__ node,
                                            float4* __restrict__ nodeQ)
{
    const float4* __restrict__ posm_in = pa.posm_in;
    const float4* __restrict__ vel_in = pa.vel_in;
    const float4* __restrict__ plo_in = pa.plo_in;
    float4* __restrict__ posm_out = pa.posm_out;
    float4* __restrict__ vel_out = pa.vel_out;
    float4* __restrict__ plo_out = pa.plo_out;
    const int compensated = pa.compensated, write_leaves = pa.write_leaves;
    // Every gather before the first store: the streams come through a
    // struct, so the compiler cannot tell the outputs from the inputs, and a
    // store placed first would hold back the loads after it.
    const float4 p = posm_in[src];
    const float4 vv = vel_in[src];
    const float4 lo = compensated ? plo_in[src] : make_float4(0.f, 0.f, 0.f, 0.f);
    posm_out[q.i] = p;
    vel_out[q.i] = vv;
    if (compensated) plo_out[q.i] = lo;
    // The group traversal reads no leaf records; an instance of one particle
    // gets its leaf (= root) record anyway, the root the COM readback copies.
    if (write_leaves || segs[q.seg].n == 1) {
        const int v = q.i + segs[q.seg].leaf_base;
        node[2 * v] = p;
        node[2 * v + 1] = make_float4(0.0f, __int_as_float(q.i), __int_as_float(q.i),
                                      __int_as_float(-1));
        nodeQ[2 * v] = make_float4(0.f, 0.f, 0.f, 0.f);
        nodeQ[2 * v + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// The sorted keys of an instance as a build block sees them: its window --
// the block's 256 nodes and 384 keys either side -- staged in LDS, the rest
// read from global memory.  Most nodes' searches stay inside the window.
constexpr int kKeyWin = 1024;
struct KeyWin {
    const uint64_t* g;  // the instance's keys (local index)
    const uint64_t* w;  // w[j - lo] for j in [lo, lo + kKeyWin) and [0, n)
    int lo, n;
};

__device__ __forceinline__ uint64_t key_at(const KeyWin& kw, int j)
{
    const unsigned r = (unsigned)(j - kw.lo);
    return r < (unsigned)kKeyWin ? kw.w[r] : kw.g[j];
}

// karras_bvh.cuh:23-31 -- common prefix of the 96-bit key (key, local index),
// with key i (ki) already loaded.
__device__ __forceinline__ int delta(const KeyWin& kw, uint64_t ki, int i, int j)
{
    if (j < 0 || j >= kw.n) return -1;
    const uint64_t b = key_at(kw, j);
    if (ki == b) return 64 + clz32s((uint32_t)(i ^ j));
    return clz64s(ki ^ b);
}

// Stages a build block's key window (all threads; ends with a barrier).
__device__ __forceinline__ KeyWin stage_keys(const TileInfo* __restrict__ tiles,
                                             const SegInfo* __restrict__ segs,
                                             const uint64_t* __restrict__ skeys, uint64_t* wk)
{
    const TileInfo t = tiles[blockIdx.x >> 2];
    const SegInfo sg = segs[t.seg];
    KeyWin kw;
    kw.g = skeys + sg.start;
    kw.w = wk;
    kw.n = sg.n;
    kw.lo = t.start - sg.start + (int)(blockIdx.x & 3) * kTileThreads -
            (kKeyWin - kTileThreads) / 2;
    // All of a thread's loads in flight before the stores (keys outside the
    // instance are never read: delta() rejects those places first).
    constexpr int kPer = kKeyWin / kTileThreads;
    uint64_t v[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int j = kw.lo + (int)threadIdx.x + r * kTileThreads;
        v[r] = j >= 0 && j < kw.n ? kw.g[j] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) wk[threadIdx.x + r * kTileThreads] = v[r];
    __syncthreads();
    return kw;
}

// Marks a subtree starting at global particle x with `count` particles and
// root node `root`, and counts it in its quarter tile.
__device__ __forceinline__ void mark_subtree(const SegInfo& sg, int x, int count, int root,
                                             const Marks& mk)
{
    mk.mark[x] = 1u;
    mk.cnt[x] = count;
    if (mk.root) mk.root[x] = root;
    const int off = x - sg.start;
    const int qt = 4 * (sg.tile0 + off / kTile) + (off % kTile) / kTileThreads;
    atomicAdd(&mk.perq[qt], 1u);
}

// karras_bvh.cuh:73-131 -- one internal node per particle but the last of
// each instance, built on instance-local indices so every instance's tree is
// the stand-alone tree.  Writes range/split, parents and clears the arrival
// flag of the node it builds.  Also marks the maximal subtrees of at most
// gmax particles (traversal groups) and of at most kATile particles
// (aggregation subtrees) -- a child of at most that many under a parent of
// more, or the root of a small instance -- at their first particle, with
// their size (and root node), counted per quarter tile for the numbering.
// The nodes of more than kATile particles of each instance (the ones above
// the aggregation subtrees), listed from the build in any order: list[start +
// k] = node, slot[node] = k, count[instance] = number listed.  list == nullptr:
// not listed.  mids (unless null): the roots of the maximal subtrees of more
// than kATile and fewer than kMidSize particles under a node of at least
// kMidSize, as (node, instance) pairs, *nmid of them, in any order.
struct TopList {
    int* list;
    int* count;
    int* slot;
    int* mids;
    int* nmid;
};

__device__ __forceinline__ void build_one(const QuarterTile& q, const SegInfo* __restrict__ segs,
                                          const KeyWin& kw,
                                          float4* __restrict__ node, int* __restrict__ parent,
                                          int* __restrict__ flag, const Marks& gm, const Marks& am,
                                          int gmax, const TopList& tl)
{
    const SegInfo sg = segs[q.seg];
    const int n = sg.n, i = q.local;
    if (i == 0) {
        parent[sg.root] = -1;
        if (n <= gmax) mark_subtree(sg, sg.start, n, sg.root, gm);
        if (n <= kATile) mark_subtree(sg, sg.start, n, sg.root, am);
    }
    if (i >= n - 1) return;
    const uint64_t ki = key_at(kw, i);
    const int dir = (delta(kw, ki, i, i + 1) - delta(kw, ki, i, i - 1)) > 0 ? 1 : -1;
    const int dmin = delta(kw, ki, i, i - dir);
    int lmax = 2;
    while (delta(kw, ki, i, i + lmax * dir) > dmin) lmax *= 2;
    int len = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(kw, ki, i, i + (len + t) * dir) > dmin) len += t;
    const int j = i + len * dir;
    const int dnode = delta(kw, ki, i, j);
    int s = 0, step = len;
    do {
        step = (step + 1) >> 1;
        if (delta(kw, ki, i, i + (s + step) * dir) > dnode) s += step;
    } while (step > 1);
    const int split = i + s * dir + (dir > 0 ? 0 : -1);
    const int lo = min(i, j), hi = max(i, j);
    const int me = sg.start + i + sg.int_base;
    const int gsplit = sg.start + split;
    const int cl = (lo == split) ? gsplit + sg.leaf_base : gsplit + sg.int_base;
    const int cr = (hi == split + 1) ? gsplit + 1 + sg.leaf_base : gsplit + 1 + sg.int_base;
    node[2 * me + 1] = make_float4(0.0f, __int_as_float(sg.start + lo),
                                   __int_as_float(sg.start + hi), __int_as_float(gsplit));
    parent[cl] = me;
    parent[cr] = me;
    flag[me] = 0;
    const int nl = split - lo + 1, nr = hi - split;
    if (hi - lo + 1 > gmax) {
        if (nl <= gmax) mark_subtree(sg, sg.start + lo, nl, cl, gm);
        if (nr <= gmax) mark_subtree(sg, gsplit + 1, nr, cr, gm);
    }
    if (hi - lo + 1 > kATile) {
        if (nl <= kATile) mark_subtree(sg, sg.start + lo, nl, cl, am);
        if (nr <= kATile) mark_subtree(sg, gsplit + 1, nr, cr, am);
        if (tl.list) {
            const int k = atomicAdd(&tl.count[q.seg], 1);
            tl.list[sg.start + k] = me;
            tl.slot[me] = k;
            if (tl.mids && hi - lo + 1 >= kMidSize) {
                if (nl > kATile && nl < kMidSize) {
                    const int j = atomicAdd(tl.nmid, 1);
                    tl.mids[2 * j] = cl;
                    tl.mids[2 * j + 1] = q.seg;
                }
                if (nr > kATile && nr < kMidSize) {
                    const int j = atomicAdd(tl.nmid, 1);
                    tl.mids[2 * j] = cr;
                    tl.mids[2 * j + 1] = q.seg;
                }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_build(const TileInfo* __restrict__ tiles,
                                               const SegInfo* __restrict__ segs,
                                               const uint64_t* __restrict__ skeys,
                                               float4* __restrict__ node, int* __restrict__ parent,
                                               int* __restrict__ flag, Marks gm, Marks am, int gmax,
                                               TopList tl)
{
    __shared__ uint64_t wk[kKeyWin];
    const KeyWin kw = stage_keys(tiles, segs, skeys, wk);
    const QuarterTile q = quarter_tile(tiles, segs);
    if (q.valid) build_one(q, segs, kw, node, parent, flag, gm, am, gmax, tl);
}

// The gather and the tree build of a step in one launch: the two only read
// the sorted keys and permutation, and write disjoint data.
__global__ __launch_bounds__(256) void k_permute_build(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs, PermuteArgs pa,
    const uint64_t* __restrict__ skeys, float4* __restrict__ node, float4* __restrict__ nodeQ,
    int* __restrict__ parent, int* __restrict__ flag, Marks gm, Marks am, int gmax, TopList tl)
{
    __shared__ uint64_t wk[kKeyWin];
    const QuarterTile q = quarter_tile(tiles, segs);
    // The particle's place in the old order goes out with the key window's
    // loads (one wait for both; a lane past the tile's end reads entry 0).
    const int src = (int)pa.perm[q.valid ? q.i : 0];
    const KeyWin kw = stage_keys(tiles, segs, skeys, wk);
    if (!q.valid) return;
    permute_one(q, src, segs, pa, node, nodeQ);
    build_one(q, segs, kw, node, parent, flag, gm, am, gmax, tl);
}

// Groups and aggregation subtrees per chunk of kGroupChunk quarter tiles.
__global__ __launch_bounds__(kGroupChunk) void k_group_chunksum(const uint32_t* __restrict__ gq,
                                                                 const uint32_t* __restrict__ aq,
                                                                 int nq,
                                                                 uint32_t* __restrict__ gchunk,
                                                                 uint32_t* __restrict__ achunk)
{
    static_assert(kGroupChunk == 2 * kWave, "two waves per chunk");
    __shared__ uint32_t w1[2];
    const int q = blockIdx.x * kGroupChunk + (int)threadIdx.x;
    // Both counts in flight together (a lane past the end reads entry 0).
    const int qc = q < nq ? q : 0;
    const int gv = (int)gq[qc], av = (int)aq[qc];
    __builtin_amdgcn_sched_barrier(0);  // keep the two loads ahead of their uses
    const uint32_t v = (uint32_t)wave_sum(q < nq ? gv : 0);
    const uint32_t u = (uint32_t)wave_sum(q < nq ? av : 0);
    if (threadIdx.x == kWave) {
        w1[0] = v;
        w1[1] = u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        gchunk[blockIdx.x] = v + w1[0];
        achunk[blockIdx.x] = u + w1[1];
    }
}

// Writes the group table and the aggregation-subtree table in particle
// (Morton) order, and their lengths.  A subtree's number is its quarter
// tile's first number plus the marks before its first particle within the
// quarter tile (block scan).
__global__ __launch_bounds__(256) void k_group_emit(const TileInfo* __restrict__ tiles,
                                                    const SegInfo* __restrict__ segs, Marks gm,
                                                    const uint32_t* __restrict__ gchunk, Marks am,
                                                    const uint32_t* __restrict__ achunk,
                                                    GroupInfo* __restrict__ groups,
                                                    int* __restrict__ ngroups,
                                                    ATileInfo* __restrict__ atiles,
                                                    int* __restrict__ natiles, int n_total)
{
    static_assert(kTileThreads == 4 * kWave, "four waves per quarter tile");
    static_assert(kGroupChunk <= kTileThreads, "a chunk's quarter tiles fit one pass");
    __shared__ uint32_t ws[4][4];
    const QuarterTile q = quarter_tile(tiles, segs);
    const int b = blockIdx.x;
    const int tid = (int)threadIdx.x, l = lane_id(), w = tid >> 6;
    // Every load first: the counts before this quarter tile (the chunks
    // before its chunk, the quarter tiles before it in its chunk), its marks
    // and the marked subtrees' sizes and roots.
    // In one round trip (places clamped to 0 where a thread has nothing to
    // read; those values are not used): the thread's first chunk sum, its
    // quarter tile count (the quarter tiles before this one in its chunk are
    // fewer than the block's threads), this quarter tile's marks and the
    // marked subtrees' sizes and roots.
    const int c = b / kGroupChunk, kq = c * kGroupChunk + tid;
    const bool hc = tid < c, hq = kq < b;
    const int qi = q.valid ? q.i : 0;
    const uint32_t gc = gchunk[hc ? tid : 0], ac = achunk[hc ? tid : 0];
    const uint32_t gp = gm.perq[hq ? kq : 0], ap = am.perq[hq ? kq : 0];
    const uint32_t gmk = gm.mark[qi], amk = am.mark[qi];
    const int gcount = gm.cnt[qi], acount = am.cnt[qi], aroot = am.root[qi];
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of their uses
    uint32_t pg = (hc ? gc : 0u) + (hq ? gp : 0u);
    uint32_t pa = (hc ? ac : 0u) + (hq ? ap : 0u);
    for (int k = tid + kTileThreads; k < c; k += kTileThreads) {
        pg += gchunk[k];
        pa += achunk[k];
    }
    const uint32_t m = q.valid ? gmk : 0u;
    const uint32_t ma = q.valid ? amk : 0u;
    // One exchange: per wave the partial counts' sums and the marks' totals.
    const uint32_t sg = (uint32_t)wave_sum((int)pg), sa = (uint32_t)wave_sum((int)pa);
    const uint32_t im = (uint32_t)wave_scan_incl_dpp((int)m);
    const uint32_t ima = (uint32_t)wave_scan_incl_dpp((int)ma);
    if (l == kWave - 1) {
        ws[0][w] = sg;
        ws[1][w] = sa;
        ws[2][w] = im;
        ws[3][w] = ima;
    }
    __syncthreads();
    uint32_t gbase = 0u, abase = 0u, bm = 0u, bma = 0u, nb = 0u, na = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        gbase += ws[0][k];
        abase += ws[1][k];
        nb += ws[2][k];
        na += ws[3][k];
        bm += k < w ? ws[2][k] : 0u;
        bma += k < w ? ws[3][k] : 0u;
    }
    if (m) {
        GroupInfo g;
        g.first = q.i;
        g.count = gcount;
        g.seg = q.seg;
        g.pad = 0;
        groups[gbase + bm + im - m] = g;
    }
    if (ma) {
        ATileInfo a;
        a.first = q.i;
        a.count = acount;
        a.seg = q.seg;
        a.root = aroot;
        atiles[abase + bma + ima - ma] = a;
    }
    if (q.valid && q.i == n_total - 1) {
        *ngroups = (int)(gbase + nb);
        *natiles = (int)(abase + na);
    }
}

// Dispatch order of the traversal.  The groups are dealt to kNumXcd queues
// in contiguous (Morton) runs; within its run each queue hands out the
// groups by decreasing size class (eighths of the largest group size) and
// in Morton order within a class, so the groups with the most targets start
// first and neighbouring groups still run close together.  The order
// changes scheduling only: every group's sums are the same whenever it runs.
__device__ __forceinline__ uint32_t size_class(int count, int gmax)
{
    return (uint32_t)min(kOrderClasses - 1, max(0, (gmax - count) * kOrderClasses / gmax));
}

__global__ __launch_bounds__(1024) void k_group_order(const GroupInfo* __restrict__ groups,
                                                      const int* __restrict__ ngroups_dev,
                                                      int gmax, int* __restrict__ gorder)
{
    __shared__ uint32_t base[kOrderClasses];
    __shared__ uint32_t wc[16][kOrderClasses];
    const int ng = *ngroups_dev;
    const int per = (ng + kNumXcd - 1) / kNumXcd;
    const int lo = blockIdx.x * per, hi = min(ng, lo + per);
    const int l = lane_id(), w = threadIdx.x >> 6;
    if (threadIdx.x < kOrderClasses) base[threadIdx.x] = 0u;
    __syncthreads();
    for (int i = lo + (int)threadIdx.x; i < hi; i += 1024)
        atomicAdd(&base[size_class(groups[i].count, gmax)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int b = 0; b < kOrderClasses; ++b) {
            const uint32_t c = base[b];
            base[b] = run;
            run += c;
        }
    }
    __syncthreads();
    // Stable scatter, 1024 groups at a time: lanes of one class are ranked by
    // lane order (ballots per class bit), waves by wave order.
    const uint64_t lt = (1ull << l) - 1ull;
    for (int c0 = lo; c0 < hi; c0 += 1024) {
        const int i = c0 + (int)threadIdx.x;
        const bool valid = i < hi;
        const uint32_t k = valid ? size_class(groups[i].count, gmax) : 0u;
        uint64_t m = ballot(valid);
#pragma unroll
        for (int b = 0; (1 << b) < kOrderClasses; ++b) {
            const uint64_t bb = ballot((k >> b) & 1u);
            m &= ((k >> b) & 1u) ? bb : ~bb;
        }
        if (l < kOrderClasses) wc[w][l] = 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && (m & lt) == 0ull) wc[w][k] = (uint32_t)__popcll(m);
        __syncthreads();
        // Per class: exclusive offsets over the waves; thread t owns class t.
        if (threadIdx.x < kOrderClasses) {
            uint32_t run = base[threadIdx.x];
            for (int v = 0; v < 16; ++v) {
                const uint32_t c = wc[v][threadIdx.x];
                wc[v][threadIdx.x] = run;
                run += c;
            }
            base[threadIdx.x] = run;
        }
        __syncthreads();
        if (valid) gorder[lo + (int)(wc[w][k] + (uint32_t)__popcll(m & lt))] = i;
        __syncthreads();
    }
}

// Aggregate of one node: COM, mass, AABB, quadrupole (xx, yy, zz, xy, xz, yz).
struct NodeData {
    float cx, cy, cz, m;
    float lx, ly, lz, hx, hy, hz;
    float q0, q1, q2, q3, q4, q5;
    float qt;  // trace of the raw second moment, sum m |x - COM|^2
};

__device__ __forceinline__ NodeData leaf_data(float4 p)
{
    NodeData d;
    d.cx = d.lx = d.hx = p.x;
    d.cy = d.ly = d.hy = p.y;
    d.cz = d.lz = d.hz = p.z;
    d.m = p.w;
    d.q0 = d.q1 = d.q2 = d.q3 = d.q4 = d.q5 = d.qt = 0.0f;
    return d;
}

__device__ __forceinline__ NodeData load_node_data(int v, const float4* node, const float4* nodeQ,
                                                   const float4* boxLo, const float4* boxHi)
{
    const float4 a = node[2 * v];
    const float4 lo = boxLo[v], hi = boxHi[v];
    const float4 qa = nodeQ[2 * v], qb = nodeQ[2 * v + 1];
    NodeData d;
    d.cx = a.x; d.cy = a.y; d.cz = a.z; d.m = a.w;
    d.lx = lo.x; d.ly = lo.y; d.lz = lo.z;
    d.hx = hi.x; d.hy = hi.y; d.hz = hi.z;
    d.q0 = qa.x; d.q1 = qa.y; d.q2 = qa.z; d.q3 = qa.w;
    d.q4 = qb.x; d.q5 = qb.y; d.qt = qb.w;
    return d;
}

// Adds a child's second moments about the parent COM (same operation order
// as the oracle): the traceless quadrupole Q += Qc + mc*(3 s s^T - |s|^2 I)
// and the trace T += Tc + mc*|s|^2, s = child COM - parent COM.
__device__ __forceinline__ void quad_shift(NodeData& o, const NodeData& c)
{
#pragma clang fp contract(off)
    const float sx = c.cx - o.cx, sy = c.cy - o.cy, sz = c.cz - o.cz;
    const float s2 = sx * sx + sy * sy + sz * sz;
    o.q0 = o.q0 + (c.q0 + c.m * (3.0f * sx * sx - s2));
    o.q1 = o.q1 + (c.q1 + c.m * (3.0f * sy * sy - s2));
    o.q2 = o.q2 + (c.q2 + c.m * (3.0f * sz * sz - s2));
    o.q3 = o.q3 + (c.q3 + c.m * (3.0f * sx * sy));
    o.q4 = o.q4 + (c.q4 + c.m * (3.0f * sx * sz));
    o.q5 = o.q5 + (c.q5 + c.m * (3.0f * sy * sz));
    o.qt = o.qt + (c.qt + c.m * s2);
}

// Merge of one internal node from its children (karras_bvh.cuh:190-221):
//   AABB = union; width = max extent; m = ml + mr;
//   com = fma(cl, ml, cr*mr)/m (the contraction nvcc applies to the
//   reference's cl*ml + cr*mr), or the AABB centre when m == 0.
// Extension: opening value = width for the reference traversal (mode 0); for
// the group traversal (width*inv_theta)^2 (mode 1) or
// (width*inv_theta + |com - AABB centre|)^2 (mode 2); quadrupole moments by the
// parallel-axis shift.
// Width and opening value of a merged node (from its AABB and COM only).
__device__ __forceinline__ void node_open(const NodeData& o, float inv_theta, int mode, float& w,
                                          float& open)
{
#pragma clang fp contract(off)
    w = fmaxf(o.hx - o.lx, fmaxf(o.hy - o.ly, o.hz - o.lz));
    if (mode == 2) {
        const float dx = o.cx - (o.lx + o.hx) * 0.5f;
        const float dy = o.cy - (o.ly + o.hy) * 0.5f;
        const float dz = o.cz - (o.lz + o.hz) * 0.5f;
        const float off = sqrtf(dx * dx + dy * dy + dz * dz);
        const float ro = w * inv_theta + off;
        open = ro * ro;
    } else if (mode == 1) {
        const float ro = w * inv_theta;
        open = ro * ro;
    } else {
        open = w;
    }
}

__device__ __forceinline__ NodeData merge(const NodeData& l, const NodeData& r)
{
#pragma clang fp contract(off)
    NodeData o;
    o.lx = fminf(l.lx, r.lx); o.ly = fminf(l.ly, r.ly); o.lz = fminf(l.lz, r.lz);
    o.hx = fmaxf(l.hx, r.hx); o.hy = fmaxf(l.hy, r.hy); o.hz = fmaxf(l.hz, r.hz);
    o.m = l.m + r.m;
    if (o.m > 0.0f) {
        // The left product fused, as nvcc contracts cl*ml + cr*mr.
        o.cx = __builtin_fmaf(l.cx, l.m, r.cx * r.m) / o.m;
        o.cy = __builtin_fmaf(l.cy, l.m, r.cy * r.m) / o.m;
        o.cz = __builtin_fmaf(l.cz, l.m, r.cz * r.m) / o.m;
    } else {
        o.cx = (o.lx + o.hx) * 0.5f;
        o.cy = (o.ly + o.hy) * 0.5f;
        o.cz = (o.lz + o.hz) * 0.5f;
    }
    o.q0 = o.q1 = o.q2 = o.q3 = o.q4 = o.q5 = o.qt = 0.0f;
    quad_shift(o, l);
    quad_shift(o, r);
    return o;
}

__device__ __forceinline__ NodeData merge(const NodeData& l, const NodeData& r, float inv_theta,
                                          int mode, float& w, float& open)
{
    const NodeData o = merge(l, r);
    node_open(o, inv_theta, mode, w, open);
    return o;
}

__device__ __forceinline__ void store_node(int v, const NodeData& d, float w, float open,
                                           float4* node, float4* nodeQ, float4* boxLo,
                                           float4* boxHi)
{
    node[2 * v] = make_float4(d.cx, d.cy, d.cz, d.m);
    node[2 * v + 1].x = open;
    nodeQ[2 * v] = make_float4(d.q0, d.q1, d.q2, d.q3);
    nodeQ[2 * v + 1] = make_float4(d.q4, d.q5, w, d.qt);
    boxLo[v] = make_float4(d.lx, d.ly, d.lz, 0.0f);
    boxHi[v] = make_float4(d.hx, d.hy, d.hz, 0.0f);
}

// karras_bvh.cuh:175-225 -- bottom-up aggregation, in two kernels.  Each
// node is merged once from its two children, by the second of the two
// climbs that reach it (the first stops there).
//
// k_aggregate: aggregation subtrees (maximal subtrees of at most kATile
// particles) are taken in work items from a counter: item 2j is subtrees 2j
// and 2j+1 together when they are of one instance and hold at most kATile
// particles between them (item 2j+1 is then empty), else subtree 2j; item
// 2j+1 is subtree 2j+1.  Consecutive subtrees cover consecutive particles,
// so an item covers particles t0 .. t0 + c - 1 and its internal nodes have
// indices in that range (Karras numbering: a node's index is one end of its
// range); an index there is a node of the item iff that node's range lies
// inside one of its subtrees.  slot = index - t0.  The climb runs in LDS:
//   1. thread k reads the node of slot k and, if its range lies inside the
//      subtree's, loads its split, leaf flags and parent slot into an LDS
//      word, and the particle data of its leaf children: a node with two
//      leaf children is merged at once and starts a climb; the one leaf child
//      of a node with one is parked in the node's LDS record;
//   2. climbs go up: at a node whose other child is a leaf, the parked leaf
//      is the sibling; at one with two internal children, an LDS arrival bit
//      (workgroup-scope acquire-release) lets the second climb through;
//   3. every node record is written to global memory once, at the end.
// No global memory access sits between two levels of a climb.  The thread
// that completes a subtree's root appends it to its instance's frontier
// for the instance pass.
__device__ __forceinline__ NodeData lds_node(const float4 (*ldat)[4], const float* lqt, int s)
{
    const float4 a4 = ldat[s][0], b4 = ldat[s][1], c4 = ldat[s][2], d4 = ldat[s][3];
    NodeData d;
    d.qt = lqt[s];
    d.cx = a4.x; d.cy = a4.y; d.cz = a4.z; d.m = a4.w;
    d.lx = b4.x; d.ly = b4.y; d.lz = b4.z; d.hx = b4.w;
    d.hy = c4.x; d.hz = c4.y; d.q0 = c4.z; d.q1 = c4.w;
    d.q2 = d4.x; d.q3 = d4.y; d.q4 = d4.z; d.q5 = d4.w;
    return d;
}

__device__ __forceinline__ void lds_put(float4 (*ldat)[4], float* lqt, int s, const NodeData& d)
{
    ldat[s][0] = make_float4(d.cx, d.cy, d.cz, d.m);
    ldat[s][1] = make_float4(d.lx, d.ly, d.lz, d.hx);
    ldat[s][2] = make_float4(d.hy, d.hz, d.q0, d.q1);
    ldat[s][3] = make_float4(d.q2, d.q3, d.q4, d.q5);
    lqt[s] = d.qt;
}

// LDS word of an internal node of a subtree: split (relative to the
// subtree's first particle, 10 bits), left child is a leaf (bit 10), right
// child is a leaf (bit 11), parent slot + 1 (bits 12..22; 0 at the root).
static_assert(kATile <= 1024, "splits fit 10 bits, parent slots + 1 fit 11");
__device__ __forceinline__ uint32_t topo_word(int split, bool lleaf, bool rleaf, int pslot)
{
    return (uint32_t)split | (lleaf ? 1u << 10 : 0u) | (rleaf ? 1u << 11 : 0u) |
           ((uint32_t)(pslot + 1) << 12);
}

__global__ __launch_bounds__(kATile) __attribute__((amdgpu_waves_per_eu(8))) void k_aggregate(
    const ATileInfo* __restrict__ atiles, const int* __restrict__ natiles_dev,
    const SegInfo* __restrict__ segs, const float4* __restrict__ posm, float4* node,
    float4* nodeQ, float4* boxLo, float4* boxHi, const int* __restrict__ parent,
    int* __restrict__ front, int* __restrict__ nfront, unsigned* __restrict__ ictr,
    float inv_theta, int mode)
{
    __shared__ uint32_t lflag[kATile / 32];
    __shared__ uint32_t ltopo[kATile];
    __shared__ float4 ldat[kATile][4];
    __shared__ float lqt[kATile];
    __shared__ int litem;
    const int tid = threadIdx.x;
    const int natiles = *natiles_dev;
    const int nitems = 2 * ((natiles + 1) / 2);
    for (;;) {
        __syncthreads();  // the previous item's LDS use is over
        if (tid == 0) litem = (int)atomicAdd(ictr, 1u);
        __syncthreads();
        const int item = litem;
        if (item >= nitems) break;
        const int j0 = item & ~1;
        const ATileInfo a0 = atiles[j0];
        const bool has1 = j0 + 1 < natiles;
        const ATileInfo a1 = has1 ? atiles[j0 + 1] : a0;
        const bool pair = has1 && a1.seg == a0.seg && a0.count + a1.count <= kATile;
        if ((item & 1) && (pair || !has1)) continue;  // an empty item
        const ATileInfo t = (item & 1) ? a1 : a0;
        const bool two = pair && !(item & 1);
        const SegInfo sg = segs[t.seg];
        const int t0 = t.first, e0 = t0 + t.count - 1;  // the first subtree's range
        const int c = two ? t.count + a1.count : t.count;
        if (tid == 0) {
            // A subtree of one particle: its root is the leaf.
            if (t.count == 1) front[2 * sg.start + atomicAdd(&nfront[t.seg], 1)] = t0 + sg.leaf_base;
            if (two && a1.count == 1)
                front[2 * sg.start + atomicAdd(&nfront[t.seg], 1)] = a1.first + sg.leaf_base;
        }
        if (tid < kATile / 32) lflag[tid] = 0u;
        const int ibase = t0 + sg.int_base;  // global index of slot 0
        bool climb = false;
        // Slot tid holds a node of the item iff the index exists in the
        // instance (internal nodes 0 .. n-2) and its range lies inside one
        // of the item's subtrees.
        const int v = ibase + tid;
        bool member = false;
        int first = 0, last = 0, split = 0, pv = -1;
        float4 pl = make_float4(0.f, 0.f, 0.f, 0.f), pr = pl;
        if (tid < c && t0 - sg.start + tid <= sg.n - 2) {
            // Two round trips: the node and its parent link, then the
            // particles at split and split + 1 (inside the node's range, so
            // inside the instance) whether or not the children are leaves.
            const float4 bb = node[2 * v + 1];
            const int pvr = parent[v];
            first = __float_as_int(bb.y);
            last = __float_as_int(bb.z);
            split = __float_as_int(bb.w);
            pl = posm[split];
            pr = posm[split + 1];
            member = (first >= t0 && last <= e0) || (two && first > e0 && last <= t0 + c - 1);
            const bool root = (t.count > 1 && v == t.root) || (two && a1.count > 1 && v == a1.root);
            if (member && !root) pv = pvr;
        }
        if (member) {
            const bool lleaf = first == split, rleaf = last == split + 1;
            ltopo[tid] = topo_word(split - t0, lleaf, rleaf, pv < 0 ? -1 : pv - ibase);
            if (lleaf && rleaf) {
                lds_put(ldat, lqt, tid, merge(leaf_data(pl), leaf_data(pr)));
                climb = true;
            } else if (lleaf || rleaf) {
                lds_put(ldat, lqt, tid, leaf_data(lleaf ? pl : pr));
            }
        }
        __syncthreads();  // topology, parked leaves and the first merges are in LDS
        int sl = tid;
        while (climb) {
            const uint32_t tw = ltopo[sl];
            const int ps = (int)(tw >> 12) - 1;
            if (ps < 0) {
                front[2 * sg.start + atomicAdd(&nfront[t.seg], 1)] = ibase + sl;
                break;
            }
            const uint32_t pw = ltopo[ps];
            const int psplit = (int)(pw & 1023u);
            const bool plleaf = (pw >> 10) & 1u, prleaf = (pw >> 11) & 1u;
            const bool mine_left = !plleaf && sl == psplit;
            const bool sleaf = mine_left ? prleaf : plleaf;
            int ss = ps;  // a parked leaf sits in the parent's own record
            if (!sleaf) {
                const uint32_t bit = 1u << (ps & 31);
                const uint32_t old = __hip_atomic_fetch_or(&lflag[ps >> 5], bit, __ATOMIC_ACQ_REL,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                if (!(old & bit)) break;
                ss = mine_left ? psplit + 1 : psplit;
            }
            // Both children from LDS (this climb's own record included).
            const int ls = mine_left ? sl : ss, rs = mine_left ? ss : sl;
            lds_put(ldat, lqt, ps, merge(lds_node(ldat, lqt, ls), lds_node(ldat, lqt, rs)));
            sl = ps;
        }
        __syncthreads();  // every node of the subtree is complete in LDS
        if (member) {
            const NodeData d = lds_node(ldat, lqt, tid);
            float w, open;
            node_open(d, inv_theta, mode, w, open);
            store_node(v, d, w, open, node, nodeQ, boxLo, boxHi);
        }
    }
}

// Arrival marks of the top pass: bits in LDS indexed by node (instances of
// up to kTopLdsNodes particles), an LDS hash of the nodes above the
// aggregation subtrees (at most kTopHashMax of them: the table stays at most
// 3/4 full), or counters in global memory.
constexpr int kTopLdsNodes = 1 << 20;
constexpr int kTopHashBits = 15;
constexpr int kTopHashMax = 3 << (kTopHashBits - 2);
enum : int { kMarksBits = 0, kMarksHash = 1, kMarksGlobal = 2 };
static_assert((1 << kTopHashBits) == kTopLdsNodes / 32, "hash and bitmap share the LDS array");

// One climber of the top pass, from the completed node `child`: at each
// parent the first of the two arrivals stops and the second merges and
// goes on, for at most `levels` levels.  Returns the last node it completed
// when it stopped because of the level limit (with more levels to go), else
// -1.  A sibling of at most kATile particles (or a leaf) was finished by the
// subtree pass; larger ones are finished in this launch.  Bit marks are
// indexed by internal node relative to the instance start.
__device__ __forceinline__ int top_climb(const SegInfo& sg, int child, int levels,
                                         const float4* __restrict__ posm, float4* node,
                                         float4* nodeQ, float4* boxLo, float4* boxHi,
                                         const int* __restrict__ parent, int* flag,
                                         uint32_t* lmem, int marks, float inv_theta, int mode)
{
    const bool cleaf = child >= sg.start + sg.leaf_base;
    NodeData mine = cleaf ? leaf_data(posm[child - sg.leaf_base])
                          : load_node_data(child, node, nodeQ, boxLo, boxHi);
    int cur = parent[child];
    if (cur == -1) return -1;
    float4 bb = node[2 * cur + 1];
    int par = parent[cur];
    for (int lv = 0;; ++lv) {
        if (lv == levels) return child;
        const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
                  split = __float_as_int(bb.w);
        const bool lleaf = first == split, rleaf = last == split + 1;
        const int lidx = lleaf ? split + sg.leaf_base : split + sg.int_base;
        const bool mine_left = child == lidx;
        const bool sleaf = mine_left ? rleaf : lleaf;
        const int sp = mine_left ? split + 1 : split;
        // A sibling of at most kATile particles (or a leaf) was finished by the
        // subtree pass: load it before the arrival check.  Any other sibling
        // is written in this launch and is read after the acquire.
        const int sf = mine_left ? split + 1 : first, sl = mine_left ? last : split;
        const bool sfinal = sleaf || sl - sf + 1 <= kATile;
        NodeData sib;
        if (sfinal)
            sib = sleaf ? leaf_data(posm[sp])
                        : load_node_data(sp + sg.int_base, node, nodeQ, boxLo, boxHi);
        float4 bb_next = make_float4(0.f, 0.f, 0.f, 0.f);
        int par_next = -1;
        if (par != -1 && lv + 1 < levels) {
            bb_next = node[2 * par + 1];
            par_next = parent[par];
        }
        // Arrival: a thread that sees the sibling's mark already set is the
        // second and continues without publishing a mark of its own; otherwise
        // it marks the node with a release (its record is then visible to the
        // sibling's thread).
        bool second;
        if (marks == kMarksBits) {
            const int slot = cur - sg.int_base - sg.start;
            uint32_t* w = &lmem[slot >> 5];
            const uint32_t bit = 1u << (slot & 31);
            second =
                (__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & bit) != 0u;
            if (!second)
                second = (__hip_atomic_fetch_or(w, bit, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_WORKGROUP) &
                          bit) != 0u;
        } else if (marks == kMarksHash) {
            // Open addressing on the node index: the first arrival claims a
            // slot with its parent's index, the second finds it.
            const uint32_t key = (uint32_t)cur;
            uint32_t slot = (key * 0x9E3779B1u) >> (32 - kTopHashBits);
            for (;;) {
                const uint32_t k = __hip_atomic_load(&lmem[slot], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_WORKGROUP);
                if (k == key) {
                    second = true;
                    break;
                }
                if (k == 0xFFFFFFFFu) {
                    uint32_t expect = 0xFFFFFFFFu;
                    if (__hip_atomic_compare_exchange_strong(
                            &lmem[slot], &expect, key, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                            __HIP_MEMORY_SCOPE_WORKGROUP)) {
                        second = false;
                        break;
                    }
                    if (expect == key) {
                        second = true;
                        break;
                    }
                }
                slot = (slot + 1u) & ((1u << kTopHashBits) - 1u);
            }
        } else {
            second = __hip_atomic_load(&flag[cur], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
            if (!second)
                second = __hip_atomic_fetch_add(&flag[cur], 1, __ATOMIC_ACQ_REL,
                                                __HIP_MEMORY_SCOPE_WORKGROUP) != 0;
        }
        if (!second) return -1;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (!sfinal) sib = load_node_data(sp + sg.int_base, node, nodeQ, boxLo, boxHi);
        float w, open;
        mine = mine_left ? merge(mine, sib, inv_theta, mode, w, open)
                         : merge(sib, mine, inv_theta, mode, w, open);
        store_node(cur, mine, w, open, node, nodeQ, boxLo, boxHi);
        if (par == -1) return -1;
        child = cur;
        cur = par;
        if (lv + 1 < levels) {
            bb = bb_next;
            par = par_next;
        }
    }
}

// Runs the climbers of list la[0, nf) (with la[n, 2n) as the second list of
// the rounds below) for the whole block.
__device__ __forceinline__ void climb_rounds(const SegInfo& sg, int* la, int n, int nf,
                                             const float4* __restrict__ posm, float4* node,
                                             float4* nodeQ, float4* boxLo, float4* boxHi,
                                             const int* __restrict__ parent, int* flag,
                                             uint32_t* lmem, int marks, float inv_theta,
                                             int mode, int* lcount)
{
    int* lb = la + n;
    int count = nf;
    while (count > (int)blockDim.x) {
        if (threadIdx.x == 0) *lcount = 0;
        __syncthreads();
        for (int k = threadIdx.x; k < count; k += blockDim.x) {
            const int v = top_climb(sg, la[k], 1, posm, node, nodeQ, boxLo, boxHi, parent, flag,
                                    lmem, marks, inv_theta, mode);
            if (v >= 0) lb[atomicAdd(lcount, 1)] = v;
        }
        __syncthreads();
        count = *lcount;
        int* t = la;
        la = lb;
        lb = t;
        __syncthreads();  // every thread has read the count before the next reset
    }
    __syncthreads();
    for (int k = threadIdx.x; k < count; k += blockDim.x)
        top_climb(sg, la[k], 1 << 30, posm, node, nodeQ, boxLo, boxHi, parent, flag, lmem,
                  marks, inv_theta, mode);
}

// The top pass in LDS, for an instance with at most kTopCap nodes above its
// aggregation subtrees (listed by the build): the same scheme as
// k_aggregate with list slots instead of index slots.  Thread k loads the
// node of slot k (and of k + blockDim): its parent's and its larger
// children's slots, and the records of its other children (aggregation
// subtree roots, finished by k_aggregate, or leaves) -- merged at once when
// both are such, else parked in the node's own LDS record.  Climbs then run
// in LDS, and every node record is written once at the end.
constexpr int kTopCap = 2048;
constexpr int kTopPhaseNodes = 8192;  // most listed nodes of the phased form
struct TopLds {
    float4 dat[kTopCap][4];
    float qt[kTopCap];
    uint16_t par[kTopCap], cl[kTopCap], cr[kTopCap];  // slot + 1; 0: none (root) / not listed
    uint32_t flag[kTopCap / 32];
};

__device__ __forceinline__ NodeData child_data(int idx, bool leaf, const SegInfo& sg,
                                               const float4* __restrict__ posm,
                                               const float4* node, const float4* nodeQ,
                                               const float4* boxLo, const float4* boxHi)
{
    // Both forms are read and one is chosen, so no branch holds back the
    // next loads: the node arrays have an entry for every node index, leaves
    // included, and an internal child reads the instance's first particle.
    const float4 P = posm[leaf ? idx - sg.leaf_base : sg.start];
    const float4 A = node[2 * idx], lo = boxLo[idx], hi = boxHi[idx];
    const float4 qa = nodeQ[2 * idx], qb = nodeQ[2 * idx + 1];
    NodeData d;
    d.cx = leaf ? P.x : A.x; d.cy = leaf ? P.y : A.y; d.cz = leaf ? P.z : A.z;
    d.m = leaf ? P.w : A.w;
    d.lx = leaf ? P.x : lo.x; d.ly = leaf ? P.y : lo.y; d.lz = leaf ? P.z : lo.z;
    d.hx = leaf ? P.x : hi.x; d.hy = leaf ? P.y : hi.y; d.hz = leaf ? P.z : hi.z;
    d.q0 = leaf ? 0.0f : qa.x; d.q1 = leaf ? 0.0f : qa.y; d.q2 = leaf ? 0.0f : qa.z;
    d.q3 = leaf ? 0.0f : qa.w; d.q4 = leaf ? 0.0f : qb.x; d.q5 = leaf ? 0.0f : qb.y;
    d.qt = leaf ? 0.0f : qb.w;
    return d;
}

// Size class of a node of more than kATile particles: eight classes per
// octave of its particle count.
constexpr int kATileShift = 10;
static_assert(kATile == 1 << kATileShift, "shift");
constexpr int kTopBuckets = 8 * (30 - kATileShift);
__device__ __forceinline__ int top_bucket(int size)
{
    const int lg = 31 - __clz(size);
    return 8 * (lg - kATileShift) + ((size >> (lg - 3)) & 7);
}

// list[0, nt): the nodes to finish (slot k = list[k], tslot[node] = k).
// phase_of == nullptr: every node of more than kATile particles is in the
// list.  Else only those whose size class c has phase_of[c] == phase: the
// smaller ones were finished by earlier phases (read like aggregation
// subtree roots), and a climb stops below a larger one.
__device__ __forceinline__ void top_lds(TopLds& T, const SegInfo& sg, int nt, const int* list,
                                        const int* tslot,
                                        const float4* __restrict__ posm, float4* node,
                                        float4* nodeQ, float4* boxLo, float4* boxHi,
                                        const int* __restrict__ parent, float inv_theta, int mode,
                                        const uint8_t* phase_of, int phase)
{
    constexpr int kPer = kTopCap / 1024;
    const int tid = threadIdx.x;
    for (int k = tid; k < kTopCap / 32; k += blockDim.x) T.flag[k] = 0u;
    int vk[kPer];
    bool start[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int k = tid + r * (int)blockDim.x;
        vk[r] = -1;
        start[r] = false;
        if (k >= nt) continue;
        const int v = list[k];
        vk[r] = v;
        const float4 bb = node[2 * v + 1];
        const int pv = parent[v];
        const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
                  split = __float_as_int(bb.w);
        const bool lleaf = first == split, rleaf = last == split + 1;
        const int lidx = lleaf ? split + sg.leaf_base : split + sg.int_base;
        const int ridx = rleaf ? split + 1 + sg.leaf_base : split + 1 + sg.int_base;
        const int nl = split - first + 1, nr = last - split;
        // Every load that needs the node's record or its parent link, in one
        // round trip before the first use: the parent's range (phase test),
        // the slots of the parent and the children (tslot has an entry for
        // every node; the unused ones are not read further), and both
        // children's records (a listed child's is taken from LDS instead).
        const int pq = pv >= 0 ? pv : v;
        const float4 pb = node[2 * pq + 1];
        const int sp = tslot[pq], sl = tslot[lidx], sr = tslot[ridx];
        const NodeData dl = child_data(lidx, lleaf, sg, posm, node, nodeQ, boxLo, boxHi);
        const NodeData dr = child_data(ridx, rleaf, sg, posm, node, nodeQ, boxLo, boxHi);
        bool ltop = nl > kATile, rtop = nr > kATile, pin = pv >= 0;
        if (phase_of) {
            ltop = ltop && phase_of[top_bucket(nl)] == phase;
            rtop = rtop && phase_of[top_bucket(nr)] == phase;
            pin = pin &&
                  phase_of[top_bucket(__float_as_int(pb.z) - __float_as_int(pb.y) + 1)] == phase;
        }
        T.par[k] = (uint16_t)(pin ? sp + 1 : 0);
        T.cl[k] = (uint16_t)(ltop ? sl + 1 : 0);
        T.cr[k] = (uint16_t)(rtop ? sr + 1 : 0);
        if (!ltop || !rtop) {
            // The child (or the left of two) that is not listed: merged with
            // the other when both are not listed, else parked.
            if (!ltop && !rtop) {
                lds_put(T.dat, T.qt, k, merge(dl, dr));
                start[r] = true;
            } else {
                lds_put(T.dat, T.qt, k, ltop ? dr : dl);
            }
        }
    }
    __syncthreads();  // topology, parked children and the first merges are in LDS
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        int sl = tid + r * (int)blockDim.x;
        // A climb passes at most nt nodes (a bound that also ends it should
        // the listed links ever be inconsistent).
        for (int steps = start[r] ? nt : 0; steps > 0; --steps) {
            const int ps = (int)T.par[sl] - 1;
            if (ps < 0) break;
            const int lc = (int)T.cl[ps] - 1, rc = (int)T.cr[ps] - 1;
            const bool mine_left = lc == sl;
            const int other = mine_left ? rc : lc;
            int ss = ps;  // a parked child sits in the parent's own record
            if (other >= 0) {
                const uint32_t bit = 1u << (ps & 31);
                const uint32_t old = __hip_atomic_fetch_or(&T.flag[ps >> 5], bit,
                                                           __ATOMIC_ACQ_REL,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                if (!(old & bit)) break;
                ss = other;
            }
            const int ls = mine_left ? sl : ss, rs = mine_left ? ss : sl;
            lds_put(T.dat, T.qt, ps, merge(lds_node(T.dat, T.qt, ls), lds_node(T.dat, T.qt, rs)));
            sl = ps;
        }
    }
    __syncthreads();  // every listed node is complete in LDS
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const int k = tid + r * (int)blockDim.x;
        if (k >= nt) continue;
        const NodeData d = lds_node(T.dat, T.qt, k);
        float w, open;
        node_open(d, inv_theta, mode, w, open);
        store_node(vk[r], d, w, open, node, nodeQ, boxLo, boxHi);
    }
}

// k_aggregate_top: one block per instance finishes the nodes of more than
// kATile particles: in LDS (top_lds) when the build listed at most kTopCap of
// them, else (at most kTopPhaseNodes) in LDS in phases of at most kTopCap
// nodes by size class (a phase's stores are visible to the next through the
// block's barrier: every access is by this block), else (more nodes, or a
// size class of more than kTopCap nodes) by climbing from the roots of the
// aggregation subtrees with
// arrival marks at workgroup scope (bits in LDS, an LDS hash, or counters in
// global memory for very large instances) -- every arrival at a node of an
// instance is in that instance's block, so no write-back or invalidate across
// XCDs is needed (the ordering the reference's plain atomicAdd, with no
// fence, leaves out); the subtrees written by k_aggregate are visible because
// it completed before this launch.
// While the frontier holds more climbers than the block has threads, every
// climber advances one level per round and the ones that go on (second
// arrivals) form the next, much shorter list; then each remaining climber
// runs to the end.  A thread thus never holds back climbers that have not
// started behind one long climb.  The lists live in the instance's 2n-entry
// frontier area: the current one in the first half, the next in the second.
// Size class of kMidSize particles: the classes below it are the first
// level of the two-level instance pass, the rest the second.
constexpr int kMidBucket = 8 * (kMidShift - kATileShift);
static_assert(kMidShift > kATileShift, "mid-level subtrees hold aggregation subtrees");
constexpr int kTopSmemWords =
    (int)(sizeof(TopLds) > 4 * (kTopLdsNodes / 32) ? sizeof(TopLds) / 4 : kTopLdsNodes / 32);
static_assert(sizeof(TopLds) % 16 == 0, "LDS words");
static_assert(kTopCap % 1024 == 0 && kTopCap <= 65535, "slots + 1 fit 16 bits");

__global__ __launch_bounds__(1024) void k_aggregate_top(
    const SegInfo* __restrict__ segs, const float4* __restrict__ posm, float4* node,
    float4* nodeQ, float4* boxLo, float4* boxHi, const int* __restrict__ parent, int* flag,
    int* front, const int* __restrict__ nfront, const int* __restrict__ tlist,
    const int* __restrict__ ntop, int* tslot, const int* __restrict__ mfail, float inv_theta,
    int mode, int two_level)
{
    __shared__ __align__(16) uint32_t smem[kTopSmemWords];
    __shared__ int lcount;
    __shared__ uint32_t bhist[kTopBuckets];
    __shared__ uint8_t bphase[kTopBuckets];
    __shared__ int nphase;
    uint32_t* const lmem = smem;
    TopLds& T = *reinterpret_cast<TopLds*>(smem);
    const int s = blockIdx.x;
    const int tid = (int)threadIdx.x;
    const SegInfo sg = segs[s];
    const bool mid_done = two_level && tlist && sg.n >= kMidSize && !mfail[s];
    if (mid_done) {
        // Second level of the two-level pass: the nodes of at least kMidSize
        // particles (their smaller children were finished by k_aggregate_mid
        // or k_aggregate), listed in the second half of the instance's
        // frontier area.
        for (int c = tid; c < kTopBuckets; c += blockDim.x) bphase[c] = c < kMidBucket ? 0 : 1;
        if (tid == 0) lcount = 0;
        __syncthreads();
        int* const plist = front + 2 * (size_t)sg.start + sg.n;
        const int nt = ntop[s];
        // Two listed nodes per thread and round, their loads in flight together.
        for (int k = tid; k < nt; k += 2 * (int)blockDim.x) {
            const int k2 = k + (int)blockDim.x;
            const int v = tlist[sg.start + k];
            const int v2 = tlist[sg.start + (k2 < nt ? k2 : k)];
            const float4 bb = node[2 * v + 1];
            const float4 bb2 = node[2 * v2 + 1];
            if (__float_as_int(bb.z) - __float_as_int(bb.y) + 1 >= kMidSize) {
                const int slot = atomicAdd(&lcount, 1);
                if (slot < kTopCap) {
                    plist[slot] = v;
                    tslot[v] = slot;
                }
            }
            if (k2 < nt && __float_as_int(bb2.z) - __float_as_int(bb2.y) + 1 >= kMidSize) {
                const int slot = atomicAdd(&lcount, 1);
                if (slot < kTopCap) {
                    plist[slot] = v2;
                    tslot[v2] = slot;
                }
            }
        }
        __syncthreads();  // the list and its slots are visible to the block
        const int nb = lcount;
        if (nb <= kTopCap) {
            top_lds(T, sg, nb, plist, tslot, posm, node, nodeQ, boxLo, boxHi, parent, inv_theta,
                    mode, bphase, 1);
            return;
        }
        // More such nodes than one LDS pass holds: the one-level paths below
        // finish every listed node (those of the first level again, alike).
        __syncthreads();
    }
    // With the two-level pass, k_aggregate_mid rewrote the slots of the
    // nodes below kMidSize particles of the instances of at least kMidSize
    // (smaller instances have no mid-level roots).  Such an instance gets
    // here only when it listed more than kTopCap nodes (a given-up root alone
    // holds more), so the LDS form below, which reads the build's slots, is
    // not taken; the phases relist their own.
    const int nt = tlist ? ntop[s] : kTopCap + 1;
    if (nt <= kTopCap) {
        if (nt > 0)
            top_lds(T, sg, nt, tlist + sg.start, tslot, posm, node, nodeQ, boxLo, boxHi, parent,
                    inv_theta, mode, nullptr, 0);
        return;
    }
    if (tlist && nt <= kTopPhaseNodes) {
        // Phases in LDS: the listed nodes by size class, consecutive classes
        // taken together while they hold at most kTopCap nodes; phase p
        // finishes its classes after phases 0 .. p-1 finished the smaller
        // ones (a node's children are smaller than it).  Not possible when
        // one class alone holds more than kTopCap nodes.  Each thread keeps
        // its nodes (k = tid + r * 1024) and their classes in registers.
        constexpr int kPerT = kTopPhaseNodes / 1024;
        int pv[kPerT];
        int pc[kPerT];
        for (int k = tid; k < kTopBuckets; k += blockDim.x) bhist[k] = 0u;
#pragma unroll
        for (int r = 0; r < kPerT; ++r) {
            const int k = tid + r * 1024;
            pv[r] = k < nt ? tlist[sg.start + k] : -1;
        }
        float4 pb[kPerT];
#pragma unroll
        for (int r = 0; r < kPerT; ++r)
            pb[r] = pv[r] >= 0 ? node[2 * pv[r] + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kPerT; ++r) {
            pc[r] = -1;
            if (pv[r] >= 0) {
                pc[r] = top_bucket(__float_as_int(pb[r].z) - __float_as_int(pb[r].y) + 1);
                atomicAdd(&bhist[pc[r]], 1u);
            }
        }
        __syncthreads();
        if (tid == 0) {
            int ph = 0;
            uint32_t cum = 0;
            bool ok = true;
#pragma unroll 16
            for (int c = 0; c < kTopBuckets; ++c) {
                const uint32_t h = bhist[c];
                if (h > (uint32_t)kTopCap) ok = false;
                if (cum + h > (uint32_t)kTopCap) {
                    ++ph;
                    cum = 0;
                }
                cum += h;
                bphase[c] = (uint8_t)ph;
            }
            nphase = ok ? ph + 1 : 0;
        }
        __syncthreads();
        const int np = nphase;
        if (np > 0) {
            int* plist = front + 2 * (size_t)sg.start;
            for (int ph = 0; ph < np; ++ph) {
                if (tid == 0) lcount = 0;
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kPerT; ++r) {
                    if (pc[r] >= 0 && bphase[pc[r]] == ph) {
                        const int slot = atomicAdd(&lcount, 1);
                        plist[slot] = pv[r];
                        tslot[pv[r]] = slot;
                    }
                }
                __syncthreads();  // the phase's list and slots are visible to the block
                top_lds(T, sg, lcount, plist, tslot, posm, node, nodeQ, boxLo, boxHi, parent,
                        inv_theta, mode, bphase, ph);
                __syncthreads();  // its records are visible to the next phase
            }
            return;
        }
    }
    const int nf = nfront[s];
    if (nf == 0) return;
    // nf - 1 nodes above the aggregation subtrees get a mark.
    const int marks = sg.n <= kTopLdsNodes ? kMarksBits
                      : nf <= kTopHashMax  ? kMarksHash
                                           : kMarksGlobal;
    if (marks == kMarksBits)
        for (int k = threadIdx.x; k < (sg.n + 31) / 32; k += blockDim.x) lmem[k] = 0u;
    else if (marks == kMarksHash)
        for (int k = threadIdx.x; k < (1 << kTopHashBits); k += blockDim.x) lmem[k] = 0xFFFFFFFFu;
    climb_rounds(sg, front + 2 * (size_t)sg.start, sg.n, nf, posm, node, nodeQ, boxLo, boxHi,
                 parent, flag, lmem, marks, inv_theta, mode, &lcount);
}

// k_aggregate_mid: the first level of the two-level instance pass.  Block b
// takes the mid-level roots b, b + gridDim.x, ... that the build recorded.
// For root r over particles [rf, rl], the instance's listed nodes whose
// particle range lies inside r's (the nodes of more than kATile particles of
// r's subtree, fewer than rl - rf + 1) are listed with their slots at
// rf - start of the second half of the instance's frontier area (the lists of
// disjoint roots do not overlap; the first half keeps the frontier), and
// top_lds finishes them as the phase of the size classes below kMidSize: r's
// parent, of at least kMidSize particles, is of the other phase, so the
// climbs stop at r.  A root with more than kTopCap such nodes is left to
// k_aggregate_top's one-level paths: mfail[instance] = 1.
__global__ __launch_bounds__(1024) void k_aggregate_mid(
    const SegInfo* __restrict__ segs, const float4* __restrict__ posm, float4* node,
    float4* nodeQ, float4* boxLo, float4* boxHi, const int* __restrict__ parent,
    const int* __restrict__ mids, const int* __restrict__ nmid, const int* __restrict__ tlist,
    const int* __restrict__ ntop, int* tslot, int* front, int* mfail, float inv_theta, int mode)
{
    __shared__ __align__(16) uint32_t smem[sizeof(TopLds) / 4];
    __shared__ uint8_t phase_of[kTopBuckets];
    __shared__ int lcount;
    TopLds& T = *reinterpret_cast<TopLds*>(smem);
    const int tid = (int)threadIdx.x;
    for (int c = tid; c < kTopBuckets; c += blockDim.x) phase_of[c] = c < kMidBucket ? 0 : 1;
    const int nm = *nmid;
    for (int j = blockIdx.x; j < nm; j += gridDim.x) {
        const int r = mids[2 * j], s = mids[2 * j + 1];
        const SegInfo sg = segs[s];
        const float4 rb = node[2 * r + 1];
        const int rf = __float_as_int(rb.y), rl = __float_as_int(rb.z);
        int* const mlist = front + 2 * (size_t)sg.start + sg.n + (rf - sg.start);
        if (tid == 0) lcount = 0;
        __syncthreads();  // the previous root's use of the LDS is over
        const int nt = ntop[s];
        // Two listed nodes per thread and round: both lookups, then both
        // ranges, in flight together.
        for (int k = tid; k < nt; k += 2 * (int)blockDim.x) {
            const int k2 = k + (int)blockDim.x;
            const int v = tlist[sg.start + k];
            const int v2 = tlist[sg.start + (k2 < nt ? k2 : k)];
            const float4 bb = node[2 * v + 1];
            const float4 bb2 = node[2 * v2 + 1];
            if (__float_as_int(bb.y) >= rf && __float_as_int(bb.z) <= rl) {
                const int slot = atomicAdd(&lcount, 1);
                mlist[slot] = v;
                tslot[v] = slot;
            }
            if (k2 < nt && __float_as_int(bb2.y) >= rf && __float_as_int(bb2.z) <= rl) {
                const int slot = atomicAdd(&lcount, 1);
                mlist[slot] = v2;
                tslot[v2] = slot;
            }
        }
        __syncthreads();  // the list and its slots are visible to the block
        const int nl = lcount;
        if (nl <= kTopCap)
            top_lds(T, sg, nl, mlist, tslot, posm, node, nodeQ, boxLo, boxHi, parent, inv_theta,
                    mode, phase_of, 0);
        else if (tid == 0)
            mfail[s] = 1;
        __syncthreads();
    }
}

__device__ __forceinline__ uint8_t to_u8(float f) { return (uint8_t)f2u_sat(f); }

// cuda_engine.cu:185-215 -- speed colour for types 0 (fire) and 2 (ice);
// other types keep their colour.  Written with the multiply-adds nvcc fuses.
__device__ __forceinline__ uint32_t speed_colour(uint32_t rgbt, float vx, float vy, float vz)
{
    const uint32_t type = rgbt & 0xFFu;
    if (type != 0u && type != 2u) return rgbt;
    const float sp = sqrtf(__builtin_fmaf(vz, vz, __builtin_fmaf(vx, vx, vy * vy)));
    const float t = fminf(sp / 150.0f, 1.0f);
    uint32_t r, g, b;
    if (type == 0u) {
        if (t < 0.5f) {
            const float u = t * 2.0f;
            r = 255u; g = to_u8(__builtin_fmaf(150.0f, u, 50.0f)); b = 0u;
        } else {
            const float u = (t - 0.5f) * 2.0f;
            r = 255u; g = to_u8(__builtin_fmaf(55.0f, u, 200.0f)); b = to_u8(150.0f * u);
        }
    } else {
        if (t < 0.5f) {
            const float u = t * 2.0f;
            r = 0u; g = to_u8(__builtin_fmaf(100.0f, u, 100.0f)); b = 255u;
        } else {
            const float u = (t - 0.5f) * 2.0f;
            r = to_u8(200.0f * u); g = to_u8(__builtin_fmaf(55.0f, u, 200.0f)); b = 255u;
        }
    }
    return pack_rgbt(type, r, g, b);
}

// cuda_engine.cu:164-216 -- kick then drift: v += a*dt, x += v*dt (the fused
// multiply-adds nvcc forms by default).  Compensated precision keeps
// x = hi + lo with the drift sum formed in FP64: X = hi + (lo + v*dt),
// hi = fl32(X), lo = fl32(X - hi).  Diagnostics (FP64, per quarter tile):
// kinetic and potential energy and momentum at the force time, with the
// velocity re-synchronised to v_old + a*dt/2.
// Largest |a| of the block's particles (all threads must call).
__device__ __forceinline__ float block_max_accel(const float4& a, bool valid)
{
    __shared__ float sm[4];
    float m2 = valid ? __builtin_fmaf(a.z, a.z, __builtin_fmaf(a.y, a.y, a.x * a.x)) : 0.0f;
    if (!(m2 == m2)) m2 = __builtin_inff();
    m2 = wave_max(m2);
    if (lane_id() == 0) sm[threadIdx.x >> 6] = m2;
    __syncthreads();
    const float r = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
    __syncthreads();
    return sqrtf(r);
}

// Per quarter tile: the largest |a| (the quantity the reference reduces with
// atomicMax at cuda_engine.cu:157-160 and never reads).
__global__ __launch_bounds__(256) void k_accmax(const TileInfo* __restrict__ tiles,
                                                const SegInfo* __restrict__ segs,
                                                const float4* __restrict__ acc,
                                                float* __restrict__ amax_part)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    const float4 a = q.valid ? acc[q.i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float m = block_max_accel(a, q.valid);
    if (threadIdx.x == 0) amax_part[blockIdx.x] = m;
}

// Adaptive step of each instance: dt_s = min(dt_max, sqrt(2 eta eps / max|a|)),
// eps the softening length.
__global__ __launch_bounds__(256) void k_adapt_dt(const SegInfo* __restrict__ segs,
                                                  const float* __restrict__ amax_part,
                                                  float dt_max, float eta, float eps,
                                                  float* __restrict__ dt_inst)
{
    const SegInfo sg = segs[blockIdx.x];
    float m = 0.0f;
    for (int b = 4 * sg.tile0 + (int)threadIdx.x; b < 4 * (sg.tile0 + sg.ntiles); b += 256)
        m = fmaxf(m, amax_part[b]);
    m = wave_max(m);
    __shared__ float sm[4];
    if (lane_id() == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
        dt_inst[blockIdx.x] = m > 0.0f ? fminf(dt_max, sqrtf(2.0f * eta * eps / m)) : dt_max;
    }
}

__global__ __launch_bounds__(256) void k_integrate(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const float4* __restrict__ acc, float4* __restrict__ posm, float4* __restrict__ vel,
    float4* __restrict__ plo, float dt_fixed, const float* __restrict__ dt_inst, int compensated,
    int colours, int diagnostics, double* __restrict__ diag_part, float* __restrict__ amax_part,
    double* __restrict__ tsim, unsigned* err, uint32_t* __restrict__ bpart)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    const float dt = dt_inst ? dt_inst[q.seg] : dt_fixed;
    float4 pn = make_float4(0.f, 0.f, 0.f, 0.f);  // the new position
    double d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // Every stream of the particle in one round trip, before the block's
    // maximum is taken (a lane past the tile's end reads particle 0 and uses
    // none of it).
    const int qi = q.valid ? q.i : 0;
    const float4 a = acc[qi];
    float4 v = vel[qi];
    float4 p = posm[qi];
    float4 lo = compensated ? plo[qi] : make_float4(0.f, 0.f, 0.f, 0.f);
    {
        const float m = block_max_accel(a, q.valid);
        if (threadIdx.x == 0) {
            amax_part[blockIdx.x] = m;
            if ((int)blockIdx.x == 4 * segs[q.seg].tile0) tsim[q.seg] += (double)dt;
        }
    }
    if (q.valid) {
        if (!(isfinite(a.x) && isfinite(a.y) && isfinite(a.z))) atomicOr(err, kErrNonFinite);
        if (diagnostics) {
            const double hdt = 0.5 * (double)dt;
            const double sx = (double)v.x + (double)a.x * hdt;
            const double sy = (double)v.y + (double)a.y * hdt;
            const double sz = (double)v.z + (double)a.z * hdt;
            const double m = (double)p.w;
            d[0] = 0.5 * m * (sx * sx + sy * sy + sz * sz);
            d[1] = 0.5 * m * (double)a.w;
            d[2] = m * sx;
            d[3] = m * sy;
            d[4] = m * sz;
        }
        v.x = __builtin_fmaf(a.x, dt, v.x);
        v.y = __builtin_fmaf(a.y, dt, v.y);
        v.z = __builtin_fmaf(a.z, dt, v.z);
        if (compensated) {
            const double X = (double)p.x + ((double)lo.x + (double)v.x * (double)dt);
            const double Y = (double)p.y + ((double)lo.y + (double)v.y * (double)dt);
            const double Z = (double)p.z + ((double)lo.z + (double)v.z * (double)dt);
            p.x = (float)X; p.y = (float)Y; p.z = (float)Z;
            lo.x = (float)(X - (double)p.x);
            lo.y = (float)(Y - (double)p.y);
            lo.z = (float)(Z - (double)p.z);
            plo[q.i] = lo;
        } else {
            p.x = __builtin_fmaf(v.x, dt, p.x);
            p.y = __builtin_fmaf(v.y, dt, p.y);
            p.z = __builtin_fmaf(v.z, dt, p.z);
        }
        if (colours)
            v.w = __uint_as_float(speed_colour(__float_as_uint(v.w), v.x, v.y, v.z));
        vel[q.i] = v;
        posm[q.i] = p;
        pn = p;
    }
    // The next step's bounds partials, from the positions just written (what
    // k_bounds would read at the start of that step).
    if (bpart) bounds_partial(pn, q.valid, bpart);
    if (diagnostics) {
        __shared__ double sd[4][5];
#pragma unroll
        for (int k = 0; k < 5; ++k) d[k] = wave_sum(d[k]);
        const int w = threadIdx.x >> 6;
        if (lane_id() == 0)
            for (int k = 0; k < 5; ++k) sd[w][k] = d[k];
        __syncthreads();
        if (threadIdx.x < 5) {
            const int k = threadIdx.x;
            diag_part[5 * (size_t)blockIdx.x + k] = ((sd[0][k] + sd[1][k]) + sd[2][k]) + sd[3][k];
        }
    }
}

// Per-instance sum of the quarter-tile diagnostics, in a fixed order.
__global__ void k_diag_final(const SegInfo* __restrict__ segs, int nseg,
                             const double* __restrict__ part, double* __restrict__ out)
{
    const int s = blockIdx.x;
    if (s >= nseg) return;
    const SegInfo sg = segs[s];
    const int k = threadIdx.x;
    if (k >= 5) return;
    double acc = 0.0;
    for (int b = 4 * sg.tile0; b < 4 * (sg.tile0 + sg.ntiles); ++b) acc += part[5 * (size_t)b + k];
    out[5 * s + k] = acc;
}

// cuda_engine.cu:54-70 -- AoS Star records (32 B) into the SoA state.
__global__ void k_unpack_stars(const uint8_t* __restrict__ stars, int64_t n, int64_t offset,
                               float4* __restrict__ posm, float4* __restrict__ vel,
                               float4* __restrict__ plo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rec = reinterpret_cast<const float4*>(stars + 32 * i);
    const float4 a = rec[0];  // x y z vx
    const float4 b = rec[1];  // vy vz mass rgbt-bytes
    const uint32_t bytes = __float_as_uint(b.w);  // type, r, g, b in memory order
    posm[offset + i] = make_float4(a.x, a.y, a.z, b.z);
    vel[offset + i] = make_float4(a.w, b.x, b.y, __uint_as_float(bytes));
    plo[offset + i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// cuda_engine.cu:73-88 -- SoA state back into Star records.
__global__ void k_pack_stars(int64_t n, int64_t offset, const float4* __restrict__ posm,
                             const float4* __restrict__ vel, uint8_t* __restrict__ stars)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[offset + i];
    const float4 v = vel[offset + i];
    float4* rec = reinterpret_cast<float4*>(stars + 32 * i);
    rec[0] = make_float4(p.x, p.y, p.z, v.x);
    rec[1] = make_float4(v.y, v.z, p.w, v.w);
}

// main.cpp:264-267 -- PlaybackStar {x, y, z, r, g, b, type}.
__global__ void k_pack_playback(int64_t n, int64_t offset, const float4* __restrict__ posm,
                                const float4* __restrict__ vel, uint8_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = posm[offset + i];
    const uint32_t rgbt = __float_as_uint(vel[offset + i].w);
    const uint32_t type = rgbt & 0xFFu, r = (rgbt >> 8) & 0xFFu, g = (rgbt >> 16) & 0xFFu,
                   b = rgbt >> 24;
    float4* rec = reinterpret_cast<float4*>(out + 16 * i);
    rec[0] = make_float4(p.x, p.y, p.z, __uint_as_float(r | (g << 8) | (b << 16) | (type << 24)));
}

}  // namespace gcs
