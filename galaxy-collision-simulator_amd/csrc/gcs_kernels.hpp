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

// Start-of-step reset: cube bounds to empty, work-queue and interaction
// counters to zero.  A kernel rather than hipMemsetAsync so the step can be
// captured into a hipGraph.
__global__ void k_reset(int nseg, uint32_t* seg_bounds, unsigned* qctr,
                        unsigned long long* counters)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nseg) {
        seg_bounds[2 * s] = 0xFFFFFFFFu;
        seg_bounds[2 * s + 1] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x < kNumXcd) qctr[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 2) counters[threadIdx.x] = 0ull;
}

// main.cpp:152-166 -- per-instance cube [min over x,y,z, max over x,y,z].
// NaN coordinates are ignored, as the reference's `<` comparisons ignore them.
__global__ __launch_bounds__(256) void k_bounds(const TileInfo* __restrict__ tiles,
                                                const SegInfo* __restrict__ segs,
                                                const float4* __restrict__ posm,
                                                uint32_t* seg_bounds)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    float lo = __builtin_inff(), hi = -__builtin_inff();
    if (q.valid) {
        const float4 p = posm[q.i];
        lo = fminf(p.x, fminf(p.y, p.z));
        hi = fmaxf(p.x, fmaxf(p.y, p.z));
        if (lo != lo) lo = __builtin_inff();
        if (hi != hi) hi = -__builtin_inff();
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    __shared__ float slo[4], shi[4];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        slo[w] = lo;
        shi[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        lo = fminf(fminf(slo[0], slo[1]), fminf(slo[2], slo[3]));
        hi = fmaxf(fmaxf(shi[0], shi[1]), fmaxf(shi[2], shi[3]));
        atomicMin(&seg_bounds[2 * q.seg], fenc(lo));
        atomicMax(&seg_bounds[2 * q.seg + 1], fenc(hi));
    }
}

// karras_bvh.cuh:34-41 -- keys of the current order, payload = global index.
__global__ __launch_bounds__(256) void k_morton(const TileInfo* __restrict__ tiles,
                                                const SegInfo* __restrict__ segs,
                                                const float4* __restrict__ posm,
                                                const uint32_t* __restrict__ seg_bounds,
                                                int use_given, float gmin, float gmax,
                                                uint64_t* __restrict__ keys,
                                                uint32_t* __restrict__ vals)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    float mn = gmin, mx = gmax;
    if (!use_given) {
        mn = fdec(seg_bounds[2 * q.seg]);
        mx = fdec(seg_bounds[2 * q.seg + 1]);
    }
    const float4 p = posm[q.i];
    keys[q.i] = morton_key(p.x, p.y, p.z, mn, mx);
    vals[q.i] = (uint32_t)q.i;
}

// karras_bvh.cuh:44-70 -- gather the state into sorted order (and, for the
// reference traversal, karras_bvh.cuh:134-164's leaf records).
__global__ __launch_bounds__(256) void k_permute(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const uint32_t* __restrict__ perm, const float4* __restrict__ posm_in,
    const float4* __restrict__ vel_in, const float4* __restrict__ plo_in,
    float4* __restrict__ posm_out, float4* __restrict__ vel_out, float4* __restrict__ plo_out,
    int compensated, int write_leaves, float4* __restrict__ node, float4* __restrict__ nodeQ,
    uint32_t* __restrict__ gmark, uint32_t* __restrict__ gidx)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    gmark[q.i] = 0u;
    gidx[q.i] = 0u;
    const int src = (int)perm[q.i];
    const float4 p = posm_in[src];
    posm_out[q.i] = p;
    vel_out[q.i] = vel_in[src];
    if (compensated) plo_out[q.i] = plo_in[src];
    if (write_leaves) {
        const int v = q.i + segs[q.seg].leaf_base;
        node[2 * v] = p;
        node[2 * v + 1] = make_float4(0.0f, __int_as_float(q.i), __int_as_float(q.i),
                                      __int_as_float(-1));
        nodeQ[2 * v] = make_float4(0.f, 0.f, 0.f, 0.f);
        nodeQ[2 * v + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// karras_bvh.cuh:23-31 -- common prefix of the 96-bit key (key, local index).
__device__ __forceinline__ int delta(int i, int j, int n, const uint64_t* __restrict__ k)
{
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + clz32s((uint32_t)(i ^ j));
    return clz64s(a ^ b);
}

// karras_bvh.cuh:73-131 -- one internal node per particle but the last of
// each instance, built on instance-local indices so every instance's tree is
// the stand-alone tree.  Writes range/split, parents and clears the arrival
// flag of the node it builds.  Also marks traversal groups: the maximal
// subtrees of at most kGroup particles (a child of at most kGroup under a
// parent of more, or the root of a small instance) are flagged at their first
// particle (gmark, gidx = 1) with their size (gcnt).
__global__ __launch_bounds__(256) void k_build(const TileInfo* __restrict__ tiles,
                                               const SegInfo* __restrict__ segs,
                                               const uint64_t* __restrict__ skeys,
                                               float4* __restrict__ node, int* __restrict__ parent,
                                               int* __restrict__ flag, uint32_t* __restrict__ gmark,
                                               uint32_t* __restrict__ gidx, int* __restrict__ gcnt)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    const SegInfo sg = segs[q.seg];
    const int n = sg.n, i = q.local;
    if (i == 0) {
        parent[sg.root] = -1;
        if (n <= kGroup) {
            gmark[sg.start] = 1u;
            gidx[sg.start] = 1u;
            gcnt[sg.start] = n;
        }
    }
    if (i >= n - 1) return;
    const uint64_t* k = skeys + sg.start;
    const int dir = (delta(i, i + 1, n, k) - delta(i, i - 1, n, k)) > 0 ? 1 : -1;
    const int dmin = delta(i, i - dir, n, k);
    int lmax = 2;
    while (delta(i, i + lmax * dir, n, k) > dmin) lmax *= 2;
    int len = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(i, i + (len + t) * dir, n, k) > dmin) len += t;
    const int j = i + len * dir;
    const int dnode = delta(i, j, n, k);
    int s = 0, step = len;
    do {
        step = (step + 1) >> 1;
        if (delta(i, i + (s + step) * dir, n, k) > dnode) s += step;
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
    if (hi - lo + 1 > kGroup) {
        const int nl = split - lo + 1, nr = hi - split;
        if (nl <= kGroup) {
            gmark[sg.start + lo] = 1u;
            gidx[sg.start + lo] = 1u;
            gcnt[sg.start + lo] = nl;
        }
        if (nr <= kGroup) {
            gmark[gsplit + 1] = 1u;
            gidx[gsplit + 1] = 1u;
            gcnt[gsplit + 1] = nr;
        }
    }
}

// Writes the group table in particle (Morton) order from the exclusive scan
// of the group marks, and the number of groups.
__global__ __launch_bounds__(256) void k_group_emit(const TileInfo* __restrict__ tiles,
                                                    const SegInfo* __restrict__ segs,
                                                    const uint32_t* __restrict__ gmark,
                                                    const uint32_t* __restrict__ gidx,
                                                    const int* __restrict__ gcnt,
                                                    GroupInfo* __restrict__ groups,
                                                    int* __restrict__ ngroups, int n_total)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    const uint32_t m = gmark[q.i];
    if (m) {
        GroupInfo g;
        g.first = q.i;
        g.count = gcnt[q.i];
        g.seg = q.seg;
        g.pad = 0;
        groups[gidx[q.i]] = g;
    }
    if (q.i == n_total - 1) *ngroups = (int)(gidx[q.i] + m);
}

struct ChildData {
    float cx, cy, cz, m;
    float lx, ly, lz, hx, hy, hz;
    float q0, q1, q2, q3, q4, q5;
};

__device__ __forceinline__ void load_child(bool leaf, int idx, const float4* __restrict__ posm,
                                           const float4* node, const float4* nodeQ,
                                           const float4* boxLo, const float4* boxHi,
                                           ChildData& c)
{
    if (leaf) {
        const float4 p = posm[idx];
        c.cx = c.lx = c.hx = p.x;
        c.cy = c.ly = c.hy = p.y;
        c.cz = c.lz = c.hz = p.z;
        c.m = p.w;
        c.q0 = c.q1 = c.q2 = c.q3 = c.q4 = c.q5 = 0.0f;
    } else {
        const float4 a = node[2 * idx];
        const float4 lo = boxLo[idx], hi = boxHi[idx];
        const float4 qa = nodeQ[2 * idx], qb = nodeQ[2 * idx + 1];
        c.cx = a.x; c.cy = a.y; c.cz = a.z; c.m = a.w;
        c.lx = lo.x; c.ly = lo.y; c.lz = lo.z;
        c.hx = hi.x; c.hy = hi.y; c.hz = hi.z;
        c.q0 = qa.x; c.q1 = qa.y; c.q2 = qa.z; c.q3 = qa.w;
        c.q4 = qb.x; c.q5 = qb.y;
    }
}

// Adds a child's traceless quadrupole about the parent COM (same operation
// order as the oracle): Q += Qc + mc*(3 s s^T - |s|^2 I), s = child COM -
// parent COM; components xx, yy, zz, xy, xz, yz.
__device__ __forceinline__ void quad_shift(float q[6], const ChildData& c, float px, float py,
                                           float pz)
{
#pragma clang fp contract(off)
    const float sx = c.cx - px, sy = c.cy - py, sz = c.cz - pz;
    const float s2 = sx * sx + sy * sy + sz * sz;
    q[0] = q[0] + (c.q0 + c.m * (3.0f * sx * sx - s2));
    q[1] = q[1] + (c.q1 + c.m * (3.0f * sy * sy - s2));
    q[2] = q[2] + (c.q2 + c.m * (3.0f * sz * sz - s2));
    q[3] = q[3] + (c.q3 + c.m * (3.0f * sx * sy));
    q[4] = q[4] + (c.q4 + c.m * (3.0f * sx * sz));
    q[5] = q[5] + (c.q5 + c.m * (3.0f * sy * sz));
}

// Merge of one internal node from its two children (karras_bvh.cuh:190-221):
//   AABB = union; width = max extent; m = ml + mr;
//   com = (cl*ml + cr*mr)/m, or the AABB centre when m == 0.
// Extension: opening value = (width*inv_theta + |com - AABB centre|)^2 for the
// group traversal (mode 1), or width for the reference traversal (mode 0);
// quadrupole moments by the parallel-axis shift.
__device__ __forceinline__ void merge_node(int cur, int int_base, const float4* __restrict__ posm,
                                           float4* node, float4* nodeQ, float4* boxLo,
                                           float4* boxHi, float inv_theta, int mode)
{
#pragma clang fp contract(off)
    const float4 bb = node[2 * cur + 1];
    const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
              split = __float_as_int(bb.w);
    ChildData l, r;
    load_child(first == split, first == split ? split : split + int_base, posm, node, nodeQ,
               boxLo, boxHi, l);
    load_child(last == split + 1, last == split + 1 ? split + 1 : split + 1 + int_base, posm,
               node, nodeQ, boxLo, boxHi, r);
    const float lx = fminf(l.lx, r.lx), ly = fminf(l.ly, r.ly), lz = fminf(l.lz, r.lz);
    const float hx = fmaxf(l.hx, r.hx), hy = fmaxf(l.hy, r.hy), hz = fmaxf(l.hz, r.hz);
    const float w = fmaxf(hx - lx, fmaxf(hy - ly, hz - lz));
    const float mt = l.m + r.m;
    float cx, cy, cz;
    if (mt > 0.0f) {
        cx = (l.cx * l.m + r.cx * r.m) / mt;
        cy = (l.cy * l.m + r.cy * r.m) / mt;
        cz = (l.cz * l.m + r.cz * r.m) / mt;
    } else {
        cx = (lx + hx) * 0.5f;
        cy = (ly + hy) * 0.5f;
        cz = (lz + hz) * 0.5f;
    }
    float open;
    if (mode == 1) {
        const float dx = cx - (lx + hx) * 0.5f;
        const float dy = cy - (ly + hy) * 0.5f;
        const float dz = cz - (lz + hz) * 0.5f;
        const float off = sqrtf(dx * dx + dy * dy + dz * dz);
        const float ro = w * inv_theta + off;
        open = ro * ro;
    } else {
        open = w;
    }
    float qq[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    quad_shift(qq, l, cx, cy, cz);
    quad_shift(qq, r, cx, cy, cz);
    node[2 * cur] = make_float4(cx, cy, cz, mt);
    node[2 * cur + 1].x = open;
    nodeQ[2 * cur] = make_float4(qq[0], qq[1], qq[2], qq[3]);
    nodeQ[2 * cur + 1] = make_float4(qq[4], qq[5], w, 0.0f);
    boxLo[cur] = make_float4(lx, ly, lz, 0.0f);
    boxHi[cur] = make_float4(hx, hy, hz, 0.0f);
}

// karras_bvh.cuh:175-225 -- bottom-up aggregation, one block per tile and one
// thread per leaf.  A thread climbs from its leaf; at each node the first of
// the two arrivals stops and the second merges the node.  Nodes whose
// particle range lies inside the block's tile are only ever reached by this
// block, so their arrival counter is an LDS word with workgroup-scope
// acquire-release; only nodes whose range crosses the tile use the global
// counter with agent-scope acquire-release, which makes the children written
// on another XCD (whose L2 is not coherent with this one) visible -- the
// ordering the reference's plain atomicAdd leaves out.
__global__ __launch_bounds__(kTile) void k_aggregate(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, float4* node, float4* nodeQ, float4* boxLo, float4* boxHi,
    const int* __restrict__ parent, int* flag, float inv_theta, int mode)
{
    __shared__ int lflag[kTile];
    const TileInfo t = tiles[blockIdx.x];
    lflag[threadIdx.x] = 0;
    __syncthreads();
    if ((int)threadIdx.x >= t.count) return;
    const SegInfo sg = segs[t.seg];
    const int t0 = t.start, t1 = t.start + t.count - 1;
    int cur = parent[t.start + (int)threadIdx.x + sg.leaf_base];
    while (cur != -1) {
        const float4 bb = node[2 * cur + 1];
        const int first = __float_as_int(bb.y), last = __float_as_int(bb.z);
        int old;
        if (first >= t0 && last <= t1) {
            // An internal node is numbered by one end of its range.
            const int slot = cur - sg.int_base - t0;
            old = __hip_atomic_fetch_add(&lflag[slot], 1, __ATOMIC_ACQ_REL,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
            old = __hip_atomic_fetch_add(&flag[cur], 1, __ATOMIC_ACQ_REL,
                                         __HIP_MEMORY_SCOPE_AGENT);
        }
        if (old == 0) return;
        merge_node(cur, sg.int_base, posm, node, nodeQ, boxLo, boxHi, inv_theta, mode);
        cur = parent[cur];
    }
}

__device__ __forceinline__ uint8_t to_u8(float f) { return (uint8_t)f2u_sat(f); }

// cuda_engine.cu:185-215 -- speed colour for types 0 (fire) and 2 (ice);
// other types keep their colour.  Written with the multiply-adds nvcc fuses.
__device__ __forceinline__ uint32_t speed_colour(uint32_t rgbt, float vx, float vy, float vz)
{
    const uint32_t type = rgbt & 0xFFu;
    if (type != 0u && type != 2u) return rgbt;
    const float sp = sqrtf(__builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx)));
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
__global__ __launch_bounds__(256) void k_integrate(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const float4* __restrict__ acc, float4* __restrict__ posm, float4* __restrict__ vel,
    float4* __restrict__ plo, float dt, int compensated, int colours, int diagnostics,
    double* __restrict__ diag_part, unsigned* err)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    double d[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (q.valid) {
        const float4 a = acc[q.i];
        float4 v = vel[q.i];
        float4 p = posm[q.i];
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
            float4 lo = plo[q.i];
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
    }
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
