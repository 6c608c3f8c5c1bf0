// gcs_traverse.hpp -- Barnes-Hut traversals.
//
// k_traverse_group: one wave per group of up to 64 consecutive (Morton-sorted)
// particles of one instance.  The wave walks the instance's tree once for the
// whole group:
//   pop up to 64 nodes (one per lane) from the wave's stack;
//   node accepted  when the squared distance from its COM to the group's
//                  bounding box exceeds its opening value (w/theta + offset)^2:
//                  appended to the LDS multipole list;
//   node opened    otherwise: each child with <= leaf particles is appended to
//                  the LDS range list (summed directly), larger children are
//                  pushed;
//   lists are evaluated when 64 entries are pending: every lane accumulates
//   its own particle against each entry, read from LDS as a broadcast.
// Persistent grid: waves take groups from eight queues, one per XCD, each
// holding a contiguous run of groups (so neighbouring groups, which touch the
// same nodes, share an L2), and steal from the other queues when theirs is
// empty.  Self-interaction contributes zero force (zero separation) and its
// potential term is removed at the end, which is index self-exclusion.
//
// k_traverse_reference: the reference's per-particle traversal
// (cuda_engine.cu:91-161) with the same decisions, for the cuda_engine.h entry
// points.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

constexpr int kNodeBuf = 128;
constexpr int kRangeBuf = 192;

struct WaveLds {
    float4 na[kNodeBuf];
    float4 nq0[kNodeBuf];
    float2 nq1[kNodeBuf];
    int2 rng[kRangeBuf];
    int rpre[kRangeBuf];
    float4 src[kWave];
};

struct Acc {
    float ax, ay, az, ph;
};

// Particle source: G m d / (|d|^2 + eps^2)^{3/2} (G applied at the end).
template <bool POT>
__device__ __forceinline__ void interact_p(Acc& a, float px, float py, float pz, float4 s,
                                           float eps2)
{
    const float dx = s.x - px, dy = s.y - py, dz = s.z - pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float minv = s.w * inv;
    const float f = minv * (inv * inv);
    a.ax = __builtin_fmaf(dx, f, a.ax);
    a.ay = __builtin_fmaf(dy, f, a.ay);
    a.az = __builtin_fmaf(dz, f, a.az);
    if (POT) a.ph += minv;
}

// Node source with its traceless quadrupole Q (softened expansion):
//   a += (M/r^3 + 5/2 (d.Qd)/r^7) d - Qd/r^5,  phi += M/r + (d.Qd)/(2 r^5).
template <bool POT, bool QUAD>
__device__ __forceinline__ void interact_n(Acc& a, float px, float py, float pz, float4 c,
                                           float4 q0, float2 q1, float eps2)
{
    const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv2 = inv * inv;
    const float minv = c.w * inv;
    if (!QUAD) {
        const float f = minv * inv2;
        a.ax = __builtin_fmaf(dx, f, a.ax);
        a.ay = __builtin_fmaf(dy, f, a.ay);
        a.az = __builtin_fmaf(dz, f, a.az);
        if (POT) a.ph += minv;
        return;
    }
    const float qx = __builtin_fmaf(q0.x, dx, __builtin_fmaf(q0.w, dy, q1.x * dz));
    const float qy = __builtin_fmaf(q0.w, dx, __builtin_fmaf(q0.y, dy, q1.y * dz));
    const float qz = __builtin_fmaf(q1.x, dx, __builtin_fmaf(q1.y, dy, q0.z * dz));
    const float dqd = __builtin_fmaf(dx, qx, __builtin_fmaf(dy, qy, dz * qz));
    const float inv3 = inv * inv2;
    const float inv5 = inv3 * inv2;
    const float t = dqd * inv5;
    const float f = __builtin_fmaf(2.5f * t, inv2, c.w * inv3);
    a.ax = __builtin_fmaf(dx, f, __builtin_fmaf(-qx, inv5, a.ax));
    a.ay = __builtin_fmaf(dy, f, __builtin_fmaf(-qy, inv5, a.ay));
    a.az = __builtin_fmaf(dz, f, __builtin_fmaf(-qz, inv5, a.az));
    if (POT) a.ph += __builtin_fmaf(0.5f, t, minv);
}

template <bool POT, bool QUAD>
__device__ __forceinline__ void eval_nodes(WaveLds& L, int k, Acc& a, float px, float py, float pz,
                                           float eps2)
{
    int j = 0;
    for (; j + 1 < k; j += 2) {
        const float4 c0 = L.na[j], c1 = L.na[j + 1];
        float4 q00 = make_float4(0, 0, 0, 0), q01 = q00;
        float2 q10 = make_float2(0, 0), q11 = q10;
        if (QUAD) {
            q00 = L.nq0[j]; q01 = L.nq0[j + 1];
            q10 = L.nq1[j]; q11 = L.nq1[j + 1];
        }
        interact_n<POT, QUAD>(a, px, py, pz, c0, q00, q10, eps2);
        interact_n<POT, QUAD>(a, px, py, pz, c1, q01, q11, eps2);
    }
    if (j < k) {
        float4 q0 = make_float4(0, 0, 0, 0);
        float2 q1 = make_float2(0, 0);
        if (QUAD) { q0 = L.nq0[j]; q1 = L.nq1[j]; }
        interact_n<POT, QUAD>(a, px, py, pz, L.na[j], q0, q1, eps2);
    }
}

// Sums the pending ranges; returns the number of source particles.
template <bool POT>
__device__ __forceinline__ int eval_ranges(WaveLds& L, int k, Acc& a, float px, float py,
                                           float pz, float eps2,
                                           const float4* __restrict__ posm)
{
    const int l = lane_id();
    int c0 = l < k ? L.rng[l].y : 0;
    int c1 = l + 64 < k ? L.rng[l + 64].y : 0;
    int c2 = l + 128 < k ? L.rng[l + 128].y : 0;
    const int i0 = wave_scan_incl(c0), i1 = wave_scan_incl(c1), i2 = wave_scan_incl(c2);
    const int t0 = __shfl(i0, 63), t1 = __shfl(i1, 63), t2 = __shfl(i2, 63);
    if (l < k) L.rpre[l] = i0 - c0;
    if (l + 64 < k) L.rpre[l + 64] = t0 + i1 - c1;
    if (l + 128 < k) L.rpre[l + 128] = t0 + t1 + i2 - c2;
    const int total = t0 + t1 + t2;
    __builtin_amdgcn_wave_barrier();
    for (int base = 0; base < total; base += kWave) {
        const int f = base + l;
        float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
        if (f < total) {
            int r = 0;
#pragma unroll
            for (int step = 128; step > 0; step >>= 1) {
                const int c = r + step;
                if (c < k && L.rpre[c] <= f) r = c;
            }
            s = posm[L.rng[r].x + (f - L.rpre[r])];
        }
        __builtin_amdgcn_wave_barrier();
        L.src[l] = s;
        __builtin_amdgcn_wave_barrier();
        const int m = min(kWave, total - base);
        int j = 0;
        for (; j + 1 < m; j += 2) {
            const float4 s0 = L.src[j], s1 = L.src[j + 1];
            interact_p<POT>(a, px, py, pz, s0, eps2);
            interact_p<POT>(a, px, py, pz, s1, eps2);
        }
        if (j < m) interact_p<POT>(a, px, py, pz, L.src[j], eps2);
        __builtin_amdgcn_wave_barrier();
    }
    return total;
}

__device__ __forceinline__ float box_dist2(float lx, float ly, float lz, float hx, float hy,
                                           float hz, float cx, float cy, float cz)
{
#pragma clang fp contract(off)
    const float ex = fmaxf(fmaxf(lx - cx, cx - hx), 0.0f);
    const float ey = fmaxf(fmaxf(ly - cy, cy - hy), 0.0f);
    const float ez = fmaxf(fmaxf(lz - cz, cz - hz), 0.0f);
    return ex * ex + ey * ey + ez * ez;
}

template <bool POT, bool QUAD, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_traverse_group(
    const GroupInfo* __restrict__ groups, int ngroups, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, const float4* __restrict__ nodeA,
    const int4* __restrict__ nodeB, const float4* __restrict__ nodeQ, float4* __restrict__ acc,
    int* __restrict__ stack_mem, unsigned* qctr, unsigned long long* counters, unsigned* err,
    float G, float eps2, int leaf)
{
    __shared__ WaveLds lds[WPB];
    const int l = lane_id();
    const int w = threadIdx.x >> 6;
    WaveLds& L = lds[w];
    int* stk = stack_mem + (size_t)(blockIdx.x * WPB + w) * kStackCap;
    const int xcd = blockIdx.x % kNumXcd;
    const int per = (ngroups + kNumXcd - 1) / kNumXcd;
    unsigned long long n_node = 0, n_part = 0;
    int q = 0;
    const float self_inv = __builtin_amdgcn_rsqf(eps2);

    for (;;) {
        int g = -1;
        while (q < kNumXcd) {
            const int c = (xcd + q) % kNumXcd;
            const int lo = c * per, hi = min(ngroups, lo + per);
            unsigned k = 0;
            if (l == 0) k = atomicAdd(&qctr[c], 1u);
            k = (unsigned)uniform((int)k);
            if (lo + (int)k < hi) {
                g = lo + (int)k;
                break;
            }
            ++q;
        }
        if (g < 0) break;

        const GroupInfo gi = groups[g];
        const SegInfo sg = segs[gi.seg];
        const int p = gi.first + min(l, gi.count - 1);
        const float4 pp = posm[p];
        const float lx = wave_min(pp.x), ly = wave_min(pp.y), lz = wave_min(pp.z);
        const float hx = wave_max(pp.x), hy = wave_max(pp.y), hz = wave_max(pp.z);
        Acc a = {0.f, 0.f, 0.f, 0.f};
        int nb = 0, nr = 0, top = 0, nnodes = 0, nparts = 0;
        if (sg.n <= leaf) {
            if (l == 0) L.rng[0] = make_int2(sg.start, sg.n);
            nr = 1;
        } else {
            if (l == 0) stk[0] = sg.root;
            top = 1;
        }
        __builtin_amdgcn_wave_barrier();
        bool overflow = false;
        while (top > 0) {
            const int cnt = min(top, kWave);
            top -= cnt;
            const bool have = l < cnt;
            int v = 0;
            if (have) v = stk[top + l];
            float4 A = make_float4(0, 0, 0, 0);
            int4 B = make_int4(0, 0, 0, 0);
            bool accept = false;
            if (have) {
                A = nodeA[v];
                B = nodeB[v];
                accept = box_dist2(lx, ly, lz, hx, hy, hz, A.x, A.y, A.z) > __int_as_float(B.x);
            }
            const bool open = have && !accept;
            const uint64_t mA = ballot(accept);
            if (accept) {
                const int s = nb + mbcnt(mA);
                L.na[s] = A;
                if (QUAD) {
                    const float4 q0 = nodeQ[2 * v], q1 = nodeQ[2 * v + 1];
                    L.nq0[s] = q0;
                    L.nq1[s] = make_float2(q1.x, q1.y);
                }
            }
            nb += __popcll(mA);
            const int lcount = B.w - B.y + 1, rcount = B.z - B.w;
            const bool lr = open && lcount <= leaf, rr = open && rcount <= leaf;
            const bool lp = open && !lr, rp = open && !rr;
            const uint64_t m1 = ballot(lr), m2 = ballot(rr);
            if (lr) L.rng[nr + mbcnt(m1)] = make_int2(B.y, lcount);
            nr += __popcll(m1);
            if (rr) L.rng[nr + mbcnt(m2)] = make_int2(B.w + 1, rcount);
            nr += __popcll(m2);
            const uint64_t m3 = ballot(lp), m4 = ballot(rp);
            const int np = __popcll(m3) + __popcll(m4);
            if (top + np > kStackCap) {
                overflow = true;
                break;
            }
            if (lp) stk[top + mbcnt(m3)] = B.w + sg.int_base;
            top += __popcll(m3);
            if (rp) stk[top + mbcnt(m4)] = B.w + 1 + sg.int_base;
            top += __popcll(m4);
            __builtin_amdgcn_wave_barrier();
            if (nb >= kWave) {
                eval_nodes<POT, QUAD>(L, nb, a, pp.x, pp.y, pp.z, eps2);
                nnodes += nb;
                nb = 0;
            }
            if (nr >= kWave) {
                nparts += eval_ranges<POT>(L, nr, a, pp.x, pp.y, pp.z, eps2, posm);
                nr = 0;
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (overflow && l == 0) atomicOr(err, kErrStack);
        if (nb > 0) {
            eval_nodes<POT, QUAD>(L, nb, a, pp.x, pp.y, pp.z, eps2);
            nnodes += nb;
        }
        if (nr > 0) nparts += eval_ranges<POT>(L, nr, a, pp.x, pp.y, pp.z, eps2, posm);
        if (l < gi.count) {
            const float ph = POT ? -G * (a.ph - pp.w * self_inv) : 0.0f;
            acc[p] = make_float4(G * a.ax, G * a.ay, G * a.az, ph);
        }
        n_node += (unsigned long long)nnodes * (unsigned long long)gi.count;
        n_part += (unsigned long long)(nparts - 1) * (unsigned long long)gi.count;
        __builtin_amdgcn_wave_barrier();
    }
    if (l == 0) {
        atomicAdd(&counters[0], n_node);
        atomicAdd(&counters[1], n_part);
    }
}

// cuda_engine.cu:91-161 -- per-particle DFS from the instance root; left child
// pushed before right; leaf taken when d^2+eps^2 > 1.01*eps^2; internal node
// taken when width^2 < d^2+eps^2; force G*m*rsqrt(d2)^3 * d.  `limit` (at most
// STACK) is the stack depth; a traversal that would exceed it stops and sets
// the stack error bit.  Decisions are
// computed exactly as written (no contraction); the accumulation uses the
// multiply-adds nvcc forms.
template <int STACK>
__global__ __launch_bounds__(256) void k_traverse_reference(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, const float4* __restrict__ nodeA,
    const int4* __restrict__ nodeB, float4* __restrict__ acc, unsigned long long* counters,
    unsigned* err, float G, float eps2, int limit)
{
#pragma clang fp contract(off)
    const TileInfo t = tiles[blockIdx.x >> 2];
    const int off = (int)(blockIdx.x & 3) * kTileThreads + (int)threadIdx.x;
    unsigned long long count = 0;
    if (off < t.count) {
        const int i = t.start + off;
        const SegInfo sg = segs[t.seg];
        const float4 p = posm[i];
        float ax = 0.f, ay = 0.f, az = 0.f;
        int stack[STACK];
        int top = 0;
        stack[top++] = sg.root;
        bool overflow = false;
        while (top > 0) {
            const int v = stack[--top];
            const float4 A = nodeA[v];
            const int4 B = nodeB[v];
            const float dx = A.x - p.x, dy = A.y - p.y, dz = A.z - p.z;
            const float d2 = dx * dx + dy * dy + dz * dz + eps2;
            bool take;
            if (B.w < 0) {
                take = d2 > eps2 * 1.01f;
            } else {
                const float width = __int_as_float(B.x);
                take = width * width < d2 * 1.0f;
                if (!take) {
                    if (top + 2 > limit) {
                        overflow = true;
                        break;
                    }
                    const int split = B.w;
                    stack[top++] = (B.y == split) ? split + sg.leaf_base : split + sg.int_base;
                    stack[top++] = (B.z == split + 1) ? split + 1 + sg.leaf_base
                                                      : split + 1 + sg.int_base;
                }
            }
            if (take) {
                const float inv = __builtin_amdgcn_rsqf(d2);
                const float inv3 = inv * inv * inv;
                const float f = G * A.w * inv3;
                ax = __builtin_fmaf(dx, f, ax);
                ay = __builtin_fmaf(dy, f, ay);
                az = __builtin_fmaf(dz, f, az);
                ++count;
            }
        }
        if (overflow) atomicOr(err, kErrStack);
        acc[i] = make_float4(ax, ay, az, 0.0f);
    }
    count = wave_sum(count);
    if (lane_id() == 0 && count) atomicAdd(&counters[1], count);
}

}  // namespace gcs
