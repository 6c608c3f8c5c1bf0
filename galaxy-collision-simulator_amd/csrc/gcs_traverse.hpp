// gcs_traverse.hpp -- Barnes-Hut traversals.
//
// k_traverse_group: one wave per group -- a maximal subtree of at most 64
// particles (consecutive in Morton order, so spatially compact).  For each
// group the wave
//   phase 1  walks the instance's tree once for the whole group, 64 nodes per
//            round from an LDS stack (spilling its oldest half to global
//            memory when full).  A popped node is
//              accepted  when the squared distance from its COM to the group's
//                        bounding box exceeds its opening value, (width/theta)^2
//                        (gravity_aggregator.cpp:37-43's s^2 < theta^2 d^2 with d
//                        taken from the group's box) or, with the offset form,
//                        (width/theta + |COM - box centre|)^2: its index goes
//                        to the node list;
//              summed    otherwise, when it holds <= leaf particles: its
//                        particles go to the particle list;
//              opened    otherwise: single-particle children go to the
//                        particle list, larger children are pushed.
//            Both lists live in a per-wave slot of global memory (L2-resident;
//            a full list is evaluated early, so no list overflows);
//   phase 2  streams each list in chunks of 64 entries: every lane gathers one
//            source (node COM, mass and quadrupole, or particle position and
//            mass) into LDS while the next chunk's gathers are in flight.
//            The lanes form S sub-waves of W lanes, each lane holding up to 3
//            of the group's particles, and sub-wave h evaluates entries h,
//            h + S, ... for its particles (LDS broadcasts); S, W and the
//            particles per lane are chosen per group so that the fewest
//            lane-steps cover the list (a group of 40 uses 3 sub-waves of 20
//            lanes with 2 particles each, 94% of the lanes instead of 63%).
//            The sub-wave sums are added through LDS in sub-wave order.
// The interaction set is a function of the tree, the group and the
// parameters only, and every sum runs in a fixed order, so results do not
// depend on scheduling.  Persistent grid: waves take groups from eight queues, one
// per XCD, each holding a contiguous run of groups (neighbouring groups touch
// the same nodes, which then stay in that XCD's L2), and steal from the other
// queues when theirs is empty.  The particle itself is in its own particle
// list: zero separation gives zero force, and its potential term is removed at
// the end, which is index self-exclusion.
//
// k_traverse_reference: the reference's per-particle traversal
// (cuda_engine.cu:91-161) with the same decisions, for the cuda_engine.h entry
// points.
#pragma once
#include "gcs_device.hpp"

namespace gcs {

constexpr int kLdsStack = 1024;    // LDS stack entries per wave
constexpr int kNodeList = 4096;    // node-list entries per wave slot
constexpr int kPartList = 16384;   // particle-list entries per wave slot
constexpr int kLocList = 4096;     // local-expansion list entries per wave slot
constexpr int kMaxLeaf = 64;       // largest direct-sum bucket
constexpr int kSlotInts = kStackCap + kNodeList + kPartList + kLocList;
constexpr int kLocTerms = 20;      // a0 (3), A (6), B (10), phi0
constexpr int kCounts = 5;         // per-wave interaction counts

constexpr int kMaxSplit = 16;               // most sub-waves a list is split into
constexpr int kMaxPer = 3;                  // most targets per lane
constexpr int kChunk = kWave + kMaxSplit;   // a chunk of 64 sources and a zero pad

struct WaveLds {
    int stk[kLdsStack + 1];  // + a spare entry that takes discarded pushes
    float4 ca[kChunk];   // COM and mass, or position and mass
    float4 cq0[kChunk];  // Qxx, Qyy, Qzz, Qxy
    float4 cq1[kChunk];  // Qxz, Qyz, (width), T
    float loc[kLocTerms + 4];  // the group's local-expansion terms (this wave's share)
};

struct Acc {
    float ax, ay, az, ph;
};

// A wave's share of a group: its lanes form S sub-waves of W lanes; lane
// (h, t) holds targets t, t + W, ..., t + (P-1) W and takes the list entries
// h, h + S, h + 2S, ...  Lanes beyond S*W idle (they repeat sub-wave 0's
// entries and are not summed).
struct Share {
    int S, W, P, h, t;
};

// Fewest lane-steps per list entry, P/S, for a group of n targets
// (W = ceil(n/P), S = min(kMaxSplit, 64/W)), ties to the smaller P.
__device__ __forceinline__ Share make_share(int n, int lane)
{
    Share sh;
    sh.P = (n + kWave - 1) / kWave;  // at least ceil(n/64) per lane
    sh.W = (n + sh.P - 1) / sh.P;
    sh.S = min(kMaxSplit, kWave / sh.W);
    for (int p = sh.P + 1; p <= kMaxPer; ++p) {
        const int w = (n + p - 1) / p;
        const int s = min(kMaxSplit, kWave / w);
        if (p * sh.S < sh.P * s) {
            sh.P = p;
            sh.W = w;
            sh.S = s;
        }
    }
    const int h = lane / sh.W;
    sh.h = h < sh.S ? h : 0;
    sh.t = lane - h * sh.W;
    return sh;
}

struct Targets {
    float x[kMaxPer], y[kMaxPer], z[kMaxPer];
};

static_assert(offsetof(WaveLds, cq0) == offsetof(WaveLds, ca) + sizeof(float4) * kChunk &&
                  offsetof(WaveLds, cq1) == offsetof(WaveLds, cq0) + sizeof(float4) * kChunk,
              "the combine scratch spans the three chunk arrays");
// S * n <= S * W * P <= 64 * kMaxPer partial sums fit the three 64-entry arrays,
// and a group of kGroupMax fits the wave at kMaxPer particles per lane.
static_assert(kMaxPer <= 3, "the combine scratch holds 3 x 64 partial sums");
static_assert(kGroupMax <= kWave * kMaxPer, "a group fits one wave");

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

// Node source with its traceless quadrupole Q and second-moment trace T,
// expanded to second order in the softened kernel 1/R, R^2 = |d|^2 + eps^2
// (the kernel's Laplacian -3 eps^2/R^5 brings in T):
//   a += (M/R^3 + 5/2 (d.Qd - T eps^2)/R^7) d - Qd/R^5
//      = ((M R^2 + 5/2 (d.Qd - T eps^2)/R^2) d - Qd) / R^5,
//   phi += M/R + (d.Qd - T eps^2)/(2 R^5).
// q1.w holds T eps^2 (the list gather forms it once per node).
template <bool POT, bool QUAD>
__device__ __forceinline__ void interact_n(Acc& a, float px, float py, float pz, float4 c,
                                           float4 q0, float4 q1, float eps2)
{
    const float dx = c.x - px, dy = c.y - py, dz = c.z - pz;
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float inv2 = inv * inv;
    if (!QUAD) {
        const float minv = c.w * inv;
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
    const float dqd = __builtin_fmaf(dx, qx, __builtin_fmaf(dy, qy, __builtin_fmaf(dz, qz, -q1.w)));
    const float inv5 = (inv2 * inv2) * inv;
    const float g = __builtin_fmaf(2.5f * dqd, inv2, c.w * r2);
    a.ax = __builtin_fmaf(inv5, __builtin_fmaf(g, dx, -qx), a.ax);
    a.ay = __builtin_fmaf(inv5, __builtin_fmaf(g, dy, -qy), a.ay);
    a.az = __builtin_fmaf(inv5, __builtin_fmaf(g, dz, -qz), a.az);
    if (POT) a.ph += __builtin_fmaf(0.5f, dqd * inv5, c.w * inv);
}

// Phase 2 over the 64-entry chunks of the node list nl[0..k) that start at
// first, first + step, ... (step a multiple of 64): gathers of one chunk are
// issued before the previous one is evaluated, and the indices one chunk
// further ahead.  Sub-wave h takes chunk entries h, h + S, h + 2S, ...;
// entries past the list end were gathered as zero-mass sources and add
// nothing, and the pad after the 64 entries holds zero-mass sources too.
template <bool POT, bool QUAD, int P>
__device__ __forceinline__ void eval_node_list(WaveLds& L, const int* nl, int k,
                                               const float4* __restrict__ node,
                                               const float4* __restrict__ nodeQ, Acc* a,
                                               const Targets& tg, float eps2, int S, int h,
                                               int first, int step)
{
    const int l = lane_id();
    float4 A = make_float4(0.f, 0.f, 0.f, 0.f), Q0 = A, Q1 = A;
    const int idx = first + l < k ? nl[first + l] : -1;
    int idx_next = first + step + l < k ? nl[first + step + l] : -1;
    if (idx >= 0) {
        A = node[2 * idx];
        if (QUAD) { Q0 = nodeQ[2 * idx]; Q1 = nodeQ[2 * idx + 1]; }
    }
    for (int base = first; base < k; base += step) {
        L.ca[l] = A;
        if (QUAD) {
            L.cq0[l] = Q0;
            L.cq1[l] = make_float4(Q1.x, Q1.y, Q1.z, Q1.w * eps2);  // T eps^2
        }
        __builtin_amdgcn_wave_barrier();
        A = make_float4(0.f, 0.f, 0.f, 0.f);
        Q0 = A;
        Q1 = A;
        if (idx_next >= 0) {
            A = node[2 * idx_next];
            if (QUAD) { Q0 = nodeQ[2 * idx_next]; Q1 = nodeQ[2 * idx_next + 1]; }
        }
        const int j2 = base + 2 * step + l;
        idx_next = j2 < k ? nl[j2] : -1;
        const int trips = uniform((min(kWave, k - base) + S - 1) / S);
        int t = 0;
        // Two entries per iteration with one target per lane; with more
        // targets the targets themselves give the independent work.
        if (P == 1) {
            for (; t + 1 < trips; t += 2) {
                const int e0 = t * S + h, e1 = e0 + S;
                const float4 c0 = L.ca[e0], c1 = L.ca[e1];
                float4 q00 = make_float4(0, 0, 0, 0), q01 = q00, q10 = q00, q11 = q00;
                if (QUAD) {
                    q00 = L.cq0[e0]; q01 = L.cq0[e1];
                    q10 = L.cq1[e0]; q11 = L.cq1[e1];
                }
                interact_n<POT, QUAD>(a[0], tg.x[0], tg.y[0], tg.z[0], c0, q00, q10, eps2);
                interact_n<POT, QUAD>(a[0], tg.x[0], tg.y[0], tg.z[0], c1, q01, q11, eps2);
            }
        }
        for (; t < trips; ++t) {
            const int e0 = t * S + h;
            const float4 c0 = L.ca[e0];
            float4 q0 = make_float4(0, 0, 0, 0), q1 = q0;
            if (QUAD) { q0 = L.cq0[e0]; q1 = L.cq1[e0]; }
#pragma unroll
            for (int u = 0; u < P; ++u)
                interact_n<POT, QUAD>(a[u], tg.x[u], tg.y[u], tg.z[u], c0, q0, q1, eps2);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// A particle-list entry: a particle index p >= 0 (its position and mass), or
// ~v < 0 for a node v taken as a monopole (its COM and mass, the same float4
// layout); kNoEntry past the list end (a zero-mass source).
constexpr int kNoEntry = (int)0x80000000;

__device__ __forceinline__ float4 point_source(int e, const float4* __restrict__ posm,
                                               const float4* __restrict__ node)
{
    if (e >= 0) return posm[e];
    if (e != kNoEntry) return node[2 * (~e)];
    return make_float4(0.f, 0.f, 0.f, 0.f);
}

// Phase 2 over the particle list pl[0..k), same chunks, pipeline and split.
// Entries are particles or far nodes taken as monopoles (point_source).
template <bool POT, int P>
__device__ __forceinline__ void eval_part_list(WaveLds& L, const int* pl, int k,
                                               const float4* __restrict__ posm,
                                               const float4* __restrict__ node, Acc* a,
                                               const Targets& tg, float eps2, int S, int h,
                                               int first, int step)
{
    const int l = lane_id();
    const int idx = first + l < k ? pl[first + l] : kNoEntry;
    int idx_next = first + step + l < k ? pl[first + step + l] : kNoEntry;
    float4 X = point_source(idx, posm, node);
    for (int base = first; base < k; base += step) {
        L.ca[l] = X;
        __builtin_amdgcn_wave_barrier();
        X = point_source(idx_next, posm, node);
        const int j2 = base + 2 * step + l;
        idx_next = j2 < k ? pl[j2] : kNoEntry;
        const int trips = uniform((min(kWave, k - base) + S - 1) / S);
        int t = 0;
        // About four interactions in flight per iteration.
        constexpr int U = P == 1 ? 4 : 2;
        for (; t + U - 1 < trips; t += U) {
            const int e0 = t * S + h;
            float4 sv[U];
#pragma unroll
            for (int v = 0; v < U; ++v) sv[v] = L.ca[e0 + v * S];
#pragma unroll
            for (int u = 0; u < P; ++u)
#pragma unroll
                for (int v = 0; v < U; ++v)
                    interact_p<POT>(a[u], tg.x[u], tg.y[u], tg.z[u], sv[v], eps2);
        }
        for (; t < trips; ++t) {
            const float4 s0 = L.ca[t * S + h];
#pragma unroll
            for (int u = 0; u < P; ++u) interact_p<POT>(a[u], tg.x[u], tg.y[u], tg.z[u], s0, eps2);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The evaluation of both lists for the share's P (uniform per group).
// first/step: the chunks this wave takes (all of them: 0 and 64).
template <bool POT, bool QUAD>
__device__ __forceinline__ void eval_nodes(WaveLds& L, const int* nl, int k, const float4* node,
                                           const float4* nodeQ, Acc* a, const Targets& tg,
                                           float eps2, const Share& sh, int first = 0,
                                           int step = kWave)
{
    if (sh.P == 1)
        eval_node_list<POT, QUAD, 1>(L, nl, k, node, nodeQ, a, tg, eps2, sh.S, sh.h, first, step);
    else if (sh.P == 2)
        eval_node_list<POT, QUAD, 2>(L, nl, k, node, nodeQ, a, tg, eps2, sh.S, sh.h, first, step);
    else
        eval_node_list<POT, QUAD, 3>(L, nl, k, node, nodeQ, a, tg, eps2, sh.S, sh.h, first, step);
}

template <bool POT>
__device__ __forceinline__ void eval_parts(WaveLds& L, const int* pl, int k, const float4* posm,
                                           const float4* node, Acc* a, const Targets& tg,
                                           float eps2, const Share& sh, int first = 0,
                                           int step = kWave)
{
    if (sh.P == 1)
        eval_part_list<POT, 1>(L, pl, k, posm, node, a, tg, eps2, sh.S, sh.h, first, step);
    else if (sh.P == 2)
        eval_part_list<POT, 2>(L, pl, k, posm, node, a, tg, eps2, sh.S, sh.h, first, step);
    else
        eval_part_list<POT, 3>(L, pl, k, posm, node, a, tg, eps2, sh.S, sh.h, first, step);
}

// ---- Far field through local expansions -------------------------------------
// An accepted node whose COM lies farther than kappa group radii from the
// centre c of the group's box is not evaluated per target: its field enters a
// second-order expansion at c -- the softened quadrupole field a0 at c and the
// softened monopole's first and second derivatives A and B -- and each target
// adds a0 + A y + B:yy/2 (y its offset from c) once, with the potential
// phi0 + a0.y + y.A.y/2 (the accumulated sign convention: phi = -G x this).
// Terms: a0 x y z, A xx yy zz xy xz yz, B xxx yyy zzz xxy xxz xyy yyz xzz yzz
// xyz, phi0.

// The group's box centre and local-expansion radius squared, rounded as the
// oracle forms them (no contraction).
__device__ __forceinline__ void local_frame(float lx, float ly, float lz, float hx, float hy,
                                            float hz, float kappa2, float& cx, float& cy,
                                            float& cz, float& lim)
{
#pragma clang fp contract(off)
    cx = (lx + hx) * 0.5f;
    cy = (ly + hy) * 0.5f;
    cz = (lz + hz) * 0.5f;
    const float ex = hx - lx, ey = hy - ly, ez = hz - lz;
    const float rg2 = 0.25f * (ex * ex + ey * ey + ez * ez);
    lim = kappa2 * rg2;
}

__device__ __forceinline__ float dist2_nc(float dx, float dy, float dz)
{
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// Adds a node's terms (r = COM - c, mass m, quadrupole q0/q1) to L.
template <bool POT>
__device__ __forceinline__ void local_terms(float rx, float ry, float rz, float m, float4 q0,
                                            float4 q1, float eps2, float* L)
{
    const float r2 = __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, eps2)));
    const float iR = __builtin_amdgcn_rsqf(r2);
    const float iR2 = iR * iR, iR3 = iR * iR2, iR5 = iR3 * iR2, iR7 = iR5 * iR2;
    const float qx = __builtin_fmaf(q0.x, rx, __builtin_fmaf(q0.w, ry, q1.x * rz));
    const float qy = __builtin_fmaf(q0.w, rx, __builtin_fmaf(q0.y, ry, q1.y * rz));
    const float qz = __builtin_fmaf(q1.x, rx, __builtin_fmaf(q1.y, ry, q0.z * rz));
    const float dqd =
        __builtin_fmaf(-q1.w, eps2, __builtin_fmaf(rx, qx, __builtin_fmaf(ry, qy, rz * qz)));
    const float t = dqd * iR5;
    const float f = __builtin_fmaf(2.5f * t, iR2, m * iR3);
    L[0] += __builtin_fmaf(rx, f, -qx * iR5);
    L[1] += __builtin_fmaf(ry, f, -qy * iR5);
    L[2] += __builtin_fmaf(rz, f, -qz * iR5);
    const float m3 = m * iR3, m5 = 3.0f * m * iR5, m7 = 15.0f * m * iR7;
    const float xx = rx * rx, yy = ry * ry, zz = rz * rz;
    L[3] += __builtin_fmaf(m5, xx, -m3);
    L[4] += __builtin_fmaf(m5, yy, -m3);
    L[5] += __builtin_fmaf(m5, zz, -m3);
    L[6] += m5 * rx * ry;
    L[7] += m5 * rx * rz;
    L[8] += m5 * ry * rz;
    const float m7x = m7 * rx, m7y = m7 * ry, m7z = m7 * rz;
    L[9] += rx * __builtin_fmaf(m7, xx, -3.0f * m5);
    L[10] += ry * __builtin_fmaf(m7, yy, -3.0f * m5);
    L[11] += rz * __builtin_fmaf(m7, zz, -3.0f * m5);
    L[12] += ry * __builtin_fmaf(m7x, rx, -m5);
    L[13] += rz * __builtin_fmaf(m7x, rx, -m5);
    L[14] += rx * __builtin_fmaf(m7y, ry, -m5);
    L[15] += rz * __builtin_fmaf(m7y, ry, -m5);
    L[16] += rx * __builtin_fmaf(m7z, rz, -m5);
    L[17] += ry * __builtin_fmaf(m7z, rz, -m5);
    L[18] += m7x * ry * rz;
    if (POT) L[19] += __builtin_fmaf(0.5f, t, m * iR);
}

// The local list ll[0..k): the chunks of 64 entries that start at first,
// first + step, ... (one node per lane), summed over the wave (a butterfly
// in a fixed pairing, identical on every lane) and added to tot (LDS).
template <bool POT>
__device__ __forceinline__ void eval_local_list(float* tot, const int* ll, int k,
                                                const float4* __restrict__ node,
                                                const float4* __restrict__ nodeQ, float cx,
                                                float cy, float cz, float eps2, int first,
                                                int step)
{
    const int l = lane_id();
    float c[kLocTerms];
#pragma unroll
    for (int t = 0; t < kLocTerms; ++t) c[t] = 0.0f;
    for (int base = first; base < k; base += step) {
        const int j = base + l;
        if (j < k) {
            const int v = ll[j];
            const float4 A = node[2 * v], q0 = nodeQ[2 * v], q1 = nodeQ[2 * v + 1];
            local_terms<POT>(A.x - cx, A.y - cy, A.z - cz, A.w, q0, q1, eps2, c);
        }
    }
#pragma unroll
    for (int t = 0; t < kLocTerms; ++t) c[t] = wave_sum(c[t]);
    if (l == 0) {
#pragma unroll
        for (int t = 0; t < kLocTerms; ++t) tot[t] += c[t];
    }
    __builtin_amdgcn_wave_barrier();
}

// The expansion L at offset y from the centre, added to r (x, y, z, phi).
template <bool POT>
__device__ __forceinline__ void local_apply(const float* L, float yx, float yy, float yz,
                                            float4& r)
{
    const float xx = yx * yx, y2 = yy * yy, zz = yz * yz, xy = yx * yy, xz = yx * yz,
                yz2 = yy * yz;
    const float ax = __builtin_fmaf(L[3], yx, __builtin_fmaf(L[6], yy, L[7] * yz));
    const float ay = __builtin_fmaf(L[6], yx, __builtin_fmaf(L[4], yy, L[8] * yz));
    const float az = __builtin_fmaf(L[7], yx, __builtin_fmaf(L[8], yy, L[5] * yz));
    const float bx = __builtin_fmaf(L[9], xx, __builtin_fmaf(L[14], y2, L[16] * zz)) +
                     2.0f * __builtin_fmaf(L[12], xy, __builtin_fmaf(L[13], xz, L[18] * yz2));
    const float by = __builtin_fmaf(L[12], xx, __builtin_fmaf(L[10], y2, L[17] * zz)) +
                     2.0f * __builtin_fmaf(L[14], xy, __builtin_fmaf(L[18], xz, L[15] * yz2));
    const float bz = __builtin_fmaf(L[13], xx, __builtin_fmaf(L[15], y2, L[11] * zz)) +
                     2.0f * __builtin_fmaf(L[18], xy, __builtin_fmaf(L[16], xz, L[17] * yz2));
    r.x += __builtin_fmaf(0.5f, bx, L[0] + ax);
    r.y += __builtin_fmaf(0.5f, by, L[1] + ay);
    r.z += __builtin_fmaf(0.5f, bz, L[2] + az);
    if (POT) {
        const float a0y = __builtin_fmaf(L[0], yx, __builtin_fmaf(L[1], yy, L[2] * yz));
        const float yAy = __builtin_fmaf(ax, yx, __builtin_fmaf(ay, yy, az * yz));
        r.w += L[19] + __builtin_fmaf(0.5f, yAy, a0y);
    }
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

template <bool POT, bool QUAD, int WPB, bool LOCAL>
__global__ __launch_bounds__(WPB * 64) void k_traverse_group(
    const GroupInfo* __restrict__ groups, const int* __restrict__ gorder,
    const int* __restrict__ ngroups_dev, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, const float4* __restrict__ node,
    const float4* __restrict__ nodeQ, float4* __restrict__ acc, int* __restrict__ slot_mem,
    unsigned* qctr, unsigned long long* counters, unsigned* err, float G, float eps2, int leaf,
    float far2, float kappa2)
{
    __shared__ WaveLds lds[WPB];
    const int l = lane_id();
    const int w = uniform(threadIdx.x >> 6);  // wave-uniform: the slot addresses stay scalar
    WaveLds& L = lds[w];
    int* const spill = slot_mem + (size_t)(blockIdx.x * WPB + w) * kSlotInts;
    int* const nl = spill + kStackCap;
    int* const pl = nl + kNodeList;
    int* const ll = pl + kPartList;
    const int xcd = blockIdx.x % kNumXcd;
    const int ngroups = uniform(*ngroups_dev);
    const int per = (ngroups + kNumXcd - 1) / kNumXcd;
    unsigned long long n_node = 0, n_part = 0, n_far = 0, n_loc = 0, n_loct = 0;
    int q = 0;
    const float self_inv = __builtin_amdgcn_rsqf(eps2);
    bool overflow = false;
    // The pads after each chunk hold zero-mass sources for the whole launch.
    if (l < kMaxSplit) {
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        L.ca[kWave + l] = z;
        L.cq0[kWave + l] = z;
        L.cq1[kWave + l] = z;
    }
    __builtin_amdgcn_wave_barrier();

    for (;;) {
        int g = -1;
        while (q < kNumXcd) {
            const int c = (xcd + q) % kNumXcd;
            const int lo = c * per, hi = min(ngroups, lo + per);
            unsigned k = 0;
            if (l == 0) k = atomicAdd(&qctr[c], 1u);
            k = (unsigned)uniform((int)k);
            if (lo + (int)k < hi) {
                g = gorder ? uniform(gorder[lo + (int)k]) : lo + (int)k;
                break;
            }
            ++q;
        }
        if (g < 0) break;

        const GroupInfo gi = groups[g];
        const SegInfo sg = segs[gi.seg];
        // The group's bounding box, then this lane's share of the targets.
        float lx = __builtin_inff(), ly = lx, lz = lx, hx = -lx, hy = -lx, hz = -lx;
        for (int j = l; j < gi.count; j += kWave) {
            const float4 pj = posm[gi.first + j];
            lx = fminf(lx, pj.x); ly = fminf(ly, pj.y); lz = fminf(lz, pj.z);
            hx = fmaxf(hx, pj.x); hy = fmaxf(hy, pj.y); hz = fmaxf(hz, pj.z);
        }
        lx = wave_min(lx); ly = wave_min(ly); lz = wave_min(lz);
        hx = wave_max(hx); hy = wave_max(hy); hz = wave_max(hz);
        const Share sh = make_share(gi.count, l);
        Targets tg;
        Acc a[kMaxPer];
#pragma unroll
        for (int u = 0; u < kMaxPer; ++u) {
            const float4 pu = posm[gi.first + min(sh.t + u * sh.W, gi.count - 1)];
            tg.x[u] = pu.x;
            tg.y[u] = pu.y;
            tg.z[u] = pu.z;
            a[u] = Acc{0.f, 0.f, 0.f, 0.f};
        }
        float gcx = 0.f, gcy = 0.f, gcz = 0.f, glim = 0.f;
        if (LOCAL) {
            local_frame(lx, ly, lz, hx, hy, hz, kappa2, gcx, gcy, gcz, glim);
            if (l < kLocTerms) L.loc[l] = 0.0f;
        }
        int nn = 0, np = 0, nloc = 0, top = 0, spilled = 0;
        long long tot_nodes = 0, tot_parts = 0, tot_far = 0, tot_loc = 0;
        // The lists are written through one buffer resource over the wave's
        // slot: a lane with nothing to write stores at kDrop.  List stores of
        // one round are issued in the next round after its node loads, so that
        // the wait for those loads does not also wait for them (vector memory
        // operations complete in issue order): the node- or local-list byte
        // offset of the lane's node and its index, and its particle run
        // [dp_f, dp_f + dp_c) at byte offset dp_off.
        const __amdgpu_buffer_rsrc_t rs = word_rsrc(spill, kSlotInts);
        constexpr int kNlOff = 4 * kStackCap, kPlOff = kNlOff + 4 * kNodeList,
                      kLlOff = kPlOff + 4 * kPartList;
        int dn_off = kDrop, dl_off = kDrop, dv = 0, dp_off = 0, dp_f = 0, dp_c = 0;
        int dp_max = 0;  // the longest run of the round (uniform)
        if (sg.n == 1) {
            if (l == 0) pl[0] = sg.start;
            np = 1;
        } else {
            if (l == 0) L.stk[0] = sg.root;
            top = 1;
        }
        __builtin_amdgcn_wave_barrier();
        while (top > 0 || spilled > 0) {
            if (top == 0) {
                // Reload the most recently spilled half.
                const int c = min(spilled, kLdsStack / 2);
                for (int k = l; k < c; k += kWave) L.stk[k] = spill[spilled - c + k];
                spilled = uniform(spilled - c);
                top = uniform(c);
                __builtin_amdgcn_wave_barrier();
            }
            // Lanes past the popped entries repeat the last one and decide
            // nothing.
            const int cnt = min(top, kWave);
            top -= cnt;
            const bool have = l < cnt;
            const int v = L.stk[top + min(l, cnt - 1)];
            const float4 A = node[2 * v];
            const float4 bb = node[2 * v + 1];
            // The previous round's list entries.
            store_word(rs, dn_off, dv);
            if (LOCAL) store_word(rs, dl_off, dv);
            for (int k = 0; k < dp_max; ++k)
                store_word(rs, k < dp_c ? dp_off + 4 * k : kDrop, dp_f + k);
            const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
                      split = __float_as_int(bb.w);
            // (bitwise logic: every lane evaluates every test, no branches)
            const float d2b = box_dist2(lx, ly, lz, hx, hy, hz, A.x, A.y, A.z);
            const float d2c = dist2_nc(A.x - gcx, A.y - gcy, A.z - gcz);
            const bool accept = have & (d2b > bb.x);
            const bool loc = LOCAL & accept & (d2c > glim);
            const bool far = accept & !loc & (d2b > far2 * bb.x);
            const bool near = accept & !far & !loc;
            const int size = last - first + 1;
            const bool direct = have & !accept & (size <= leaf);
            const bool open = have & !accept & !direct;
            const bool lleaf = open & (first == split), rleaf = open & (last == split + 1);
            // accepted nodes far from the group -> local list; beyond far x the
            // opening radius -> particle list (monopole); the others -> node
            // list (quadrupole)
            if (LOCAL) {
                const uint64_t mL = ballot(loc);
                const int nlo = __popcll(mL);
                if (nloc + nlo > kLocList) {
                    eval_local_list<POT>(L.loc, ll, nloc, node, nodeQ, gcx, gcy, gcz, eps2, 0,
                                         kWave);
                    nloc = 0;
                }
                dl_off = loc ? kLlOff + 4 * (nloc + mbcnt(mL)) : kDrop;
                nloc = uniform(nloc + nlo);
                tot_loc += nlo;
            }
            const uint64_t mA = ballot(near);
            const int na = __popcll(mA);
            if (nn + na > kNodeList) {
                eval_nodes<POT, QUAD>(L, nl, nn, node, nodeQ, a, tg, eps2, sh);
                tot_nodes += nn;
                nn = 0;
            }
            dn_off = near ? kNlOff + 4 * (nn + mbcnt(mA)) : kDrop;
            dv = v;
            nn = uniform(nn + na);
            tot_far += __popcll(ballot(far));
            // particles: the run [pf, pf + pc): the node's own (summed), its
            // single-particle children, or the node itself as a monopole (~v)
            int pc = far ? 1 : (int)lleaf + (int)rleaf;
            pc = direct ? size : pc;
            int pf = lleaf ? split : split + 1;
            pf = far ? ~v : pf;
            pf = direct ? first : pf;
            const int incl = wave_scan_incl_dpp(pc);
            const int tp = __builtin_amdgcn_readlane(incl, kWave - 1);
            if (np + tp > kPartList) {
                eval_parts<POT>(L, pl, np, posm, node, a, tg, eps2, sh);
                tot_parts += np;
                np = 0;
            }
            dp_off = kPlOff + 4 * (np + incl - pc);
            dp_f = pf;
            dp_c = pc;
            dp_max = 0;
            while (ballot(dp_max < pc)) ++dp_max;
            np = uniform(np + tp);
            // larger children -> stack (lanes with nothing to push write the
            // spare entry past the stack)
            const bool lp = open & !lleaf, rp = open & !rleaf;
            const uint64_t m3 = ballot(lp), m4 = ballot(rp);
            const int c3 = __popcll(m3), npush = c3 + __popcll(m4);
            if (top + npush > kLdsStack) {
                // Spill the oldest half of the LDS stack to global memory.
                const int h = kLdsStack / 2;
                if (spilled + h > kStackCap) {
                    overflow = true;
                    break;
                }
                for (int k = l; k < h; k += kWave) spill[spilled + k] = L.stk[k];
                __builtin_amdgcn_wave_barrier();
                for (int k = l; k < top - h; k += kWave) {
                    const int t = L.stk[k + h];
                    __builtin_amdgcn_wave_barrier();
                    L.stk[k] = t;
                }
                spilled = uniform(spilled + h);
                top -= h;
            }
            L.stk[lp ? top + mbcnt(m3) : kLdsStack] = split + sg.int_base;
            L.stk[rp ? top + c3 + mbcnt(m4) : kLdsStack] = split + 1 + sg.int_base;
            top = uniform(top + npush);
            __builtin_amdgcn_wave_barrier();
        }
        if (overflow) {
            if (l == 0) atomicOr(err, kErrStack);
            break;
        }
        store_word(rs, dn_off, dv);
        if (LOCAL) store_word(rs, dl_off, dv);
        for (int k = 0; k < dp_max; ++k) store_word(rs, k < dp_c ? dp_off + 4 * k : kDrop, dp_f + k);
        if (LOCAL && nloc > 0)
            eval_local_list<POT>(L.loc, ll, nloc, node, nodeQ, gcx, gcy, gcz, eps2, 0, kWave);
        if (nn > 0) eval_nodes<POT, QUAD>(L, nl, nn, node, nodeQ, a, tg, eps2, sh);
        if (np > 0) eval_parts<POT>(L, pl, np, posm, node, a, tg, eps2, sh);
        tot_nodes += nn;
        tot_parts += np;
        // Combine the sub-wave partial sums through LDS (the chunk arrays, 3 x 64
        // entries), in sub-wave order.
        float4* const scr = L.ca;
        if (l < sh.S * sh.W) {
#pragma unroll
            for (int u = 0; u < kMaxPer; ++u) {
                const int j = sh.t + u * sh.W;
                if (u < sh.P && j < gi.count) {
                    const int i = sh.h * gi.count + j;
                    scr[(i >> 6) * kChunk + (i & 63)] =
                        make_float4(a[u].ax, a[u].ay, a[u].az, a[u].ph);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (int j = l; j < gi.count; j += kWave) {
            float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int hh = 0; hh < sh.S; ++hh) {
                const int i = hh * gi.count + j;
                const float4 v = scr[(i >> 6) * kChunk + (i & 63)];
                r.x += v.x;
                r.y += v.y;
                r.z += v.z;
                r.w += v.w;
            }
            const float4 pj = posm[gi.first + j];
            if (LOCAL && tot_loc > 0) local_apply<POT>(L.loc, pj.x - gcx, pj.y - gcy, pj.z - gcz, r);
            const float ph = POT ? -G * (r.w - pj.w * self_inv) : 0.0f;
            acc[gi.first + j] = make_float4(G * r.x, G * r.y, G * r.z, ph);
        }
        n_node += (unsigned long long)(tot_nodes + tot_far) * (unsigned long long)gi.count;
        n_far += (unsigned long long)tot_far * (unsigned long long)gi.count;
        n_part += (unsigned long long)(tot_parts - tot_far - 1) * (unsigned long long)gi.count;
        n_loc += (unsigned long long)tot_loc;
        n_loct += (unsigned long long)tot_loc * (unsigned long long)gi.count;
        __builtin_amdgcn_wave_barrier();
    }
    // This wave's counts in its own slot (plain stores: no contention on one
    // address at the end of the launch); the host adds the launch's slots.
    if (l == 0) {
        unsigned long long* c = counters + kCounts * ((size_t)blockIdx.x * WPB + w);
        c[0] = n_node;
        c[1] = n_part;
        c[2] = n_far;
        c[3] = n_loc;
        c[4] = n_loct;
    }
}

// Sub-wave partial sums of a wave's targets, added through LDS (the chunk
// arrays, 3 x 64 entries) in sub-wave order: r[u] is the sum for target
// l + 64 u (targets < n).
__device__ __forceinline__ void combine_share(WaveLds& L, const Acc* a, const Share& sh, int n,
                                              float4* r)
{
    const int l = lane_id();
    float4* const scr = L.ca;
    __builtin_amdgcn_wave_barrier();
    if (l < sh.S * sh.W) {
#pragma unroll
        for (int u = 0; u < kMaxPer; ++u) {
            const int j = sh.t + u * sh.W;
            if (u < sh.P && j < n) {
                const int i = sh.h * n + j;
                scr[(i >> 6) * kChunk + (i & 63)] = make_float4(a[u].ax, a[u].ay, a[u].az, a[u].ph);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < kMaxPer; ++u) {
        const int j = l + u * kWave;
        float4 v4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (j < n) {
            for (int hh = 0; hh < sh.S; ++hh) {
                const int i = hh * n + j;
                const float4 v = scr[(i >> 6) * kChunk + (i & 63)];
                v4.x += v.x;
                v4.y += v.y;
                v4.z += v.z;
                v4.w += v.w;
            }
        }
        r[u] = v4;
    }
    __builtin_amdgcn_wave_barrier();
}

// k_traverse_team: the group traversal with a team of TEAM waves (one block)
// per group.  The team walks the tree together from one LDS stack, TEAM x 64
// nodes per round; each round's list entries and pushes are placed in wave
// order (wave 0's first), so the lists are a function of the tree and the
// group only.  Each wave then evaluates the 64-entry chunks w, w + TEAM, ...
// of the lists for all of the group's targets, and the TEAM partial sums of a
// target are added through LDS in wave order.  A group's wall time is about
// 1/TEAM of the one-wave kernel's, which evens out the work when there are
// few groups per wave.  Lists that would overflow are evaluated early into
// partial sums (kept unscaled in acc until the final evaluation).

template <bool POT, bool QUAD, int TEAM, bool LOCAL>
__global__ __launch_bounds__(TEAM * 64) void k_traverse_team(
    const GroupInfo* __restrict__ groups, const int* __restrict__ gorder,
    const int* __restrict__ ngroups_dev, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, const float4* __restrict__ node,
    const float4* __restrict__ nodeQ, float4* __restrict__ acc, int* __restrict__ slot_mem,
    unsigned* qctr, unsigned long long* counters, unsigned* err, float G, float eps2, int leaf,
    float far2, float kappa2)
{
    __shared__ WaveLds lds[TEAM];  // the team's stack is lds[0].stk
    // per wave and round: accepted (quadrupole), particle-list entries, left
    // pushes, right pushes (one 16-byte row per wave); far nodes (monopole
    // entries) and local-list entries apart, so the rows stay 16-byte reads
    __shared__ __align__(16) int xch[TEAM][4];
    __shared__ int xfar[TEAM];
    __shared__ int xloc[TEAM];
    __shared__ int ctl[4];          // claimed group; stack top; spilled entries; overflow
    const int l = lane_id();
    const int w = threadIdx.x >> 6;
    const int tid = threadIdx.x;
    WaveLds& L = lds[w];
    int* const stk = lds[0].stk;
    // The block owns TEAM wave slots: one stack spill area and TEAM times each
    // list, written through one buffer resource (kDrop: nothing to write).
    int* const spill = slot_mem + (size_t)blockIdx.x * TEAM * kSlotInts;
    int* const nl = spill + kStackCap;
    int* const pl = nl + TEAM * kNodeList;
    int* const ll = pl + TEAM * kPartList;
    const __amdgpu_buffer_rsrc_t rs = word_rsrc(spill, TEAM * kSlotInts);
    constexpr int kNlOff = 4 * kStackCap, kPlOff = kNlOff + 4 * TEAM * kNodeList,
                  kLlOff = kPlOff + 4 * TEAM * kPartList;
    const int xcd = blockIdx.x % kNumXcd;
    const int ngroups = uniform(*ngroups_dev);
    const int per = (ngroups + kNumXcd - 1) / kNumXcd;
    const bool use_far = far2 < __builtin_inff();
    unsigned long long n_node = 0, n_part = 0, n_far = 0, n_loc = 0, n_loct = 0;
    int q = 0;
    const float self_inv = __builtin_amdgcn_rsqf(eps2);
    // A round adds at most TEAM*64 node entries and TEAM*64*(leaf+2)
    // particle entries; lists are evaluated before a round could overflow them
    // (each room is positive for every leaf size up to kMaxLeaf).
    const int node_room = TEAM * (kNodeList - kWave);
    const int part_room = TEAM * (kPartList - kWave * (leaf + 2));
    const int loc_room = TEAM * (kLocList - kWave);
    if (l < kMaxSplit) {
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        L.ca[kWave + l] = z;
        L.cq0[kWave + l] = z;
        L.cq1[kWave + l] = z;
    }
    bool overflow = false;
    for (;;) {
        __syncthreads();  // the previous group's use of ctl is over
        if (tid == 0) {
            int g = -1;
            while (q < kNumXcd) {
                const int c = (xcd + q) % kNumXcd;
                const int lo = c * per, hi = min(ngroups, lo + per);
                const unsigned k = atomicAdd(&qctr[c], 1u);
                if (lo + (int)k < hi) {
                    g = gorder ? gorder[lo + (int)k] : lo + (int)k;
                    break;
                }
                ++q;
            }
            ctl[0] = g;
        }
        __syncthreads();
        const int g = ctl[0];
        if (g < 0) break;
        const GroupInfo gi = groups[g];
        const SegInfo sg = segs[gi.seg];
        // The group's bounding box (each wave computes it).
        float lx = __builtin_inff(), ly = lx, lz = lx, hx = -lx, hy = -lx, hz = -lx;
        for (int j = l; j < gi.count; j += kWave) {
            const float4 pj = posm[gi.first + j];
            lx = fminf(lx, pj.x); ly = fminf(ly, pj.y); lz = fminf(lz, pj.z);
            hx = fmaxf(hx, pj.x); hy = fmaxf(hy, pj.y); hz = fmaxf(hz, pj.z);
        }
        lx = wave_min(lx); ly = wave_min(ly); lz = wave_min(lz);
        hx = wave_max(hx); hy = wave_max(hy); hz = wave_max(hz);
        const int n = gi.count;
        int nn = 0, np = 0, nfar = 0, nloc = 0;
        int top = 0, spilled = 0;
        bool flushed = false;
        float gcx = 0.f, gcy = 0.f, gcz = 0.f, glim = 0.f;
        long long tot_loc = 0;
        if (LOCAL) {
            local_frame(lx, ly, lz, hx, hy, hz, kappa2, gcx, gcy, gcz, glim);
            if (l < kLocTerms) L.loc[l] = 0.0f;
        }
        if (sg.n == 1) {
            if (tid == 0) pl[0] = sg.start;
            np = 1;
        } else if (tid == 0) {
            stk[0] = sg.root;
        }
        if (sg.n != 1) top = 1;
        for (;;) {
            // Walk until the stack is empty or a list could overflow.
            bool need = false;
            while (top > 0 || spilled > 0) {
                if (nn > node_room || np > part_room || (LOCAL && nloc > loc_room)) {
                    need = true;
                    break;
                }
                __syncthreads();  // the previous round's stack and list writes are done
                if (top == 0) {
                    // Reload the most recently spilled half.
                    const int c = min(spilled, kLdsStack / 2);
                    for (int k = tid; k < c; k += TEAM * kWave) stk[k] = spill[spilled - c + k];
                    spilled -= c;
                    top = c;
                    __syncthreads();
                }
                // Lanes past the popped entries repeat the last one and
                // decide nothing (bitwise logic: no branches).
                const int cnt = min(top, TEAM * kWave);
                const int e = w * kWave + l;
                const bool have = e < cnt;
                const int v = stk[top - cnt + min(e, cnt - 1)];
                const float4 A = node[2 * v];
                const float4 bb = node[2 * v + 1];
                const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
                          split = __float_as_int(bb.w);
                const float d2b = box_dist2(lx, ly, lz, hx, hy, hz, A.x, A.y, A.z);
                const float d2c = dist2_nc(A.x - gcx, A.y - gcy, A.z - gcz);
                const bool accept = have & (d2b > bb.x);
                // far from the group: the local list; beyond far * the opening
                // radius: a monopole in the particle list
                const bool loc = LOCAL & accept & (d2c > glim);
                const bool far = accept & !loc & (d2b > far2 * bb.x);
                const bool near = accept & !far & !loc;
                const int size = last - first + 1;
                const bool direct = have & !accept & (size <= leaf);
                const bool open = have & !accept & !direct;
                const bool lleaf = open & (first == split), rleaf = open & (last == split + 1);
                // the particle-list run [pf, pf + pc): the node's own particles,
                // its single-particle children, or the node as a monopole (~v)
                int pc = far ? 1 : (int)lleaf + (int)rleaf;
                pc = direct ? size : pc;
                int pf = lleaf ? split : split + 1;
                pf = far ? ~v : pf;
                pf = direct ? first : pf;
                const bool lp = open & !lleaf, rp = open & !rleaf;
                const uint64_t mA = ballot(near), m3 = ballot(lp), m4 = ballot(rp);
                const int incl = wave_scan_incl_dpp(pc);
                const int tp = __builtin_amdgcn_readlane(incl, kWave - 1);
                int nfw = 0;
                if (use_far) nfw = __popcll(ballot(far));
                const uint64_t mL = LOCAL ? ballot(loc) : 0ull;
                if (l == 0) {
                    xch[w][0] = __popcll(mA);
                    xch[w][1] = tp;
                    xch[w][2] = __popcll(m3);
                    xch[w][3] = __popcll(m4);
                    if (use_far) xfar[w] = nfw;
                    if (LOCAL) xloc[w] = __popcll(mL);
                }
                __syncthreads();  // every entry of the round has been read; counts exchanged
                int na_before = 0, tp_before = 0, nl_before = 0, nr_before = 0;
                int na_all = 0, tp_all = 0, nl_all = 0, nr_all = 0, nf_all = 0;
                int lo_before = 0, lo_all = 0;
#pragma unroll
                for (int k = 0; k < TEAM; ++k) {
                    const int a0 = xch[k][0], a1 = xch[k][1], a2 = xch[k][2], a3 = xch[k][3];
                    if (k < w) {
                        na_before += a0;
                        tp_before += a1;
                        nl_before += a2;
                        nr_before += a3;
                    }
                    na_all += a0;
                    tp_all += a1;
                    nl_all += a2;
                    nr_all += a3;
                    if (use_far) nf_all += xfar[k];
                    if (LOCAL) {
                        const int a4 = xloc[k];
                        if (k < w) lo_before += a4;
                        lo_all += a4;
                    }
                }
                // accepted nodes -> node list; particles and far nodes ->
                // particle list
                store_word(rs, near ? kNlOff + 4 * (nn + na_before + mbcnt(mA)) : kDrop, v);
                if (LOCAL)
                    store_word(rs, loc ? kLlOff + 4 * (nloc + lo_before + mbcnt(mL)) : kDrop, v);
                const int o = kPlOff + 4 * (np + tp_before + incl - pc);
                int pmax = 0;
                while (ballot(pmax < pc)) ++pmax;
                for (int k = 0; k < pmax; ++k) store_word(rs, k < pc ? o + 4 * k : kDrop, pf + k);
                nn += na_all;
                np += tp_all;
                nfar += nf_all;
                nloc += lo_all;
                // larger children -> stack, in place of the popped entries
                // (lanes with nothing to push write the spare entry past it)
                int base = top - cnt;
                const int npush = nl_all + nr_all;
                if (base + npush > kLdsStack) {
                    // Spill the oldest half of the LDS stack to global memory.
                    const int h = kLdsStack / 2;
                    if (spilled + h > kStackCap) {
                        overflow = true;
                        break;
                    }
                    for (int k = tid; k < h; k += TEAM * kWave) spill[spilled + k] = stk[k];
                    int keep[kLdsStack / 2 / (TEAM * kWave)];
#pragma unroll
                    for (int r = 0; r < kLdsStack / 2 / (TEAM * kWave); ++r) {
                        const int k = tid + r * TEAM * kWave;
                        keep[r] = k < base - h ? stk[k + h] : 0;
                    }
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < kLdsStack / 2 / (TEAM * kWave); ++r) {
                        const int k = tid + r * TEAM * kWave;
                        if (k < base - h) stk[k] = keep[r];
                    }
                    spilled += h;
                    base -= h;
                }
                stk[lp ? base + nl_before + mbcnt(m3) : kLdsStack] = split + sg.int_base;
                stk[rp ? base + nl_all + nr_before + mbcnt(m4) : kLdsStack] =
                    split + 1 + sg.int_base;
                top = base + npush;
            }
            if (overflow) break;
            __syncthreads();  // the lists are complete in memory
            // Evaluate this wave's targets over the lists: partial sums, or
            // the final sums once the walk is done.
            // This wave's chunks of the lists, for all targets; then the
            // waves' sums of each target in wave order.
            {
                // This wave's chunks of the local list first (its running
                // share of the group's expansion, in its own LDS row).
                if (LOCAL && nloc > w * kWave)
                    eval_local_list<POT>(L.loc, ll, nloc, node, nodeQ, gcx, gcy, gcz, eps2,
                                         w * kWave, TEAM * kWave);
                const Share sh = make_share(n, l);
                Targets tg;
                Acc a[kMaxPer];
#pragma unroll
                for (int u = 0; u < kMaxPer; ++u) {
                    const float4 pu = posm[gi.first + min(sh.t + u * sh.W, n - 1)];
                    tg.x[u] = pu.x;
                    tg.y[u] = pu.y;
                    tg.z[u] = pu.z;
                    a[u] = Acc{0.f, 0.f, 0.f, 0.f};
                }
                if (nn > w * kWave)
                    eval_nodes<POT, QUAD>(L, nl, nn, node, nodeQ, a, tg, eps2, sh, w * kWave,
                                          TEAM * kWave);
                if (np > w * kWave)
                    eval_parts<POT>(L, pl, np, posm, node, a, tg, eps2, sh, w * kWave,
                                    TEAM * kWave);
                float4 r[kMaxPer];
                combine_share(L, a, sh, n, r);
                float4* const scr = L.ca;
#pragma unroll
                for (int u = 0; u < kMaxPer; ++u) {
                    const int t = l + u * kWave;
                    if (t < n) scr[(t >> 6) * kChunk + (t & 63)] = r[u];
                }
            }
            __syncthreads();  // every wave's sums are in LDS
            tot_loc += nloc;
            if (w == 0) {
                // The group's local expansion: the waves' shares added in wave
                // order into wave 0's row (applied with the final sums).
                const bool apply = LOCAL && !need && tot_loc > 0;
                if (apply) {
                    if (l < kLocTerms) {
                        float v = lds[0].loc[l];
#pragma unroll
                        for (int ww = 1; ww < TEAM; ++ww) v += lds[ww].loc[l];
                        lds[0].loc[l] = v;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
#pragma unroll
                for (int u = 0; u < kMaxPer; ++u) {
                    const int t = l + u * kWave;
                    if (t < n) {
                        const int slot = (t >> 6) * kChunk + (t & 63);
                        float4 v4 = lds[0].ca[slot];
#pragma unroll
                        for (int ww = 1; ww < TEAM; ++ww) {
                            const float4 o4 = lds[ww].ca[slot];
                            v4.x += o4.x;
                            v4.y += o4.y;
                            v4.z += o4.z;
                            v4.w += o4.w;
                        }
                        const int i = gi.first + t;
                        if (flushed) {
                            const float4 p4 = acc[i];
                            v4 = make_float4(p4.x + v4.x, p4.y + v4.y, p4.z + v4.z, p4.w + v4.w);
                        }
                        if (!need) {
                            const float4 pi = posm[i];
                            if (apply)
                                local_apply<POT>(lds[0].loc, pi.x - gcx, pi.y - gcy, pi.z - gcz, v4);
                            const float ph = POT ? -G * (v4.w - pi.w * self_inv) : 0.0f;
                            v4 = make_float4(G * v4.x, G * v4.y, G * v4.z, ph);
                        }
                        acc[i] = v4;
                    }
                }
                n_node += (unsigned long long)(nn + nfar) * (unsigned long long)n;
                n_far += (unsigned long long)nfar * (unsigned long long)n;
                n_part += (unsigned long long)(np - nfar) * (unsigned long long)n;
                n_loc += (unsigned long long)nloc;
                n_loct += (unsigned long long)nloc * (unsigned long long)n;
            }
            nn = 0;
            np = 0;
            nfar = 0;
            nloc = 0;
            flushed = true;
            if (!need) break;
            __syncthreads();  // the waves are done with the lists and LDS sums before reuse
        }
        if (overflow) {
            if (tid == 0) atomicOr(err, kErrStack);
            break;
        }
        if (w == 0) n_part -= (unsigned long long)n;  // the particle itself, in its own list
    }
    // The team's counts (accumulated by wave 0) in the block's slot.
    if (tid == 0) {
        unsigned long long* c = counters + kCounts * (size_t)blockIdx.x;
        c[0] = n_node;
        c[1] = n_part;
        c[2] = n_far;
        c[3] = n_loc;
        c[4] = n_loct;
    }
}

// cuda_engine.cu:91-161 -- per-particle DFS from the instance root; left child
// pushed before right; leaf taken when d^2+eps^2 > 1.01*eps^2; internal node
// taken when width^2 < d^2+eps^2; force G*m*rsqrt(d2)^3 * d.  `limit` (at most
// STACK) is the stack depth; a traversal that would exceed it stops and sets
// the stack error bit.  Every expression rounds as nvcc compiles the
// reference (--fmad=true): d^2 = fma(dz,dz, fma(dx,dx, dy*dy)) + eps^2, and
// the accumulation's multiply-adds; nothing else is contracted.
//
// One wave walks for its 64 particles (consecutive in Morton order) at once:
// an LDS stack of (node, lane mask) entries, the mask holding the lanes whose
// own walk reaches the node.  A popped node is read once (scalar loads) and
// decided by the lanes of its mask; the lanes that open it push its children
// with their mask.  Every lane thus meets its own nodes in the order of its own
// DFS (the tree's preorder, right child first), so its sums are those of a
// per-particle walk.  The pending entries are left children along the current
// path, their masks nested, so the stack never holds more entries than the
// deepest lane's own stack: at most `limit`.  Each lane counts its own depth
// and stops, with the error bit, where its own walk would overflow.
template <int STACK>
__global__ __launch_bounds__(256) void k_traverse_reference(
    const TileInfo* __restrict__ tiles, const SegInfo* __restrict__ segs,
    const float4* __restrict__ posm, const float4* __restrict__ node,
    float4* __restrict__ acc, unsigned long long* counters, unsigned* err, float G, float eps2,
    int limit)
{
#pragma clang fp contract(off)
    __shared__ int snode[4][STACK];
    __shared__ uint64_t smask[4][STACK];
    const TileInfo t = tiles[blockIdx.x >> 2];
    const int w = (int)threadIdx.x >> 6, l = lane_id();
    const int off = (int)(blockIdx.x & 3) * kTileThreads + (int)threadIdx.x;
    const bool valid = off < t.count;
    const int i = t.start + (valid ? off : 0);
    const SegInfo sg = segs[t.seg];
    const float4 p = posm[i];
    float ax = 0.f, ay = 0.f, az = 0.f;
    unsigned long long count = 0;
    int depth = 1;  // this lane's own stack depth
    uint64_t live = ballot(valid);  // lanes still walking
    bool overflow = false;
    int top = 0;
    if (live) {
        if (l == 0) {
            snode[w][0] = sg.root;
            smask[w][0] = live;
        }
        top = 1;
    }
    __builtin_amdgcn_wave_barrier();
    while (top > 0) {
        --top;
        const int v = __builtin_amdgcn_readfirstlane(snode[w][top]);
        const uint64_t m = smask[w][top] & live;
        __builtin_amdgcn_wave_barrier();
        if (m == 0ull) continue;
        const bool act = (m >> l) & 1ull;
        if (act) --depth;
        const float4 A = node[2 * v];
        const float4 bb = node[2 * v + 1];
        const int first = __float_as_int(bb.y), last = __float_as_int(bb.z),
                  split = __float_as_int(bb.w);
        const float dx = A.x - p.x, dy = A.y - p.y, dz = A.z - p.z;
        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dx, dx, dy * dy)) + eps2;
        bool take = false, open = false;
        if (act) {
            if (split < 0) {
                take = d2 > eps2 * 1.01f;
            } else {
                const float width = bb.x;
                take = width * width < d2 * 1.0f;
                if (!take) {
                    if (depth + 2 > limit) {
                        overflow = true;
                    } else {
                        open = true;
                        depth += 2;
                    }
                }
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
        live &= ~ballot(overflow);
        const uint64_t om = ballot(open);
        if (om != 0ull) {
            if (l == 0) {
                snode[w][top] = (first == split) ? split + sg.leaf_base : split + sg.int_base;
                smask[w][top] = om;
                snode[w][top + 1] =
                    (last == split + 1) ? split + 1 + sg.leaf_base : split + 1 + sg.int_base;
                smask[w][top + 1] = om;
            }
            top += 2;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (valid) {
        if (overflow) atomicOr(err, kErrStack);
        acc[i] = make_float4(ax, ay, az, 0.0f);
    }
    count = wave_sum(count);
    if (l == 0 && count) atomicAdd(&counters[1], count);
}

}  // namespace gcs
