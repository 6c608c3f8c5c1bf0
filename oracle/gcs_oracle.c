/*
 * gcs_oracle.c -- CPU oracle for the per-frame N-body step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
 * Never linked into or called by the product library.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Contraction
 * is off so every float expression rounds exactly as written; where the
 * reference's nvcc build fuses a multiply-add (nvcc defaults to --fmad=true)
 * the fused form is written out with fmaf().  A sum of two products a*b + c*d
 * is written fmaf(a, b, c*d) -- the left product fused, the right one rounded
 * -- and a*a + b*b + c*c as fmaf(c, c, fmaf(a, a, b*b)): the operand choice
 * of LLVM's DAG combiner (visitFADDForFMACombine, left operand first when
 * both multiplies have one use), the code generator nvcc's NVVM derives
 * from.  nvcc itself cannot run here, so which product nvcc rounds is parity
 * unpinned (a last-bit difference in a COM coordinate, a traversal distance
 * or a colour speed).
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_expand_bits / orc_morton   naive per-bit interleave over all 2^21
 *                                  inputs, known-answer values, edge values,
 *                                  the reference's compiled get_morton_code
 *                                  (tests/golden/morton_ref.npz)
 *   orc_stable_sort_pairs          numpy stable argsort
 *   orc_karras_build               orc_radix_tree_definition (split at the
 *                                  highest differing bit), tree invariants
 *   orc_aggregate                  FP64 subtree sums, exact AABB min/max
 *   orc_traverse_reference         theta->0 limit equals the direct sum;
 *                                  every mass taken or skipped once; the
 *                                  opening rule w^2 < d^2 + eps^2 and the
 *                                  leaf skip d^2 + eps^2 <= 1.01 eps^2 on
 *                                  cases just either side of each threshold
 *   orc_traverse_group             theta=0 and leaf>=n equal the direct sum;
 *                                  Plummer two-body closed form; every mass
 *                                  counted once; error bound vs FP64; the
 *                                  group opening threshold either side of
 *                                  (w/theta)^2 (and of the offset form); the
 *                                  softened quadrupole's second-order
 *                                  convergence (halving a cluster cuts the
 *                                  error ~8x); far monopoles: the threshold
 *                                  either side of mono_far x the opening
 *                                  radius against the monopole's closed form,
 *                                  unchanged lists; local expansions: third-
 *                                  order convergence in the target offset
 *                                  (~8x per halving; a dropped B term gives
 *                                  4x, a dropped A term 2x) for force and
 *                                  potential, the threshold either side of
 *                                  kappa x the group half-diagonal, every
 *                                  mass kept, the per-target counts split
 *                                  exactly between node and local terms
 *   orc_tree_groups                maximal-subtree definition by brute force
 *   orc_direct_sum_f64             Plummer two-body closed form, Newton's
 *                                  third law (zero total force)
 *   orc_integrate_reference        constant-acceleration closed form, colour
 *                                  table spot values
 *   (compensated positions have no oracle function: the GPU test pins them
 *    to the free-drift closed form x0 + k v dt in FP64)
 *   orc_octree_step                leaf_cap >= n reduces to the direct sum;
 *                                  error bound vs FP64 at theta=0.5;
 *                                  bit-exact against the compiled reference
 *                                  (oracle/_ref) and its golden frame
 *                                  (tests/golden/octree_ref.npz)
 */
#include "gcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ======================================================================== */
/* Morton codes                                                             */
/* ======================================================================== */

/* morton.h:11-19 -- spread the low 21 bits of v to every third bit. */
uint64_t orc_expand_bits(uint32_t v)
{
    uint64_t w = (uint64_t)(v & 0x1FFFFFu);
    w = (w | (w << 32)) & 0x001F00000000FFFFull;
    w = (w | (w << 16)) & 0x001F0000FF0000FFull;
    w = (w | (w << 8)) & 0x100F00F00F00F00Full;
    w = (w | (w << 4)) & 0x10C30C30C30C30C3ull;
    w = (w | (w << 2)) & 0x1249249249249249ull;
    return w;
}

/* morton.h:24-26 casts float to uint32_t.  On the reference's GPU that cast
 * is the saturating round-toward-zero conversion: NaN and values below 1
 * give 0, values at or above 2^32 give 0xFFFFFFFF, everything else is
 * truncated.  x86 leaves out-of-range casts undefined, so the conversion is
 * spelled out here (and identically in the HIP kernel). */
uint32_t orc_f2u_sat(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}

/* morton.h:22-28 -- one IEEE float divide for the scale, then per axis
 * (v - min) * scale, converted and masked to 21 bits; x is most significant. */
uint64_t orc_morton(float x, float y, float z, float mn, float mx)
{
    float scale = 2097151.0f / (mx - mn);
    uint32_t ix = orc_f2u_sat((x - mn) * scale);
    uint32_t iy = orc_f2u_sat((y - mn) * scale);
    uint32_t iz = orc_f2u_sat((z - mn) * scale);
    return (orc_expand_bits(ix) << 2) | (orc_expand_bits(iy) << 1) |
           orc_expand_bits(iz);
}

/* karras_bvh.cuh:34-41 */
void orc_morton_n(const float* x, const float* y, const float* z, int64_t n,
                  float mn, float mx, uint64_t* keys)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) keys[i] = orc_morton(x[i], y[i], z[i], mn, mx);
}

/* main.cpp:152-166 -- per-axis min/max, then one scalar cube. */
void orc_bounds_cube(const float* x, const float* y, const float* z, int64_t n,
                     float* mn, float* mx)
{
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int64_t i = 0; i < n; ++i) {
        if (x[i] < lo[0]) lo[0] = x[i];
        if (y[i] < lo[1]) lo[1] = y[i];
        if (z[i] < lo[2]) lo[2] = z[i];
        if (x[i] > hi[0]) hi[0] = x[i];
        if (y[i] > hi[1]) hi[1] = y[i];
        if (z[i] > hi[2]) hi[2] = z[i];
    }
    float a = lo[0] < lo[1] ? lo[0] : lo[1];
    a = a < lo[2] ? a : lo[2];
    float b = hi[0] > hi[1] ? hi[0] : hi[1];
    b = b > hi[2] ? b : hi[2];
    *mn = a;
    *mx = b;
}

/* ======================================================================== */
/* Stable sort by key (CUB SortPairs: ascending, stable, all 64 bits)       */
/* ======================================================================== */

void orc_stable_sort_pairs(const uint64_t* keys, int64_t n, uint64_t* keys_out,
                           int32_t* perm_out)
{
    /* Bottom-up merge sort of (key, original index); ties keep input order. */
    uint64_t* ka = (uint64_t*)malloc((size_t)(n > 0 ? n : 1) * 8);
    uint64_t* kb = (uint64_t*)malloc((size_t)(n > 0 ? n : 1) * 8);
    int32_t* va = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * 4);
    int32_t* vb = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * 4);
    for (int64_t i = 0; i < n; ++i) {
        ka[i] = keys[i];
        va[i] = (int32_t)i;
    }
    for (int64_t width = 1; width < n; width *= 2) {
        for (int64_t lo = 0; lo < n; lo += 2 * width) {
            int64_t mid = lo + width < n ? lo + width : n;
            int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            int64_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                if (ka[j] < ka[i]) { kb[o] = ka[j]; vb[o++] = va[j++]; }
                else { kb[o] = ka[i]; vb[o++] = va[i++]; }
            }
            while (i < mid) { kb[o] = ka[i]; vb[o++] = va[i++]; }
            while (j < hi) { kb[o] = ka[j]; vb[o++] = va[j++]; }
        }
        uint64_t* tk = ka; ka = kb; kb = tk;
        int32_t* tv = va; va = vb; vb = tv;
    }
    for (int64_t i = 0; i < n; ++i) {
        keys_out[i] = ka[i];
        perm_out[i] = va[i];
    }
    free(ka); free(kb); free(va); free(vb);
}

/* ======================================================================== */
/* Karras radix tree                                                        */
/* ======================================================================== */

static int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
static int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

/* karras_bvh.cuh:23-31 -- common-prefix length of the 96-bit key
 * (key, position); -1 outside [0, n). */
static int delta(int32_t i, int32_t j, int32_t n, const uint64_t* k)
{
    if (j < 0 || j >= n) return -1;
    if (k[i] == k[j]) return 64 + clz32((uint32_t)(i ^ j));
    return clz64(k[i] ^ k[j]);
}

/* karras_bvh.cuh:73-131 -- one internal node per i in [0, n-1).  Internal
 * nodes are 0..n-2, leaves n-1..2n-2; parent of the root is -1. */
void orc_karras_build(const uint64_t* skeys, int32_t n, int32_t* left,
                      int32_t* right, int32_t* parent)
{
    int32_t nodes = n > 0 ? 2 * n - 1 : 0;
    for (int32_t v = 0; v < nodes; ++v) {
        left[v] = -1;
        right[v] = -1;
        parent[v] = -1;
    }
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n - 1; ++i) {
        int dir = delta(i, i + 1, n, skeys) - delta(i, i - 1, n, skeys) > 0 ? 1 : -1;
        int dmin = delta(i, i - dir, n, skeys);
        int64_t lmax = 2;
        while (delta(i, (int32_t)(i + lmax * dir), n, skeys) > dmin) lmax *= 2;
        int64_t len = 0;
        for (int64_t t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, (int32_t)(i + (len + t) * dir), n, skeys) > dmin) len += t;
        int32_t j = (int32_t)(i + len * dir);
        int dnode = delta(i, j, n, skeys);
        int64_t s = 0, step = len;
        do {
            step = (step + 1) >> 1;
            if (delta(i, (int32_t)(i + (s + step) * dir), n, skeys) > dnode) s += step;
        } while (step > 1);
        int32_t split = (int32_t)(i + s * dir + (dir > 0 ? 0 : -1));
        int32_t lo = i < j ? i : j, hi = i < j ? j : i;
        int32_t cl = (lo == split) ? (n - 1) + split : split;
        int32_t cr = (hi == split + 1) ? (n - 1) + split + 1 : split + 1;
        left[i] = cl;
        right[i] = cr;
        parent[cl] = i;
        parent[cr] = i;
    }
}

/* Bit p (0 = most significant) of the 96-bit key (skeys[i], i). */
static int aug_bit(const uint64_t* k, int32_t i, int p)
{
    if (p < 64) return (int)((k[i] >> (63 - p)) & 1u);
    return (int)(((uint32_t)i >> (95 - p)) & 1u);
}

/* Radix tree by definition: the node over [lo,hi] splits where bit p, the
 * first bit on which the 96-bit keys of lo and hi differ, turns from 0 to 1.
 * The node over [lo,hi] is numbered by the Karras convention: the root is 0,
 * a left child over [lo,s] is internal node s, a right child over [s+1,hi]
 * is internal node s+1, and single-element ranges are leaves n-1+index. */
int orc_radix_tree_definition(const uint64_t* skeys, int32_t n, int32_t* left,
                              int32_t* right, int32_t* parent)
{
    int32_t nodes = n > 0 ? 2 * n - 1 : 0;
    for (int32_t v = 0; v < nodes; ++v) {
        left[v] = -1;
        right[v] = -1;
        parent[v] = -1;
    }
    if (n < 2) return 0;
    int32_t* st_lo = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* st_hi = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* st_id = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int top = 0;
    st_lo[top] = 0; st_hi[top] = n - 1; st_id[top] = 0; ++top;
    int rc = 0;
    while (top > 0) {
        --top;
        int32_t lo = st_lo[top], hi = st_hi[top], id = st_id[top];
        int p = 0;
        while (p < 96 && aug_bit(skeys, lo, p) == aug_bit(skeys, hi, p)) ++p;
        if (p == 96) { rc = -1; break; } /* impossible: positions differ */
        int32_t s = lo;
        while (s + 1 <= hi && aug_bit(skeys, s + 1, p) == 0) ++s;
        int32_t cl = (s == lo) ? (n - 1) + s : s;
        int32_t cr = (s + 1 == hi) ? (n - 1) + s + 1 : s + 1;
        left[id] = cl;
        right[id] = cr;
        parent[cl] = id;
        parent[cr] = id;
        if (s > lo) { st_lo[top] = lo; st_hi[top] = s; st_id[top] = s; ++top; }
        if (s + 1 < hi) { st_lo[top] = s + 1; st_hi[top] = hi; st_id[top] = s + 1; ++top; }
    }
    free(st_lo); free(st_hi); free(st_id);
    return rc;
}

/* Post-order list of internal nodes (children before parents). */
static int32_t* post_order(int32_t n, const int32_t* left, const int32_t* right,
                           int32_t* count)
{
    *count = 0;
    if (n < 2) return NULL;
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n - 1));
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n));
    int32_t* state = (int32_t*)calloc((size_t)(n - 1), sizeof(int32_t));
    int top = 0, oc = 0;
    stack[top++] = 0;
    while (top > 0) {
        int32_t v = stack[top - 1];
        if (!state[v]) {
            state[v] = 1;
            if (right[v] < n - 1) stack[top++] = right[v];
            if (left[v] < n - 1) stack[top++] = left[v];
        } else {
            --top;
            order[oc++] = v;
        }
    }
    free(stack);
    free(state);
    *count = oc;
    return order;
}

void orc_node_ranges(int32_t n, const int32_t* left, const int32_t* right,
                     int32_t* first, int32_t* last)
{
    for (int32_t i = 0; i < n; ++i) {
        first[n - 1 + i] = i;
        last[n - 1 + i] = i;
    }
    int32_t oc;
    int32_t* order = post_order(n, left, right, &oc);
    for (int32_t k = 0; k < oc; ++k) {
        int32_t v = order[k];
        first[v] = first[left[v]];
        last[v] = last[right[v]];
    }
    free(order);
}

/* Adds a child's second moments about the parent's COM: the traceless
 * quadrupole Q += Qc + mc*(3 s s^T - |s|^2 I) and the trace of the raw second
 * moment T += Tc + mc*|s|^2, s = child COM - parent COM, with the components
 * ordered xx, yy, zz, xy, xz, yz, T.  Same operation order as the HIP
 * aggregation kernel. */
static void quad_shift(float* q, const float* qc, const float* cc, const float* cp)
{
    float sx = cc[0] - cp[0], sy = cc[1] - cp[1], sz = cc[2] - cp[2];
    float mc = cc[3];
    float s2 = sx * sx + sy * sy + sz * sz;
    q[0] = q[0] + (qc[0] + mc * (3.0f * sx * sx - s2));
    q[1] = q[1] + (qc[1] + mc * (3.0f * sy * sy - s2));
    q[2] = q[2] + (qc[2] + mc * (3.0f * sz * sz - s2));
    q[3] = q[3] + (qc[3] + mc * (3.0f * sx * sy));
    q[4] = q[4] + (qc[4] + mc * (3.0f * sx * sz));
    q[5] = q[5] + (qc[5] + mc * (3.0f * sy * sz));
    q[6] = q[6] + (qc[6] + mc * s2);
}

/* karras_bvh.cuh:134-164 (leaves) and :175-225 (internal merge).  The merge
 * is order-independent given the children, so a post-order walk reproduces
 * the atomic climb exactly.  Children read: AABB, mass, COM.  Written:
 *   AABB = union; width = max extent (:206); m = ml + mr;
 *   com = fmaf(cl, ml, cr*mr)/m (nvcc's contraction of cl*ml + cr*mr), or the
 *   AABB centre when m == 0 (:208-221).
 * Extension (not in the reference): crit = (width*inv_theta)^2, or with
 * offset != 0 (width*inv_theta + |com - c|)^2 with c the AABB centre, the
 * opening radius squared used by the group traversal. */
void orc_aggregate(int32_t n, const int32_t* left, const int32_t* right,
                   const float* x, const float* y, const float* z,
                   const float* m, float inv_theta, int32_t offset, float* com4, float* width,
                   float* bmin, float* bmax, float* crit, float* quad)
{
    for (int32_t i = 0; i < n; ++i) {
        int32_t v = n - 1 + i;
        com4[4 * v + 0] = x[i];
        com4[4 * v + 1] = y[i];
        com4[4 * v + 2] = z[i];
        com4[4 * v + 3] = m[i];
        width[v] = 0.0f;
        crit[v] = 0.0f;
        if (quad)
            for (int c = 0; c < 7; ++c) quad[7 * v + c] = 0.0f;
        bmin[3 * v + 0] = bmax[3 * v + 0] = x[i];
        bmin[3 * v + 1] = bmax[3 * v + 1] = y[i];
        bmin[3 * v + 2] = bmax[3 * v + 2] = z[i];
    }
    int32_t oc;
    int32_t* order = post_order(n, left, right, &oc);
    for (int32_t k = 0; k < oc; ++k) {
        int32_t v = order[k], l = left[v], r = right[v];
        float lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(bmin[3 * l + a], bmin[3 * r + a]);
            hi[a] = fmaxf(bmax[3 * l + a], bmax[3 * r + a]);
            bmin[3 * v + a] = lo[a];
            bmax[3 * v + a] = hi[a];
        }
        float w = fmaxf(hi[0] - lo[0], fmaxf(hi[1] - lo[1], hi[2] - lo[2]));
        width[v] = w;
        float ml = com4[4 * l + 3], mr = com4[4 * r + 3];
        float mt = ml + mr;
        com4[4 * v + 3] = mt;
        for (int a = 0; a < 3; ++a) {
            if (mt > 0.0f)
                com4[4 * v + a] = fmaf(com4[4 * l + a], ml, com4[4 * r + a] * mr) / mt;
            else
                com4[4 * v + a] = (lo[a] + hi[a]) * 0.5f;
        }
        float dx = com4[4 * v + 0] - (lo[0] + hi[0]) * 0.5f;
        float dy = com4[4 * v + 1] - (lo[1] + hi[1]) * 0.5f;
        float dz = com4[4 * v + 2] - (lo[2] + hi[2]) * 0.5f;
        float off = sqrtf(dx * dx + dy * dy + dz * dz);
        float ro = offset ? w * inv_theta + off : w * inv_theta;
        crit[v] = ro * ro;
        if (quad) {
            float q[7] = {0, 0, 0, 0, 0, 0, 0};
            quad_shift(q, &quad[7 * l], &com4[4 * l], &com4[4 * v]);
            quad_shift(q, &quad[7 * r], &com4[4 * r], &com4[4 * v]);
            for (int c = 0; c < 7; ++c) quad[7 * v + c] = q[c];
        }
    }
    free(order);
}

/* ======================================================================== */
/* Traversals                                                               */
/* ======================================================================== */

/* cuda_engine.cu:91-161.  Stack order as written (left pushed, then right). */
int64_t orc_traverse_reference(int32_t n, const float* x, const float* y,
                               const float* z, const int32_t* left,
                               const int32_t* right, const float* com4,
                               const float* width, float G, float eps2,
                               double* acc, double* msum, double* mskip)
{
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int32_t i = 0; i < n; ++i) {
        int32_t stack[512];
        int top = 0;
        double ax = 0, ay = 0, az = 0, ms = 0, mk = 0;
        stack[top++] = 0;
        while (top > 0) {
            int32_t v = stack[--top];
            float dx = com4[4 * v + 0] - x[i];
            float dy = com4[4 * v + 1] - y[i];
            float dz = com4[4 * v + 2] - z[i];
            float d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy)) + eps2;
            int take = 0;
            if (v >= n - 1) {
                take = d2 > eps2 * 1.01f;
                if (!take) mk += com4[4 * v + 3];
            } else if (width[v] * width[v] < d2 * 1.0f) {
                take = 1;
            } else {
                if (left[v] != -1) stack[top++] = left[v];
                if (right[v] != -1) stack[top++] = right[v];
                if (top > 510) abort();
            }
            if (take) {
                double inv = 1.0 / sqrt((double)d2);
                double f = (double)G * (double)com4[4 * v + 3] * inv * inv * inv;
                ax += dx * f;
                ay += dy * f;
                az += dz * f;
                ms += com4[4 * v + 3];
                ++total;
            }
        }
        if (msum) msum[i] = ms;
        if (mskip) mskip[i] = mk;
        acc[4 * i + 0] = ax;
        acc[4 * i + 1] = ay;
        acc[4 * i + 2] = az;
        acc[4 * i + 3] = 0.0;
    }
    return total;
}

/* Squared distance from point c to the box [lo,hi] in float, as the HIP
 * kernel computes it (no contraction). */
static float box_dist2(const float* lo, const float* hi, float cx, float cy,
                       float cz)
{
    float ex = fmaxf(fmaxf(lo[0] - cx, cx - hi[0]), 0.0f);
    float ey = fmaxf(fmaxf(lo[1] - cy, cy - hi[1]), 0.0f);
    float ez = fmaxf(fmaxf(lo[2] - cz, cz - hi[2]), 0.0f);
    return ex * ex + ey * ey + ez * ez;
}

/* Second-order local expansion at a group's centre c of a node at r = COM - c
 * (mass m, quadrupole and trace q[0..6]): its field at c + y is
 *   a(y) = a0 + A y + (1/2) B:yy,   phi(y) = phi0 - a0.y - (1/2) y.A.y,
 * with a0 and phi0 the softened quadrupole field at c and, from the softened
 * monopole (R^2 = |r|^2 + eps^2),
 *   A_ij  = m (3 r_i r_j / R^5 - delta_ij / R^3),
 *   B_ijk = m (15 r_i r_j r_k / R^7 - 3 (delta_ij r_k + delta_ik r_j + delta_jk r_i) / R^5).
 * L[20] accumulates a0 (3), A (xx yy zz xy xz yz), B (xxx yyy zzz xxy xxz xyy
 * yyz xzz yzz xyz) and phi0, each a sum over the group's far nodes. */
static void local_terms(double rx, double ry, double rz, double m, const double* q, double e2,
                        double* L)
{
    const double R2 = rx * rx + ry * ry + rz * rz + e2;
    const double iR = 1.0 / sqrt(R2), iR2 = iR * iR, iR3 = iR * iR2, iR5 = iR3 * iR2,
                 iR7 = iR5 * iR2;
    const double qx = q[0] * rx + q[3] * ry + q[4] * rz;
    const double qy = q[3] * rx + q[1] * ry + q[5] * rz;
    const double qz = q[4] * rx + q[5] * ry + q[2] * rz;
    const double dqd = rx * qx + ry * qy + rz * qz - q[6] * e2;
    const double f = m * iR3 + 2.5 * dqd * iR7;
    L[0] += rx * f - qx * iR5;
    L[1] += ry * f - qy * iR5;
    L[2] += rz * f - qz * iR5;
    const double m3 = m * iR3, m5 = 3.0 * m * iR5, m7 = 15.0 * m * iR7;
    L[3] += m5 * rx * rx - m3;
    L[4] += m5 * ry * ry - m3;
    L[5] += m5 * rz * rz - m3;
    L[6] += m5 * rx * ry;
    L[7] += m5 * rx * rz;
    L[8] += m5 * ry * rz;
    L[9] += m7 * rx * rx * rx - 3.0 * m5 * rx;
    L[10] += m7 * ry * ry * ry - 3.0 * m5 * ry;
    L[11] += m7 * rz * rz * rz - 3.0 * m5 * rz;
    L[12] += m7 * rx * rx * ry - m5 * ry;
    L[13] += m7 * rx * rx * rz - m5 * rz;
    L[14] += m7 * rx * ry * ry - m5 * rx;
    L[15] += m7 * ry * ry * rz - m5 * rz;
    L[16] += m7 * rx * rz * rz - m5 * rx;
    L[17] += m7 * ry * rz * rz - m5 * ry;
    L[18] += m7 * rx * ry * rz;
    L[19] += -(m * iR + 0.5 * dqd * iR5);
}

static int cloc_any(const double* L)
{
    for (int c = 0; c < 20; ++c)
        if (L[c] != 0.0) return 1;
    return 0;
}

/* Adds G x (the accumulated expansion at offset y from the centre) to a[0..3]. */
static void local_apply(const double* L, double yx, double yy, double yz, double G, double* a)
{
    const double xx = yx * yx, yy2 = yy * yy, zz = yz * yz, xy = yx * yy, xz = yx * yz,
                 yz2 = yy * yz;
    const double Ax = L[3] * yx + L[6] * yy + L[7] * yz;
    const double Ay = L[6] * yx + L[4] * yy + L[8] * yz;
    const double Az = L[7] * yx + L[8] * yy + L[5] * yz;
    const double bx = L[9] * xx + L[14] * yy2 + L[16] * zz +
                      2.0 * (L[12] * xy + L[13] * xz + L[18] * yz2);
    const double by = L[12] * xx + L[10] * yy2 + L[17] * zz +
                      2.0 * (L[14] * xy + L[18] * xz + L[15] * yz2);
    const double bz = L[13] * xx + L[15] * yy2 + L[11] * zz +
                      2.0 * (L[18] * xy + L[16] * xz + L[17] * yz2);
    a[0] += G * (L[0] + Ax + 0.5 * bx);
    a[1] += G * (L[1] + Ay + 0.5 * by);
    a[2] += G * (L[2] + Az + 0.5 * bz);
    a[3] += G * (L[19] - (L[0] * yx + L[1] * yy + L[2] * yz) -
                 0.5 * (Ax * yx + Ay * yy + Az * yz));
}

void orc_traverse_group(int32_t n, const float* x, const float* y,
                        const float* z, const float* m, const int32_t* left,
                        const int32_t* right, const int32_t* first,
                        const int32_t* last, const float* com4,
                        const float* crit, const float* quad, int32_t group,
                        int32_t leaf, float G, float eps2, double* acc,
                        int64_t* counts, double* msum, double* aabs,
                        int64_t* gcost, int32_t ngl, const int32_t* glist,
                        float far2, float kappa2)
{
    if (n <= 0) return;
    int32_t ngroups = glist ? ngl : (n + group - 1) / group;
    int64_t cnode = 0, cpart = 0, cfar = 0, cloc = 0, cloct = 0;
    const double Gd = G, e2 = eps2;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : cnode, cpart, cfar, cloc, cloct)
    for (int32_t g = 0; g < ngroups; ++g) {
        int32_t g0 = glist ? glist[2 * g] : g * group;
        int32_t g1 = glist ? g0 + glist[2 * g + 1] : (g0 + group < n ? g0 + group : n);
        float lo[3] = {x[g0], y[g0], z[g0]}, hi[3] = {x[g0], y[g0], z[g0]};
        for (int32_t i = g0; i < g1; ++i) {
            lo[0] = fminf(lo[0], x[i]); hi[0] = fmaxf(hi[0], x[i]);
            lo[1] = fminf(lo[1], y[i]); hi[1] = fmaxf(hi[1], y[i]);
            lo[2] = fminf(lo[2], z[i]); hi[2] = fmaxf(hi[2], z[i]);
        }
        /* The group's box centre and squared half-diagonal, in float as the
         * kernel forms them; the local-expansion radius squared. */
        const float gc[3] = {(lo[0] + hi[0]) * 0.5f, (lo[1] + hi[1]) * 0.5f,
                             (lo[2] + hi[2]) * 0.5f};
        const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
        const float rg2 = 0.25f * (ex * ex + ey * ey + ez * ez);
        const float lim = kappa2 * rg2;
        double L[20];
        for (int c = 0; c < 20; ++c) L[c] = 0.0;
        int32_t cap = 4096;
        int64_t gc_nodes = 0, gc_parts = 0;
        int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
        int top = 0;
        stack[top++] = 0;
        for (int32_t i = g0; i < g1; ++i) {
            acc[4 * i + 0] = acc[4 * i + 1] = acc[4 * i + 2] = acc[4 * i + 3] = 0.0;
            if (msum) msum[i] = 0.0;
            if (aabs) aabs[i] = 0.0;
        }
        while (top > 0) {
            int32_t v = stack[--top];
            int32_t cnt = last[v] - first[v] + 1;
            int internal = v < n - 1;
            float bd2 = internal ? box_dist2(lo, hi, com4[4 * v], com4[4 * v + 1],
                                             com4[4 * v + 2])
                                 : 0.0f;
            if (internal && bd2 > crit[v]) {
                double cx = com4[4 * v], cy = com4[4 * v + 1], cz = com4[4 * v + 2];
                double mm = com4[4 * v + 3];
                const float dlx = com4[4 * v] - gc[0], dly = com4[4 * v + 1] - gc[1],
                            dlz = com4[4 * v + 2] - gc[2];
                if (dlx * dlx + dly * dly + dlz * dlz > lim) {
                    /* Far from the group: its field in a second-order local
                     * expansion at the group's centre. */
                    double qn[7] = {0, 0, 0, 0, 0, 0, 0};
                    if (quad)
                        for (int c = 0; c < 7; ++c) qn[c] = quad[7 * (int64_t)v + c];
                    local_terms(cx - gc[0], cy - gc[1], cz - gc[2], mm, qn, e2, L);
                    cloc += 1;
                    cloct += g1 - g0;
                    cnode += 0;
                    gc_nodes += 1;
                    for (int32_t i = g0; i < g1; ++i) {
                        if (msum) msum[i] += mm;
                        if (aabs) {
                            double dx = cx - (double)x[i], dy = cy - (double)y[i],
                                   dz = cz - (double)z[i];
                            double r2 = dx * dx + dy * dy + dz * dz + e2;
                            aabs[i] += fabs(Gd) * mm / r2;
                        }
                    }
                    continue;
                }
                double q[7] = {0, 0, 0, 0, 0, 0, 0};
                const float far_r2 = far2 * crit[v];
                const int mono = bd2 > far_r2;
                if (quad && !mono)
                    for (int c = 0; c < 7; ++c) q[c] = quad[7 * (int64_t)v + c];
                if (mono) cfar += g1 - g0;
                for (int32_t i = g0; i < g1; ++i) {
                    double dx = cx - (double)x[i], dy = cy - (double)y[i], dz = cz - (double)z[i];
                    double r2 = dx * dx + dy * dy + dz * dz + e2;
                    double inv = 1.0 / sqrt(r2);
                    double inv2 = inv * inv, inv3 = inv * inv2, inv5 = inv3 * inv2;
                    double qx = q[0] * dx + q[3] * dy + q[4] * dz;
                    double qy = q[3] * dx + q[1] * dy + q[5] * dz;
                    double qz = q[4] * dx + q[5] * dy + q[2] * dz;
                    /* Softened kernel: the Laplacian of 1/sqrt(r^2+eps^2) is
                     * -3 eps^2/R^5, so the trace of the second moment adds
                     * -T eps^2 next to d.Q.d. */
                    double dqd = dx * qx + dy * qy + dz * qz - q[6] * e2;
                    double f = mm * inv3 + 2.5 * dqd * inv5 * inv2;
                    acc[4 * i + 0] += Gd * (dx * f - qx * inv5);
                    acc[4 * i + 1] += Gd * (dy * f - qy * inv5);
                    acc[4 * i + 2] += Gd * (dz * f - qz * inv5);
                    acc[4 * i + 3] -= Gd * (mm * inv + 0.5 * dqd * inv5);
                    if (msum) msum[i] += mm;
                    if (aabs)
                        aabs[i] += fabs(Gd) * sqrt((dx * f - qx * inv5) * (dx * f - qx * inv5) +
                                                   (dy * f - qy * inv5) * (dy * f - qy * inv5) +
                                                   (dz * f - qz * inv5) * (dz * f - qz * inv5));
                }
                cnode += g1 - g0;
                gc_nodes += 1;
            } else if (!internal || cnt <= leaf) {
                gc_parts += cnt;
                for (int32_t j = first[v]; j <= last[v]; ++j) {
                    for (int32_t i = g0; i < g1; ++i) {
                        if (i == j) continue;
                        double dx = (double)x[j] - x[i], dy = (double)y[j] - y[i],
                               dz = (double)z[j] - z[i];
                        double r2 = dx * dx + dy * dy + dz * dz + e2;
                        double inv = 1.0 / sqrt(r2);
                        double f = Gd * m[j] * inv * inv * inv;
                        acc[4 * i + 0] += dx * f;
                        acc[4 * i + 1] += dy * f;
                        acc[4 * i + 2] += dz * f;
                        acc[4 * i + 3] -= Gd * m[j] * inv;
                        if (msum) msum[i] += m[j];
                        if (aabs) aabs[i] += fabs(f) * sqrt(dx * dx + dy * dy + dz * dz);
                    }
                }
                /* Self-interactions are excluded from the count. */
                int32_t ov0 = first[v] > g0 ? first[v] : g0;
                int32_t ov1 = last[v] + 1 < g1 ? last[v] + 1 : g1;
                cpart += (int64_t)cnt * (g1 - g0) - (ov1 > ov0 ? ov1 - ov0 : 0);
            } else {
                if (top + 2 > cap) {
                    cap *= 2;
                    stack = (int32_t*)realloc(stack, sizeof(int32_t) * (size_t)cap);
                }
                stack[top++] = right[v];
                stack[top++] = left[v];
            }
        }
        free(stack);
        if (cloc_any(L))
            for (int32_t i = g0; i < g1; ++i)
                local_apply(L, (double)x[i] - gc[0], (double)y[i] - gc[1], (double)z[i] - gc[2], Gd,
                            &acc[4 * i]);
        if (gcost) {
            gcost[2 * g] = gc_nodes;
            gcost[2 * g + 1] = gc_parts;
        }
    }
    counts[0] += cnode;
    counts[1] += cpart;
    counts[2] += cfar;
    counts[3] += cloc;
    counts[4] += cloct;
}

/* Groups taken from the tree: the maximal subtrees holding at most gmax
 * particles, i.e. every node with count <= gmax whose parent holds more (or
 * the root when n <= gmax), in particle order.  out: (first, count) pairs.
 * Returns the number of groups. */
int32_t orc_tree_groups(int32_t n, const int32_t* left, const int32_t* right,
                        const int32_t* first, const int32_t* last, int32_t gmax,
                        int32_t* out)
{
    if (n <= 0) return 0;
    int32_t ng = 0;
    int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * n));
    int top = 0;
    stack[top++] = 0;
    while (top > 0) {
        int32_t v = stack[--top];
        int32_t cnt = last[v] - first[v] + 1;
        if (cnt <= gmax) {
            out[2 * ng] = first[v];
            out[2 * ng + 1] = cnt;
            ++ng;
        } else {
            stack[top++] = right[v];
            stack[top++] = left[v];
        }
    }
    free(stack);
    return ng;
}

void orc_direct_sum_f64(int64_t n, const double* x, const double* y,
                        const double* z, const double* m, int64_t nt,
                        const int64_t* tgt, double G, double eps2, double* acc)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < nt; ++t) {
        int64_t i = tgt[t];
        double ax = 0, ay = 0, az = 0, ph = 0;
        for (int64_t j = 0; j < n; ++j) {
            if (j == i) continue;
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            double inv = 1.0 / sqrt(r2);
            double f = G * m[j] * inv * inv * inv;
            ax += dx * f;
            ay += dy * f;
            az += dz * f;
            ph -= G * m[j] * inv;
        }
        acc[4 * t + 0] = ax;
        acc[4 * t + 1] = ay;
        acc[4 * t + 2] = az;
        acc[4 * t + 3] = ph;
    }
}

/* ======================================================================== */
/* Integrate (cuda_engine.cu:164-216)                                       */
/* ======================================================================== */

static uint8_t to_u8(float f) { return (uint8_t)orc_f2u_sat(f); }

/* Kick then drift with dt (:176-182), written with the fused multiply-add
 * that nvcc's default --fmad=true forms for `v += a*dt` and `x += v*dt`.
 * Colour from speed t = min(|v|/150, 1) (:185-215): fire palette for type 0,
 * ice palette for type 2, other types untouched; `50 + 150*u` and the like
 * are the fused multiply-adds nvcc forms, |v|^2 = fma(vz,vz, fma(vx,vx,
 * vy*vy)). */
void orc_integrate_reference(int64_t n, float* x, float* y, float* z,
                             float* vx, float* vy, float* vz, uint8_t* r,
                             uint8_t* g, uint8_t* b, const uint8_t* type,
                             const float* ax, const float* ay, const float* az,
                             float dt)
{
    for (int64_t i = 0; i < n; ++i) {
        vx[i] = fmaf(ax[i], dt, vx[i]);
        vy[i] = fmaf(ay[i], dt, vy[i]);
        vz[i] = fmaf(az[i], dt, vz[i]);
        x[i] = fmaf(vx[i], dt, x[i]);
        y[i] = fmaf(vy[i], dt, y[i]);
        z[i] = fmaf(vz[i], dt, z[i]);
        if (type[i] != 0 && type[i] != 2) continue;
        float sp = sqrtf(fmaf(vz[i], vz[i], fmaf(vx[i], vx[i], vy[i] * vy[i])));
        float t = fminf(sp / 150.0f, 1.0f);
        if (type[i] == 0) {
            if (t < 0.5f) {
                float u = t * 2.0f;
                r[i] = 255; g[i] = to_u8(fmaf(150.0f, u, 50.0f)); b[i] = 0;
            } else {
                float u = (t - 0.5f) * 2.0f;
                r[i] = 255; g[i] = to_u8(fmaf(55.0f, u, 200.0f)); b[i] = to_u8(150.0f * u);
            }
        } else {
            if (t < 0.5f) {
                float u = t * 2.0f;
                r[i] = 0; g[i] = to_u8(fmaf(100.0f, u, 100.0f)); b[i] = 255;
            } else {
                float u = (t - 0.5f) * 2.0f;
                r[i] = to_u8(200.0f * u); g[i] = to_u8(fmaf(55.0f, u, 200.0f)); b[i] = 255;
            }
        }
    }
}

/* ======================================================================== */
/* CPU octree step                                                          */
/* ======================================================================== */

typedef struct oct_node {
    float cx, cy, cz, hw;     /* cube centre and half width */
    int32_t first_child;      /* index of 8 contiguous children, or -1 */
    uint32_t start, count;    /* leaf slots in leaf_idx */
    int32_t is_leaf;
} oct_node;

typedef struct oct_tree {
    oct_node* nodes;
    int64_t nnodes, cap_nodes;
    uint32_t* leaf_idx;
    int64_t next_slot, cap_slots;
    int32_t leaf_cap;
    int failed;
} oct_tree;

static int64_t oct_alloc8(oct_tree* t)
{
    if (t->nnodes + 8 > t->cap_nodes) {
        t->cap_nodes = t->cap_nodes * 2 + 64;
        t->nodes = (oct_node*)realloc(t->nodes, sizeof(oct_node) * (size_t)t->cap_nodes);
    }
    int64_t c = t->nnodes;
    t->nnodes += 8;
    return c;
}

static uint32_t oct_slots(oct_tree* t)
{
    if (t->next_slot + t->leaf_cap > t->cap_slots) {
        t->cap_slots = t->cap_slots * 2 + t->leaf_cap;
        t->leaf_idx = (uint32_t*)realloc(t->leaf_idx, sizeof(uint32_t) * (size_t)t->cap_slots);
    }
    uint32_t s = (uint32_t)t->next_slot;
    t->next_slot += t->leaf_cap;
    return s;
}

/* engine.cpp:16-22 -- child octant by >= comparisons against the centre. */
static int oct_child_index(const oct_node* nd, float x, float y, float z)
{
    return (x >= nd->cx ? 1 : 0) | (y >= nd->cy ? 2 : 0) | (z >= nd->cz ? 4 : 0);
}

/* engine.cpp:28-64 -- insert; a full leaf (leaf_cap entries) splits into 8
 * cube children (engine.cpp:5-14), re-routes its stars in slot order, then
 * routes the new one.  Depth is limited so coincident points fail instead of
 * recursing without end. */
static void oct_insert(oct_tree* t, int64_t v, uint32_t s, const orc_star* st, int depth)
{
    if (t->failed) return;
    if (depth > 200) { t->failed = 1; return; }
    if (t->nodes[v].is_leaf) {
        if ((int32_t)t->nodes[v].count < t->leaf_cap) {
            t->leaf_idx[t->nodes[v].start + t->nodes[v].count] = s;
            t->nodes[v].count++;
            return;
        }
        uint32_t old_count = t->nodes[v].count, old_start = t->nodes[v].start;
        int64_t c = oct_alloc8(t);
        oct_node* nd = &t->nodes[v];
        nd->is_leaf = 0;
        nd->first_child = (int32_t)c;
        float q = nd->hw * 0.5f;
        for (int k = 0; k < 8; ++k) {
            oct_node* ch = &t->nodes[c + k];
            ch->hw = q;
            ch->cx = nd->cx + ((k & 1) ? q : -q);
            ch->cy = nd->cy + ((k & 2) ? q : -q);
            ch->cz = nd->cz + ((k & 4) ? q : -q);
            ch->is_leaf = 1;
            ch->count = 0;
            ch->first_child = -1;
            ch->start = oct_slots(t);
        }
        for (uint32_t k = 0; k < old_count; ++k) {
            uint32_t o = t->leaf_idx[old_start + k];
            int ci = oct_child_index(&t->nodes[v], st[o].x, st[o].y, st[o].z);
            oct_insert(t, c + ci, o, st, depth + 1);
        }
        int ci = oct_child_index(&t->nodes[v], st[s].x, st[s].y, st[s].z);
        oct_insert(t, c + ci, s, st, depth + 1);
    } else {
        int ci = oct_child_index(&t->nodes[v], st[s].x, st[s].y, st[s].z);
        oct_insert(t, t->nodes[v].first_child + ci, s, st, depth + 1);
    }
}

/* engine.cpp:67-127 -- recursive mass and COM (float, slot / child order). */
static void oct_mass(oct_tree* t, int64_t v, const orc_star* st, float* nm,
                     float* ncx, float* ncy, float* ncz)
{
    oct_node* nd = &t->nodes[v];
    float tm = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
    if (nd->is_leaf) {
        for (uint32_t k = 0; k < nd->count; ++k) {
            uint32_t s = t->leaf_idx[nd->start + k];
            float cm = st[s].mass;
            tm += cm;
            wx += st[s].x * cm;
            wy += st[s].y * cm;
            wz += st[s].z * cm;
        }
    } else {
        for (int k = 0; k < 8; ++k) {
            int64_t c = nd->first_child + k;
            oct_mass(t, c, st, nm, ncx, ncy, ncz);
            float cm = nm[c];
            tm += cm;
            wx += ncx[c] * cm;
            wy += ncy[c] * cm;
            wz += ncz[c] * cm;
        }
        nd = &t->nodes[v];
    }
    nm[v] = tm;
    if (tm > 0.0f) {
        ncx[v] = wx / tm;
        ncy[v] = wy / tm;
        ncz[v] = wz / tm;
    } else {
        ncx[v] = nd->cx;
        ncy[v] = nd->cy;
        ncz[v] = nd->cz;
    }
}

/* gravity_aggregator.cpp:8-83 -- stack query.  theta^2 in float; nodes of
 * zero mass skipped; accept when (2*hw)^2 < theta^2 * |com - p|^2 with the
 * unsoftened distance; leaves summed directly with index self-exclusion;
 * accepted nodes as softened monopoles; 1/sqrt in float. */
static int oct_query_full(const oct_tree* t, const orc_star* st, uint32_t ti,
                          const float* nm, const float* ncx, const float* ncy,
                          const float* ncz, float G, float eps2, float theta,
                          float* oax, float* oay, float* oaz)
{
    int32_t stack[2048];
    int top = 0;
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    const float th2 = theta * theta;
    const float px = st[ti].x, py = st[ti].y, pz = st[ti].z;
    stack[top++] = 0;
    while (top > 0) {
        int32_t v = stack[--top];
        const oct_node* nd = &t->nodes[v];
        if (nm[v] == 0.0f) continue;
        float s = nd->hw * 2.0f;
        float dx = ncx[v] - px, dy = ncy[v] - py, dz = ncz[v] - pz;
        float d2 = dx * dx + dy * dy + dz * dz;
        if (nd->is_leaf) {
            for (uint32_t k = 0; k < nd->count; ++k) {
                uint32_t j = t->leaf_idx[nd->start + k];
                if (j == ti) continue;
                float sx = st[j].x - px, sy = st[j].y - py, sz = st[j].z - pz;
                float r2 = sx * sx + sy * sy + sz * sz + eps2;
                float inv = 1.0f / sqrtf(r2);
                float inv3 = inv * inv * inv;
                float am = G * st[j].mass * inv3;
                ax += am * sx;
                ay += am * sy;
                az += am * sz;
            }
        } else if (s * s < th2 * d2) {
            float r2 = d2 + eps2;
            float inv = 1.0f / sqrtf(r2);
            float inv3 = inv * inv * inv;
            float am = G * nm[v] * inv3;
            ax += am * dx;
            ay += am * dy;
            az += am * dz;
        } else {
            if (top + 8 > 2048) return -1;
            for (int k = 0; k < 8; ++k) stack[top++] = nd->first_child + k;
        }
    }
    *oax = ax; *oay = ay; *oaz = az;
    return 0;
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* main.cpp:152-253 -- one CPU frame. */
int orc_octree_step(orc_star* stars, int64_t n, float G, float eps2, float dt,
                    float theta, int32_t leaf_cap, int32_t do_integrate,
                    float* acc_out, double* phase_s)
{
    double t0 = now_s();
    if (n <= 0) return 0;
    /* main.cpp:152-166 */
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
    for (int64_t i = 0; i < n; ++i) {
        if (stars[i].x < lo[0]) lo[0] = stars[i].x;
        if (stars[i].y < lo[1]) lo[1] = stars[i].y;
        if (stars[i].z < lo[2]) lo[2] = stars[i].z;
        if (stars[i].x > hi[0]) hi[0] = stars[i].x;
        if (stars[i].y > hi[1]) hi[1] = stars[i].y;
        if (stars[i].z > hi[2]) hi[2] = stars[i].z;
    }
    float gmin = fminf(lo[0], fminf(lo[1], lo[2]));
    float gmax = fmaxf(hi[0], fmaxf(hi[1], hi[2]));
    double t1 = now_s();

    /* main.cpp:170-192 -- Morton keys and a sort by key (stable here; the
     * reference's parallel sort leaves equal keys in unspecified order). */
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t* skeys = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)n);
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    orc_star* tmp = (orc_star*)malloc(sizeof(orc_star) * (size_t)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        keys[i] = orc_morton(stars[i].x, stars[i].y, stars[i].z, gmin, gmax);
    orc_stable_sort_pairs(keys, n, skeys, perm);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) tmp[i] = stars[perm[i]];
    memcpy(stars, tmp, sizeof(orc_star) * (size_t)n);
    free(keys); free(skeys); free(perm); free(tmp);
    double t2 = now_s();

    /* main.cpp:208-226 -- root cube on the AABB centre, half the max span. */
    oct_tree t;
    memset(&t, 0, sizeof(t));
    t.leaf_cap = leaf_cap;
    t.cap_nodes = 1024;
    t.nodes = (oct_node*)malloc(sizeof(oct_node) * (size_t)t.cap_nodes);
    t.cap_slots = 1024 * (int64_t)leaf_cap;
    t.leaf_idx = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)t.cap_slots);
    t.nnodes = 1;
    float span = fmaxf(hi[0] - lo[0], fmaxf(hi[1] - lo[1], hi[2] - lo[2]));
    t.nodes[0].cx = (lo[0] + hi[0]) * 0.5f;
    t.nodes[0].cy = (lo[1] + hi[1]) * 0.5f;
    t.nodes[0].cz = (lo[2] + hi[2]) * 0.5f;
    t.nodes[0].hw = span * 0.5f;
    t.nodes[0].is_leaf = 1;
    t.nodes[0].count = 0;
    t.nodes[0].first_child = -1;
    t.nodes[0].start = oct_slots(&t);
    for (int64_t i = 0; i < n && !t.failed; ++i) oct_insert(&t, 0, (uint32_t)i, stars, 0);
    double t3 = now_s();
    if (t.failed) {
        free(t.nodes); free(t.leaf_idx);
        return -1;
    }

    /* main.cpp:230 */
    float* nm = (float*)malloc(sizeof(float) * (size_t)t.nnodes);
    float* ncx = (float*)malloc(sizeof(float) * (size_t)t.nnodes);
    float* ncy = (float*)malloc(sizeof(float) * (size_t)t.nnodes);
    float* ncz = (float*)malloc(sizeof(float) * (size_t)t.nnodes);
    oct_mass(&t, 0, stars, nm, ncx, ncy, ncz);
    double t4 = now_s();

    /* main.cpp:233-240 */
    float* acc = (float*)malloc(sizeof(float) * 3 * (size_t)n);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad)
    for (int64_t i = 0; i < n; ++i) {
        if (oct_query_full(&t, stars, (uint32_t)i, nm, ncx, ncy, ncz, G, eps2, theta,
                           &acc[3 * i], &acc[3 * i + 1], &acc[3 * i + 2]) != 0)
            bad = 1;
    }
    double t5 = now_s();

    /* main.cpp:243-252 */
    if (do_integrate) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            stars[i].vx += acc[3 * i] * dt;
            stars[i].vy += acc[3 * i + 1] * dt;
            stars[i].vz += acc[3 * i + 2] * dt;
            stars[i].x += stars[i].vx * dt;
            stars[i].y += stars[i].vy * dt;
            stars[i].z += stars[i].vz * dt;
        }
    }
    double t6 = now_s();
    if (acc_out) memcpy(acc_out, acc, sizeof(float) * 3 * (size_t)n);
    if (phase_s) {
        phase_s[0] = t1 - t0; phase_s[1] = t2 - t1; phase_s[2] = t3 - t2;
        phase_s[3] = t4 - t3; phase_s[4] = t5 - t4; phase_s[5] = t6 - t5;
    }
    free(acc); free(nm); free(ncx); free(ncy); free(ncz);
    free(t.nodes); free(t.leaf_idx);
    return bad ? -1 : 0;
}
