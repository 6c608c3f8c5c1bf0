/*
 * gcs_oracle.h -- CPU oracle for the per-frame N-body step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; it is never linked into, loaded by, or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may use it.
 *
 * Every function restates one piece of the reference
 * (lakshya076/Galaxy-Collision-Simulator) in plain C; each definition cites the
 * reference file:line it follows.  All arrays are caller-owned.
 */
#ifndef GCS_ORACLE_H
#define GCS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Morton (morton.h:11-28) ------------------------------------------ */
uint64_t orc_expand_bits(uint32_t v);
uint32_t orc_f2u_sat(float f);
uint64_t orc_morton(float x, float y, float z, float mn, float mx);
void orc_morton_n(const float* x, const float* y, const float* z, int64_t n,
                  float mn, float mx, uint64_t* keys);
/* main.cpp:152-166: one scalar cube [min,max] over all three axes. */
void orc_bounds_cube(const float* x, const float* y, const float* z, int64_t n,
                     float* mn, float* mx);

/* ---- Sort (cub::DeviceRadixSort::SortPairs semantics, cuda_engine.cu:397) */
void orc_stable_sort_pairs(const uint64_t* keys, int64_t n, uint64_t* keys_out,
                           int32_t* perm_out);

/* ---- Karras 2012 radix tree (karras_bvh.cuh:23-131) -------------------- */
void orc_karras_build(const uint64_t* skeys, int32_t n, int32_t* left,
                      int32_t* right, int32_t* parent);
/* Same tree from the definition (highest differing bit of the 96-bit key
 * (key, index)), by recursive splitting.  Returns 0 on success. */
int orc_radix_tree_definition(const uint64_t* skeys, int32_t n, int32_t* left,
                              int32_t* right, int32_t* parent);
/* Leaf range [first,last] and split of every node (from left/right). */
void orc_node_ranges(int32_t n, const int32_t* left, const int32_t* right,
                     int32_t* first, int32_t* last);

/* ---- Bottom-up aggregation (karras_bvh.cuh:134-225) ---------------------
 * com4: (2n-1)*4 floats (com x,y,z, mass); width: 2n-1; bmin/bmax: (2n-1)*3;
 * crit: (2n-1) opening radius squared, (w*inv_theta)^2, or with offset != 0
 * (w*inv_theta + |com - box centre|)^2; 0 for leaves.  quad (optional, (2n-1)*7): traceless quadrupole moment
 * about the COM, components xx, yy, zz, xy, xz, yz, then the trace T of the
 * raw second moment sum m |x - COM|^2 (0 for leaves). */
void orc_aggregate(int32_t n, const int32_t* left, const int32_t* right,
                   const float* x, const float* y, const float* z,
                   const float* m, float inv_theta, int32_t offset, float* com4,
                   float* width, float* bmin, float* bmax, float* crit, float* quad);

/* ---- Reference GPU traversal semantics (cuda_engine.cu:91-161) ---------
 * Per-particle DFS from root 0; softened distance; internal accepted when
 * width^2 < d^2+eps^2; leaves skipped when d^2+eps^2 <= 1.01*eps^2.
 * Decisions in float exactly as written; sums in double.  acc: n*4 doubles
 * (ax, ay, az, 0); msum and mskip (optional, n doubles each) receive the
 * mass each target interacted with and the mass of the leaves it skipped by
 * the distance rule.  Returns the number of interactions. */
int64_t orc_traverse_reference(int32_t n, const float* x, const float* y,
                               const float* z, const int32_t* left,
                               const int32_t* right, const float* com4,
                               const float* width, float G, float eps2,
                               double* acc, double* msum, double* mskip);

/* ---- Group traversal with an opening radius (the MI355X kernel's rule) --
 * Groups are runs of `group` consecutive sorted particles.  A node is taken
 * as a monopole when the squared distance from its COM to the group's
 * bounding box exceeds crit[node]; otherwise it is summed directly when it
 * holds <= leaf particles, else opened.  Index self-exclusion.  acc: n*4
 * doubles (ax, ay, az, phi).  With quad non-null accepted nodes add their
 * quadrupole term of the softened potential, whose second-order part is
 * (d.Q.d - T eps^2)/(2 R^5), R^2 = |d|^2 + eps^2.  counts[0] += node (multipole) interactions,
 * counts[1] += particle-particle interactions (per target particle, self
 * excluded);
 * msum (optional, n doubles) receives the mass each target interacted with
 * (the total mass less its own when every particle is counted once); aabs
 * (optional, n doubles) the sum of the magnitudes of the acceleration terms,
 * the scale of the rounding error of any summation order; gcost
 * (optional, 2 per group) the node and particle list lengths of each group.
 * far2: an accepted node whose box distance squared also exceeds
 * far2 * crit[node] (float product) is taken as a plain softened monopole,
 * its quadrupole and trace left out; counts[2] += those interactions (they
 * are in counts[0] too); +inf: none.
 * kappa2: an accepted node whose COM lies farther from the group's box
 * centre c than sqrt(kappa2) times the box's half-diagonal (squares and
 * products in float) enters a second-order local expansion at c instead
 * (the quadrupole field at c plus the monopole's first and second
 * derivatives), which every target of the group evaluates at its offset from
 * c; counts[3] += 1 per such (group, node) pair, counts[4] += the group's
 * targets (not in counts[0]); +inf: none.  counts has 5 entries. */
void orc_traverse_group(int32_t n, const float* x, const float* y,
                        const float* z, const float* m, const int32_t* left,
                        const int32_t* right, const int32_t* first,
                        const int32_t* last, const float* com4,
                        const float* crit, const float* quad, int32_t group,
                        int32_t leaf, float G, float eps2, double* acc,
                        int64_t* counts, double* msum, double* aabs,
                        int64_t* gcost, int32_t ngl, const int32_t* glist,
                        float far2, float kappa2);
/* With glist non-null the groups are the ngl (first, count) pairs of glist
 * instead of runs of `group` particles.
 * Tree groups: the maximal subtrees of at most gmax particles (every node of
 * count <= gmax whose parent holds more), in particle order, as (first,
 * count) pairs.  Returns their number. */
int32_t orc_tree_groups(int32_t n, const int32_t* left, const int32_t* right,
                        const int32_t* first, const int32_t* last, int32_t gmax,
                        int32_t* out);

/* ---- FP64 direct sum (index self-exclusion) ------------------------------
 * Targets tgt[0..nt) (indices into the n sources).  acc: nt*4 doubles
 * (ax, ay, az, phi). */
void orc_direct_sum_f64(int64_t n, const double* x, const double* y,
                        const double* z, const double* m, int64_t nt,
                        const int64_t* tgt, double G, double eps2,
                        double* acc);

/* ---- Kick-drift and colour (cuda_engine.cu:164-216) -------------------- */
void orc_integrate_reference(int64_t n, float* x, float* y, float* z,
                             float* vx, float* vy, float* vz, uint8_t* r,
                             uint8_t* g, uint8_t* b, const uint8_t* type,
                             const float* ax, const float* ay, const float* az,
                             float dt);

/* ---- CPU octree step (engine.cpp:5-127, gravity_aggregator.cpp:8-83,
 *      main.cpp:152-253) -----------------------------------------------
 * Stars are AoS records of 32 bytes laid out as types.h:5-12.  One full
 * frame: bounds, Morton sort (stable), octree insert with `leaf_cap`,
 * mass/COM, Barnes-Hut query at `theta`, kick-drift.  When `acc_out` is
 * non-null the accelerations (in post-sort order) are written there as n*3
 * floats.  phase_s[6] receives seconds for bounds, sort, insert, mass, query,
 * integrate.  Returns 0, or -1 when the octree cannot be built (more than
 * leaf_cap coincident points) or a query stack overflows. */
typedef struct orc_star {
    float x, y, z, vx, vy, vz, mass;
    uint8_t type, r, g, b;
} orc_star;

int orc_octree_step(orc_star* stars, int64_t n, float G, float eps2, float dt,
                    float theta, int32_t leaf_cap, int32_t do_integrate,
                    float* acc_out, double* phase_s);

#ifdef __cplusplus
}
#endif
#endif
