/*
 * gcs.h -- C ABI of the MI355X (gfx950) per-frame N-body step.
 *
 * One engine holds the device-resident state of one or more independent
 * collisions ("instances").  Each step runs, on the engine's HIP stream:
 *
 *   cube bounds -> 63-bit Morton keys -> stable LSD radix sort -> permute
 *   -> Karras radix tree per instance -> bottom-up mass/COM/AABB
 *   -> Barnes-Hut traversal -> kick-drift (+ colours, + diagnostics)
 *
 * The entry points replace the reference's cuda_engine.h (cuda_engine.h:5-32)
 * and the kernels behind it (cuda_engine.cu, karras_bvh.cuh, morton.h); the
 * ten reference-named C++ functions are provided over this ABI by
 * gcs_cuda_engine.h.  All functions return GCS_OK (0) or a negative error
 * code; gcs_last_error() gives the message of the calling thread's last
 * failure.  No function throws.  An engine is not thread-safe: use one engine
 * per host thread (or serialise calls).  Pointers marked "host" are ordinary
 * host memory; the library copies through its own staging buffers.
 */
#ifndef GCS_H
#define GCS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Records -------------------------------------------------------------
 * Byte-identical to the reference's Star (types.h:5-12, 32 bytes) and
 * PlaybackStar (types.h:14-18, 16 bytes, the simulation.bin record). */
typedef struct gcs_star {
    float x, y, z;
    float vx, vy, vz;
    float mass;
    uint8_t type, r, g, b;
} gcs_star;

typedef struct gcs_playback_star {
    float x, y, z;
    uint8_t r, g, b;
    uint8_t padding; /* carries `type` (main.cpp:266) */
} gcs_playback_star;

/* ---- Status codes ------------------------------------------------------ */
enum {
    GCS_OK = 0,
    GCS_ERR_INVALID = -1,       /* bad argument */
    GCS_ERR_NO_DEVICE = -2,     /* no HIP device */
    GCS_ERR_OOM = -3,           /* device or host allocation failed */
    GCS_ERR_HIP = -4,           /* a HIP runtime call failed */
    GCS_ERR_CAPACITY = -5,      /* request exceeds the engine's capacity */
    GCS_ERR_STACK = -6,         /* a traversal stack overflowed */
    GCS_ERR_NONFINITE = -7,     /* a non-finite acceleration was produced */
    GCS_ERR_NOT_FOUND = -8      /* unknown instance id */
};

/* ---- Opening criteria ----------------------------------------------------
 * GCS_MAC_REFERENCE  the reference's GPU traversal (cuda_engine.cu:91-161):
 *                    one particle per thread, DFS from the root, internal
 *                    node accepted when width^2 < |com-p|^2 + eps^2 (theta is
 *                    ignored: the reference's rule is theta = 1 on the
 *                    softened distance), leaves skipped when within 0.1*eps.
 * GCS_MAC_GROUP      64 consecutive sorted particles share one traversal; a
 *                    node is taken as a multipole (mass and quadrupole) for
 *                    the group when the distance from its COM to the group's
 *                    bounding box exceeds width/theta + |COM - box centre|;
 *                    otherwise a node of at most leaf_size particles is
 *                    summed directly and a larger one is opened; index
 *                    self-exclusion.  Needs eps2 > 0 and theta <= 1. */
enum { GCS_MAC_REFERENCE = 0, GCS_MAC_GROUP = 1 };

/* Position precision: plain FP32, or a compensated FP32 hi/lo pair whose
 * drift carries are formed in FP64. */
enum { GCS_PREC_PLAIN = 0, GCS_PREC_COMPENSATED = 1 };

typedef struct gcs_params {
    float G;            /* gravitational constant */
    float eps2;         /* Plummer softening squared (epsilon_sq) */
    float dt;           /* time step */
    float theta;        /* opening angle (GCS_MAC_GROUP) */
    int32_t mac;        /* GCS_MAC_* */
    int32_t precision;  /* GCS_PREC_* */
    int32_t leaf_size;  /* direct-sum bucket size, 1..64 (GCS_MAC_GROUP) */
    int32_t block_size; /* traversal threads per block: 64, 128, 256 */
    int32_t stack_depth;/* per-thread stack (GCS_MAC_REFERENCE), 16..1024 */
    int32_t use_bounds; /* 1: use global_min/global_max below for every
                           instance (cuda_physics_step semantics);
                           0: per-instance cube bounds from the device */
    float global_min, global_max;
    int32_t colours;    /* 1: update r,g,b from speed (cuda_engine.cu:185-215) */
    int32_t diagnostics;/* 1: FP64 energy/momentum per instance each step */
    int32_t use_graph;  /* 1: replay steps from a captured hipGraph */
    int32_t waves_per_simd; /* resident traversal waves per SIMD, 1..4 (GCS_MAC_GROUP) */
    float dt_eta;       /* > 0: adaptive step per instance, min(dt, sqrt(2 dt_eta eps /
                           max|a|)) with eps = sqrt(eps2); 0: fixed dt */
    int32_t group_size; /* largest traversal group (maximal subtree), 1..192
                           (GCS_MAC_GROUP) */
    int32_t open_offset;/* GCS_MAC_GROUP opening radius: 0 = w/theta, the
                           reference's theta test taken from the group's box;
                           1 = w/theta + |COM - box centre| (stricter) */
    int32_t team;       /* GCS_MAC_GROUP waves per group: 0 or 1 = one wave walks
                           and evaluates a group; 2, 4 or 8 = a block of that many
                           waves walks it together and splits its targets
                           (block_size is then 64 x team) */
    float mono_far;     /* GCS_MAC_GROUP: 0 = every accepted node is taken as a
                           quadrupole; >= 1 = a node whose distance from the group's
                           box exceeds mono_far times its opening radius is taken
                           as a softened monopole (its COM and mass), the rest as
                           quadrupoles */
    float local_kappa;  /* GCS_MAC_GROUP: 0 = off; >= 1 = an accepted node whose COM
                           lies farther than local_kappa times the half-diagonal of
                           the group's box from the box centre is not evaluated per
                           target: its softened quadrupole field at the centre and
                           its monopole's first and second derivatives enter a
                           second-order local expansion that every target of the
                           group evaluates once (a cell-cell far field) */
    int32_t instance_pass; /* aggregation of the nodes above the subtrees of at most
                           1,024 particles: 1 = one block per instance; 2 = for
                           instances of at least 32,768 particles in two levels,
                           the maximal subtrees of fewer than 32,768 particles one
                           block each, then the rest one block per instance; 0 =
                           automatic: two levels when an instance has at least
                           32,768 particles and the instances number at most half
                           the CUs.  The node records are the same either way */
    int32_t sort;          /* key sort of a step: 1 = LSD radix sort (8 passes); 2 =
                           bucket sort (splitters from the keys at every 1,536th
                           place of the current order, buckets sorted in LDS), for
                           keys in nearly sorted order; 0 = automatic: buckets for
                           the instances stored in the last step's sorted order,
                           radix for instances loaded or joined since.  The result
                           is the same stable sort either way */
} gcs_params;

/* Defaults: G=1, eps2=0.0025, dt=0.01, theta=0.5, GCS_MAC_GROUP,
 * GCS_PREC_PLAIN, leaf 8, block 256, stack 256, device bounds, colours on,
 * diagnostics off, graph on, 3 traversal waves per SIMD, fixed dt, groups of
 * at most 64. */
void gcs_params_default(gcs_params* p);

typedef struct gcs_engine gcs_engine;

/* ---- Lifetime (cuda_engine.cu:219-285 cuda_init, :303-340 cuda_cleanup) --
 * capacity_particles: total particles over all instances, 1 .. 2^29 (larger
 * capacities return GCS_ERR_INVALID); capacity_instances: the most instances
 * held at once.  device: HIP ordinal.  The engine creates its own stream
 * unless gcs_set_stream is called. */
int gcs_create(int64_t capacity_particles, int32_t capacity_instances,
               int32_t device, gcs_engine** out);
int gcs_destroy(gcs_engine* e);
const char* gcs_last_error(void);
/* hip_stream: a hipStream_t, or NULL for the engine's own stream. */
int gcs_set_stream(gcs_engine* e, void* hip_stream);
/* Bytes of device memory an engine of this capacity allocates. */
int64_t gcs_device_bytes(int64_t capacity_particles, int32_t capacity_instances);

/* ---- State in / out (cuda_engine.cu:342-352 upload, :441-449 read back) --
 * gcs_load replaces the whole state with k instances whose particles are
 * consecutive in `stars` (host), counts[i] each.  Instances keep their
 * order; ids 0..k-1 are assigned. */
int gcs_load(gcs_engine* e, const gcs_star* stars, const int32_t* counts,
             int32_t k);
/* Adds one instance at the end (continuous batching); *id receives its id. */
int gcs_instance_add(gcs_engine* e, const gcs_star* stars, int32_t n,
                     int32_t* id);
/* Removes an instance; the others keep their ids and order. */
int gcs_instance_remove(gcs_engine* e, int32_t id);
int gcs_instance_count(gcs_engine* e, int32_t* k, int64_t* n_total);
/* Particle count and the current offset of an instance in the state. */
int gcs_instance_info(gcs_engine* e, int32_t id, int32_t* n, int64_t* offset);
/* Reads every particle (host, n_total records), instance by instance, each
 * in the device's current (Morton-sorted) order. */
int gcs_read_stars(gcs_engine* e, gcs_star* out);
int gcs_read_instance(gcs_engine* e, int32_t id, gcs_star* out);
/* Compensated low words (x_lo, y_lo, z_lo per particle, host n_total*3). */
int gcs_read_position_lo(gcs_engine* e, float* out);
/* Checkpoint / resume of one instance (no counterpart in the reference:
 * job preemption and restart).  A checkpoint holds the instance's records in
 * its current order and the low parts of its compensated positions; an
 * instance restored from it (appended to any engine) continues bit for bit
 * as the original would.  gcs_checkpoint_bytes gives the buffer size. */
int64_t gcs_checkpoint_bytes(int32_t n);
int gcs_checkpoint(gcs_engine* e, int32_t id, void* buf, int64_t cap);
int gcs_restore(gcs_engine* e, const void* buf, int64_t bytes, int32_t* id);
/* PlaybackStar frame of all particles (simulation.bin record, main.cpp:266). */
int gcs_read_playback(gcs_engine* e, gcs_playback_star* out);
/* Bake (main.cpp:133-297): `frames` steps, each frame's PlaybackStar records
 * of every particle (device order, as gcs_read_playback gives them after that
 * step) appended to the file at `path` -- the reference's headerless
 * simulation.bin.  Records are packed on the device and copied 16 bytes per
 * particle into pinned double buffers on a second stream while the next
 * frame steps; writer threads append them in order.  *seconds (optional):
 * wall time of the frames, the last write included. */
int gcs_bake(gcs_engine* e, const gcs_params* p, int32_t frames, const char* path,
             double* seconds);

/* ---- Stepping (cuda_engine.cu:379-451) -----------------------------------
 * Advances every instance by nsteps full steps.  Returns when the work is
 * queued on the stream; gcs_sync waits and reports device-side errors. */
int gcs_step(gcs_engine* e, const gcs_params* p, int32_t nsteps);
int gcs_sync(gcs_engine* e);
/* Runs bounds, Morton, sort, permute, tree build, aggregation and traversal
 * without integrating; accelerations are then readable. */
int gcs_compute_accels(gcs_engine* e, const gcs_params* p);
/* Autotuner (no counterpart in the reference).  Times accelerations-only
 * passes of the current state (positions are not advanced; the state is
 * re-sorted as every step does) for each candidate setting and writes the
 * fastest into *best (base with those fields replaced) and its mean time per
 * pass in ms into *best_ms.  Group traversal: group_size {48, 64, 96, 128} x
 * leaf_size {4, 8, 12, 16} x (block_size {64, 128, 256} with one wave per
 * group, or teams of 2 or 4 waves) x waves_per_simd {2, 3, 4}; reference
 * traversal:
 * stack_depth {64, 128, 256, 1024}, skipping depths that overflow.  reps
 * timed passes per candidate after one warm-up pass, 1 <= reps <= 100. */
int gcs_autotune(gcs_engine* e, const gcs_params* base, int32_t reps,
                 gcs_params* best, float* best_ms);
/* The autotuner keeps each result per (traversal kind, theta, opening form,
 * precision, floor(log2 particles), floor(log2 instances), mono_far,
 * local_kappa), process-wide:
 * gcs_autotune returns a kept result without timing anything (its best_ms is
 * the time measured when it was found).  _clear forgets them; _save writes
 * them as a text table, _load adds a saved table (an earlier run's tuning). */
int gcs_autotune_cache_clear(void);
int gcs_autotune_cache_save(const char* path);
int gcs_autotune_cache_load(const char* path);
/* Host n_total*4 floats: ax, ay, az, potential (potential only for
 * GCS_MAC_GROUP), in current particle order. */
int gcs_read_accels(gcs_engine* e, float* out);

/* ---- Inspection for parity tests -----------------------------------------
 * Keys as encoded this step (before the sort, in pre-sort particle order),
 * sorted keys, and the sort permutation (sorted position -> pre-sort
 * position), each n_total entries, instance-local values for the
 * permutation. */
int gcs_read_keys(gcs_engine* e, uint64_t* keys_unsorted, uint64_t* keys_sorted,
                  int32_t* perm);
/* Tree of one instance in the reference's numbering (karras_bvh.cuh:114-124:
 * internal nodes 0..n-2, leaves n-1..2n-2), 2n-1 entries each:
 * left/right children (-1 for leaves), parent (-1 for the root), com4
 * (com x,y,z, mass), width (max AABB extent), bmin/bmax (3 floats).  Leaf
 * entries are taken from the current positions: after gcs_compute_accels
 * they are the positions the tree was built from, after gcs_step the
 * drifted ones (the internal nodes keep the built tree's values). */
int gcs_read_tree(gcs_engine* e, int32_t id, int32_t* left, int32_t* right,
                  int32_t* parent, float* com4, float* width, float* bmin,
                  float* bmax);
/* Root record of instance `id` in the last built tree: com x, y, z and mass
 * (4 floats), one 16-byte device copy.  Replaces cuda_get_center_of_mass's
 * copy of node 0 (cuda_engine.cu:365-376); a one-particle instance gives its
 * leaf, the position the tree was built from. */
int gcs_read_root(gcs_engine* e, int32_t id, float* com4);
/* Cube bounds used for instance `id` in the last step. */
int gcs_read_bounds(gcs_engine* e, int32_t id, float* mn, float* mx);
/* Interaction counts of the last traversal: [0] particle-node interactions
 * (quadrupole or monopole), [1] particle-particle interactions, summed over
 * targets. */
int gcs_read_interactions(gcs_engine* e, int64_t* out2);
/* Of [0] above, the nodes taken as monopoles (mono_far). */
int gcs_read_monopole_interactions(gcs_engine* e, int64_t* out);
/* All counts of the last group traversal: [0] node interactions evaluated
 * per target, [1] particle interactions, [2] of [0] the monopoles, [3]
 * (group, node) pairs taken through local expansions (local_kappa), [4] the
 * target interactions those stand for (not in [0]). */
int gcs_read_interaction_counts(gcs_engine* e, int64_t* out5);
/* FP64 diagnostics of instance `id` from the last step run with
 * diagnostics=1: kinetic energy, potential energy, momentum x, y, z, all at
 * the force time (velocities re-synchronised by half a kick). */
int gcs_read_diagnostics(gcs_engine* e, int32_t id, double* out5);
/* Largest |a| over instance `id` in the last step (the value the reference
 * reduces at cuda_engine.cu:157-160), and the instance's simulated time (the
 * sum of its steps' dt since it was loaded or added). */
int gcs_read_max_accel(gcs_engine* e, int32_t id, float* amax);
int gcs_read_sim_time(gcs_engine* e, int32_t id, double* t);
/* Engine facts: [0] compute units, [1] persistent traversal blocks, [2]
 * traversal wave slots, [3] bytes of per-wave traversal memory, [4] particle
 * capacity, [5] instance capacity, [6] traversal groups, [7] sort tiles. */
int gcs_engine_info(gcs_engine* e, int64_t* out8);
/* Milliseconds of the last step's stages on the device (needs
 * use_graph = 0): [0] bounds+morton, [1] sort, [2] permute, [3] build,
 * [4] aggregate, [5] traverse, [6] integrate.  The permute runs in the
 * build's launch, so [3] holds both and [2] only the event overhead. */
int gcs_read_stage_ms(gcs_engine* e, float* out7);

/* ---- Stand-alone stage entry points (host buffers in and out) ----------- */
int gcs_morton_keys(const float* x, const float* y, const float* z, int64_t n,
                    float mn, float mx, uint64_t* keys);
/* Stable ascending sort of 64-bit keys with 32-bit payload (the payload
 * starts as 0..n-1): the semantics of cub::DeviceRadixSort::SortPairs
 * (cuda_engine.cu:397) on all 64 bits. */
int gcs_sort_pairs(const uint64_t* keys, int64_t n, uint64_t* keys_out,
                   int32_t* perm_out);
/* Karras 2012 radix tree over sorted keys (karras_bvh.cuh:73-131). */
int gcs_karras_tree(const uint64_t* sorted_keys, int32_t n, int32_t* left,
                    int32_t* right, int32_t* parent);

/* Allocation-free variant of gcs_sort_pairs.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same stable ascending sort of n 64-bit keys with the payload 0..n-1,
 * on device buffers of the current device: keys (n), keys_out (n; may equal
 * keys), perm_out (n int32, must not overlap the others).  Queued on
 * hip_stream (a hipStream_t, NULL = the null stream) and returns at once;
 * results are ready when that stream is.  No device or host memory is
 * allocated: the working memory is a static device array of 53 MiB per GPU,
 * held until the process ends, shared by every call on that GPU -- calls
 * that overlap on one GPU race.  n above GCS_NOALLOC_MAX_KEYS returns
 * GCS_ERR_CAPACITY and queues nothing. */
enum { GCS_NOALLOC_MAX_KEYS = 1 << 22 };
int gcs_sort_pairs_noalloc(const uint64_t* keys, int64_t n, uint64_t* keys_out,
                           int32_t* perm_out, void* hip_stream);

/* ---- Initial conditions --------------------------------------------------
 * Two-galaxy collision: per galaxy a Miyamoto-Nagai disc (a=3, b=0.3,
 * M=1), a Hernquist bulge (a=0.5, M=0.25) and a Hernquist dark halo (a=12,
 * M=5) truncated at r<60 (the bulge too), split 70/10/20 of each galaxy's
 * count (disc = round(0.7 n), bulge = round(0.1 n), halo = the rest), equal
 * mass per particle within a component, circular velocities from the
 * enclosed mass plus a 5% Gaussian dispersion.  Galaxy A at rest at the
 * origin; galaxy B tilted about x, offset and moving as given. */
typedef struct gcs_collision_spec {
    int32_t n_a, n_b;     /* particles per galaxy */
    float offset[3];      /* position of galaxy B */
    float velocity[3];    /* velocity of galaxy B */
    float tilt_deg;       /* rotation of galaxy B about the x axis */
    uint64_t seed;
} gcs_collision_spec;

/* The single-collision configuration: n per galaxy, B at (40,10,0) moving
 * (-0.5,0,0), tilted 30 degrees. */
void gcs_spec_default(int32_t n_per_galaxy, uint64_t seed, gcs_collision_spec* s);
/* Randomised collision k of a stream: total N log-uniform in [n_min, n_max]
 * split 50/50, |offset| uniform in [20,60] in a random direction, approach
 * speed uniform in [0.2,0.8] toward galaxy A, tilt uniform in [0,90]. */
void gcs_spec_random(uint64_t seed, int64_t k, int32_t n_min, int32_t n_max,
                     gcs_collision_spec* s);
/* Writes n_a + n_b records (galaxy A then B; each halo, disc, bulge). */
int gcs_generate(const gcs_collision_spec* s, gcs_star* out);
/* Generates k collisions on the device straight into the engine's state, as
 * gcs_load would load gcs_generate's records (instance ids 0..k-1, instance s
 * = specs[s] with n_a + n_b particles).  The records equal gcs_generate's
 * except where the device's log, cos, sin or pow rounds differently in the
 * last bit (about one float value in 10^8).  No counterpart in the reference
 * (its generator is host code, generator.cpp:3-196). */
int gcs_load_generated(gcs_engine* e, const gcs_collision_spec* specs, int32_t k);
/* Adds one generated collision at the end, as gcs_instance_add would add
 * gcs_generate's records; only the new instance is generated and sorted.  No
 * counterpart in the reference; the job server admits jobs this way when it
 * has no generator threads. */
int gcs_instance_add_generated(gcs_engine* e, const gcs_collision_spec* spec, int32_t* id);

/* ---- Served form: continuous batching -----------------------------------
 * A server owns one engine.  Jobs (a randomised collision of n particles
 * advanced `steps` steps) arrive at given times; worker threads generate
 * their initial conditions on arrival; the stepping loop admits generated
 * jobs in arrival order while the resident particles fit the HBM state
 * budget and the instance cap, advances every resident job together (one
 * launch sequence per step, hipGraph-replayed) in rounds that end when the
 * next job finishes or `steps_per_round` steps have run, and retires finished
 * jobs.  With max_jobs = 1 the same loop runs jobs one at a time through the
 * plain path. */
typedef struct gcs_job {
    int64_t id;
    int32_t n;          /* particles (split 50/50 between the two galaxies) */
    int32_t steps;
    uint64_t seed;      /* gcs_spec_random(seed, id, n, n) fixes the scenario */
    double arrival_s;   /* server-clock seconds (the clock starts with the first
                           gcs_server_run call and runs on across calls) */
} gcs_job;

typedef struct gcs_job_result {
    int64_t id;
    int32_t n, steps;
    double arrival_s, admit_s, finish_s; /* seconds after the start */
} gcs_job_result;

typedef struct gcs_server_config {
    int64_t hbm_budget_bytes;     /* device memory for the engine (state + trees) */
    int64_t max_resident;         /* resident-particle cap, 0 = what the budget holds */
    int32_t max_jobs;             /* resident-job cap (1 = one at a time) */
    int32_t steps_per_round;      /* steps between admission checks */
    int32_t gen_threads;          /* initial-condition worker threads; 0: each job is
                                     generated on the device when admitted */
    int32_t keep_final;           /* 1: keep every finished job's final state */
    gcs_params params;
} gcs_server_config;

typedef struct gcs_server gcs_server;

int gcs_server_create(const gcs_server_config* c, gcs_server** out);
int gcs_server_destroy(gcs_server* s);
/* Adds jobs to the arrival schedule (before or between runs). */
int gcs_server_submit(gcs_server* s, const gcs_job* jobs, int32_t n);
/* Runs the loop for `duration_s` seconds of wall time, arrivals following the
 * server clock, or (duration_s <= 0) until every submitted job has finished.
 * A later call continues where the last stopped (same clock, same jobs). */
int gcs_server_run(gcs_server* s, double duration_s);
/* Finished jobs in completion order. */
int gcs_server_results(gcs_server* s, gcs_job_result* out, int32_t cap, int32_t* n_out);
/* Up to n of: [0] particle-steps advanced, [1] run seconds, [2] jobs
 * finished, [3] jobs admitted, [4] jobs waiting at the end, [5] mean resident
 * particles per step, [6] rounds, [7] resident particle cap, [8] seconds
 * spent stepping (device work and its waits), [9] particle-steps and [10]
 * seconds from the first round that found the resident set full (admission
 * held back by the cap) to the end -- the steady state of an overloaded
 * stream; 0 if it never filled. */
int gcs_server_stats(gcs_server* s, double* out, int32_t n);
/* Final state of a finished job (keep_final = 1), n records. */
int gcs_server_final(gcs_server* s, int64_t id, gcs_star* out, int32_t n);

#ifdef __cplusplus
}
#endif
#endif
