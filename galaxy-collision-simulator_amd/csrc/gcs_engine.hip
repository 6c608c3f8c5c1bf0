// gcs_engine.hip -- host engine and C ABI of the MI355X per-frame N-body step.
//
// Replaces the reference's cuda_engine.cu:219-451 (init, upload, the
// per-frame step and read-back) with an engine object that owns its device
// buffers and HIP stream, holds any number of independent instances back to
// back, and queues each step without host synchronisation (optionally as a
// replayed hipGraph).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <chrono>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gcs.h"
#include "gcs_error.hpp"
#include "gcs_common.hpp"
#include "gcs_kernels.hpp"
#include "gcs_sort.hpp"
#include "gcs_traverse.hpp"
#include "gcs_gen_device.hpp"

using namespace gcs;

static_assert(sizeof(gcs_star) == 32, "gcs_star must match the reference Star (types.h:5-12)");
static_assert(sizeof(gcs_playback_star) == 16, "gcs_playback_star must match PlaybackStar");
static_assert(kRadix == kTileThreads, "one digit per scatter thread");

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

}  // namespace

int gcs_detail::set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }

namespace {

#define HIPCHK(expr)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? GCS_ERR_OOM : GCS_ERR_HIP,           \
                        std::string(#expr) + ": " + hipGetErrorString(e_));              \
    } while (0)

#define LAUNCHCHK()                                                                      \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess)                                                            \
            return fail(GCS_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
};

struct Instance {
    int32_t id;
    int32_t n;
    // Its particles are stored in the order of the last step's sort, so its
    // next keys are nearly sorted (the bucket sort applies).
    bool ordered = false;
};

size_t max_tiles(int64_t cap_n, int32_t cap_k) { return (size_t)(cap_n / kTile) + (size_t)cap_k + 1; }
size_t hist_entries(int64_t cap_n, int32_t cap_k) { return (size_t)kRadix * max_tiles(cap_n, cap_k); }
size_t max_buckets(int64_t cap_n, int32_t cap_k) { return (size_t)(cap_n / kBucketKeys) + (size_t)cap_k + 1; }
// Sort-histogram chunks: at most one partial chunk per instance.
size_t max_chunks(int64_t cap_n, int32_t cap_k)
{
    return max_tiles(cap_n, cap_k) / kChunkTiles + (size_t)cap_k + 1;
}
size_t max_gchunks(int64_t cap_n, int32_t cap_k)
{
    return (4 * max_tiles(cap_n, cap_k) + kGroupChunk - 1) / kGroupChunk;
}

}  // namespace

struct gcs_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int64_t cap_n = 0;
    int32_t cap_k = 0;
    int num_cu = 256;
    int trav_blocks_cap = 0;  // persistent traversal blocks resident at once
    int trav_slots = 0;       // traversal waves the per-wave memory covers

    // state (double-buffered by the permute)
    float4* posm[2] = {nullptr, nullptr};
    float4* vel[2] = {nullptr, nullptr};
    float4* plo[2] = {nullptr, nullptr};
    int cur = 0;
    float4* acc = nullptr;
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* vals[2] = {nullptr, nullptr};
    uint64_t* keys_raw = nullptr;
    // tree (node: 2 float4 per node -- COM/mass, opening value/first/last/split)
    float4* node = nullptr;
    float4* nodeQ = nullptr;
    float4* boxLo = nullptr;
    float4* boxHi = nullptr;
    int* parent = nullptr;
    int* flag = nullptr;
    // tables
    TileInfo* tiles = nullptr;
    TileInfo* stiles = nullptr;  // radix-sort tiles (at most stile_keys keys)
    int stile_keys = kTile;      // kSortTileBig for engines of at least kSortBigMin
    SegInfo* segs = nullptr;
    GroupInfo* groups = nullptr;  // traversal groups, rebuilt on the device each step
    int* gorder = nullptr;        // their dispatch order (size classes within each XCD queue)
    uint32_t* gmark = nullptr;
    uint32_t* gq = nullptr;      // groups per quarter tile
    uint32_t* gchunk = nullptr;  // groups per kGroupChunk quarter tiles
    int* gcnt = nullptr;
    int* ngroups = nullptr;
    int* front = nullptr;   // aggregation frontier: instance s uses [2*start_s, 2*(start_s+n_s))
    int* nfront = nullptr;
    // Nodes above the aggregation subtrees, listed by the build: instance s
    // uses tlist[start_s, start_s + ntop[s]); tslot[node] = its list slot.
    int* tlist = nullptr;
    int* ntop = nullptr;
    int* tslot = nullptr;
    // Roots of the mid-level subtrees (node, instance) and their count, for
    // the two-level instance pass.
    int* mids = nullptr;
    int* nmid = nullptr;
    int* mfail = nullptr;  // per instance: a mid-level root left to the one-level pass
    // Bucket sort: buckets [bkt0[s], bkt0[s+1]) of instance s, their
    // splitters, counts, offsets and fill counters, and each key's bucket.
    int* bkt0 = nullptr;
    int* bseg = nullptr;  // per bucket: its instance
    uint64_t* bsplit = nullptr;
    uint64_t* bcand = nullptr;
    uint32_t* bcnt = nullptr;
    uint32_t* boff = nullptr;
    uint32_t* bfill = nullptr;
    uint32_t* bid = nullptr;
    std::vector<int> h_bkt0;
    // The instances the captured step graphs sort with buckets.
    int graph_bucket_segs = -1;
    // aggregation subtrees (maximal subtrees of at most kATile particles)
    uint32_t* amark = nullptr;
    uint32_t* aq = nullptr;      // subtrees per quarter tile
    uint32_t* achunk = nullptr;  // subtrees per kGroupChunk quarter tiles
    int* acnt = nullptr;
    int* aroot = nullptr;
    ATileInfo* atiles = nullptr;
    int* natiles = nullptr;
    uint32_t* hist = nullptr;   // per-tile digit counts of the current pass
    uint32_t* chist = nullptr;  // per-chunk digit counts, one block per pass
    uint32_t* seg_bounds = nullptr;
    uint32_t* bpart = nullptr;  // per quarter tile (min, max) encodings
    // bpart holds the current positions' partials (written by the last
    // k_integrate or k_bounds; any change of the composition clears it).
    bool bpart_valid = false;
    bool capturing = false;  // enqueue_step is recording a step graph
    unsigned* qctr = nullptr;
    unsigned long long* counters = nullptr;  // reference traversal: [1] interactions
    // Group traversal: node, particle and monopole interactions per wave slot
    // (3 per slot) of the last launch, which wrote the first wslots slots.
    unsigned long long* wcount = nullptr;
    int wslots = 0;
    int graph_wslots = 0;  // wslots of the captured step graphs
    unsigned* err = nullptr;
    double* diag_part = nullptr;
    double* diag_out = nullptr;
    float* amax_part = nullptr;  // per quarter tile: largest |a|
    float* dt_inst = nullptr;    // per instance: adaptive step
    double* tsim = nullptr;      // per instance: simulated time
    int* slot_mem = nullptr;  // per traversal wave: stack spill, node and particle lists
    uint8_t* staging = nullptr;
    double* mn_dev = nullptr;  // disc sampler table on the device (generated loads)

    // host mirror of the composition
    std::vector<Instance> inst;
    std::vector<SegInfo> h_segs;
    std::vector<TileInfo> h_tiles;
    std::vector<TileInfo> h_stiles;
    std::vector<int> h_stile0;  // first sort tile of each instance
    int64_t n_total = 0;
    int32_t next_id = 0;
    int nchunks = 0;  // sort-histogram chunks in use
    bool have_tree = false;
    bool capture_keys = false;
    int last_mac = GCS_MAC_GROUP;

    // graphs, one per state parity
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    gcs_params gparams;
    bool graph_valid = false;

    hipEvent_t ev[8] = {};
    bool timed = false;
};

namespace {

template <typename T>
int dalloc(T** p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(GCS_ERR_OOM, std::string("hipMalloc of ") + std::to_string(sizeof(T) * count) +
                                     " bytes: " + hipGetErrorString(e));
    }
    return GCS_OK;
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

void free_all(gcs_engine* e)
{
    for (int k = 0; k < 2; ++k) {
        dfree(e->posm[k]); dfree(e->vel[k]); dfree(e->plo[k]);
        dfree(e->keys[k]); dfree(e->vals[k]);
        if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
        e->gexec[k] = nullptr;
    }
    dfree(e->acc); dfree(e->keys_raw);
    dfree(e->node); dfree(e->nodeQ); dfree(e->boxLo); dfree(e->boxHi);
    dfree(e->parent); dfree(e->flag);
    dfree(e->tiles); dfree(e->stiles); dfree(e->segs); dfree(e->amark); dfree(e->aq); dfree(e->achunk); dfree(e->acnt); dfree(e->aroot); dfree(e->atiles); dfree(e->natiles); dfree(e->groups); dfree(e->hist); dfree(e->chist);
    dfree(e->gmark); dfree(e->gq); dfree(e->gorder); dfree(e->gchunk); dfree(e->gcnt); dfree(e->ngroups);
    dfree(e->front); dfree(e->nfront);
    dfree(e->tlist); dfree(e->ntop); dfree(e->tslot); dfree(e->mids); dfree(e->nmid); dfree(e->mfail);
    dfree(e->bkt0); dfree(e->bseg); dfree(e->bsplit); dfree(e->bcand); dfree(e->bcnt); dfree(e->boff); dfree(e->bfill); dfree(e->bid);
    dfree(e->seg_bounds); dfree(e->bpart); dfree(e->qctr); dfree(e->counters); dfree(e->wcount); dfree(e->err);
    dfree(e->diag_part); dfree(e->diag_out); dfree(e->amax_part); dfree(e->dt_inst); dfree(e->tsim); dfree(e->slot_mem); dfree(e->staging); dfree(e->mn_dev);
    for (auto& v : e->ev)
        if (v) (void)hipEventDestroy(v), v = nullptr;
}

// Rebuilds the instance, tile and group tables after the composition changed.
int rebuild_tables(gcs_engine* e)
{
    e->h_segs.clear();
    e->h_tiles.clear();
    e->h_stiles.clear();
    e->h_stile0.clear();
    int64_t n_all = 0;
    for (const Instance& in : e->inst) n_all += in.n;
    const int st_keys = n_all >= kSortBigMin ? kSortTileBig : kTile;
    e->stile_keys = st_keys;
    int start = 0, chunks = 0;
    for (size_t s = 0; s < e->inst.size(); ++s) {
        const int n = e->inst[s].n;
        SegInfo sg;
        sg.start = start;
        sg.n = n;
        sg.int_base = start - (int)s;
        sg.leaf_base = start + n - (int)s - 1;
        sg.root = 2 * start - (int)s;
        sg.tile0 = (int)e->h_tiles.size();
        sg.ntiles = (n + kTile - 1) / kTile;
        sg.chunk0 = chunks;
        chunks += (sg.ntiles + kChunkTiles - 1) / kChunkTiles;
        for (int j = 0; j < sg.ntiles; ++j) {
            TileInfo t;
            t.start = start + j * kTile;
            t.count = std::min(kTile, n - j * kTile);
            t.seg = (int)s;
            t.tile_in_seg = j;
            t.pad0 = t.pad1 = t.pad2 = t.pad3 = 0;
            e->h_tiles.push_back(t);
        }
        e->h_stile0.push_back((int)e->h_stiles.size());
        for (int j = 0; j * st_keys < n; ++j) {
            TileInfo t;
            t.start = start + j * st_keys;
            t.count = std::min(st_keys, n - j * st_keys);
            t.seg = (int)s;
            t.tile_in_seg = j;
            t.pad0 = t.pad1 = t.pad2 = t.pad3 = 0;
            e->h_stiles.push_back(t);
        }
        e->h_segs.push_back(sg);
        start += n;
    }
    e->n_total = start;
    e->nchunks = chunks;
    e->h_bkt0.assign(1, 0);
    std::vector<int> bseg;
    for (size_t s = 0; s < e->h_segs.size(); ++s) {
        const int nb = std::max(1, (e->h_segs[s].n + kBucketKeys - 1) / kBucketKeys);
        e->h_bkt0.push_back(e->h_bkt0.back() + nb);
        bseg.insert(bseg.end(), (size_t)nb, (int)s);
    }
    if (!e->h_segs.empty()) {
        HIPCHK(hipMemcpy(e->bkt0, e->h_bkt0.data(), sizeof(int) * e->h_bkt0.size(),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->bseg, bseg.data(), sizeof(int) * bseg.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->segs, e->h_segs.data(), sizeof(SegInfo) * e->h_segs.size(),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->tiles, e->h_tiles.data(), sizeof(TileInfo) * e->h_tiles.size(),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(e->stiles, e->h_stiles.data(), sizeof(TileInfo) * e->h_stiles.size(),
                         hipMemcpyHostToDevice));
    }
    e->graph_valid = false;
    e->have_tree = false;
    e->bpart_valid = false;
    return GCS_OK;
}

int check_params(const gcs_params* p)
{
    if (!p) return fail(GCS_ERR_INVALID, "params is null");
    if (!std::isfinite(p->G) || !std::isfinite(p->eps2) || !std::isfinite(p->dt) ||
        !std::isfinite(p->theta))
        return fail(GCS_ERR_INVALID, "non-finite parameter");
    if (p->eps2 < 0.0f) return fail(GCS_ERR_INVALID, "eps2 must be >= 0");
    if (p->mac != GCS_MAC_REFERENCE && p->mac != GCS_MAC_GROUP)
        return fail(GCS_ERR_INVALID, "unknown opening criterion");
    if (p->mac == GCS_MAC_GROUP) {
        if (!(p->eps2 > 0.0f)) return fail(GCS_ERR_INVALID, "group traversal needs eps2 > 0");
        if (p->theta < 0.0f || p->theta > 1.0f)
            return fail(GCS_ERR_INVALID, "theta must be in [0, 1]");
        if (p->leaf_size < 1 || p->leaf_size > kMaxLeaf)
            return fail(GCS_ERR_INVALID, "leaf_size must be in [1, 64]");
        if (p->block_size != 64 && p->block_size != 128 && p->block_size != 256 &&
            !(p->team == 8 && p->block_size == 512))
            return fail(GCS_ERR_INVALID, "block_size must be 64, 128 or 256 (512 for a team of 8)");
        if (p->waves_per_simd < 1 || p->waves_per_simd > 4)
            return fail(GCS_ERR_INVALID, "waves_per_simd must be in [1, 4]");
        if (p->group_size < 1 || p->group_size > kGroupMax)
            return fail(GCS_ERR_INVALID, "group_size must be in [1, 192]");
        if (p->open_offset != 0 && p->open_offset != 1)
            return fail(GCS_ERR_INVALID, "open_offset must be 0 or 1");
        if (p->team != 0 && p->team != 1 && p->team != 2 && p->team != 4 && p->team != 8)
            return fail(GCS_ERR_INVALID, "team must be 0, 1, 2, 4 or 8");
        if (p->team > 1 && p->block_size != p->team * kWave)
            return fail(GCS_ERR_INVALID, "a team of w waves takes block_size 64 w");
        if (!std::isfinite(p->mono_far) || (p->mono_far != 0.0f && p->mono_far < 1.0f))
            return fail(GCS_ERR_INVALID, "mono_far must be 0 (off) or >= 1");
        if (!std::isfinite(p->local_kappa) || (p->local_kappa != 0.0f && p->local_kappa < 1.0f))
            return fail(GCS_ERR_INVALID, "local_kappa must be 0 (off) or >= 1");
    } else {
        if (p->stack_depth < 2 || p->stack_depth > 1024)
            return fail(GCS_ERR_INVALID, "stack_depth must be in [2, 1024]");
    }
    if (p->instance_pass < 0 || p->instance_pass > 2)
        return fail(GCS_ERR_INVALID, "instance_pass must be 0 (automatic), 1 or 2 (levels)");
    if (p->sort < 0 || p->sort > 2)
        return fail(GCS_ERR_INVALID, "sort must be 0 (automatic), 1 (radix) or 2 (buckets)");
    if (p->precision != GCS_PREC_PLAIN && p->precision != GCS_PREC_COMPENSATED)
        return fail(GCS_ERR_INVALID, "unknown precision mode");
    if (!std::isfinite(p->dt_eta) || p->dt_eta < 0.0f)
        return fail(GCS_ERR_INVALID, "dt_eta must be >= 0");
    if (p->dt_eta > 0.0f && !(p->eps2 > 0.0f))
        return fail(GCS_ERR_INVALID, "the adaptive step needs eps2 > 0");
    if (p->use_bounds && !(std::isfinite(p->global_min) && std::isfinite(p->global_max)))
        return fail(GCS_ERR_INVALID, "non-finite bounds");
    return GCS_OK;
}

// Waves of the persistent group traversal: waves_per_simd x 4 SIMDs x CUs,
// within the per-wave memory, and no more blocks than particles need.
int group_waves(const gcs_engine* e, const gcs_params& p)
{
    const int wpb = p.team > 1 ? p.team : p.block_size / kWave;
    const int waves = std::min(e->trav_slots, e->num_cu * 4 * p.waves_per_simd);
    return std::max(1, std::min(waves / wpb, (int)((e->n_total + wpb - 1) / wpb))) * wpb;
}

// Squared monopole radius factor of the group traversal: an accepted node
// farther than mono_far times its opening radius is taken as a monopole
// (+inf: every accepted node is a quadrupole).
float far2_of(const gcs_params& p)
{
    return p.mono_far > 0.0f ? p.mono_far * p.mono_far : __builtin_inff();
}

// Squared local-expansion radius factor (+inf: off; the kernels are then the
// LOCAL = false instantiations and never read it).
float kappa2_of(const gcs_params& p)
{
    return p.local_kappa > 0.0f ? p.local_kappa * p.local_kappa : __builtin_inff();
}

template <bool POT, bool QUAD, int WPB>
void launch_group(gcs_engine* e, int blocks, const gcs_params& p, float4* posm, bool order)
{
    if (p.local_kappa > 0.0f)
        k_traverse_group<POT, QUAD, WPB, true><<<blocks, WPB * kWave, 0, e->stream>>>(
            e->groups, order ? e->gorder : nullptr, e->ngroups, e->segs, posm, e->node, e->nodeQ,
            e->acc, e->slot_mem, e->qctr, e->wcount, e->err, p.G, p.eps2, p.leaf_size,
            far2_of(p), kappa2_of(p));
    else
        k_traverse_group<POT, QUAD, WPB, false><<<blocks, WPB * kWave, 0, e->stream>>>(
            e->groups, order ? e->gorder : nullptr, e->ngroups, e->segs, posm, e->node, e->nodeQ,
            e->acc, e->slot_mem, e->qctr, e->wcount, e->err, p.G, p.eps2, p.leaf_size,
            far2_of(p), kappa2_of(p));
}

template <bool POT, int TEAM>
void launch_team(gcs_engine* e, int blocks, const gcs_params& p, float4* posm, bool order)
{
    if (p.local_kappa > 0.0f)
        k_traverse_team<POT, true, TEAM, true><<<blocks, TEAM * kWave, 0, e->stream>>>(
            e->groups, order ? e->gorder : nullptr, e->ngroups, e->segs, posm, e->node, e->nodeQ,
            e->acc, e->slot_mem, e->qctr, e->wcount, e->err, p.G, p.eps2, p.leaf_size,
            far2_of(p), kappa2_of(p));
    else
        k_traverse_team<POT, true, TEAM, false><<<blocks, TEAM * kWave, 0, e->stream>>>(
            e->groups, order ? e->gorder : nullptr, e->ngroups, e->segs, posm, e->node, e->nodeQ,
            e->acc, e->slot_mem, e->qctr, e->wcount, e->err, p.G, p.eps2, p.leaf_size,
            far2_of(p), kappa2_of(p));
}

template <int STACK>
void launch_reference(gcs_engine* e, const gcs_params& p, float4* posm, int nblk)
{
    k_traverse_reference<STACK><<<nblk, kTileThreads, 0, e->stream>>>(
        e->tiles, e->segs, posm, e->node, e->acc, e->counters, e->err, p.G, p.eps2,
        std::min(p.stack_depth, STACK));
}

void mark(gcs_engine* e, int k)
{
    if (e->timed) (void)hipEventRecord(e->ev[k], e->stream);
}

Marks group_marks(gcs_engine* e) { return Marks{e->gmark, e->gq, e->gcnt, nullptr}; }
Marks atile_marks(gcs_engine* e) { return Marks{e->amark, e->aq, e->acnt, e->aroot}; }

// Queues one step on the stream.  Reads state[cur], leaves the result in
// state[cur ^ 1] and flips cur.
// The eight sort passes over keys[0]/vals[0] (sorted result back in
// buffer 0).  The chunk histograms of all passes must be zero on entry.
// Sorts instances seg0 .. seg0 + nseg - 1; the other instances' keys are
// left alone.
void enqueue_sort_range(gcs_engine* e, int seg0, int nseg)
{
    if (nseg <= 0) return;
    const size_t cstride = (size_t)e->nchunks * kRadix;
    const int st0 = e->h_stile0[(size_t)seg0];
    const int st1 = seg0 + nseg < (int)e->h_stile0.size() ? e->h_stile0[(size_t)(seg0 + nseg)]
                                                          : (int)e->h_stiles.size();
    const TileInfo* tiles = e->stiles + st0;
    const int ntiles = st1 - st0;
    // Tile rows are indexed by the launch's block index: the rows of the
    // range start at row 0, the chunk rows stay global.
    for (int pass = 0; pass < kSortPasses; ++pass) {
        const int shift = pass * kRadixBits;
        const int src = pass & 1, dst = src ^ 1;
        uint32_t* ch = e->chist + (size_t)pass * cstride;
        if (e->stile_keys == kSortTileBig) {
            k_sort_count<1024><<<ntiles, 1024, 0, e->stream>>>(tiles, e->segs, e->keys[src],
                                                                e->hist, ch, shift);
            k_sort_scatter<1024><<<ntiles, 1024, 0, e->stream>>>(
                tiles, e->segs, e->keys[src], e->vals[src], e->keys[dst], e->vals[dst],
                e->hist, ch, shift);
        } else {
            k_sort_count<kTileThreads><<<ntiles, kTileThreads, 0, e->stream>>>(
                tiles, e->segs, e->keys[src], e->hist, ch, shift);
            k_sort_scatter<kTileThreads><<<ntiles, kTileThreads, 0, e->stream>>>(
                tiles, e->segs, e->keys[src], e->vals[src], e->keys[dst], e->vals[dst],
                e->hist, ch, shift);
        }
    }
    static_assert(kSortPasses % 2 == 0, "sorted keys end in buffer 0");
}

void enqueue_sort(gcs_engine* e) { enqueue_sort_range(e, 0, (int)e->h_segs.size()); }

// The bucket sort of instances 0 .. nseg - 1 over keys[0]/vals[0] (sorted
// result back in buffer 0, buffer 1 the buckets); the other instances' keys
// are left alone.
void enqueue_bucket_sort(gcs_engine* e, int nseg)
{
    if (nseg <= 0) return;
    const int ntiles = nseg < (int)e->h_segs.size() ? e->h_segs[(size_t)nseg].tile0
                                                    : (int)e->h_tiles.size();
    const int nbkt = e->h_bkt0[(size_t)nseg];
    k_bucket_cand<<<(nbkt + 255) / 256, 256, 0, e->stream>>>(e->segs, e->bkt0, e->bseg, nbkt,
                                                             e->keys[0], e->bcand, e->bcnt,
                                                             e->bfill);
    k_bucket_split<<<nseg, 1024, 0, e->stream>>>(e->bkt0, e->bcand, e->bsplit);
    k_bucket_count<<<ntiles, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->bkt0, e->keys[0],
                                                           e->bsplit, e->bid, e->bcnt);
    k_bucket_scan<<<nseg, 1024, 0, e->stream>>>(e->segs, e->bkt0, e->bcnt, e->boff);
    k_bucket_scatter<<<ntiles, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->bkt0,
                                                             e->keys[0], e->bid, e->boff, e->bfill,
                                                             e->keys[1], e->vals[1]);
    k_bucket_sort<<<nbkt, kBucketThreads, 0, e->stream>>>(e->bcnt, e->boff, e->keys[1],
                                                          e->vals[1], e->keys[0], e->vals[0]);
}

// Instances 0 .. k - 1 get the bucket sort, the rest the radix sort: with
// sort 0 the leading instances whose particles are in the last step's sorted
// order (instances joined since then come after them).
int bucket_segs(const gcs_engine* e, const gcs_params& p)
{
    const int nseg = (int)e->h_segs.size();
    if (p.sort == 1) return 0;
    if (p.sort == 2) return nseg;
    int k = 0;
    while (k < nseg && e->inst[(size_t)k].ordered) ++k;
    return k;
}

// The disc sampler table on the device (uploaded on first use).
int mn_table_dev(gcs_engine* e, gcs_gen::MnTableView* mt)
{
    constexpr int kMnWords =
        (gcs_gen::kNR + 1) + (gcs_gen::kNZ + 1) + 2 * gcs_gen::kNR * gcs_gen::kNZ;
    if (!e->mn_dev) {
        const gcs_gen::MnTableView& ht = gcs_gen::host_mn_table();
        int rc = dalloc(&e->mn_dev, (size_t)kMnWords);
        if (rc) return rc;
        std::vector<double> w;
        w.reserve(kMnWords);
        w.insert(w.end(), ht.re, ht.re + gcs_gen::kNR + 1);
        w.insert(w.end(), ht.ze, ht.ze + gcs_gen::kNZ + 1);
        w.insert(w.end(), ht.bound, ht.bound + gcs_gen::kNR * gcs_gen::kNZ);
        w.insert(w.end(), ht.cdf, ht.cdf + gcs_gen::kNR * gcs_gen::kNZ);
        HIPCHK(hipMemcpy(e->mn_dev, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    }
    mt->re = e->mn_dev;
    mt->ze = mt->re + gcs_gen::kNR + 1;
    mt->bound = mt->ze + gcs_gen::kNZ + 1;
    mt->cdf = mt->bound + gcs_gen::kNR * gcs_gen::kNZ;
    return GCS_OK;
}

GenInst gen_inst(const gcs_collision_spec& sp)
{
    GenInst g;
    g.seed = sp.seed;
    g.n_a = sp.n_a;
    g.n_b = sp.n_b;
    for (int d = 0; d < 3; ++d) {
        g.off[d] = sp.offset[d];
        g.vel[d] = sp.velocity[d];
    }
    const double t = sp.tilt_deg * gcs_gen::kPi / 180.0;
    g.ct = std::cos(t);
    g.st = std::sin(t);
    return g;
}

// Generates instances seg0 .. seg0 + k - 1 (the tables already list them)
// into the current state, from their collisions insts[0..k).
int generate_into(gcs_engine* e, const GenInst* insts, int seg0, int k)
{
    gcs_gen::MnTableView mt;
    int rc = mn_table_dev(e, &mt);
    if (rc) return rc;
    GenInst* di = nullptr;
    rc = dalloc(&di, (size_t)k);
    if (rc) return rc;
    const int tile0 = e->h_segs[(size_t)seg0].tile0;
    const int ntiles = e->h_segs[(size_t)(seg0 + k - 1)].tile0 +
                       e->h_segs[(size_t)(seg0 + k - 1)].ntiles - tile0;
    const TileInfo* tiles = e->tiles + tile0;
    double4* xyz = reinterpret_cast<double4*>(e->staging);  // 32 B per particle
    const int c = e->cur;
    hipError_t le = hipMemcpyAsync(di, insts, sizeof(GenInst) * (size_t)k, hipMemcpyHostToDevice,
                                   e->stream);
    if (le == hipSuccess)
        le = hipMemsetAsync(e->chist, 0, sizeof(uint32_t) * kSortPasses * kRadix * e->nchunks,
                            e->stream);
    if (le == hipSuccess) {
        k_gen_pos<<<4 * ntiles, kTileThreads, 0, e->stream>>>(tiles, e->segs, di, seg0, mt, xyz,
                                                               e->keys[0], e->vals[0]);
        enqueue_sort_range(e, seg0, k);
        k_gen_count<<<ntiles, kTileThreads, 0, e->stream>>>(tiles, e->segs, di, seg0, e->vals[0],
                                                             e->hist);
        k_gen_scan<<<(k + 255) / 256, 256, 0, e->stream>>>(e->segs, seg0, k, tile0, e->hist);
        k_gen_menc<<<ntiles, kTileThreads, 0, e->stream>>>(tiles, e->segs, di, seg0, e->vals[0],
                                                            e->hist, xyz);
        k_gen_state<<<4 * ntiles, kTileThreads, 0, e->stream>>>(tiles, e->segs, di, seg0, xyz,
                                                                 e->posm[c], e->vel[c],
                                                                 e->plo[c]);
        le = hipGetLastError();
    }
    if (le == hipSuccess) le = hipStreamSynchronize(e->stream);
    dfree(di);
    if (le != hipSuccess) return fail(GCS_ERR_HIP, std::string("generation: ") + hipGetErrorString(le));
    return GCS_OK;
}

int enqueue_step(gcs_engine* e, const gcs_params& p, bool integrate)
{
    const int ntiles = (int)e->h_tiles.size();
    const int nseg = (int)e->h_segs.size();
    if (ntiles == 0) return GCS_OK;
    const int qblocks = 4 * ntiles;
    const int c = e->cur;
    const bool comp = p.precision == GCS_PREC_COMPENSATED;
    const bool group = p.mac == GCS_MAC_GROUP;

    mark(e, 0);
    const int64_t cwords = (int64_t)kSortPasses * e->nchunks * kRadix;
    const int ngchunk = (4 * ntiles + kGroupChunk - 1) / kGroupChunk;
    if (!p.use_bounds) {
        // A recorded step relies on the previous step's k_integrate (gcs_step
        // runs k_bounds once before replays when the partials are stale).
        if (!e->bpart_valid && !e->capturing) {
            k_bounds<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->posm[c],
                                                               e->bpart);
            e->bpart_valid = true;
        }
        k_bounds_reduce<<<nseg, 1024, 0, e->stream>>>(e->segs, e->bpart, e->seg_bounds);
    }
    StepReset rs;
    rs.nseg = nseg;
    rs.nfront = e->nfront;
    rs.ntop = e->ntop;
    rs.nmid = e->nmid;
    rs.mfail = e->mfail;
    rs.qctr = e->qctr;
    rs.counters = e->counters;
    rs.chist = e->chist;
    rs.chist_words = cwords;
    rs.gm = group_marks(e);
    rs.am = atile_marks(e);
    // Instances 0 .. nbs - 1 are sorted with buckets, the rest by radix.
    const int nbs = bucket_segs(e, p);
    k_morton<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->posm[c],
                                                       e->seg_bounds, p.use_bounds, p.global_min,
                                                       p.global_max, e->keys[0], e->vals[0], nbs,
                                                       rs);
    LAUNCHCHK();
    if (e->capture_keys)
        HIPCHK(hipMemcpyAsync(e->keys_raw, e->keys[0], sizeof(uint64_t) * e->n_total,
                              hipMemcpyDeviceToDevice, e->stream));
    mark(e, 1);
    enqueue_bucket_sort(e, nbs);
    enqueue_sort_range(e, nbs, nseg - nbs);
    LAUNCHCHK();
    mark(e, 2);
    // (The permute is timed with the build: one launch does both.)
    mark(e, 3);
    // The two-level instance pass when an instance has at least kMidSize
    // particles and the instances number at most half the CUs: one block per
    // instance would then leave most CUs idle behind the largest instance's
    // single-block pass (2 x 1,048,576 particles: tree stage 0.656 -> 0.557
    // ms; 2 x 131,072: 0.218 -> 0.211 ms).  With an instance per CU the pass
    // is spread already and the extra launch costs more (256 collisions of
    // 4,096-65,536 particles: 1.201 -> 1.215 ms).
    bool mid_any = false;
    for (const SegInfo& sg : e->h_segs) mid_any = mid_any || sg.n >= kMidSize;
    const bool two_level = p.instance_pass == 2 ||
                           (p.instance_pass == 0 && mid_any && 2 * nseg <= e->num_cu);
    PermuteArgs pa;
    pa.perm = e->vals[0];
    pa.posm_in = e->posm[c];
    pa.vel_in = e->vel[c];
    pa.plo_in = e->plo[c];
    pa.posm_out = e->posm[c ^ 1];
    pa.vel_out = e->vel[c ^ 1];
    pa.plo_out = e->plo[c ^ 1];
    pa.compensated = comp ? 1 : 0;
    pa.write_leaves = group ? 0 : 1;
    k_permute_build<<<qblocks, kTileThreads, 0, e->stream>>>(
        e->tiles, e->segs, pa, e->keys[0], e->node, e->nodeQ, e->parent, e->flag, group_marks(e),
        atile_marks(e), group ? p.group_size : kGroup,
        TopList{e->tlist, e->ntop, e->tslot, two_level ? e->mids : nullptr, e->nmid});
    e->cur = c ^ 1;
    float4* posm = e->posm[e->cur];
    // Number the traversal groups and the aggregation subtrees in particle
    // order and write their tables.
    k_group_chunksum<<<ngchunk, kGroupChunk, 0, e->stream>>>(e->gq, e->aq, qblocks, e->gchunk,
                                                             e->achunk);
    k_group_emit<<<qblocks, kTileThreads, 0, e->stream>>>(
        e->tiles, e->segs, group_marks(e), e->gchunk, atile_marks(e), e->achunk, e->groups,
        e->ngroups, e->atiles, e->natiles, (int)e->n_total);
    // Size-class dispatch order, when the groups (at least n / group_size)
    // outnumber the traversal's waves four times over but not eight: with about
    // one group per wave the extra launch costs more than the order saves, and
    // with many groups per wave the tail it shortens is small while the order
    // spreads a queue's concurrent groups over many instances' trees (L2
    // misses): 256 collisions of 4,096-65,536 particles ran 3.5% faster
    // without it, 2 x 1,048,576 particles 2.3% slower.  Launched with the
    // traversal, after the aggregation (the order is the traversal's).
    const int64_t gw = (int64_t)p.group_size * group_waves(e, p);
    const bool order = group && e->n_total >= 4 * gw && e->n_total < 8 * gw;
    mark(e, 4);
    const float inv_theta = p.theta > 0.0f ? 1.0f / p.theta : __builtin_inff();
    // Opening value of the merge: width (reference traversal), (w/theta)^2 or
    // (w/theta + |COM - box centre|)^2 (group traversal).
    const int amode = group ? (p.open_offset ? 2 : 1) : 0;
    // Persistent: as many blocks as fit on the CUs at once (2048 threads per
    // CU; the LDS of two 1024-thread blocks) take the work items (one or two
    // aggregation subtrees each) from a counter.
    const int ablocks = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)e->num_cu * (2048 / kATile), e->n_total));
    k_aggregate<<<ablocks, kATile, 0, e->stream>>>(e->atiles, e->natiles, e->segs, posm, e->node,
                                                   e->nodeQ, e->boxLo, e->boxHi, e->parent,
                                                   e->front, e->nfront, e->qctr + kNumXcd,
                                                   inv_theta, amode);
    if (two_level)
        k_aggregate_mid<<<e->num_cu, 1024, 0, e->stream>>>(
            e->segs, posm, e->node, e->nodeQ, e->boxLo, e->boxHi, e->parent, e->mids, e->nmid,
            e->tlist, e->ntop, e->tslot, e->front, e->mfail, inv_theta, amode);
    k_aggregate_top<<<nseg, 1024, 0, e->stream>>>(e->segs, posm, e->node, e->nodeQ, e->boxLo,
                                                  e->boxHi, e->parent, e->flag, e->front,
                                                  e->nfront, e->tlist, e->ntop, e->tslot,
                                                  e->mfail, inv_theta, amode, two_level ? 1 : 0);
    LAUNCHCHK();
    mark(e, 5);
    if (order)
        k_group_order<<<kNumXcd, 1024, 0, e->stream>>>(e->groups, e->ngroups, p.group_size,
                                                        e->gorder);
    if (group) {
        // Persistent: as many blocks as the per-wave memory covers (blocks *
        // waves per block <= trav_slots); waves that find no group left exit
        // at once.
        const int wpb = p.team > 1 ? p.team : p.block_size / kWave;
        const int blocks = group_waves(e, p) / wpb;
        const bool pot = p.diagnostics != 0;
        e->wslots = p.team > 1 ? blocks : blocks * wpb;  // count slots the launch writes
        if (p.team == 2) {
            if (pot) launch_team<true, 2>(e, blocks, p, posm, order);
            else launch_team<false, 2>(e, blocks, p, posm, order);
        } else if (p.team == 4) {
            if (pot) launch_team<true, 4>(e, blocks, p, posm, order);
            else launch_team<false, 4>(e, blocks, p, posm, order);
        } else if (p.team == 8) {
            if (pot) launch_team<true, 8>(e, blocks, p, posm, order);
            else launch_team<false, 8>(e, blocks, p, posm, order);
        } else if (wpb == 4) {
            if (pot) launch_group<true, true, 4>(e, blocks, p, posm, order);
            else launch_group<false, true, 4>(e, blocks, p, posm, order);
        } else if (wpb == 2) {
            if (pot) launch_group<true, true, 2>(e, blocks, p, posm, order);
            else launch_group<false, true, 2>(e, blocks, p, posm, order);
        } else {
            if (pot) launch_group<true, true, 1>(e, blocks, p, posm, order);
            else launch_group<false, true, 1>(e, blocks, p, posm, order);
        }
    } else {
        if (p.stack_depth <= 64) launch_reference<64>(e, p, posm, qblocks);
        else if (p.stack_depth <= 128) launch_reference<128>(e, p, posm, qblocks);
        else if (p.stack_depth <= 256) launch_reference<256>(e, p, posm, qblocks);
        else launch_reference<1024>(e, p, posm, qblocks);
    }
    LAUNCHCHK();
    mark(e, 6);
    if (integrate) {
        const bool adaptive = p.dt_eta > 0.0f;
        if (adaptive) {
            k_accmax<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->acc,
                                                               e->amax_part);
            k_adapt_dt<<<nseg, 256, 0, e->stream>>>(e->segs, e->amax_part, p.dt, p.dt_eta,
                                                     std::sqrt(p.eps2), e->dt_inst);
        }
        k_integrate<<<qblocks, kTileThreads, 0, e->stream>>>(
            e->tiles, e->segs, e->acc, posm, e->vel[e->cur], e->plo[e->cur], p.dt,
            adaptive ? e->dt_inst : nullptr, comp ? 1 : 0, p.colours, p.diagnostics,
            e->diag_part, e->amax_part, e->tsim, e->err, e->bpart);
        if (!e->capturing) e->bpart_valid = true;
        if (p.diagnostics)
            k_diag_final<<<nseg, 64, 0, e->stream>>>(e->segs, nseg, e->diag_part, e->diag_out);
        LAUNCHCHK();
    }
    mark(e, 7);
    e->have_tree = true;
    e->last_mac = p.mac;
    if (!e->capturing)
        for (Instance& in : e->inst) in.ordered = true;  // the permute stored the sorted order
    return GCS_OK;
}

int capture_graphs(gcs_engine* e, const gcs_params& p)
{
    for (int k = 0; k < 2; ++k) {
        if (e->gexec[k]) (void)hipGraphExecDestroy(e->gexec[k]);
        e->gexec[k] = nullptr;
    }
    const int saved = e->cur;
    for (int k = 0; k < 2; ++k) {
        e->cur = k;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
        e->capturing = true;
        int rc = enqueue_step(e, p, true);
        e->capturing = false;
        hipError_t ec = hipStreamEndCapture(e->stream, &g);
        if (rc != GCS_OK) {
            if (g) (void)hipGraphDestroy(g);
            e->cur = saved;
            return rc;
        }
        if (ec != hipSuccess) {
            e->cur = saved;
            return fail(GCS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&e->gexec[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            e->cur = saved;
            return fail(GCS_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    e->cur = saved;
    e->gparams = p;
    e->graph_valid = true;
    e->graph_wslots = e->wslots;
    e->graph_bucket_segs = bucket_segs(e, p);
    return GCS_OK;
}

// Autotuner results per (traversal kind, theta, opening form, precision,
// log2 particle bucket, log2 instance bucket, far-monopole factor, local
// expansion factor), shared by every engine in the process and
// readable/writable as a text table.
struct TuneKey {
    int mac, open_offset, precision, nbucket, kbucket;
    uint32_t theta_bits, far_bits, kappa_bits;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(mac, open_offset, precision, nbucket, kbucket, theta_bits, far_bits,
                        kappa_bits) < std::tie(o.mac, o.open_offset, o.precision, o.nbucket,
                                               o.kbucket, o.theta_bits, o.far_bits, o.kappa_bits);
    }
};
struct TuneVal {
    int group, leaf, block, team, wps, stack;
    float ms;
};
std::map<TuneKey, TuneVal> g_tune;
std::mutex g_tune_mu;

int log2_bucket(int64_t v)
{
    int b = 0;
    while (v > 1) {
        v >>= 1;
        ++b;
    }
    return b;
}

TuneKey tune_key(const gcs_engine* e, const gcs_params& p)
{
    TuneKey k;
    k.mac = p.mac;
    k.open_offset = p.mac == GCS_MAC_GROUP ? p.open_offset : 0;
    k.precision = p.precision;
    k.nbucket = log2_bucket(e->n_total);
    k.kbucket = log2_bucket((int64_t)e->inst.size());
    const bool grp = p.mac == GCS_MAC_GROUP;
    const float th = grp ? p.theta : 0.0f, far = grp ? p.mono_far : 0.0f,
                kap = grp ? p.local_kappa : 0.0f;
    std::memcpy(&k.theta_bits, &th, 4);
    std::memcpy(&k.far_bits, &far, 4);
    std::memcpy(&k.kappa_bits, &kap, 4);
    return k;
}

void apply_tune(const TuneVal& v, gcs_params* q)
{
    q->group_size = v.group;
    q->leaf_size = v.leaf;
    q->block_size = v.block;
    q->team = v.team;
    q->waves_per_simd = v.wps;
    q->stack_depth = v.stack;
}

int find_slot(gcs_engine* e, int32_t id)
{
    for (size_t s = 0; s < e->inst.size(); ++s)
        if (e->inst[s].id == id) return (int)s;
    return -1;
}

int upload_at(gcs_engine* e, const gcs_star* stars, int64_t n, int64_t offset)
{
    if (n <= 0) return GCS_OK;
    HIPCHK(hipMemcpyAsync(e->staging, stars, sizeof(gcs_star) * (size_t)n, hipMemcpyHostToDevice,
                          e->stream));
    const int c = e->cur;
    k_unpack_stars<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(e->staging, n, offset,
                                                                       e->posm[c], e->vel[c],
                                                                       e->plo[c]);
    LAUNCHCHK();
    HIPCHK(hipStreamSynchronize(e->stream));
    return GCS_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================

extern "C" {

const char* gcs_last_error(void) { return g_err.c_str(); }

void gcs_params_default(gcs_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->G = 1.0f;
    p->eps2 = 0.0025f;
    p->dt = 0.01f;
    p->theta = 0.5f;
    p->mac = GCS_MAC_GROUP;
    p->precision = GCS_PREC_PLAIN;
    p->leaf_size = 8;
    p->block_size = 256;
    p->stack_depth = 256;
    p->use_bounds = 0;
    p->colours = 1;
    p->diagnostics = 0;
    p->use_graph = 1;
    p->waves_per_simd = 3;
    p->group_size = kGroup;
}

int64_t gcs_device_bytes(int64_t cap_n, int32_t cap_k)
{
    if (cap_n < 1 || cap_k < 1) return 0;
    const int64_t n = cap_n;
    int64_t b = 0;
    b += 6 * 16 * n;           // posm, vel, plo x2
    b += 16 * n;               // acc
    b += 3 * 8 * n + 2 * 4 * n; // keys x2 + raw, vals x2
    b += 2 * n * (32 + 32 + 16 + 16 + 4 + 4);  // nodes
    b += (int64_t)(2 * sizeof(TileInfo) * max_tiles(cap_n, cap_k) + sizeof(SegInfo) * cap_k +
                   sizeof(GroupInfo) * (cap_n + 1) + 12 * cap_n);
    b += (int64_t)(4 * hist_entries(cap_n, cap_k) +
                   4 * (size_t)kSortPasses * kRadix * max_chunks(cap_n, cap_k));
    b += (int64_t)(4 * 4 * max_tiles(cap_n, cap_k) + 4 * max_gchunks(cap_n, cap_k));
    b += 8 * cap_k + 64 + 5 * 8 * 4 * (int64_t)max_tiles(cap_n, cap_k) + 5 * 8 * cap_k;
    b += 12 * n + 16 * (n + 1) + 4 * 4 * (int64_t)max_tiles(cap_n, cap_k) +
         4 * (int64_t)max_gchunks(cap_n, cap_k) + 4;  // aggregation subtrees
    b += 4 * 4 * (int64_t)max_tiles(cap_n, cap_k) + 12 * (int64_t)cap_k;  // amax, dt, time
    b += 4 * n + 8 * n + 4 * (int64_t)cap_k;  // nodes above the aggregation subtrees
    b += 8 * (n / kATile + (int64_t)cap_k + 1) + 4 + 4 * (int64_t)cap_k;  // mid-level roots
    b += 4 * ((int64_t)cap_k + 1) + 32 * (int64_t)max_buckets(cap_n, cap_k) + 4 * n;  // buckets
    b += 32 * n;               // staging (also the device generator's scratch)
    b += 8 * ((gcs_gen::kNR + 1) + (gcs_gen::kNZ + 1) + 2 * gcs_gen::kNR * gcs_gen::kNZ);
    return b;
}

int gcs_create(int64_t cap_n, int32_t cap_k, int32_t device, gcs_engine** out)
{
    if (!out) return fail(GCS_ERR_INVALID, "out is null");
    *out = nullptr;
    // Node records are indexed as node[2 * v] with a 32-bit v < 2 cap_n.
    if (cap_n < 1 || cap_k < 1 || cap_n > kMaxParticles)
        return fail(GCS_ERR_INVALID, "capacity out of range (1 .. 2^29 particles)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(GCS_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(GCS_ERR_NO_DEVICE, "device ordinal out of range");
    HIPCHK(hipSetDevice(device));
    gcs_engine* e = new (std::nothrow) gcs_engine();
    if (!e) return fail(GCS_ERR_OOM, "host allocation failed");
    e->device = device;
    e->cap_n = cap_n;
    e->cap_k = cap_k;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        e->num_cu = prop.multiProcessorCount;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_traverse_group<true, true, 4, false>, 256, 0) !=
            hipSuccess ||
        occ < 1)
        occ = 2;
    e->trav_blocks_cap = e->num_cu * occ;
    const size_t nt = max_tiles(cap_n, cap_k);
    const size_t he = hist_entries(cap_n, cap_k);
    int rc = GCS_OK;
    const size_t n = (size_t)cap_n;
    for (int k = 0; k < 2 && rc == GCS_OK; ++k) {
        if ((rc = dalloc(&e->posm[k], n))) break;
        if ((rc = dalloc(&e->vel[k], n))) break;
        if ((rc = dalloc(&e->plo[k], n))) break;
        if ((rc = dalloc(&e->keys[k], n))) break;
        if ((rc = dalloc(&e->vals[k], n))) break;
    }
    if (!rc) rc = dalloc(&e->acc, n);
    if (!rc) rc = dalloc(&e->keys_raw, n);
    if (!rc) rc = dalloc(&e->node, 4 * n);
    if (!rc) rc = dalloc(&e->nodeQ, 4 * n);
    if (!rc) rc = dalloc(&e->boxLo, 2 * n);
    if (!rc) rc = dalloc(&e->boxHi, 2 * n);
    if (!rc) rc = dalloc(&e->parent, 2 * n);
    if (!rc) rc = dalloc(&e->flag, 2 * n);
    if (!rc) rc = dalloc(&e->tiles, nt);
    if (!rc) rc = dalloc(&e->stiles, nt);
    if (!rc) rc = dalloc(&e->segs, (size_t)cap_k);
    if (!rc) rc = dalloc(&e->groups, n + 1);
    if (!rc) rc = dalloc(&e->gmark, n);
    if (!rc) rc = dalloc(&e->gorder, n + 1);
    if (!rc) rc = dalloc(&e->gq, 4 * nt);
    if (!rc) rc = dalloc(&e->gchunk, max_gchunks(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->gcnt, n);
    if (!rc) rc = dalloc(&e->ngroups, 1);
    if (!rc) rc = dalloc(&e->front, 2 * n + 2);
    if (!rc) rc = dalloc(&e->nfront, (size_t)cap_k);
    if (!rc) rc = dalloc(&e->tlist, n);
    if (!rc) rc = dalloc(&e->ntop, (size_t)cap_k);
    if (!rc) rc = dalloc(&e->tslot, 2 * n);
    if (!rc) rc = dalloc(&e->mids, 2 * (n / kATile + (size_t)cap_k + 1));
    if (!rc) rc = dalloc(&e->nmid, 1);
    if (!rc) rc = dalloc(&e->mfail, (size_t)cap_k);
    if (!rc) rc = dalloc(&e->bkt0, (size_t)cap_k + 1);
    if (!rc) rc = dalloc(&e->bseg, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->bsplit, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->bcand, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->bcnt, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->boff, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->bfill, max_buckets(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->bid, n);
    if (!rc) rc = dalloc(&e->amark, n);
    if (!rc) rc = dalloc(&e->aq, 4 * nt);
    if (!rc) rc = dalloc(&e->achunk, max_gchunks(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->acnt, n);
    if (!rc) rc = dalloc(&e->aroot, n);
    if (!rc) rc = dalloc(&e->atiles, n + 1);
    if (!rc) rc = dalloc(&e->natiles, 1);
    if (!rc) rc = dalloc(&e->hist, he);
    if (!rc) rc = dalloc(&e->chist, (size_t)kSortPasses * kRadix * max_chunks(cap_n, cap_k));
    if (!rc) rc = dalloc(&e->seg_bounds, 2 * (size_t)cap_k);
    if (!rc) rc = dalloc(&e->bpart, 8 * nt);
    if (!rc) rc = dalloc(&e->qctr, kNumXcd + 1);  // traversal queues; aggregation items
    if (!rc) rc = dalloc(&e->counters, 3);
    if (!rc) rc = dalloc(&e->err, 1);
    if (!rc) rc = dalloc(&e->diag_part, 5 * 4 * nt);
    if (!rc) rc = dalloc(&e->diag_out, 5 * (size_t)cap_k);
    if (!rc) rc = dalloc(&e->amax_part, 4 * nt);
    if (!rc) rc = dalloc(&e->dt_inst, (size_t)cap_k);
    if (!rc) rc = dalloc(&e->tsim, (size_t)cap_k);
    // One slot per resident traversal wave, and no more than the groups of
    // cap_n particles can use.
    e->trav_slots = (int)std::max<int64_t>(4, std::min<int64_t>((int64_t)e->trav_blocks_cap * 4,
                                                              (cap_n + 3) / 4 * 4));
    if (!rc) rc = dalloc(&e->slot_mem, (size_t)e->trav_slots * kSlotInts);
    if (!rc) rc = dalloc(&e->wcount, (size_t)kCounts * e->trav_slots);
    if (!rc) rc = dalloc(&e->staging, 32 * n);
    if (rc) {
        free_all(e);
        delete e;
        return rc;
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        free_all(e);
        delete e;
        return fail(GCS_ERR_HIP, "hipStreamCreate failed");
    }
    e->own_stream = true;
    for (auto& v : e->ev)
        if (hipEventCreate(&v) != hipSuccess) {
            (void)hipStreamDestroy(e->stream);
            free_all(e);
            delete e;
            return fail(GCS_ERR_HIP, "hipEventCreate failed");
        }
    if (hipMemset(e->err, 0, sizeof(unsigned)) != hipSuccess ||
        hipMemset(e->counters, 0, 3 * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipStreamDestroy(e->stream);
        free_all(e);
        delete e;
        return fail(GCS_ERR_HIP, "hipMemset failed");
    }
    gcs_params_default(&e->gparams);
    *out = e;
    return GCS_OK;
}

int gcs_destroy(gcs_engine* e)
{
    if (!e) return GCS_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_all(e);
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return GCS_OK;
}

int gcs_set_stream(gcs_engine* e, void* s)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    HIPCHK(hipStreamSynchronize(e->stream));
    if (s) {
        if (e->own_stream) (void)hipStreamDestroy(e->stream);
        e->stream = (hipStream_t)s;
        e->own_stream = false;
    } else if (!e->own_stream) {
        HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
        e->own_stream = true;
    }
    e->graph_valid = false;
    return GCS_OK;
}

int gcs_load_generated(gcs_engine* e, const gcs_collision_spec* specs, int32_t k)
{
    if (!e || (!specs && k > 0)) return fail(GCS_ERR_INVALID, "null argument");
    if (k < 0 || k > e->cap_k) return fail(GCS_ERR_CAPACITY, "instance count exceeds capacity");
    int64_t total = 0;
    for (int32_t s = 0; s < k; ++s) {
        if (specs[s].n_a < 0 || specs[s].n_b < 0 || (int64_t)specs[s].n_a + specs[s].n_b < 1)
            return fail(GCS_ERR_INVALID, "every collision needs at least 1 particle");
        total += (int64_t)specs[s].n_a + specs[s].n_b;
    }
    if (total > e->cap_n) return fail(GCS_ERR_CAPACITY, "particles exceed capacity");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->inst.clear();
    for (int32_t s = 0; s < k; ++s) e->inst.push_back({s, specs[s].n_a + specs[s].n_b});
    e->next_id = k;
    e->cur = 0;
    int rc = rebuild_tables(e);
    if (rc || k == 0) return rc;
    HIPCHK(hipMemset(e->tsim, 0, sizeof(double) * (size_t)k));
    std::vector<GenInst> hi((size_t)k);
    for (int32_t s = 0; s < k; ++s) hi[(size_t)s] = gen_inst(specs[s]);
    return generate_into(e, hi.data(), 0, k);
}

int gcs_instance_add_generated(gcs_engine* e, const gcs_collision_spec* spec, int32_t* id)
{
    if (!e || !spec) return fail(GCS_ERR_INVALID, "bad argument");
    const int64_t n = (int64_t)spec->n_a + spec->n_b;
    if (spec->n_a < 0 || spec->n_b < 0 || n < 1)
        return fail(GCS_ERR_INVALID, "a collision needs at least 1 particle");
    if ((int32_t)e->inst.size() + 1 > e->cap_k) return fail(GCS_ERR_CAPACITY, "instance slots full");
    if (e->n_total + n > e->cap_n) return fail(GCS_ERR_CAPACITY, "particle capacity full");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    const int slot = (int)e->inst.size();
    HIPCHK(hipMemset(e->tsim + slot, 0, sizeof(double)));
    const int32_t nid = e->next_id++;
    e->inst.push_back({nid, (int32_t)n});
    int rc = rebuild_tables(e);
    if (!rc) {
        const GenInst g = gen_inst(*spec);
        rc = generate_into(e, &g, slot, 1);
    }
    if (rc) {
        // Leave the engine as it was.
        e->inst.pop_back();
        (void)rebuild_tables(e);
        return rc;
    }
    if (id) *id = nid;
    return GCS_OK;
}

int gcs_load(gcs_engine* e, const gcs_star* stars, const int32_t* counts, int32_t k)
{
    if (!e || (!stars && k > 0) || (!counts && k > 0)) return fail(GCS_ERR_INVALID, "null argument");
    if (k < 0 || k > e->cap_k) return fail(GCS_ERR_CAPACITY, "instance count exceeds capacity");
    int64_t total = 0;
    for (int32_t s = 0; s < k; ++s) {
        if (counts[s] < 1) return fail(GCS_ERR_INVALID, "every instance needs at least 1 particle");
        total += counts[s];
    }
    if (total > e->cap_n) return fail(GCS_ERR_CAPACITY, "particles exceed capacity");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->inst.clear();
    for (int32_t s = 0; s < k; ++s) e->inst.push_back({s, counts[s]});
    e->next_id = k;
    e->cur = 0;
    int rc = upload_at(e, stars, total, 0);
    if (rc) return rc;
    if (k > 0) HIPCHK(hipMemset(e->tsim, 0, sizeof(double) * (size_t)k));
    return rebuild_tables(e);
}

int gcs_instance_add(gcs_engine* e, const gcs_star* stars, int32_t n, int32_t* id)
{
    if (!e || !stars || n < 1) return fail(GCS_ERR_INVALID, "bad argument");
    if ((int32_t)e->inst.size() + 1 > e->cap_k) return fail(GCS_ERR_CAPACITY, "instance slots full");
    if (e->n_total + n > e->cap_n) return fail(GCS_ERR_CAPACITY, "particle capacity full");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    int rc = upload_at(e, stars, n, e->n_total);
    if (rc) return rc;
    HIPCHK(hipMemset(e->tsim + e->inst.size(), 0, sizeof(double)));
    const int32_t nid = e->next_id++;
    e->inst.push_back({nid, n});
    if (id) *id = nid;
    return rebuild_tables(e);
}

int gcs_instance_remove(gcs_engine* e, int32_t id)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (slot + 1 < (int)e->inst.size()) {
        // Compact into the other state buffer, then switch to it.
        const int c = e->cur, d = c ^ 1;
        int64_t src = 0, dst = 0;
        for (size_t s = 0; s < e->inst.size(); ++s) {
            const int64_t n = e->inst[s].n;
            if ((int)s != slot) {
                HIPCHK(hipMemcpyAsync(e->posm[d] + dst, e->posm[c] + src, 16 * n,
                                      hipMemcpyDeviceToDevice, e->stream));
                HIPCHK(hipMemcpyAsync(e->vel[d] + dst, e->vel[c] + src, 16 * n,
                                      hipMemcpyDeviceToDevice, e->stream));
                HIPCHK(hipMemcpyAsync(e->plo[d] + dst, e->plo[c] + src, 16 * n,
                                      hipMemcpyDeviceToDevice, e->stream));
                dst += n;
            }
            src += n;
        }
        HIPCHK(hipStreamSynchronize(e->stream));
        e->cur = d;
    }
    if (slot + 1 < (int)e->inst.size()) {
        std::vector<double> t(e->inst.size());
        HIPCHK(hipMemcpy(t.data(), e->tsim, sizeof(double) * t.size(), hipMemcpyDeviceToHost));
        t.erase(t.begin() + slot);
        HIPCHK(hipMemcpy(e->tsim, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    }
    e->inst.erase(e->inst.begin() + slot);
    return rebuild_tables(e);
}

int gcs_instance_count(gcs_engine* e, int32_t* k, int64_t* n_total)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    if (k) *k = (int32_t)e->inst.size();
    if (n_total) *n_total = e->n_total;
    return GCS_OK;
}

int gcs_instance_info(gcs_engine* e, int32_t id, int32_t* n, int64_t* offset)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    if (n) *n = e->h_segs[slot].n;
    if (offset) *offset = e->h_segs[slot].start;
    return GCS_OK;
}

static int read_stars_range(gcs_engine* e, int64_t off, int64_t n, gcs_star* out)
{
    if (n <= 0) return GCS_OK;
    HIPCHK(hipSetDevice(e->device));
    const int c = e->cur;
    k_pack_stars<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(n, off, e->posm[c], e->vel[c],
                                                                      e->staging);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(out, e->staging, 32 * (size_t)n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return GCS_OK;
}

int gcs_read_stars(gcs_engine* e, gcs_star* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    return read_stars_range(e, 0, e->n_total, out);
}

int gcs_read_instance(gcs_engine* e, int32_t id, gcs_star* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    return read_stars_range(e, e->h_segs[slot].start, e->h_segs[slot].n, out);
}

// ---- Checkpoint / resume of one instance ----------------------------------
// Layout: 32-byte header (magic, version, n, simulated time, 2 reserved
// words), n gcs_star
// records in the instance's current order, then n float4 low parts of the
// compensated positions (x_lo, y_lo, z_lo, 0).  The current order and the low
// parts are the whole state a step carries over (the tree, keys and bounds
// are rebuilt every step), so a restored instance continues bit for bit.
namespace {
constexpr char kCkptMagic[8] = {'G', 'C', 'S', 'C', 'K', 'P', 'T', '1'};
struct CkptHeader {
    char magic[8];
    int32_t version;
    int32_t n;
    double sim_time;
    int32_t reserved[2];
};
static_assert(sizeof(CkptHeader) == 32, "checkpoint header is 32 bytes");
}  // namespace

int64_t gcs_checkpoint_bytes(int32_t n)
{
    if (n < 1) return 0;
    return (int64_t)sizeof(CkptHeader) + (int64_t)n * (32 + 16);
}

int gcs_checkpoint(gcs_engine* e, int32_t id, void* buf, int64_t cap)
{
    if (!e || !buf) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    const int32_t n = e->h_segs[slot].n;
    if (cap < gcs_checkpoint_bytes(n)) return fail(GCS_ERR_CAPACITY, "checkpoint buffer too small");
    CkptHeader h;
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.n = n;
    std::memset(h.reserved, 0, sizeof(h.reserved));
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(&h.sim_time, e->tsim + slot, sizeof(double), hipMemcpyDeviceToHost));
    uint8_t* out = static_cast<uint8_t*>(buf);
    std::memcpy(out, &h, sizeof(h));
    int rc = read_stars_range(e, e->h_segs[slot].start, n, reinterpret_cast<gcs_star*>(out + 32));
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(out + 32 + 32 * (size_t)n, e->plo[e->cur] + e->h_segs[slot].start,
                          16 * (size_t)n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return GCS_OK;
}

int gcs_restore(gcs_engine* e, const void* buf, int64_t bytes, int32_t* id)
{
    if (!e || !buf) return fail(GCS_ERR_INVALID, "null argument");
    if (bytes < (int64_t)sizeof(CkptHeader)) return fail(GCS_ERR_INVALID, "truncated checkpoint");
    CkptHeader h;
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1 || h.n < 1)
        return fail(GCS_ERR_INVALID, "not a checkpoint");
    if (bytes < gcs_checkpoint_bytes(h.n)) return fail(GCS_ERR_INVALID, "truncated checkpoint");
    const uint8_t* in = static_cast<const uint8_t*>(buf);
    int32_t nid = -1;
    int rc = gcs_instance_add(e, reinterpret_cast<const gcs_star*>(in + 32), h.n, &nid);
    if (rc) return rc;
    const int slot = find_slot(e, nid);
    hipError_t he = hipMemcpyAsync(e->plo[e->cur] + e->h_segs[slot].start,
                                   in + 32 + 32 * (size_t)h.n, 16 * (size_t)h.n,
                                   hipMemcpyHostToDevice, e->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(e->tsim + slot, &h.sim_time, sizeof(double), hipMemcpyHostToDevice,
                            e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) {
        // Leave the engine as it was: no half-restored instance.
        const std::string msg = std::string("restore: ") + hipGetErrorString(he);
        gcs_instance_remove(e, nid);
        return fail(he == hipErrorOutOfMemory ? GCS_ERR_OOM : GCS_ERR_HIP, msg);
    }
    if (id) *id = nid;
    return GCS_OK;
}

int gcs_read_position_lo(gcs_engine* e, float* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->device));
    std::vector<float4> tmp((size_t)e->n_total);
    if (e->n_total)
        HIPCHK(hipMemcpyAsync(tmp.data(), e->plo[e->cur], 16 * (size_t)e->n_total,
                              hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int64_t i = 0; i < e->n_total; ++i) {
        out[3 * i] = tmp[(size_t)i].x;
        out[3 * i + 1] = tmp[(size_t)i].y;
        out[3 * i + 2] = tmp[(size_t)i].z;
    }
    return GCS_OK;
}

int gcs_read_playback(gcs_engine* e, gcs_playback_star* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    const int64_t n = e->n_total;
    if (n == 0) return GCS_OK;
    HIPCHK(hipSetDevice(e->device));
    const int c = e->cur;
    k_pack_playback<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(n, 0, e->posm[c],
                                                                         e->vel[c], e->staging);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(out, e->staging, 16 * (size_t)n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return GCS_OK;
}

// ---- Bake output pipeline (main.cpp:133-297) ---------------------------------
// Per frame: one step on the engine stream; the PlaybackStar records of every
// particle (16 bytes, current device order) packed on the device into one half
// of the staging buffer; a 16-byte-per-particle copy into one of two pinned
// host buffers on a second stream, which overlaps the next frame's step; a
// writer thread appending the frame to the file once its copy has landed.
// Frame f uses staging half and pinned buffer f % 2.  As in the reference's
// loop one writer is in flight: frame f's copy is queued after frame f - 1's
// writer has been joined, so the pinned buffer of frame f - 2 is free, and
// frame f's pack waits (on the device) for the copy of frame f - 2 out of the
// same staging half.
int gcs_bake(gcs_engine* e, const gcs_params* p, int32_t frames, const char* path,
             double* seconds)
{
    if (!e || !path) return fail(GCS_ERR_INVALID, "null argument");
    int rc = check_params(p);
    if (rc) return rc;
    if (frames < 0) return fail(GCS_ERR_INVALID, "frames < 0");
    if (seconds) *seconds = 0.0;
    const int64_t n = e->n_total;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(GCS_ERR_INVALID, std::string("cannot open ") + path);
    if (n == 0 || frames == 0) {
        std::fclose(f);
        return GCS_OK;
    }
    const size_t bytes = 16 * (size_t)n;  // two halves fit the 32 B/particle staging
    uint8_t* host[2] = {nullptr, nullptr};
    hipStream_t cs = nullptr;
    hipEvent_t packed[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
    std::thread writer;
    bool write_failed = false;
    hipError_t he = hipSuccess;
    for (int b = 0; b < 2 && he == hipSuccess; ++b) {
        he = hipHostMalloc((void**)&host[b], bytes, hipHostMallocDefault);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&packed[b], hipEventDisableTiming);
        if (he == hipSuccess) he = hipEventCreateWithFlags(&copied[b], hipEventDisableTiming);
    }
    if (he == hipSuccess) he = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    const auto t0 = std::chrono::steady_clock::now();
    for (int32_t fr = 0; fr < frames && he == hipSuccess; ++fr) {
        const int b = fr & 1;
        rc = gcs_step(e, p, 1);
        if (rc) break;
        uint8_t* stage = e->staging + (size_t)b * bytes;
        if (fr >= 2) he = hipStreamWaitEvent(e->stream, copied[b], 0);
        if (he != hipSuccess) break;
        const int c = e->cur;
        k_pack_playback<<<(unsigned)((n + 255) / 256), 256, 0, e->stream>>>(n, 0, e->posm[c],
                                                                             e->vel[c], stage);
        he = hipGetLastError();
        if (he == hipSuccess) he = hipEventRecord(packed[b], e->stream);
        if (writer.joinable()) writer.join();  // frame fr - 1 is in the file
        if (write_failed || he != hipSuccess) break;
        he = hipStreamWaitEvent(cs, packed[b], 0);
        if (he == hipSuccess) he = hipMemcpyAsync(host[b], stage, bytes, hipMemcpyDeviceToHost, cs);
        if (he == hipSuccess) he = hipEventRecord(copied[b], cs);
        if (he != hipSuccess) break;
        hipEvent_t ev = copied[b];
        uint8_t* buf = host[b];
        writer = std::thread([f, ev, buf, bytes, &write_failed]() {
            write_failed = hipEventSynchronize(ev) != hipSuccess ||
                           std::fwrite(buf, 1, bytes, f) != bytes;
        });
    }
    if (writer.joinable()) writer.join();
    if (he == hipSuccess && rc == GCS_OK) he = hipStreamSynchronize(e->stream);
    if (he == hipSuccess && rc == GCS_OK) he = hipStreamSynchronize(cs);
    if (seconds)
        *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const bool closed = std::fclose(f) == 0;
    if (cs) (void)hipStreamDestroy(cs);
    for (int b = 0; b < 2; ++b) {
        if (packed[b]) (void)hipEventDestroy(packed[b]);
        if (copied[b]) (void)hipEventDestroy(copied[b]);
        if (host[b]) (void)hipHostFree(host[b]);
    }
    if (rc) return rc;
    if (he != hipSuccess) return fail(GCS_ERR_HIP, std::string("bake: ") + hipGetErrorString(he));
    if (write_failed || !closed) return fail(GCS_ERR_INVALID, std::string("writing ") + path + " failed");
    return gcs_sync(e);
}

int gcs_step(gcs_engine* e, const gcs_params* p, int32_t nsteps)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    int rc = check_params(p);
    if (rc) return rc;
    if (nsteps < 0) return fail(GCS_ERR_INVALID, "nsteps < 0");
    if (e->n_total == 0 || nsteps == 0) return GCS_OK;
    HIPCHK(hipSetDevice(e->device));
    e->capture_keys = false;
    if (p->use_graph) {
        if (p->sort == 0 && bucket_segs(e, *p) < (int)e->h_segs.size()) {
            // Instances not in a sorted order (loaded or joined since the
            // last step): one step runs eagerly, its radix sort orders them,
            // and the graphs then sort every instance with buckets.
            e->timed = false;
            rc = enqueue_step(e, *p, true);
            if (rc) return rc;
            if (--nsteps == 0) return GCS_OK;
        }
        if (!e->graph_valid || std::memcmp(&e->gparams, p, sizeof(gcs_params)) != 0 ||
            e->graph_bucket_segs != bucket_segs(e, *p)) {
            e->timed = false;
            rc = capture_graphs(e, *p);
            if (rc) return rc;
        }
        if (!p->use_bounds && !e->bpart_valid) {
            const int qblocks = 4 * (int)e->h_tiles.size();
            k_bounds<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs,
                                                               e->posm[e->cur], e->bpart);
            LAUNCHCHK();
        }
        for (int32_t s = 0; s < nsteps; ++s) {
            HIPCHK(hipGraphLaunch(e->gexec[e->cur], e->stream));
            e->cur ^= 1;
        }
        e->bpart_valid = true;  // every replayed step ends with k_integrate
        e->have_tree = true;
        e->last_mac = p->mac;
        e->wslots = e->graph_wslots;
        return GCS_OK;
    }
    for (int32_t s = 0; s < nsteps; ++s) {
        e->timed = (s == nsteps - 1);
        rc = enqueue_step(e, *p, true);
        e->timed = false;
        if (rc) return rc;
    }
    return GCS_OK;
}

int gcs_sync(gcs_engine* e)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    unsigned h = 0;
    HIPCHK(hipMemcpy(&h, e->err, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (h) {
        HIPCHK(hipMemset(e->err, 0, sizeof(unsigned)));
        if (h & kErrStack) return fail(GCS_ERR_STACK, "traversal stack overflow");
        return fail(GCS_ERR_NONFINITE, "non-finite acceleration");
    }
    return GCS_OK;
}

int gcs_compute_accels(gcs_engine* e, const gcs_params* p)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    int rc = check_params(p);
    if (rc) return rc;
    if (e->n_total == 0) return GCS_OK;
    HIPCHK(hipSetDevice(e->device));
    e->capture_keys = true;
    e->timed = true;
    rc = enqueue_step(e, *p, false);
    e->capture_keys = false;
    e->timed = false;
    if (rc) return rc;
    return gcs_sync(e);
}

int gcs_autotune(gcs_engine* e, const gcs_params* base, int32_t reps, gcs_params* best,
                 float* best_ms)
{
    if (!e || !best) return fail(GCS_ERR_INVALID, "null argument");
    int rc = check_params(base);
    if (rc) return rc;
    if (reps < 1 || reps > 100) return fail(GCS_ERR_INVALID, "reps must be in [1, 100]");
    *best = *base;
    if (best_ms) *best_ms = 0.0f;
    if (e->n_total == 0) return GCS_OK;
    const TuneKey key = tune_key(e, *base);
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            apply_tune(it->second, best);
            if (best_ms) *best_ms = it->second.ms;
            return check_params(best);
        }
    }
    HIPCHK(hipSetDevice(e->device));
    rc = gcs_sync(e);
    if (rc) return rc;
    // Candidate settings: the group traversal's group size, leaf size, block
    // size (or teams of two or four waves per group) and waves per SIMD; the reference
    // traversal's stack depth.  The leaf size barely moves the error (2 x
    // 1,048,576 particles at theta = 0.5: 2.7e-4 relative RMS at leaf 4,
    // 2.6e-4 at leaf 16).
    std::vector<gcs_params> cand;
    if (base->mac == GCS_MAC_GROUP) {
        for (int group : {48, 64, 96, 128})
            for (int leaf : {4, 8, 12, 16})
                for (int block : {-4, -2, 64, 128, 256})  // -t: teams of t waves
                    for (int wps : {2, 3, 4}) {
                        gcs_params q = *base;
                        q.group_size = group;
                        q.leaf_size = leaf;
                        q.team = block < 0 ? -block : 1;
                        q.block_size = block < 0 ? -block * kWave : block;
                        q.waves_per_simd = wps;
                        cand.push_back(q);
                    }
    } else {
        for (int depth : {64, 128, 256, 1024}) {
            gcs_params q = *base;
            q.stack_depth = depth;
            cand.push_back(q);
        }
    }
    float best_t = 0.0f;
    bool found = false;
    for (gcs_params q : cand) {
        q.use_graph = 0;
        bool ok = true;
        float sum = 0.0f;
        for (int r = 0; r <= reps && ok; ++r) {
            e->capture_keys = false;
            e->timed = true;
            rc = enqueue_step(e, q, false);
            e->timed = false;
            if (rc) return rc;
            HIPCHK(hipStreamSynchronize(e->stream));
            unsigned h = 0;
            HIPCHK(hipMemcpy(&h, e->err, sizeof(unsigned), hipMemcpyDeviceToHost));
            if (h) {
                HIPCHK(hipMemset(e->err, 0, sizeof(unsigned)));
                ok = false;
                break;
            }
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[6]));
            if (r > 0) sum += ms;  // the first pass warms up
        }
        if (!ok) continue;
        const float t = sum / (float)reps;
        if (!found || t < best_t) {
            best_t = t;
            *best = q;
            best->use_graph = base->use_graph;
            found = true;
        }
    }
    if (!found) return fail(GCS_ERR_STACK, "every candidate overflowed its traversal stack");
    if (best_ms) *best_ms = best_t;
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tune[key] = TuneVal{best->group_size, best->leaf_size, best->block_size, best->team,
                              best->waves_per_simd, best->stack_depth, best_t};
    }
    return GCS_OK;
}

int gcs_autotune_cache_clear(void)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
    return GCS_OK;
}

int gcs_autotune_cache_save(const char* path)
{
    if (!path) return fail(GCS_ERR_INVALID, "null path");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(GCS_ERR_INVALID, std::string("cannot write ") + path);
    std::fprintf(f, "gcs-autotune 2\n");
    std::fprintf(f, "# mac open_offset precision log2_n log2_k theta_bits mono_far_bits "
                    "local_kappa_bits group leaf block team waves_per_simd stack_depth ms\n");
    for (const auto& kv : g_tune) {
        const TuneKey& k = kv.first;
        const TuneVal& v = kv.second;
        std::fprintf(f, "%d %d %d %d %d %08x %08x %08x %d %d %d %d %d %d %.6g\n", k.mac,
                     k.open_offset, k.precision, k.nbucket, k.kbucket, k.theta_bits, k.far_bits,
                     k.kappa_bits, v.group, v.leaf, v.block, v.team, v.wps, v.stack,
                     (double)v.ms);
    }
    const bool ok = std::fclose(f) == 0;
    return ok ? GCS_OK : fail(GCS_ERR_INVALID, std::string("cannot write ") + path);
}

int gcs_autotune_cache_load(const char* path)
{
    if (!path) return fail(GCS_ERR_INVALID, "null path");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(GCS_ERR_NOT_FOUND, std::string("cannot read ") + path);
    char line[512];
    std::map<TuneKey, TuneVal> got;
    bool header = false, bad = false;
    while (std::fgets(line, sizeof(line), f)) {
        if (!header) {
            header = std::strncmp(line, "gcs-autotune 2", 14) == 0;
            if (!header) { bad = true; break; }
            continue;
        }
        if (line[0] == '#' || line[0] == '\n') continue;
        TuneKey k;
        TuneVal v;
        double ms = 0.0;
        if (std::sscanf(line, "%d %d %d %d %d %x %x %x %d %d %d %d %d %d %lf", &k.mac,
                        &k.open_offset, &k.precision, &k.nbucket, &k.kbucket, &k.theta_bits,
                        &k.far_bits, &k.kappa_bits, &v.group, &v.leaf, &v.block, &v.team, &v.wps,
                        &v.stack, &ms) != 15) {
            bad = true;
            break;
        }
        v.ms = (float)ms;
        got[k] = v;
    }
    std::fclose(f);
    if (bad || !header) return fail(GCS_ERR_INVALID, std::string("not an autotune table: ") + path);
    std::lock_guard<std::mutex> lock(g_tune_mu);
    for (const auto& kv : got) g_tune[kv.first] = kv.second;
    return GCS_OK;
}

int gcs_read_accels(gcs_engine* e, float* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    HIPCHK(hipSetDevice(e->device));
    if (e->n_total)
        HIPCHK(hipMemcpyAsync(out, e->acc, 16 * (size_t)e->n_total, hipMemcpyDeviceToHost,
                              e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return GCS_OK;
}

int gcs_read_keys(gcs_engine* e, uint64_t* keys_unsorted, uint64_t* keys_sorted, int32_t* perm)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    HIPCHK(hipSetDevice(e->device));
    const size_t n = (size_t)e->n_total;
    if (n == 0) return GCS_OK;
    if (keys_unsorted)
        HIPCHK(hipMemcpyAsync(keys_unsorted, e->keys_raw, 8 * n, hipMemcpyDeviceToHost, e->stream));
    if (keys_sorted)
        HIPCHK(hipMemcpyAsync(keys_sorted, e->keys[0], 8 * n, hipMemcpyDeviceToHost, e->stream));
    if (perm)
        HIPCHK(hipMemcpyAsync(perm, e->vals[0], 4 * n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (perm)
        for (const SegInfo& sg : e->h_segs)
            for (int i = sg.start; i < sg.start + sg.n; ++i) perm[i] -= sg.start;
    return GCS_OK;
}

int gcs_read_tree(gcs_engine* e, int32_t id, int32_t* left, int32_t* right, int32_t* parent,
                  float* com4, float* width, float* bmin, float* bmax)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    if (!e->have_tree) return fail(GCS_ERR_INVALID, "no tree has been built");
    HIPCHK(hipSetDevice(e->device));
    const SegInfo sg = e->h_segs[slot];
    const int n = sg.n, nodes = 2 * n - 1, base = sg.root;
    std::vector<float4> AB((size_t)2 * nodes), Q((size_t)2 * nodes), lo((size_t)nodes),
        hi((size_t)nodes), P((size_t)n);
    std::vector<int> par((size_t)nodes);
    HIPCHK(hipMemcpyAsync(AB.data(), e->node + 2 * (size_t)base, 32 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(Q.data(), e->nodeQ + 2 * (size_t)base, 32 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(lo.data(), e->boxLo + base, 16 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(hi.data(), e->boxHi + base, 16 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(par.data(), e->parent + base, 4 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(P.data(), e->posm[e->cur] + sg.start, 16 * (size_t)n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int v = 0; v < nodes; ++v) {
        const bool leaf = v >= n - 1;
        if (parent) parent[v] = par[(size_t)v] < 0 ? -1 : par[(size_t)v] - base;
        if (leaf) {
            const float4 p = P[(size_t)(v - (n - 1))];
            if (left) left[v] = -1;
            if (right) right[v] = -1;
            if (com4) { com4[4 * v] = p.x; com4[4 * v + 1] = p.y; com4[4 * v + 2] = p.z; com4[4 * v + 3] = p.w; }
            if (width) width[v] = 0.0f;
            if (bmin) { bmin[3 * v] = p.x; bmin[3 * v + 1] = p.y; bmin[3 * v + 2] = p.z; }
            if (bmax) { bmax[3 * v] = p.x; bmax[3 * v + 1] = p.y; bmax[3 * v + 2] = p.z; }
            continue;
        }
        const float4 bf = AB[(size_t)2 * v + 1];
        int4 b;
        std::memcpy(&b, &bf, sizeof(b));
        const int f = b.y - sg.start, l = b.z - sg.start, s = b.w - sg.start;
        if (left) left[v] = (f == s) ? (n - 1) + s : s;
        if (right) right[v] = (l == s + 1) ? (n - 1) + s + 1 : s + 1;
        if (com4) { const float4 a = AB[(size_t)2 * v]; com4[4 * v] = a.x; com4[4 * v + 1] = a.y; com4[4 * v + 2] = a.z; com4[4 * v + 3] = a.w; }
        if (width) width[v] = Q[(size_t)2 * v + 1].z;
        if (bmin) { bmin[3 * v] = lo[(size_t)v].x; bmin[3 * v + 1] = lo[(size_t)v].y; bmin[3 * v + 2] = lo[(size_t)v].z; }
        if (bmax) { bmax[3 * v] = hi[(size_t)v].x; bmax[3 * v + 1] = hi[(size_t)v].y; bmax[3 * v + 2] = hi[(size_t)v].z; }
    }
    return GCS_OK;
}

int gcs_read_root(gcs_engine* e, int32_t id, float* com4)
{
    if (!e || !com4) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    if (!e->have_tree) return fail(GCS_ERR_INVALID, "no tree has been built");
    HIPCHK(hipSetDevice(e->device));
    // One 16-byte copy of the root's (COM, mass): an internal node's merge,
    // or for a single particle its leaf record as the build wrote it.
    float4 a;
    HIPCHK(hipMemcpyAsync(&a, e->node + 2 * (size_t)e->h_segs[slot].root, sizeof(float4),
                          hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    com4[0] = a.x;
    com4[1] = a.y;
    com4[2] = a.z;
    com4[3] = a.w;
    return GCS_OK;
}

int gcs_read_bounds(gcs_engine* e, int32_t id, float* mn, float* mx)
{
    if (!e) return fail(GCS_ERR_INVALID, "engine is null");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    uint32_t b[2];
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(b, e->seg_bounds + 2 * slot, 8, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (mn) *mn = fdec(b[0]);
    if (mx) *mx = fdec(b[1]);
    return GCS_OK;
}

// Interaction counts of the last traversal: per-target node, particle, of
// the nodes the monopoles, local-expansion (group, node) pairs, the target
// interactions those stand for.
static int read_counts(gcs_engine* e, unsigned long long* c5)
{
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int k = 0; k < kCounts; ++k) c5[k] = 0;
    if (e->last_mac == GCS_MAC_GROUP) {
        std::vector<unsigned long long> w((size_t)kCounts * e->wslots);
        if (!w.empty())
            HIPCHK(hipMemcpy(w.data(), e->wcount, sizeof(unsigned long long) * w.size(),
                             hipMemcpyDeviceToHost));
        for (size_t i = 0; i < w.size(); i += kCounts)
            for (int k = 0; k < kCounts; ++k) c5[k] += w[i + k];
    } else {
        HIPCHK(hipMemcpy(c5, e->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    }
    return GCS_OK;
}

int gcs_read_interaction_counts(gcs_engine* e, int64_t* out5)
{
    if (!e || !out5) return fail(GCS_ERR_INVALID, "null argument");
    unsigned long long c[kCounts];
    int rc = read_counts(e, c);
    if (rc) return rc;
    for (int k = 0; k < kCounts; ++k) out5[k] = (int64_t)c[k];
    return GCS_OK;
}

int gcs_read_interactions(gcs_engine* e, int64_t* out2)
{
    if (!e || !out2) return fail(GCS_ERR_INVALID, "null argument");
    unsigned long long c[kCounts];
    int rc = read_counts(e, c);
    if (rc) return rc;
    out2[0] = (int64_t)c[0];
    out2[1] = (int64_t)c[1];
    return GCS_OK;
}

int gcs_read_monopole_interactions(gcs_engine* e, int64_t* out)
{
    if (!e || !out) return fail(GCS_ERR_INVALID, "null argument");
    unsigned long long c[kCounts];
    int rc = read_counts(e, c);
    if (rc) return rc;
    *out = (int64_t)c[2];
    return GCS_OK;
}

int gcs_read_diagnostics(gcs_engine* e, int32_t id, double* out5)
{
    if (!e || !out5) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(out5, e->diag_out + 5 * slot, 5 * sizeof(double), hipMemcpyDeviceToHost));
    return GCS_OK;
}

int gcs_read_max_accel(gcs_engine* e, int32_t id, float* amax)
{
    if (!e || !amax) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    const SegInfo& sg = e->h_segs[slot];
    std::vector<float> part((size_t)4 * sg.ntiles);
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(part.data(), e->amax_part + 4 * (size_t)sg.tile0, sizeof(float) * part.size(),
                     hipMemcpyDeviceToHost));
    float m = 0.0f;
    for (float v : part) m = std::max(m, v);
    *amax = m;
    return GCS_OK;
}

int gcs_read_sim_time(gcs_engine* e, int32_t id, double* t)
{
    if (!e || !t) return fail(GCS_ERR_INVALID, "null argument");
    const int slot = find_slot(e, id);
    if (slot < 0) return fail(GCS_ERR_NOT_FOUND, "unknown instance id");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipMemcpy(t, e->tsim + slot, sizeof(double), hipMemcpyDeviceToHost));
    return GCS_OK;
}

int gcs_engine_info(gcs_engine* e, int64_t* out8)
{
    if (!e || !out8) return fail(GCS_ERR_INVALID, "null argument");
    out8[0] = e->num_cu;
    out8[1] = e->trav_blocks_cap;
    out8[2] = (int64_t)e->trav_slots;
    out8[3] = (int64_t)e->trav_slots * kSlotInts * 4;
    out8[4] = e->cap_n;
    out8[5] = e->cap_k;
    int ng = 0;
    if (e->have_tree) {
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipMemcpy(&ng, e->ngroups, sizeof(int), hipMemcpyDeviceToHost));
    }
    out8[6] = ng;
    out8[7] = (int64_t)e->h_tiles.size();
    return GCS_OK;
}

int gcs_read_stage_ms(gcs_engine* e, float* out7)
{
    if (!e || !out7) return fail(GCS_ERR_INVALID, "null argument");
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int k = 0; k < 7; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e->ev[k], e->ev[k + 1]) != hipSuccess) ms = -1.0f;
        out7[k] = ms;
    }
    return GCS_OK;
}

// ---- Stand-alone stages ------------------------------------------------------

int gcs_morton_keys(const float* x, const float* y, const float* z, int64_t n, float mn, float mx,
                    uint64_t* keys)
{
    if (n < 0 || (n > 0 && (!x || !y || !z || !keys))) return fail(GCS_ERR_INVALID, "bad argument");
    if (n == 0) return GCS_OK;
    if (n > kMaxParticles) return fail(GCS_ERR_CAPACITY, "too many keys (at most 2^29)");
    gcs_engine* e = nullptr;
    int rc = gcs_create(n, 1, 0, &e);
    if (rc) return rc;
    std::vector<gcs_star> st((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        std::memset(&st[(size_t)i], 0, sizeof(gcs_star));
        st[(size_t)i].x = x[i]; st[(size_t)i].y = y[i]; st[(size_t)i].z = z[i];
        st[(size_t)i].mass = 1.0f;
    }
    int32_t cnt = (int32_t)n;
    rc = gcs_load(e, st.data(), &cnt, 1);
    if (!rc) {
        const int qblocks = 4 * (int)e->h_tiles.size();
        StepReset none;
        std::memset(&none, 0, sizeof(none));
        k_morton<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->posm[e->cur],
                                                           e->seg_bounds, 1, mn, mx, e->keys[0],
                                                           e->vals[0], 0, none);
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = fail(GCS_ERR_HIP, hipGetErrorString(le));
        else if (hipMemcpyAsync(keys, e->keys[0], 8 * (size_t)n, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                 hipStreamSynchronize(e->stream) != hipSuccess)
            rc = fail(GCS_ERR_HIP, "copy of keys failed");
    }
    gcs_destroy(e);
    return rc;
}

int gcs_sort_pairs(const uint64_t* keys, int64_t n, uint64_t* keys_out, int32_t* perm_out)
{
    if (n < 0 || (n > 0 && (!keys || !keys_out || !perm_out))) return fail(GCS_ERR_INVALID, "bad argument");
    if (n == 0) return GCS_OK;
    if (n > kMaxParticles) return fail(GCS_ERR_CAPACITY, "too many keys (at most 2^29)");
    gcs_engine* e = nullptr;
    int rc = gcs_create(n, 1, 0, &e);
    if (rc) return rc;
    e->inst.assign(1, Instance{0, (int32_t)n});
    rc = rebuild_tables(e);
    if (!rc) {
        std::vector<uint32_t> iota((size_t)n);
        for (int64_t i = 0; i < n; ++i) iota[(size_t)i] = (uint32_t)i;
        if (hipMemcpy(e->keys[0], keys, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(e->vals[0], iota.data(), 4 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
            rc = fail(GCS_ERR_HIP, "upload failed");
    }
    if (!rc) {
        hipError_t le = hipMemsetAsync(e->chist, 0,
                                       sizeof(uint32_t) * kSortPasses * kRadix * e->nchunks,
                                       e->stream);
        if (le == hipSuccess) {
            enqueue_sort(e);
            le = hipGetLastError();
        }
        if (le != hipSuccess) rc = fail(GCS_ERR_HIP, hipGetErrorString(le));
        else if (hipMemcpyAsync(keys_out, e->keys[0], 8 * (size_t)n, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                 hipMemcpyAsync(perm_out, e->vals[0], 4 * (size_t)n, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
                 hipStreamSynchronize(e->stream) != hipSuccess)
            rc = fail(GCS_ERR_HIP, "copy back failed");
    }
    gcs_destroy(e);
    return rc;
}

int gcs_karras_tree(const uint64_t* sorted_keys, int32_t n, int32_t* left, int32_t* right,
                    int32_t* parent)
{
    if (n < 1 || !sorted_keys || !left || !right || !parent) return fail(GCS_ERR_INVALID, "bad argument");
    gcs_engine* e = nullptr;
    int rc = gcs_create(n, 1, 0, &e);
    if (rc) return rc;
    e->inst.assign(1, Instance{0, n});
    rc = rebuild_tables(e);
    if (!rc && hipMemcpy(e->keys[0], sorted_keys, 8 * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(GCS_ERR_HIP, "upload failed");
    if (!rc) {
        const int qblocks = 4 * (int)e->h_tiles.size();
        k_build<<<qblocks, kTileThreads, 0, e->stream>>>(e->tiles, e->segs, e->keys[0], e->node,
                                                         e->parent, e->flag, group_marks(e),
                                                         atile_marks(e), kGroup,
                                                         TopList{nullptr, nullptr, nullptr,
                                                                 nullptr, nullptr});
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) rc = fail(GCS_ERR_HIP, hipGetErrorString(le));
    }
    if (!rc) {
        const int nodes = 2 * n - 1;
        std::vector<int4> B((size_t)2 * nodes);
        if (hipMemcpyAsync(B.data(), e->node, 32 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipMemcpyAsync(parent, e->parent, 4 * (size_t)nodes, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) {
            rc = fail(GCS_ERR_HIP, "copy back failed");
        } else {
            for (int v = 0; v < nodes; ++v) {
                if (v >= n - 1) {
                    left[v] = right[v] = -1;
                    continue;
                }
                const int f = B[(size_t)2 * v + 1].y, l = B[(size_t)2 * v + 1].z,
                          s = B[(size_t)2 * v + 1].w;
                left[v] = (f == s) ? (n - 1) + s : s;
                right[v] = (l == s + 1) ? (n - 1) + s + 1 : s + 1;
            }
        }
    }
    gcs_destroy(e);
    return rc;
}

}  // extern "C"

// ============================================================================
// Allocation-free sort (gcs_sort_pairs_noalloc)
// ============================================================================
// The sort's working memory -- the second key/payload buffer of the
// ping-pong, the per-tile and per-chunk digit histograms and the tile table --
// sits in one file-scope device array, so a call allocates nothing.  Its
// name is a host-side stand-in: host code uses only the address
// hipGetSymbolAddress returns for the current device (kept per device).
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two threads) race on it.
namespace {

constexpr int64_t kNaMaxKeys = GCS_NOALLOC_MAX_KEYS;
constexpr int kNaMaxTiles = (int)(kNaMaxKeys / kTile);
constexpr int kNaMaxChunks = (kNaMaxTiles + kChunkTiles - 1) / kChunkTiles;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaKeys = 0;
constexpr size_t kNaVals = kNaKeys + na_align(sizeof(uint64_t) * kNaMaxKeys);
constexpr size_t kNaHist = kNaVals + na_align(sizeof(uint32_t) * kNaMaxKeys);
constexpr size_t kNaChist = kNaHist + na_align(sizeof(uint32_t) * kRadix * kNaMaxTiles);
constexpr size_t kNaTiles =
    kNaChist + na_align(sizeof(uint32_t) * kSortPasses * kRadix * kNaMaxChunks);
constexpr size_t kNaSegs = kNaTiles + na_align(sizeof(TileInfo) * kNaMaxTiles);
constexpr size_t kNaBytes = kNaSegs + na_align(sizeof(SegInfo));
static_assert(kNaMaxKeys % kSortTileBig == 0, "whole sort tiles");

}  // namespace

// 53 MiB per device, held until the process ends.
static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
void* g_scratch_dev[kNaMaxDevices] = {};
std::mutex g_scratch_mu;

// The scratch array's device address on the current device (looked up once
// per device, with that device current).
int scratch_address(uint8_t** out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNaMaxDevices) return fail(GCS_ERR_NO_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GCS_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                         hipGetErrorString(e));
        g_scratch_dev[dev] = p;
    }
    *out = static_cast<uint8_t*>(g_scratch_dev[dev]);
    return GCS_OK;
}

// Set-up of one sort of n keys: the sort-tile table and the instance record
// of a single instance, the chunk histograms of all passes cleared, and the
// payload 0 .. n-1 written into perm (the first pass's payload input).  A
// kernel rather than hipMemsetAsync, so the call can be captured into a graph.
__global__ __launch_bounds__(256) void k_noalloc_setup(int n, int ntiles, int nstiles,
                                                       int st_keys, int chist_words,
                                                       TileInfo* __restrict__ tiles,
                                                       SegInfo* __restrict__ segs,
                                                       uint32_t* __restrict__ chist,
                                                       uint32_t* __restrict__ perm)
{
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int i = tid; i < n; i += stride) perm[i] = (uint32_t)i;
    for (int k = tid; k < chist_words; k += stride) chist[k] = 0u;
    for (int t = tid; t < nstiles; t += stride) {
        TileInfo ti;
        ti.start = t * st_keys;
        ti.count = min(st_keys, n - t * st_keys);
        ti.seg = 0;
        ti.tile_in_seg = t;
        ti.pad0 = ti.pad1 = ti.pad2 = ti.pad3 = 0;
        tiles[t] = ti;
    }
    if (tid == 0) {
        SegInfo sg;
        sg.start = 0;
        sg.n = n;
        sg.int_base = 0;
        sg.leaf_base = n - 1;
        sg.root = 0;
        sg.tile0 = 0;
        sg.ntiles = ntiles;
        sg.chunk0 = 0;
        segs[0] = sg;
    }
}

}  // namespace

extern "C" int gcs_sort_pairs_noalloc(const uint64_t* keys, int64_t n, uint64_t* keys_out,
                                      int32_t* perm_out, void* hip_stream)
{
    if (n < 0) return fail(GCS_ERR_INVALID, "n < 0");
    if (n > kNaMaxKeys)
        return fail(GCS_ERR_CAPACITY, "gcs_sort_pairs_noalloc sorts at most " +
                                          std::to_string(kNaMaxKeys) + " keys");
    if (n == 0) return GCS_OK;
    if (!keys || !keys_out || !perm_out) return fail(GCS_ERR_INVALID, "null argument");
    uint8_t* base = nullptr;
    int rc = scratch_address(&base);
    if (rc) return rc;
    hipStream_t st = static_cast<hipStream_t>(hip_stream);
    const int nk = (int)n;
    const int ntiles = (nk + kTile - 1) / kTile;
    const bool big = nk >= kSortBigMin;
    const int st_keys = big ? kSortTileBig : kTile;
    const int nstiles = (nk + st_keys - 1) / st_keys;
    const int nchunks = (ntiles + kChunkTiles - 1) / kChunkTiles;
    const size_t cstride = (size_t)nchunks * kRadix;
    uint64_t* sk = reinterpret_cast<uint64_t*>(base + kNaKeys);
    uint32_t* sv = reinterpret_cast<uint32_t*>(base + kNaVals);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + kNaHist);
    uint32_t* chist = reinterpret_cast<uint32_t*>(base + kNaChist);
    TileInfo* tiles = reinterpret_cast<TileInfo*>(base + kNaTiles);
    SegInfo* segs = reinterpret_cast<SegInfo*>(base + kNaSegs);
    uint32_t* pv = reinterpret_cast<uint32_t*>(perm_out);
    k_noalloc_setup<<<std::min(1024, (nk + 255) / 256), 256, 0, st>>>(
        nk, ntiles, nstiles, st_keys, (int)(kSortPasses * cstride), tiles, segs, chist, pv);
    LAUNCHCHK();
    // Even passes read the caller's buffers (the input keys first, then
    // keys_out) and write the scratch pair; odd passes write keys_out and
    // perm_out, so the eighth pass leaves the result there.
    static_assert(kSortPasses % 2 == 0, "the last pass writes the caller's buffers");
    for (int pass = 0; pass < kSortPasses; ++pass) {
        const int shift = pass * kRadixBits;
        const bool even = (pass & 1) == 0;
        const uint64_t* kin = even ? (pass == 0 ? keys : keys_out) : sk;
        const uint32_t* vin = even ? pv : sv;
        uint64_t* kdst = even ? sk : keys_out;
        uint32_t* vdst = even ? sv : pv;
        uint32_t* ch = chist + (size_t)pass * cstride;
        if (big) {
            k_sort_count<1024><<<nstiles, 1024, 0, st>>>(tiles, segs, kin, hist, ch, shift);
            k_sort_scatter<1024><<<nstiles, 1024, 0, st>>>(tiles, segs, kin, vin, kdst, vdst,
                                                            hist, ch, shift);
        } else {
            k_sort_count<kTileThreads><<<nstiles, kTileThreads, 0, st>>>(tiles, segs, kin, hist,
                                                                        ch, shift);
            k_sort_scatter<kTileThreads><<<nstiles, kTileThreads, 0, st>>>(
                tiles, segs, kin, vin, kdst, vdst, hist, ch, shift);
        }
    }
    LAUNCHCHK();
    return GCS_OK;
}
