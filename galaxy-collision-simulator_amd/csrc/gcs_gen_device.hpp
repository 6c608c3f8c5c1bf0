// gcs_gen_device.hpp -- the seeded two-galaxy initial conditions generated on
// the device, straight into an engine's state.
//
// The model is gcs_gen_common.hpp's, shared with the host generator; the
// launches are:
//   k_gen_pos    one thread per particle: its position and mass (FP64, its
//                own stream), kept in the staging buffer, and a sort key
//                (galaxy, radius bits) with the particle index as payload
//   (the engine's segmented radix sort orders every instance by galaxy, then
//    radius, ties by index -- the host's stable sort by radius)
//   k_gen_count  per sorted tile: particles of each (galaxy, component)
//   k_gen_scan   per instance: exclusive prefix of those counts over its tiles
//   k_gen_menc   per sorted tile: each particle's counts at smaller radius,
//                hence its enclosed mass
//   k_gen_state  one thread per particle: velocity, galaxy B's rotation and
//                offset, colour and type, written as the engine's state.
#pragma once
#include "gcs_device.hpp"
#include "gcs_gen_common.hpp"

namespace gcs {

// One collision of a generated load.
struct GenInst {
    uint64_t seed;
    int32_t n_a, n_b;
    float off[3], vel[3];
    double ct, st;  // cosine and sine of galaxy B's tilt
};

// Galaxy, index within it and component of particle `local` of an instance.
struct GenWho {
    int g;
    int64_t i;
    int comp;
    gcs_gen::GalaxySizes gs;
};

__device__ __forceinline__ GenWho gen_who(const GenInst& I, int local)
{
    GenWho w;
    w.g = local < I.n_a ? 0 : 1;
    w.i = w.g ? local - I.n_a : local;
    w.gs = gcs_gen::galaxy_sizes(w.g ? I.n_b : I.n_a);
    w.comp = w.i < w.gs.nh ? 0 : (w.i < w.gs.nh + w.gs.nd ? 1 : 2);
    return w;
}

// Every kernel takes the tiles of the instances generated (from the first
// tile of the first of them) and their collisions as insts[slot - seg_base].
__global__ __launch_bounds__(256) void k_gen_pos(const TileInfo* __restrict__ tiles,
                                                 const SegInfo* __restrict__ segs,
                                                 const GenInst* __restrict__ insts, int seg_base,
                                                 gcs_gen::MnTableView mt,
                                                 double4* __restrict__ xyz,
                                                 uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    const GenInst I = insts[q.seg - seg_base];
    const GenWho w = gen_who(I, q.local);
    const gcs_gen::GenParticle p = gcs_gen::gen_position(I.seed, w.g, w.i, w.gs, mt);
    const double r = gcs_gen::gen_radius(p);
    xyz[q.i] = make_double4(p.x, p.y, p.z, 0.0);
    keys[q.i] = ((uint64_t)w.g << 63) | (uint64_t)__double_as_longlong(r);
    vals[q.i] = (uint32_t)q.i;
}

// Counts of each (galaxy, component) among a sorted tile's particles, in
// the tile's row of `cnt` (entries 0..5 = galaxy * 3 + component).
__global__ __launch_bounds__(kTileThreads) void k_gen_count(const TileInfo* __restrict__ tiles,
                                                            const SegInfo* __restrict__ segs,
                                                            const GenInst* __restrict__ insts,
                                                            int seg_base,
                                                            const uint32_t* __restrict__ vals,
                                                            uint32_t* __restrict__ cnt)
{
    __shared__ uint32_t c[6];
    const TileInfo t = tiles[blockIdx.x];
    if (threadIdx.x < 6) c[threadIdx.x] = 0u;
    __syncthreads();
    const GenInst I = insts[t.seg - seg_base];
    const int start = segs[t.seg].start;
    for (int k = threadIdx.x; k < t.count; k += kTileThreads) {
        const GenWho w = gen_who(I, (int)vals[t.start + k] - start);
        atomicAdd(&c[w.g * 3 + w.comp], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 6) cnt[(size_t)blockIdx.x * kRadix + threadIdx.x] = c[threadIdx.x];
}

// Per instance (one thread each, instances seg_base .. seg_base + nseg - 1):
// exclusive prefix of the tile counts over its tiles, written to entries
// 8..13 of each tile's row (rows from the first generated tile, tile_base).
__global__ void k_gen_scan(const SegInfo* __restrict__ segs, int seg_base, int nseg,
                           int tile_base, uint32_t* __restrict__ cnt)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    const SegInfo sg = segs[seg_base + s];
    uint32_t run[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    for (int t = sg.tile0; t < sg.tile0 + sg.ntiles; ++t) {
        uint32_t* row = cnt + (size_t)(t - tile_base) * kRadix;
        for (int k = 0; k < 6; ++k) {
            const uint32_t v = row[k];
            row[8 + k] = run[k];
            run[k] += v;
        }
    }
}

// Enclosed mass of every particle: the counts of its galaxy's components at
// smaller sorted position (tile prefix + earlier threads + earlier items).
__global__ __launch_bounds__(kTileThreads) void k_gen_menc(const TileInfo* __restrict__ tiles,
                                                           const SegInfo* __restrict__ segs,
                                                           const GenInst* __restrict__ insts,
                                                           int seg_base,
                                                           const uint32_t* __restrict__ vals,
                                                           const uint32_t* __restrict__ cnt,
                                                           double4* __restrict__ xyz)
{
    __shared__ uint32_t ws[kTileThreads / kWave];
    const TileInfo t = tiles[blockIdx.x];
    const GenInst I = insts[t.seg - seg_base];
    const int start = segs[t.seg].start;
    // Thread k holds sorted items 4k .. 4k+3 of the tile.
    GenWho who[kTileItems];
    uint32_t mine[6] = {0u, 0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        const int k = (int)threadIdx.x * kTileItems + r;
        who[r].g = -1;
        if (k < t.count) {
            who[r] = gen_who(I, (int)vals[t.start + k] - start);
            ++mine[who[r].g * 3 + who[r].comp];
        }
    }
    uint32_t before[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        uint32_t total;
        before[c] = block_exclusive_scan(mine[c], &total, ws) +
                    cnt[(size_t)blockIdx.x * kRadix + 8 + c];
    }
#pragma unroll
    for (int r = 0; r < kTileItems; ++r) {
        if (who[r].g < 0) continue;
        const uint32_t* b = before + who[r].g * 3;
        const int k = (int)threadIdx.x * kTileItems + r;
        const int idx = (int)vals[t.start + k];
        xyz[idx].w = gcs_gen::enclosed_mass(who[r].gs, b[0], b[1], b[2]);
        ++before[who[r].g * 3 + who[r].comp];
    }
}

// Velocities and the records, as the engine's state (current order = the
// generation order, galaxy A then B, each halo, disc, bulge).
__global__ __launch_bounds__(256) void k_gen_state(const TileInfo* __restrict__ tiles,
                                                   const SegInfo* __restrict__ segs,
                                                   const GenInst* __restrict__ insts, int seg_base,
                                                   const double4* __restrict__ xyz,
                                                   float4* __restrict__ posm,
                                                   float4* __restrict__ vel,
                                                   float4* __restrict__ plo)
{
    const QuarterTile q = quarter_tile(tiles, segs);
    if (!q.valid) return;
    const GenInst I = insts[q.seg - seg_base];
    const GenWho w = gen_who(I, q.local);
    const double4 d = xyz[q.i];
    gcs_gen::GenParticle p;
    p.x = d.x;
    p.y = d.y;
    p.z = d.z;
    p.comp = w.comp;
    p.m = w.comp == 0 ? gcs_gen::kHaloM / w.gs.nh
                      : (w.comp == 1 ? gcs_gen::kDiscM / w.gs.nd : gcs_gen::kBulgeM / w.gs.nb);
    double vx, vy, vz;
    gcs_gen::gen_velocity(I.seed, w.g, w.i, p, gcs_gen::gen_radius(p), d.w, vx, vy, vz);
    const gcs_gen::GenStar o = gcs_gen::gen_star(p, w.g, vx, vy, vz, I.ct, I.st, I.off, I.vel);
    posm[q.i] = make_float4(o.x, o.y, o.z, o.m);
    vel[q.i] = make_float4(o.vx, o.vy, o.vz, __uint_as_float(pack_rgbt(o.type, o.r, o.g, o.b)));
    plo[q.i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

}  // namespace gcs
