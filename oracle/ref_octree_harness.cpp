// ref_octree_harness.cpp -- drives the reference's own CPU octree code.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (engine.cpp, gravity_aggregator.cpp and
// the headers they include), into oracle/_ref/.  Nothing of the reference is
// copied here: this file only calls insert_star, node_physics and
// query_gravity in the order of the reference's CPU frame (main.cpp:152-253),
// so the oracle's restatement (gcs_oracle.c, orc_octree_step) can be checked
// against the reference's compiled code, and bench.py can time it as the CPU
// baseline.
//
// One change from main.cpp:184: the Morton sort is std::stable_sort rather
// than the unstable __gnu_parallel::sort, so that equal keys keep input order
// and the frame is reproducible.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <utility>
#include <vector>

#include "arena_allocator.h"
#include "engine.h"
#include "gravity_aggregator.h"
#include "morton.h"
#include "types.h"

namespace {

struct Scratch {
    // main.cpp:102-108 sizes.
    ArenaAllocator arena{1024u * 1024u * 64u};
    size_t max_nodes = (1024u * 1024u * 64u) / sizeof(OctreeNode);
    std::vector<uint32_t> leaf_idx;
    std::vector<float> nm, ncx, ncy, ncz;
    Scratch()
        : leaf_idx(max_nodes * 32), nm(max_nodes), ncx(max_nodes), ncy(max_nodes), ncz(max_nodes)
    {
    }
};

double secs(std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b)
{
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

extern "C" int ref_star_size() { return (int)sizeof(Star); }

// The reference's Morton encoding (morton.h get_morton_code), n points.
extern "C" void ref_morton_codes(const float* x, const float* y, const float* z, int n,
                                 float min_val, float max_val, uint64_t* out)
{
    for (int i = 0; i < n; ++i) out[i] = get_morton_code(x[i], y[i], z[i], min_val, max_val);
}

// One CPU frame of the reference (main.cpp:152-253) on `stars` in place.
// acc_out (n*3 floats, post-sort order) and phase_s[6] are optional.
// Returns 0, or -1 when the arena runs out.
extern "C" int ref_octree_step(Star* stars, int n, float G, float eps2, float dt,
                               int do_integrate, float* acc_out, double* phase_s)
{
    static Scratch* s = nullptr;
    if (!s) s = new Scratch();
    auto t0 = std::chrono::steady_clock::now();
    float min_x = std::numeric_limits<float>::max(), min_y = min_x, min_z = min_x;
    float max_x = std::numeric_limits<float>::lowest(), max_y = max_x, max_z = max_x;
    for (int i = 0; i < n; i++) {
        if (stars[i].x < min_x) min_x = stars[i].x;
        if (stars[i].y < min_y) min_y = stars[i].y;
        if (stars[i].z < min_z) min_z = stars[i].z;
        if (stars[i].x > max_x) max_x = stars[i].x;
        if (stars[i].y > max_y) max_y = stars[i].y;
        if (stars[i].z > max_z) max_z = stars[i].z;
    }
    float gmin = std::min({min_x, min_y, min_z});
    float gmax = std::max({max_x, max_y, max_z});
    auto t1 = std::chrono::steady_clock::now();

    std::vector<std::pair<uint64_t, Star>> buf(n);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        buf[i].first = get_morton_code(stars[i].x, stars[i].y, stars[i].z, gmin, gmax);
        buf[i].second = stars[i];
    }
    std::stable_sort(buf.begin(), buf.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) stars[i] = buf[i].second;
    auto t2 = std::chrono::steady_clock::now();

    try {
        s->arena.reset();
        OctreeNode* root = s->arena.alloc<OctreeNode>();
        float span = std::max({max_x - min_x, max_y - min_y, max_z - min_z});
        root->center_x = (min_x + max_x) * 0.5f;
        root->center_y = (min_y + max_y) * 0.5f;
        root->center_z = (min_z + max_z) * 0.5f;
        root->half_width = span * 0.5f;
        root->is_leaf = true;
        root->start_star_index = 0;
        root->star_count = 0;
        uint32_t next_free = 32;
        for (uint32_t i = 0; i < (uint32_t)n; i++)
            insert_star(root, i, stars, s->arena, s->leaf_idx.data(), next_free);
        auto t3 = std::chrono::steady_clock::now();
        node_physics(root, root, stars, s->nm.data(), s->ncx.data(), s->ncy.data(), s->ncz.data(),
                     s->leaf_idx.data());
        auto t4 = std::chrono::steady_clock::now();
        std::vector<float> acc(3 * (size_t)n);
#pragma omp parallel for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            query_gravity(root, stars[i], i, stars, s->nm.data(), s->ncx.data(), s->ncy.data(),
                          s->ncz.data(), s->leaf_idx.data(), G, eps2, acc[3 * i], acc[3 * i + 1],
                          acc[3 * i + 2]);
        }
        auto t5 = std::chrono::steady_clock::now();
        if (do_integrate) {
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i) {
                stars[i].vx += acc[3 * i] * dt;
                stars[i].vy += acc[3 * i + 1] * dt;
                stars[i].vz += acc[3 * i + 2] * dt;
                stars[i].x += stars[i].vx * dt;
                stars[i].y += stars[i].vy * dt;
                stars[i].z += stars[i].vz * dt;
            }
        }
        auto t6 = std::chrono::steady_clock::now();
        if (acc_out) std::memcpy(acc_out, acc.data(), sizeof(float) * 3 * (size_t)n);
        if (phase_s) {
            phase_s[0] = secs(t0, t1);
            phase_s[1] = secs(t1, t2);
            phase_s[2] = secs(t2, t3);
            phase_s[3] = secs(t3, t4);
            phase_s[4] = secs(t4, t5);
            phase_s[5] = secs(t5, t6);
        }
    } catch (...) {
        return -1;
    }
    return 0;
}
