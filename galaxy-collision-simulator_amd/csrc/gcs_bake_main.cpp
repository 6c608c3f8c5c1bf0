// gcs_bake_main.cpp -- Bake driver: writes simulation.bin (main.cpp:119-297).
//
//   gcs_bake [frames] [--n PER_GALAXY] [--seed S] [--eps2 E] [--dt DT]
//            [--out FILE] [--via-shim]
//
// Initial conditions come from gcs_generate (two galaxies of PER_GALAXY).
// Either way the frames are the reference's: headerless frames of N x 16-byte
// PlaybackStar records in device (Morton-sorted) order, the GPU traversal
// rule of cuda_engine.cu:91-161, cube bounds over all coordinates.
//
// Default: the MI355X pipeline, gcs_bake -- device bounds, graph-replayed
// steps, PlaybackStar records packed on the device and copied 16 bytes per
// particle into pinned double buffers while the next frame steps.
//
// --via-shim: the reference's own frame, written against the cuda_engine.h
// entry points as main.cpp:150-273 does it -- host cube bounds over the host
// copy, the host Morton sort that cuda_physics_step then overwrites
// (main.cpp:168-192), cuda_physics_step with its 32-byte-per-particle copy
// back, the PlaybackStar pack on the host and a writer thread per frame.
//
// Prints one JSON line: mode, particles, frames, seconds, ms per frame.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "gcs.h"
#include "gcs_cuda_engine.h"

namespace {

// 21-bit interleave of the host Morton sort (morton.h:11-28).
uint64_t spread21(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFu;
    x = (x | x << 32) & 0x1F00000000FFFFull;
    x = (x | x << 16) & 0x1F0000FF0000FFull;
    x = (x | x << 8) & 0x100F00F00F00F00Full;
    x = (x | x << 4) & 0x10C30C30C30C30C3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

uint32_t to_cell(float v, float mn, float scale)
{
    const float q = (v - mn) * scale;
    if (!(q >= 1.0f)) return 0u;
    return q >= 4294967296.0f ? 0xFFFFFFFFu : (uint32_t)q;
}

// The reference's frame through the ten cuda_engine.h functions.
int bake_via_shim(Star* stars, int n, int frames, float G, float eps2, float dt, FILE* f,
                  double* total)
{
    PlaybackStar* buf_a = nullptr;
    PlaybackStar* buf_b = nullptr;
    cuda_allocate_pinned_playback(&buf_a, n);
    cuda_allocate_pinned_playback(&buf_b, n);
    if (!buf_a || !buf_b) return 6;
    std::vector<std::pair<uint64_t, Star>> sort_buffer((size_t)n);
    std::thread writer;
    PlaybackStar* cur = buf_a;
    PlaybackStar* nxt = buf_b;
    *total = 0.0;
    for (int frame = 0; frame < frames; ++frame) {
        auto t0 = std::chrono::steady_clock::now();
        float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                       std::numeric_limits<float>::max()};
        float hi[3] = {std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
                       std::numeric_limits<float>::lowest()};
        for (int i = 0; i < n; ++i) {
            lo[0] = std::min(lo[0], stars[i].x); hi[0] = std::max(hi[0], stars[i].x);
            lo[1] = std::min(lo[1], stars[i].y); hi[1] = std::max(hi[1], stars[i].y);
            lo[2] = std::min(lo[2], stars[i].z); hi[2] = std::max(hi[2], stars[i].z);
        }
        const float gmin = std::min({lo[0], lo[1], lo[2]});
        const float gmax = std::max({hi[0], hi[1], hi[2]});
        // The host sort the reference runs every frame (its result is
        // overwritten by the step's copy back).
        const float scale = 2097151.0f / (gmax - gmin);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) {
            const uint64_t k = (spread21(to_cell(stars[i].x, gmin, scale)) << 2) |
                               (spread21(to_cell(stars[i].y, gmin, scale)) << 1) |
                               spread21(to_cell(stars[i].z, gmin, scale));
            sort_buffer[(size_t)i] = {k, stars[i]};
        }
        std::sort(sort_buffer.begin(), sort_buffer.end(),
                  [](const auto& a, const auto& b) { return a.first < b.first; });
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i) stars[i] = sort_buffer[(size_t)i].second;
        cuda_physics_step(stars, n, G, eps2, dt, gmin, gmax, false);
        if (writer.joinable()) writer.join();
        std::swap(cur, nxt);
        for (int i = 0; i < n; ++i)
            cur[i] = {stars[i].x, stars[i].y, stars[i].z, stars[i].r, stars[i].g, stars[i].b,
                      stars[i].type};
        writer = std::thread([f, cur, n]() { std::fwrite(cur, sizeof(PlaybackStar), n, f); });
        *total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (writer.joinable()) writer.join();
    cuda_free_pinned_playback(buf_a);
    cuda_free_pinned_playback(buf_b);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    int frames = 10;
    int per_galaxy = 8192;
    unsigned long long seed = 1001;
    float eps2 = 0.0025f, dt = 0.01f, G = 1.0f;
    bool via_shim = false;
    std::string out = "simulation.bin";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--n") per_galaxy = std::atoi(next());
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--eps2") eps2 = (float)std::atof(next());
        else if (a == "--dt") dt = (float)std::atof(next());
        else if (a == "--out") out = next();
        else if (a == "--via-shim") via_shim = true;
        else frames = std::atoi(a.c_str());
    }
    const int n = 2 * per_galaxy;
    gcs_collision_spec spec;
    gcs_spec_default(per_galaxy, seed, &spec);
    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) return 5;
    double total = 0.0;
    if (via_shim) {
        if (!cuda_init(n)) {
            std::fprintf(stderr, "cuda_init failed\n");
            return 2;
        }
        Star* stars = nullptr;
        cuda_allocate_pinned_stars(&stars, n);
        if (!stars) return 3;
        if (gcs_generate(&spec, reinterpret_cast<gcs_star*>(stars)) != GCS_OK) return 4;
        cuda_upload_initial_stars(stars, n);
        const int rc = bake_via_shim(stars, n, frames, G, eps2, dt, f, &total);
        std::fclose(f);
        if (rc) return rc;
        cuda_free_pinned_stars(stars);
        cuda_cleanup();
    } else {
        std::fclose(f);
        gcs_engine* e = nullptr;
        std::vector<gcs_star> st((size_t)n);
        int32_t cnt = n;
        if (gcs_create(n, 1, 0, &e) != GCS_OK || gcs_generate(&spec, st.data()) != GCS_OK ||
            gcs_load(e, st.data(), &cnt, 1) != GCS_OK) {
            std::fprintf(stderr, "engine set-up failed: %s\n", gcs_last_error());
            return 2;
        }
        gcs_params p;
        gcs_params_default(&p);
        p.G = G;
        p.eps2 = eps2;
        p.dt = dt;
        p.mac = GCS_MAC_REFERENCE;
        p.use_bounds = 0;  // the same cube as the host bounds, from the device
        p.colours = 1;
        p.use_graph = 1;
        if (gcs_bake(e, &p, frames, out.c_str(), &total) != GCS_OK) {
            std::fprintf(stderr, "bake failed: %s\n", gcs_last_error());
            gcs_destroy(e);
            return 7;
        }
        gcs_destroy(e);
    }
    std::printf("{\"mode\": \"%s\", \"particles\": %d, \"frames\": %d, \"seconds\": %.6f, "
                "\"ms_per_frame\": %.4f}\n",
                via_shim ? "cuda_engine.h shim, reference frame" : "gcs_bake pipeline", n,
                frames, total, frames > 0 ? 1e3 * total / frames : 0.0);
    return 0;
}
