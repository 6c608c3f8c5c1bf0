// gcs_bake_main.cpp -- Bake driver over the cuda_engine.h entry points.
//
// The reference's Bake loop (main.cpp:119-297) written against
// gcs_cuda_engine.h: per frame it takes the host cube bounds
// (main.cpp:152-166), calls cuda_physics_step (main.cpp:196-199), packs the
// PlaybackStar records (main.cpp:264-267) and hands them to a writer thread
// that appends them to simulation.bin (main.cpp:269-272), double-buffered.
// The file format is unchanged: headerless frames of N x 16-byte records in
// device (Morton-sorted) order.
//
//   gcs_bake [frames] [--n PER_GALAXY] [--seed S] [--eps2 E] [--dt DT]
//            [--out FILE]
// Initial conditions come from gcs_generate (two galaxies of PER_GALAXY).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "gcs.h"
#include "gcs_cuda_engine.h"

int main(int argc, char** argv)
{
    int frames = 10;
    int per_galaxy = 8192;
    unsigned long long seed = 1001;
    float eps2 = 0.0025f, dt = 0.01f, G = 1.0f;
    std::string out = "simulation.bin";
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (a == "--n") per_galaxy = std::atoi(next());
        else if (a == "--seed") seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--eps2") eps2 = (float)std::atof(next());
        else if (a == "--dt") dt = (float)std::atof(next());
        else if (a == "--out") out = next();
        else frames = std::atoi(a.c_str());
    }
    const int n = 2 * per_galaxy;
    if (!cuda_init(n)) {
        std::fprintf(stderr, "cuda_init failed\n");
        return 2;
    }
    Star* stars = nullptr;
    cuda_allocate_pinned_stars(&stars, n);
    if (!stars) return 3;
    gcs_collision_spec spec;
    gcs_spec_default(per_galaxy, seed, &spec);
    if (gcs_generate(&spec, reinterpret_cast<gcs_star*>(stars)) != GCS_OK) return 4;
    cuda_upload_initial_stars(stars, n);

    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) return 5;
    PlaybackStar* buf_a = nullptr;
    PlaybackStar* buf_b = nullptr;
    cuda_allocate_pinned_playback(&buf_a, n);
    cuda_allocate_pinned_playback(&buf_b, n);
    if (!buf_a || !buf_b) return 6;
    std::thread writer;
    PlaybackStar* cur = buf_a;
    PlaybackStar* nxt = buf_b;
    double total = 0.0;
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
        cuda_physics_step(stars, n, G, eps2, dt, gmin, gmax, false);
        if (writer.joinable()) writer.join();
        std::swap(cur, nxt);
        for (int i = 0; i < n; ++i)
            cur[i] = {stars[i].x, stars[i].y, stars[i].z, stars[i].r, stars[i].g, stars[i].b,
                      stars[i].type};
        writer = std::thread([f, cur, n]() { std::fwrite(cur, sizeof(PlaybackStar), n, f); });
        total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    if (writer.joinable()) writer.join();
    std::fclose(f);
    float cx, cy, cz;
    cuda_get_center_of_mass(&cx, &cy, &cz);
    std::printf("baked %d frames of %d stars in %.4f s; root com %.6g %.6g %.6g\n", frames, n,
                total, cx, cy, cz);
    cuda_free_pinned_playback(buf_a);
    cuda_free_pinned_playback(buf_b);
    cuda_free_pinned_stars(stars);
    cuda_cleanup();
    return 0;
}
