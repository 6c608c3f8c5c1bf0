// gcs_cuda_engine.cpp -- the reference's ten cuda_engine.h entry points over
// the gcs_* C ABI (one process-wide engine, as cuda_engine.cu:17-44 keeps one
// simulation per process).
#include "gcs_cuda_engine.h"

#include <hip/hip_runtime.h>

#include <cstdio>

#include "gcs.h"

static_assert(sizeof(Star) == sizeof(gcs_star), "Star layout");
static_assert(sizeof(PlaybackStar) == sizeof(gcs_playback_star), "PlaybackStar layout");

namespace {
gcs_engine* g_engine = nullptr;
int g_capacity = 0;
}  // namespace

bool cuda_init(int num_stars)
{
    if (g_engine) cuda_cleanup();
    if (num_stars < 1) return false;
    if (gcs_create(num_stars, 1, 0, &g_engine) != GCS_OK) {
        std::fprintf(stderr, "HIP Error: %s\n", gcs_last_error());
        g_engine = nullptr;
        return false;
    }
    g_capacity = num_stars;
    return true;
}

void cuda_cleanup()
{
    gcs_destroy(g_engine);
    g_engine = nullptr;
    g_capacity = 0;
}

void cuda_allocate_pinned_stars(Star** stars, int num_stars)
{
    if (!stars) return;
    *stars = nullptr;
    if (hipHostMalloc((void**)stars, sizeof(Star) * (size_t)num_stars, hipHostMallocDefault) !=
        hipSuccess)
        *stars = nullptr;
}

void cuda_free_pinned_stars(Star* stars)
{
    if (stars) (void)hipHostFree(stars);
}

void cuda_allocate_pinned_playback(PlaybackStar** buf, int size)
{
    if (!buf) return;
    *buf = nullptr;
    if (hipHostMalloc((void**)buf, sizeof(PlaybackStar) * (size_t)size, hipHostMallocDefault) !=
        hipSuccess)
        *buf = nullptr;
}

void cuda_free_pinned_playback(PlaybackStar* buf)
{
    if (buf) (void)hipHostFree(buf);
}

void cuda_upload_initial_stars(const Star* host_stars, int num_stars)
{
    if (!g_engine || !host_stars || num_stars < 1 || num_stars > g_capacity) return;
    int32_t n = num_stars;
    if (gcs_load(g_engine, reinterpret_cast<const gcs_star*>(host_stars), &n, 1) != GCS_OK)
        std::fprintf(stderr, "HIP Error: %s\n", gcs_last_error());
}

void cuda_register_vbo(unsigned int vbo_id) { (void)vbo_id; }

void cuda_unregister_vbo() {}

void cuda_get_center_of_mass(float* cx, float* cy, float* cz)
{
    // cuda_engine.cu:365-376: the root record of the last built tree (one
    // small copy), zeros when nothing has been built.
    float c[4] = {0.f, 0.f, 0.f, 0.f};
    if (g_engine && gcs_read_root(g_engine, 0, c) != GCS_OK) c[0] = c[1] = c[2] = 0.f;
    if (cx) *cx = c[0];
    if (cy) *cy = c[1];
    if (cz) *cz = c[2];
}

float cuda_physics_step(Star* host_stars, int num_stars, float G, float epsilon_sq, float dt,
                        float global_min, float global_max, bool use_interop)
{
    (void)use_interop;
    if (!g_engine) return dt;
    gcs_params p;
    gcs_params_default(&p);
    p.G = G;
    p.eps2 = epsilon_sq;
    p.dt = dt;
    p.mac = GCS_MAC_REFERENCE;
    p.precision = GCS_PREC_PLAIN;
    p.stack_depth = 256;
    p.use_bounds = 1;
    p.global_min = global_min;
    p.global_max = global_max;
    p.colours = 1;
    p.diagnostics = 0;
    p.use_graph = 0;  // bounds change every frame
    int rc = gcs_step(g_engine, &p, 1);
    if (rc == GCS_OK) rc = gcs_sync(g_engine);
    if (rc != GCS_OK) std::fprintf(stderr, "HIP Error: %s\n", gcs_last_error());
    int32_t k = 0;
    int64_t n = 0;
    if (host_stars && gcs_instance_count(g_engine, &k, &n) == GCS_OK && n <= num_stars &&
        gcs_read_stars(g_engine, reinterpret_cast<gcs_star*>(host_stars)) != GCS_OK)
        std::fprintf(stderr, "HIP Error: %s\n", gcs_last_error());
    return dt;
}
