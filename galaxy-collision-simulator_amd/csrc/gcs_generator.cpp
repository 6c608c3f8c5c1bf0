// gcs_generator.cpp -- seeded two-galaxy initial conditions on the host.
//
// The model and every sampling step are in gcs_gen_common.hpp, shared with
// the device generator; this file builds the disc envelope table, runs the
// per-particle draws on host threads and finds each particle's enclosed mass
// (a stable sort by radius, then counts of each component at smaller
// radius).
//
// Every particle draws from its own counter-based stream (SplitMix64 keyed
// by seed, galaxy, component and index) so the output does not depend on the
// number of threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "gcs.h"
#include "gcs_gen_common.hpp"

namespace gcs_gen {

namespace {

// Edges dense near the centre: v_i = s * sinh(k i / N) / sinh(k).
struct MnTable {
    std::vector<double> re, ze, bound, cdf;
    MnTableView view;
    MnTable()
    {
        re.resize(kNR + 1);
        ze.resize(kNZ + 1);
        for (int i = 0; i <= kNR; ++i) re[i] = kDiscRmax * std::sinh(4.0 * i / kNR) / std::sinh(4.0);
        for (int j = 0; j <= kNZ; ++j) ze[j] = kDiscZmax * std::sinh(5.0 * j / kNZ) / std::sinh(5.0);
        bound.resize((size_t)kNR * kNZ);
        cdf.resize((size_t)kNR * kNZ);
        double acc = 0.0;
        for (int i = 0; i < kNR; ++i)
            for (int j = 0; j < kNZ; ++j) {
                const double b = mn_density(re[i], ze[j]);
                const double vol = kPi * (re[i + 1] * re[i + 1] - re[i] * re[i]) * (ze[j + 1] - ze[j]);
                bound[(size_t)i * kNZ + j] = b;
                acc += b * vol;
                cdf[(size_t)i * kNZ + j] = acc;
            }
        for (double& c : cdf) c /= acc;
        view = MnTableView{re.data(), ze.data(), bound.data(), cdf.data()};
    }
};

template <typename F>
void parallel_for(int64_t n, F&& f)
{
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1, 16);
    if (n < 4096 || nt <= 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<int64_t> next(0);
    const int64_t chunk = 1024;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (;;) {
                int64_t s = next.fetch_add(chunk);
                if (s >= n) break;
                int64_t e = std::min(n, s + chunk);
                for (int64_t i = s; i < e; ++i) f(i);
            }
        });
    for (auto& t : th) t.join();
}

// One galaxy: positions, then enclosed masses, then velocities (in the
// galaxy's frame).
void galaxy(uint64_t seed, int g, int32_t n, std::vector<GenParticle>& out,
            std::vector<double>& vel)
{
    const GalaxySizes gs = galaxy_sizes(n);
    out.resize((size_t)n);
    vel.resize(3 * (size_t)n);
    const MnTableView& mt = host_mn_table();
    parallel_for(n, [&](int64_t i) { out[(size_t)i] = gen_position(seed, g, i, gs, mt); });
    std::vector<double> rad((size_t)n);
    for (int32_t i = 0; i < n; ++i) rad[(size_t)i] = gen_radius(out[(size_t)i]);
    std::vector<int32_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return rad[(size_t)a] < rad[(size_t)b]; });
    std::vector<double> menc((size_t)n);
    int64_t cnt[3] = {0, 0, 0};
    for (int32_t k = 0; k < n; ++k) {
        const int32_t i = ord[(size_t)k];
        menc[(size_t)i] = enclosed_mass(gs, cnt[0], cnt[1], cnt[2]);
        ++cnt[out[(size_t)i].comp];
    }
    parallel_for(n, [&](int64_t i) {
        gen_velocity(seed, g, i, out[(size_t)i], rad[(size_t)i], menc[(size_t)i], vel[3 * i],
                     vel[3 * i + 1], vel[3 * i + 2]);
    });
}

}  // namespace

const MnTableView& host_mn_table()
{
    static MnTable t;
    return t.view;
}

}  // namespace gcs_gen

using gcs_gen::Rng;
using gcs_gen::mix64;
using gcs_gen::unit_vector;

extern "C" void gcs_spec_default(int32_t n_per_galaxy, uint64_t seed, gcs_collision_spec* s)
{
    std::memset(s, 0, sizeof(*s));
    s->n_a = n_per_galaxy;
    s->n_b = n_per_galaxy;
    s->offset[0] = 40.0f; s->offset[1] = 10.0f; s->offset[2] = 0.0f;
    s->velocity[0] = -0.5f; s->velocity[1] = 0.0f; s->velocity[2] = 0.0f;
    s->tilt_deg = 30.0f;
    s->seed = seed;
}

extern "C" void gcs_spec_random(uint64_t seed, int64_t k, int32_t n_min, int32_t n_max,
                                gcs_collision_spec* s)
{
    std::memset(s, 0, sizeof(*s));
    Rng r(mix64(mix64(seed) ^ mix64((uint64_t)k * 0xD1B54A32D192ED03ull + 1)));
    double ln0 = std::log((double)n_min), ln1 = std::log((double)n_max);
    int64_t n = std::llround(std::exp(ln0 + (ln1 - ln0) * r.uniform()));
    n = std::max<int64_t>(n_min, std::min<int64_t>(n_max, n));
    s->n_a = (int32_t)(n / 2);
    s->n_b = (int32_t)(n - n / 2);
    double d = 20.0 + 40.0 * r.uniform();
    double ux, uy, uz;
    unit_vector(r, ux, uy, uz);
    double sp = 0.2 + 0.6 * r.uniform();
    s->offset[0] = (float)(d * ux); s->offset[1] = (float)(d * uy); s->offset[2] = (float)(d * uz);
    s->velocity[0] = (float)(-sp * ux); s->velocity[1] = (float)(-sp * uy);
    s->velocity[2] = (float)(-sp * uz);
    s->tilt_deg = (float)(90.0 * r.uniform());
    s->seed = r.next();
}

extern "C" int gcs_generate(const gcs_collision_spec* s, gcs_star* out)
{
    if (!s || !out || s->n_a < 0 || s->n_b < 0) return GCS_ERR_INVALID;
    try {
        const double t = s->tilt_deg * gcs_gen::kPi / 180.0, ct = std::cos(t), st = std::sin(t);
        const int32_t ns[2] = {s->n_a, s->n_b};
        gcs_star* dst = out;
        for (int g = 0; g < 2; ++g) {
            std::vector<gcs_gen::GenParticle> gp;
            std::vector<double> vel;
            gcs_gen::galaxy(s->seed, g, ns[g], gp, vel);
            for (int32_t i = 0; i < ns[g]; ++i) {
                const gcs_gen::GenStar o =
                    gcs_gen::gen_star(gp[(size_t)i], g, vel[3 * (size_t)i], vel[3 * (size_t)i + 1],
                                      vel[3 * (size_t)i + 2], ct, st, s->offset, s->velocity);
                gcs_star& r = dst[i];
                r.x = o.x; r.y = o.y; r.z = o.z;
                r.vx = o.vx; r.vy = o.vy; r.vz = o.vz;
                r.mass = o.m;
                r.type = o.type; r.r = o.r; r.g = o.g; r.b = o.b;
            }
            dst += ns[g];
        }
    } catch (...) {
        return GCS_ERR_OOM;
    }
    return GCS_OK;
}
