// gcs_generator.cpp -- seeded two-galaxy initial conditions (host C++).
//
// The reference's generator (generator.cpp:3-196) draws from an unseeded
// std::mt19937 with implementation-defined std distributions and different
// constants; this one is written from the model description instead:
//   disc   Miyamoto-Nagai, a=3, b=0.3, M=1, truncated at R<30, |z|<3
//   bulge  Hernquist, a=0.5, M=0.25, truncated at r<60
//   halo   Hernquist, a=12,  M=5,    truncated at r<60
// M is the mass inside the truncation; particle mass is M/count per
// component.  Velocities are circular, v = sqrt(G M(<r) / r) with M(<r) the
// mass of the galaxy's own particles at smaller spherical radius, tangential
// in the disc plane for disc stars and perpendicular to a random direction
// for bulge and halo particles, plus a Gaussian of 5% of v per component.
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

namespace {

constexpr double kPi = 3.14159265358979323846;

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double gauss()
    {
        double u1 = 1.0 - uniform();
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
};

uint64_t mix64(uint64_t v)
{
    Rng r(v);
    return r.next();
}

uint64_t stream_key(uint64_t seed, int galaxy, int comp, int64_t idx)
{
    uint64_t tag = ((uint64_t)galaxy << 56) ^ ((uint64_t)comp << 48) ^ (uint64_t)idx;
    return mix64(mix64(seed) ^ mix64(tag + 0x5851F42D4C957F2Dull));
}

// ---- Miyamoto-Nagai sampler ------------------------------------------------
constexpr double kDiscA = 3.0, kDiscB = 0.3, kDiscM = 1.0;
constexpr double kDiscRmax = 30.0, kDiscZmax = 3.0;
constexpr double kBulgeA = 0.5, kBulgeM = 0.25;
constexpr double kHaloA = 12.0, kHaloM = 5.0;
constexpr double kRtrunc = 60.0;

// Density up to a constant factor.  It decreases in R at fixed z and in |z|
// at fixed R, so over a cell its maximum is at the corner nearest the origin.
double mn_density(double R, double z)
{
    const double a = kDiscA, b = kDiscB;
    double zeta = std::sqrt(z * z + b * b);
    double az = a + zeta;
    double num = a * R * R + (a + 3.0 * zeta) * az * az;
    double den = std::pow(R * R + az * az, 2.5) * zeta * zeta * zeta;
    return num / den;
}

// Piecewise-constant envelope over an (R, |z|) grid; exact rejection.
struct MnTable {
    static constexpr int NR = 192, NZ = 96;
    std::vector<double> re, ze;     // edges
    std::vector<double> bound;      // per cell
    std::vector<double> cdf;        // cumulative envelope mass
    MnTable()
    {
        re.resize(NR + 1);
        ze.resize(NZ + 1);
        // Edges dense near the centre: v_i = s * sinh(k i / N) / sinh(k).
        for (int i = 0; i <= NR; ++i)
            re[i] = kDiscRmax * std::sinh(4.0 * i / NR) / std::sinh(4.0);
        for (int j = 0; j <= NZ; ++j)
            ze[j] = kDiscZmax * std::sinh(5.0 * j / NZ) / std::sinh(5.0);
        bound.resize((size_t)NR * NZ);
        cdf.resize((size_t)NR * NZ);
        double acc = 0.0;
        for (int i = 0; i < NR; ++i)
            for (int j = 0; j < NZ; ++j) {
                double b = mn_density(re[i], ze[j]);
                double vol = kPi * (re[i + 1] * re[i + 1] - re[i] * re[i]) * (ze[j + 1] - ze[j]);
                bound[(size_t)i * NZ + j] = b;
                acc += b * vol;
                cdf[(size_t)i * NZ + j] = acc;
            }
        for (double& c : cdf) c /= acc;
    }
    void sample(Rng& r, double& x, double& y, double& z) const
    {
        for (;;) {
            double u = r.uniform();
            size_t c = (size_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
            if (c >= cdf.size()) c = cdf.size() - 1;
            int i = (int)(c / NZ), j = (int)(c % NZ);
            double r0 = re[i], r1 = re[i + 1];
            double R = std::sqrt(r0 * r0 + r.uniform() * (r1 * r1 - r0 * r0));
            double az = ze[j] + r.uniform() * (ze[j + 1] - ze[j]);
            if (r.uniform() * bound[c] >= mn_density(R, az)) continue;
            double phi = 2.0 * kPi * r.uniform();
            x = R * std::cos(phi);
            y = R * std::sin(phi);
            z = r.uniform() < 0.5 ? -az : az;
            return;
        }
    }
};

const MnTable& mn_table()
{
    static MnTable t;
    return t;
}

// Truncated Hernquist radius by inverse CDF: M(<r)/M(<rt) = u.
double hernquist_radius(Rng& r, double a)
{
    double ft = (kRtrunc / (kRtrunc + a)) * (kRtrunc / (kRtrunc + a));
    double su = std::sqrt(r.uniform() * ft);
    return a * su / (1.0 - su);
}

void unit_vector(Rng& r, double& x, double& y, double& z)
{
    double c = 2.0 * r.uniform() - 1.0;
    double s = std::sqrt(std::max(0.0, 1.0 - c * c));
    double p = 2.0 * kPi * r.uniform();
    x = s * std::cos(p);
    y = s * std::sin(p);
    z = c;
}

struct P {
    double x, y, z, vx, vy, vz, m;
    int comp;  // 0 halo, 1 disc, 2 bulge
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

void galaxy(uint64_t seed, int g, int32_t n, std::vector<P>& out)
{
    const int32_t nd = (int32_t)std::lround(0.7 * n);
    const int32_t nb = (int32_t)std::lround(0.1 * n);
    const int32_t nh = n - nd - nb;
    out.resize((size_t)n);
    const MnTable& mt = mn_table();
    parallel_for(n, [&](int64_t i) {
        P& p = out[(size_t)i];
        if (i < nh) {
            Rng r(stream_key(seed, g, 0, i));
            double rr = hernquist_radius(r, kHaloA), ux, uy, uz;
            unit_vector(r, ux, uy, uz);
            p = {rr * ux, rr * uy, rr * uz, 0, 0, 0, kHaloM / nh, 0};
        } else if (i < nh + nd) {
            Rng r(stream_key(seed, g, 1, i - nh));
            double x, y, z;
            mt.sample(r, x, y, z);
            p = {x, y, z, 0, 0, 0, kDiscM / nd, 1};
        } else {
            Rng r(stream_key(seed, g, 2, i - nh - nd));
            double rr = hernquist_radius(r, kBulgeA), ux, uy, uz;
            unit_vector(r, ux, uy, uz);
            p = {rr * ux, rr * uy, rr * uz, 0, 0, 0, kBulgeM / nb, 2};
        }
    });
    // Enclosed mass by spherical radius.
    std::vector<double> rad((size_t)n);
    for (int32_t i = 0; i < n; ++i)
        rad[(size_t)i] = std::sqrt(out[(size_t)i].x * out[(size_t)i].x +
                                   out[(size_t)i].y * out[(size_t)i].y +
                                   out[(size_t)i].z * out[(size_t)i].z);
    std::vector<int32_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int32_t a, int32_t b) { return rad[(size_t)a] < rad[(size_t)b]; });
    std::vector<double> menc((size_t)n);
    double cum = 0.0;
    for (int32_t k = 0; k < n; ++k) {
        menc[(size_t)ord[(size_t)k]] = cum;
        cum += out[(size_t)ord[(size_t)k]].m;
    }
    parallel_for(n, [&](int64_t i) {
        P& p = out[(size_t)i];
        Rng r(stream_key(seed, g, 3, i));
        double rr = rad[(size_t)i];
        double vc = rr > 0.0 ? std::sqrt(menc[(size_t)i] / rr) : 0.0;
        double tx, ty, tz;
        if (p.comp == 1) {
            double R = std::sqrt(p.x * p.x + p.y * p.y);
            if (R > 0.0) { tx = -p.y / R; ty = p.x / R; tz = 0.0; }
            else { tx = 1.0; ty = 0.0; tz = 0.0; }
        } else {
            // Random direction perpendicular to the radius vector.
            double ux, uy, uz;
            unit_vector(r, ux, uy, uz);
            if (rr > 0.0) {
                double nx = p.x / rr, ny = p.y / rr, nz = p.z / rr;
                double d = ux * nx + uy * ny + uz * nz;
                ux -= d * nx; uy -= d * ny; uz -= d * nz;
            }
            double l = std::sqrt(ux * ux + uy * uy + uz * uz);
            if (l > 1e-12) { tx = ux / l; ty = uy / l; tz = uz / l; }
            else { tx = 1.0; ty = 0.0; tz = 0.0; }
        }
        p.vx = vc * tx + 0.05 * vc * r.gauss();
        p.vy = vc * ty + 0.05 * vc * r.gauss();
        p.vz = vc * tz + 0.05 * vc * r.gauss();
    });
}

}  // namespace

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
        std::vector<P> ga, gb;
        galaxy(s->seed, 0, s->n_a, ga);
        galaxy(s->seed, 1, s->n_b, gb);
        const double t = s->tilt_deg * kPi / 180.0, ct = std::cos(t), st = std::sin(t);
        // Colours and types as the reference assigns them (generator.cpp:158-195):
        // halo type 1 black; galaxy A type 0, galaxy B type 2.
        static const uint8_t col[2][3][3] = {
            {{0, 0, 0}, {255, 100, 50}, {255, 255, 150}},
            {{0, 0, 0}, {50, 100, 255}, {200, 255, 255}}};
        auto emit = [&](const std::vector<P>& src, int g, gcs_star* dst) {
            for (size_t i = 0; i < src.size(); ++i) {
                const P& p = src[i];
                double x = p.x, y = p.y, z = p.z, vx = p.vx, vy = p.vy, vz = p.vz;
                if (g == 1) {
                    double y2 = y * ct - z * st, z2 = y * st + z * ct;
                    double vy2 = vy * ct - vz * st, vz2 = vy * st + vz * ct;
                    x += s->offset[0]; y = y2 + s->offset[1]; z = z2 + s->offset[2];
                    vx += s->velocity[0]; vy = vy2 + s->velocity[1]; vz = vz2 + s->velocity[2];
                }
                gcs_star& o = dst[i];
                o.x = (float)x; o.y = (float)y; o.z = (float)z;
                o.vx = (float)vx; o.vy = (float)vy; o.vz = (float)vz;
                o.mass = (float)p.m;
                o.type = p.comp == 0 ? 1 : (g == 0 ? 0 : 2);
                o.r = col[g][p.comp][0]; o.g = col[g][p.comp][1]; o.b = col[g][p.comp][2];
            }
        };
        emit(ga, 0, out);
        emit(gb, 1, out + s->n_a);
    } catch (...) {
        return GCS_ERR_OOM;
    }
    return GCS_OK;
}
