// gcs_gen_common.hpp -- the seeded two-galaxy model, shared by the host
// generator (gcs_generator.cpp) and the device generator (gcs_gen_device.hpp).
//
// One source for both, so the two produce the same records: every step of a
// particle's sampling is a __host__ __device__ function of its own
// counter-based stream, and the arithmetic is written out with no
// multiply-add contraction.  The only host/device differences left are the
// last-bit results of log, cos, sin and pow (glibc vs the device library),
// which change a float record in about one value in 10^8.
//
// The reference's generator (generator.cpp:3-196) draws from an unseeded
// std::mt19937 with implementation-defined std distributions and different
// constants; this model is written from the description instead:
//   disc   Miyamoto-Nagai, a=3, b=0.3, M=1, truncated at R<30, |z|<3
//   bulge  Hernquist, a=0.5, M=0.25, truncated at r<60
//   halo   Hernquist, a=12,  M=5,    truncated at r<60
// M is the mass inside the truncation; particle mass is M/count per
// component.  Velocities are circular, v = sqrt(G M(<r) / r) with M(<r) the
// mass of the galaxy's own particles at smaller spherical radius (ties by
// index), tangential in the disc plane for disc stars and perpendicular to a
// random direction for bulge and halo particles, plus a Gaussian of 5% of v
// per component.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace gcs_gen {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDiscA = 3.0, kDiscB = 0.3, kDiscM = 1.0;
constexpr double kDiscRmax = 30.0, kDiscZmax = 3.0;
constexpr double kBulgeA = 0.5, kBulgeM = 0.25;
constexpr double kHaloA = 12.0, kHaloM = 5.0;
constexpr double kRtrunc = 60.0;
// Disc sampler envelope: NR x NZ cells over (R, |z|).
constexpr int kNR = 192, kNZ = 96;

// SplitMix64 stream.
struct Rng {
    uint64_t s;
    __host__ __device__ explicit Rng(uint64_t seed) : s(seed) {}
    __host__ __device__ uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    __host__ __device__ double uniform()
    {
        return (double)(next() >> 11) * (1.0 / 9007199254740992.0);
    }
    // Box-Muller.
    __host__ __device__ double gauss()
    {
#pragma clang fp contract(off)
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(2.0 * kPi * u2);
    }
};

__host__ __device__ inline uint64_t mix64(uint64_t v)
{
    Rng r(v);
    return r.next();
}

// The stream of one particle draw: seed, galaxy, draw kind (0 halo, 1 disc,
// 2 bulge positions; 3 velocities) and index within the galaxy.
__host__ __device__ inline uint64_t stream_key(uint64_t seed, int galaxy, int comp, int64_t idx)
{
    const uint64_t tag = ((uint64_t)galaxy << 56) ^ ((uint64_t)comp << 48) ^ (uint64_t)idx;
    return mix64(mix64(seed) ^ mix64(tag + 0x5851F42D4C957F2Dull));
}

// Miyamoto-Nagai density up to a constant factor.  It decreases in R at
// fixed z and in |z| at fixed R, so over a cell its maximum is at the corner
// nearest the origin.
__host__ __device__ inline double mn_density(double R, double z)
{
#pragma clang fp contract(off)
    const double a = kDiscA, b = kDiscB;
    const double zeta = sqrt(z * z + b * b);
    const double az = a + zeta;
    const double num = a * R * R + (a + 3.0 * zeta) * az * az;
    const double den = pow(R * R + az * az, 2.5) * zeta * zeta * zeta;
    return num / den;
}

// The disc envelope: cell edges re[kNR+1], ze[kNZ+1], the density bound of
// each cell and the normalised cumulative envelope mass, cell (i, j) at
// i*kNZ + j.
struct MnTableView {
    const double* re;
    const double* ze;
    const double* bound;
    const double* cdf;
};

// The host's table (built once, on first use, in gcs_generator.cpp).
const MnTableView& host_mn_table();

// First cell whose cumulative mass exceeds u (std::upper_bound), clamped.
__host__ __device__ inline int mn_cell(const double* cdf, double u)
{
    int lo = 0, hi = kNR * kNZ;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    return lo < kNR * kNZ ? lo : kNR * kNZ - 1;
}

// Exact rejection from the piecewise-constant envelope.
__host__ __device__ inline void mn_sample(const MnTableView& t, Rng& r, double& x, double& y,
                                          double& z)
{
#pragma clang fp contract(off)
    for (;;) {
        const int c = mn_cell(t.cdf, r.uniform());
        const int i = c / kNZ, j = c % kNZ;
        const double r0 = t.re[i], r1 = t.re[i + 1];
        const double R = sqrt(r0 * r0 + r.uniform() * (r1 * r1 - r0 * r0));
        const double az = t.ze[j] + r.uniform() * (t.ze[j + 1] - t.ze[j]);
        if (r.uniform() * t.bound[c] >= mn_density(R, az)) continue;
        const double phi = 2.0 * kPi * r.uniform();
        x = R * cos(phi);
        y = R * sin(phi);
        z = r.uniform() < 0.5 ? -az : az;
        return;
    }
}

// Truncated Hernquist radius by inverse CDF: M(<r)/M(<rt) = u.
__host__ __device__ inline double hernquist_radius(Rng& r, double a)
{
#pragma clang fp contract(off)
    const double ft = (kRtrunc / (kRtrunc + a)) * (kRtrunc / (kRtrunc + a));
    const double su = sqrt(r.uniform() * ft);
    return a * su / (1.0 - su);
}

__host__ __device__ inline void unit_vector(Rng& r, double& x, double& y, double& z)
{
#pragma clang fp contract(off)
    const double c = 2.0 * r.uniform() - 1.0;
    const double cc = 1.0 - c * c;
    const double s = sqrt(cc > 0.0 ? cc : 0.0);
    const double p = 2.0 * kPi * r.uniform();
    x = s * cos(p);
    y = s * sin(p);
    z = c;
}

// Component sizes of a galaxy of n particles: 70% disc, 10% bulge, the rest
// halo (rounded half away from zero, as llround).
struct GalaxySizes {
    int32_t nh, nd, nb;
};

__host__ __device__ inline GalaxySizes galaxy_sizes(int32_t n)
{
    GalaxySizes s;
    s.nd = (int32_t)llround(0.7 * n);
    s.nb = (int32_t)llround(0.1 * n);
    s.nh = n - s.nd - s.nb;
    return s;
}

// Position, mass and component (0 halo, 1 disc, 2 bulge) of particle i of
// galaxy g (particles ordered halo, disc, bulge).
struct GenParticle {
    double x, y, z, m;
    int comp;
};

__host__ __device__ inline GenParticle gen_position(uint64_t seed, int g, int64_t i,
                                                    const GalaxySizes& gs,
                                                    const MnTableView& mt)
{
    GenParticle p;
    if (i < gs.nh) {
        Rng r(stream_key(seed, g, 0, i));
        const double rr = hernquist_radius(r, kHaloA);
        double ux, uy, uz;
        unit_vector(r, ux, uy, uz);
        p.x = rr * ux; p.y = rr * uy; p.z = rr * uz;
        p.m = kHaloM / gs.nh;
        p.comp = 0;
    } else if (i < gs.nh + gs.nd) {
        Rng r(stream_key(seed, g, 1, i - gs.nh));
        mn_sample(mt, r, p.x, p.y, p.z);
        p.m = kDiscM / gs.nd;
        p.comp = 1;
    } else {
        Rng r(stream_key(seed, g, 2, i - gs.nh - gs.nd));
        const double rr = hernquist_radius(r, kBulgeA);
        double ux, uy, uz;
        unit_vector(r, ux, uy, uz);
        p.x = rr * ux; p.y = rr * uy; p.z = rr * uz;
        p.m = kBulgeM / gs.nb;
        p.comp = 2;
    }
    return p;
}

__host__ __device__ inline double gen_radius(const GenParticle& p)
{
#pragma clang fp contract(off)
    return sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
}

// Mass at smaller radius from the counts of each component there.
__host__ __device__ inline double enclosed_mass(const GalaxySizes& gs, int64_t ch, int64_t cd,
                                                int64_t cb)
{
#pragma clang fp contract(off)
    const double mh = gs.nh > 0 ? kHaloM / gs.nh : 0.0;
    const double md = gs.nd > 0 ? kDiscM / gs.nd : 0.0;
    const double mb = gs.nb > 0 ? kBulgeM / gs.nb : 0.0;
    return mh * (double)ch + md * (double)cd + mb * (double)cb;
}

// Velocity of particle i of galaxy g at radius rr with enclosed mass menc.
__host__ __device__ inline void gen_velocity(uint64_t seed, int g, int64_t i,
                                             const GenParticle& p, double rr, double menc,
                                             double& vx, double& vy, double& vz)
{
#pragma clang fp contract(off)
    Rng r(stream_key(seed, g, 3, i));
    const double vc = rr > 0.0 ? sqrt(menc / rr) : 0.0;
    double tx, ty, tz;
    if (p.comp == 1) {
        const double R = sqrt(p.x * p.x + p.y * p.y);
        if (R > 0.0) { tx = -p.y / R; ty = p.x / R; tz = 0.0; }
        else { tx = 1.0; ty = 0.0; tz = 0.0; }
    } else {
        // Random direction perpendicular to the radius vector.
        double ux, uy, uz;
        unit_vector(r, ux, uy, uz);
        if (rr > 0.0) {
            const double nx = p.x / rr, ny = p.y / rr, nz = p.z / rr;
            const double d = ux * nx + uy * ny + uz * nz;
            ux -= d * nx; uy -= d * ny; uz -= d * nz;
        }
        const double l = sqrt(ux * ux + uy * uy + uz * uz);
        if (l > 1e-12) { tx = ux / l; ty = uy / l; tz = uz / l; }
        else { tx = 1.0; ty = 0.0; tz = 0.0; }
    }
    vx = vc * tx + 0.05 * vc * r.gauss();
    vy = vc * ty + 0.05 * vc * r.gauss();
    vz = vc * tz + 0.05 * vc * r.gauss();
}

// The record of a particle: galaxy B rotated by the tilt (cos ct, sin st)
// about x, then offset; colours and types as the reference assigns them
// (generator.cpp:158-195): halo type 1 black; galaxy A type 0, galaxy B
// type 2.
struct GenStar {
    float x, y, z, vx, vy, vz, m;
    uint8_t type, r, g, b;
};

__host__ __device__ inline GenStar gen_star(const GenParticle& p, int g, double vx, double vy,
                                            double vz, double ct, double st, const float* off,
                                            const float* vel)
{
#pragma clang fp contract(off)
    double x = p.x, y = p.y, z = p.z;
    if (g == 1) {
        const double y2 = y * ct - z * st, z2 = y * st + z * ct;
        const double vy2 = vy * ct - vz * st, vz2 = vy * st + vz * ct;
        x += off[0]; y = y2 + off[1]; z = z2 + off[2];
        vx += vel[0]; vy = vy2 + vel[1]; vz = vz2 + vel[2];
    }
    GenStar o;
    o.x = (float)x; o.y = (float)y; o.z = (float)z;
    o.vx = (float)vx; o.vy = (float)vy; o.vz = (float)vz;
    o.m = (float)p.m;
    o.type = (uint8_t)(p.comp == 0 ? 1 : (g == 0 ? 0 : 2));
    // halo, disc, bulge of galaxy A, then of galaxy B
    const uint8_t col[2][3][3] = {{{0, 0, 0}, {255, 100, 50}, {255, 255, 150}},
                                  {{0, 0, 0}, {50, 100, 255}, {200, 255, 255}}};
    o.r = col[g][p.comp][0];
    o.g = col[g][p.comp][1];
    o.b = col[g][p.comp][2];
    return o;
}

}  // namespace gcs_gen
