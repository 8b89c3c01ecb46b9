// nbx_ics.cpp -- seeded initial conditions (host only).
//
// The reference draws its disks from an unseeded std::default_random_engine
// (src/BarnesHut.cu:356-415), whose distributions are implementation-defined;
// these generators instead use xoshiro256** seeded through splitmix64 and
// plain libm, so a seed gives the same bodies on every host of one image.
// Quantities are computed in FP64 and rounded to float once.  G = 1.
#include <cmath>
#include <cstdint>

#include "../../include/nbx.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (auto &w : s) {  // splitmix64
            z += 0x9e3779b97f4a7c15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ull;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebull;
            w = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {  // xoshiro256**
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }               // [0, 1)
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    double normal(double sigma)
    {  // Box-Muller, u1 in (0, 1]
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

bool bad(int n, const float *a, const float *b, const float *c, const float *d, const float *e, const float *f,
         const float *g)
{
    return n < 0 || !a || !b || !c || !d || !e || !f || !g;
}

// Keplerian disk around a central mass mc: r ~ U[r0, r1], phi ~ U[0, 2pi),
// z ~ N(0, sigma_z), v = sqrt(mc / r) along +phi (counter-clockwise).
void disk(Rng &rng, int lo, int hi, double r0, double r1, double sigma_z, double mb, double mc, float *x,
          float *y, float *z, float *vx, float *vy, float *vz, float *mass)
{
    for (int i = lo; i < hi; ++i) {
        const double r = rng.uniform(r0, r1);
        const double ph = rng.uniform(0.0, 2.0 * M_PI);
        const double zz = rng.normal(sigma_z);
        const double v = std::sqrt(mc / r);
        const double c = std::cos(ph), s = std::sin(ph);
        x[i] = (float)(r * c); y[i] = (float)(r * s); z[i] = (float)zz;
        vx[i] = (float)(-v * s); vy[i] = (float)(v * c); vz[i] = 0.f;
        mass[i] = (float)mb;
    }
}

}  // namespace

extern "C" {

// Plain-method CPU case: star of mass 1 at rest at the origin plus n-1 disk bodies of
// mass 1e-6, r ~ U[0.1, 1], z ~ N(0, 0.01) (0.01 read as the standard deviation).
int nbx_ic_star_disk(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy, float *vz,
                     float *mass)
{
    if (bad(n, x, y, z, vx, vy, vz, mass)) return NBX_ERR_INVALID;
    if (n == 0) return NBX_OK;
    Rng rng(seed);
    x[0] = y[0] = z[0] = 0.f;
    vx[0] = vy[0] = vz[0] = 0.f;
    mass[0] = 1.f;
    disk(rng, 1, n, 0.1, 1.0, 0.01, 1e-6, 1.0, x, y, z, vx, vy, vz, mass);
    return NBX_OK;
}

// Plummer sphere, total mass 1, scale radius 1, equal masses.  Radii from the
// inverted cumulative mass r = (U^(-2/3) - 1)^(-1/2) with U truncated to
// (0, 0.999] (the customary mass cut of Aarseth, Henon & Wielen 1974; without it
// a million-body sample reaches r ~ 10^3); speeds from the rejection sampling of
// q^2 (1 - q^2)^3.5 with v = q sqrt(2) (1 + r^2)^(-1/4); isotropic directions;
// then shifted to the centre-of-mass frame.
int nbx_ic_plummer(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy, float *vz,
                   float *mass)
{
    if (bad(n, x, y, z, vx, vy, vz, mass)) return NBX_ERR_INVALID;
    if (n == 0) return NBX_OK;
    Rng rng(seed);
    double *buf = new double[6 * (size_t)n];
    double cm[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < n; ++i) {
        double u;
        do { u = rng.uniform(); } while (u <= 0.0 || u > 0.999);
        const double r = 1.0 / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
        double ct = rng.uniform(-1.0, 1.0), ph = rng.uniform(0.0, 2.0 * M_PI), st = std::sqrt(1.0 - ct * ct);
        double *b = buf + 6 * (size_t)i;
        b[0] = r * st * std::cos(ph); b[1] = r * st * std::sin(ph); b[2] = r * ct;
        double q, g;
        do {
            q = rng.uniform();
            g = rng.uniform(0.0, 0.1);
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double v = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        ct = rng.uniform(-1.0, 1.0); ph = rng.uniform(0.0, 2.0 * M_PI); st = std::sqrt(1.0 - ct * ct);
        b[3] = v * st * std::cos(ph); b[4] = v * st * std::sin(ph); b[5] = v * ct;
        for (int k = 0; k < 6; ++k) cm[k] += b[k];
    }
    for (int k = 0; k < 6; ++k) cm[k] /= n;
    const float mb = (float)(1.0 / n);
    for (int i = 0; i < n; ++i) {
        const double *b = buf + 6 * (size_t)i;
        x[i] = (float)(b[0] - cm[0]); y[i] = (float)(b[1] - cm[1]); z[i] = (float)(b[2] - cm[2]);
        vx[i] = (float)(b[3] - cm[3]); vy[i] = (float)(b[4] - cm[4]); vz[i] = (float)(b[5] - cm[5]);
        mass[i] = mb;
    }
    delete[] buf;
    return NBX_OK;
}

// Precision-form case: circular binary (0.5 + 0.5 at separation 0.2, each star at
// speed 0.5 sqrt(1/0.2) about the origin) plus n-2 disk bodies of mass 1e-7 with
// r ~ U[0.3, 2], z ~ N(0, 0.02), circular speed about the central mass 1.
int nbx_ic_binary_disk(int n, uint64_t seed, float *x, float *y, float *z, float *vx, float *vy, float *vz,
                       float *mass)
{
    if (bad(n, x, y, z, vx, vy, vz, mass) || n < 2) return NBX_ERR_INVALID;
    Rng rng(seed);
    const double vb = 0.5 * std::sqrt(1.0 / 0.2);
    x[0] = 0.1f; y[0] = 0.f; z[0] = 0.f; vx[0] = 0.f; vy[0] = (float)vb; vz[0] = 0.f; mass[0] = 0.5f;
    x[1] = -0.1f; y[1] = 0.f; z[1] = 0.f; vx[1] = 0.f; vy[1] = (float)-vb; vz[1] = 0.f; mass[1] = 0.5f;
    disk(rng, 2, n, 0.3, 2.0, 0.02, 1e-7, 1.0, x, y, z, vx, vy, vz, mass);
    return NBX_OK;
}

}  // extern "C"
