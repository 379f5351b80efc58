// nb_gen.cpp -- seeded synthetic inputs (host code, part of libnbody_hip.so).
//
// Random numbers: splitmix64 seeds xoshiro256** (Blackman & Vigna 2018);
// uniform doubles are (u >> 11) * 2^-53.  Draws happen in a fixed order per
// body, so a seed defines the input completely (given the C library's
// sin/cos/atan2/pow).
#include <cmath>
#include <cstdint>

#include "nbody_capi.h"

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (auto &w : s) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

constexpr double kSolarMass = 1.98847e30;  // kg, dataset_converter/src/conversion.cpp:96-97
constexpr double kPi = 3.14159265358979323846;

double g_ref() {
    nb_params p;
    nb_default_params(&p);
    return p.G;
}

}  // namespace

extern "C" {

int nb_gen_star_disk(uint32_t n, uint64_t seed, double *x, double *y, double *z, double *vx,
                     double *vy, double *vz, double *m) {
    if (n == 0 || !x || !y || !z || !vx || !vy || !vz || !m) return NB_E_INVALID_ARG;
    Rng rng(seed);
    x[0] = y[0] = z[0] = vx[0] = vy[0] = vz[0] = 0.0;
    m[0] = kSolarMass;
    const double mu = g_ref() * kSolarMass;
    for (uint32_t b = 1; b < n; ++b) {
        const double a = 2.1 + 1.2 * rng.uniform();
        const double e = 0.3 * rng.uniform();
        const double inc = (20.0 * kPi / 180.0) * rng.uniform();
        const double om = 2.0 * kPi * rng.uniform();  // longitude of the ascending node
        const double w = 2.0 * kPi * rng.uniform();   // argument of periapsis
        const double ma = 2.0 * kPi * rng.uniform();  // mean anomaly
        m[b] = std::pow(10.0, -20.0 + 8.0 * rng.uniform()) * kSolarMass;
        // Kepler's equation by 30 Newton steps, then the orbital-frame state
        // rotated into the inertial frame (the converter's maths,
        // dataset_converter/src/conversion.cpp:24-71, at t = t0).
        double E = ma;
        for (int it = 0; it < 30; ++it) E = E - (E - e * std::sin(E) - ma) / (1.0 - e * std::cos(E));
        const double nu = 2.0 * std::atan2(std::sqrt(1.0 + e) * std::sin(E / 2.0),
                                           std::sqrt(1.0 - e) * std::cos(E / 2.0));
        const double rc = a * (1.0 - e * std::cos(E));
        const double ox = rc * std::cos(nu), oy = rc * std::sin(nu);
        const double vfac = std::sqrt(mu * a) / rc;
        const double ovx = vfac * -std::sin(E);
        const double ovy = vfac * (std::sqrt(1.0 - e * e) * std::cos(E));
        const double cw = std::cos(w), sw = std::sin(w), co = std::cos(om), so = std::sin(om);
        const double ci = std::cos(inc), si = std::sin(inc);
        const double r11 = cw * co - sw * ci * so, r12 = -(sw * co + cw * ci * so);
        const double r21 = cw * so + sw * ci * co, r22 = cw * ci * co - sw * so;
        const double r31 = sw * si, r32 = cw * si;
        x[b] = ox * r11 + oy * r12;
        y[b] = ox * r21 + oy * r22;
        z[b] = ox * r31 + oy * r32;
        vx[b] = ovx * r11 + ovy * r12;
        vy[b] = ovx * r21 + ovy * r22;
        vz[b] = ovx * r31 + ovy * r32;
    }
    return NB_OK;
}

int nb_gen_plummer(uint32_t n, uint64_t seed, double *x, double *y, double *z, double *vx,
                   double *vy, double *vz, double *m) {
    if (n == 0 || !x || !y || !z || !vx || !vy || !vz || !m) return NB_E_INVALID_ARG;
    Rng rng(seed);
    const double mass = (1.0 / (double)n) / g_ref();
    double sx = 0, sy = 0, sz = 0, svx = 0, svy = 0, svz = 0;
    for (uint32_t b = 0; b < n; ++b) {
        double r;
        do {  // cumulative mass X -> radius, truncated at r = 100
            double X;
            do X = rng.uniform(); while (X <= 0.0);
            r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        } while (!(r <= 100.0));
        double cz = 1.0 - 2.0 * rng.uniform();
        double ph = 2.0 * kPi * rng.uniform();
        double sz1 = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        x[b] = r * sz1 * std::cos(ph);
        y[b] = r * sz1 * std::sin(ph);
        z[b] = r * cz;
        // speed fraction q from g(q) = q^2 (1 - q^2)^(7/2) by von Neumann rejection
        double q, g;
        do {
            q = rng.uniform();
            g = 0.1 * rng.uniform();
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double v = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        cz = 1.0 - 2.0 * rng.uniform();
        ph = 2.0 * kPi * rng.uniform();
        sz1 = std::sqrt(std::max(0.0, 1.0 - cz * cz));
        vx[b] = v * sz1 * std::cos(ph);
        vy[b] = v * sz1 * std::sin(ph);
        vz[b] = v * cz;
        m[b] = mass;
        sx += x[b]; sy += y[b]; sz += z[b];
        svx += vx[b]; svy += vy[b]; svz += vz[b];
    }
    // equal masses: centre of mass = mean; put it at rest at the origin
    sx /= n; sy /= n; sz /= n; svx /= n; svy /= n; svz /= n;
    for (uint32_t b = 0; b < n; ++b) {
        x[b] -= sx; y[b] -= sy; z[b] -= sz;
        vx[b] -= svx; vy[b] -= svy; vz[b] -= svz;
    }
    return NB_OK;
}

}  // extern "C"
