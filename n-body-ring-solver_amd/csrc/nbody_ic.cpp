// nbody_ic.cpp -- host-side system set-up: seeded initial conditions and the
// Yoshida composition coefficients.  Deterministic for (kind, N, seed,
// params): the random stream is splitmix64-seeded xoshiro256** (Blackman &
// Vigna 2018) turned into 53-bit uniforms, so any caller that implements the
// same published generator draws the same particles.
//
// Compiled with -ffp-contract=off so no fused multiply-add changes rounding.
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "nbody_host.h"

namespace nbh {

namespace {

struct Rng {
    uint64_t s[4];

    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto& w : s) w = splitmix(x);
    }
    static uint64_t splitmix(uint64_t& x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t out = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return out;
    }
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }   // [0, 1)

    // isotropic vector of length len: cos(theta) uniform in [-1, 1), azimuth uniform
    void iso(double len, double& ox, double& oy, double& oz) {
        const double cz = 2.0 * uniform() - 1.0;
        const double ph = 2.0 * M_PI * uniform();
        const double sz = std::sqrt(1.0 - cz * cz);
        ox = len * sz * std::cos(ph);
        oy = len * sz * std::sin(ph);
        oz = len * cz;
    }
};

}  // namespace

void ic_sphere(size_t n, uint64_t seed, double* m, double* const pos[3], double* const vel[3]) {
    Rng r(seed);
    for (size_t i = 0; i < n; i++) {
        double px, py, pz;
        do {
            px = 2.0 * r.uniform() - 1.0;
            py = 2.0 * r.uniform() - 1.0;
            pz = 2.0 * r.uniform() - 1.0;
        } while (px * px + py * py + pz * pz >= 1.0);
        m[i] = 1.0 / double(n);
        pos[0][i] = px; pos[1][i] = py; pos[2][i] = pz;
        vel[0][i] = vel[1][i] = vel[2][i] = 0.0;
    }
}

// Plummer sphere (G = M = 1, scale a), radii from the inverted cumulative mass
// (Aarseth, Henon & Wielen 1974) resampled until r < rmax, speeds q v_esc with
// q from von Neumann rejection on q^2 (1-q^2)^(7/2) under the bound 0.1, and
// v_esc = sqrt(2) (r^2 + a^2)^(-1/4).  The mass-weighted centre of mass and
// its velocity are subtracted afterwards.
void ic_plummer(size_t n, uint64_t seed, double a, double rmax, double* m, double* const pos[3],
                double* const vel[3]) {
    Rng r(seed);
    for (size_t i = 0; i < n; i++) {
        double rad;
        for (;;) {
            const double u = r.uniform();
            if (u <= 0.0) continue;
            rad = a / std::sqrt(std::pow(u, -2.0 / 3.0) - 1.0);
            if (rad < rmax) break;
        }
        r.iso(rad, pos[0][i], pos[1][i], pos[2][i]);
        double q, g;
        do {
            q = r.uniform();
            g = 0.1 * r.uniform();
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double vesc = std::sqrt(2.0) * std::pow(rad * rad + a * a, -0.25);
        r.iso(q * vesc, vel[0][i], vel[1][i], vel[2][i]);
        m[i] = 1.0 / double(n);
    }
    double cm[6] = {0, 0, 0, 0, 0, 0}, mt = 0.0;
    for (size_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            cm[k] += m[i] * pos[k][i];
            cm[3 + k] += m[i] * vel[k][i];
        }
        mt += m[i];
    }
    for (double& v : cm) v /= mt;
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            pos[k][i] -= cm[k];
            vel[k][i] -= cm[3 + k];
        }
}

// Ring of n-1 equal masses rotating rigidly at omega around a central mass
// ratio*m (the last particle), masses chosen so that gravity supplies exactly
// the centripetal force; also the analytic force.  Same construction as
// NBody_system::init_stable_orbiting_particles (CPU/nbody_system.cpp:235-300),
// whose normalisation sums the half-ring terms 1/sqrt(2-2cos(theta_i)) written
// as (2+2cos(pi(n-2i)/n))/(2-2cos(2 pi i/n))^(3/2).  G = 1.
void ic_ring(size_t n_total, double R, double omega, double ratio, double* m, double* const pos[3],
             double* const vel[3], double* const force[3]) {
    const double G = 1.0;
    const size_t n = n_total - 1;
    double total_mass = R * R * R * omega * omega / G;
    double sum = 0.0;
    for (size_t i = 1; i <= (n - 1) / 2; i++)
        sum += (2.0 + 2.0 * std::cos((M_PI * double(n - 2 * i)) / double(n))) /
               std::pow(2.0 - 2.0 * std::cos(double(i) * 2 * M_PI / double(n)), 3.0 / 2.0);
    if (n % 2 == 0) sum += 1.0 / 4.0;
    sum += ratio;
    total_mass /= sum;
    for (size_t i = 0; i < n; i++) {
        // one sincos per angle, as the reference's compiled code evaluates it
        // (its sin/cos pairs are merged by the compiler): identical bits
        const double th = double(i) * 2 * M_PI / double(n);
        double sn, cs;
        ::sincos(th, &sn, &cs);
        m[i] = total_mass;
        pos[0][i] = R * cs;
        pos[1][i] = R * sn;
        pos[2][i] = 0.0;
        vel[0][i] = -R * omega * sn;
        vel[1][i] = R * omega * cs;
        vel[2][i] = 0.0;
        if (force) {
            force[0][i] = -G * m[i] * R * omega * omega * cs;
            force[1][i] = -G * m[i] * R * omega * omega * sn;
            force[2][i] = 0.0;
        }
    }
    m[n] = ratio * total_mass;
    for (int k = 0; k < 3; k++) {
        pos[k][n] = 0.0;
        vel[k][n] = 0.0;
        if (force) force[k][n] = 0.0;
    }
}

// Yoshida (1990) composition for DKD leapfrog, order 2/4/6/8: w from the
// paper's solutions (order 6: solution A; order 8: solution D, as the
// reference uses), with the central weight w0 = 1 - 2 sum_{i>0} w_i; then
// c[0] = c[2s-1] = w[s-1]/2, c[i] = c[2s-1-i] = (w[s-i] + w[s-1-i])/2,
// d[i] = d[2s-2-i] = w[s-1-i]  (CPU/omp_solver.cpp:677-726).
size_t yoshida(int order, double* c, double* d) {
    // 2^(1/3) correctly rounded (what a compiler folds std::cbrt(2.0) to; the
    // runtime libm cbrt returns the next double up)
    constexpr double cbrt2 = 0x1.428a2f98d728bp+0;
    double w2[1] = {1.0};
    double w4[2] = {-cbrt2 / (2.0 - cbrt2), 1.0 / (2.0 - cbrt2)};
    double w6[4] = {1.0, -1.17767998417887, 0.235573213359357, 0.78451361047756};
    for (int i = 1; i < 4; i++) w6[0] -= 2.0 * w6[i];
    double w8[8] = {1.0,                  -0.161582374150097e1, -0.244699182370524e1,
                    -0.716989419708120e-2, 0.244002732616735e1, 0.157739928123617,
                    0.182020630970714e1,  0.104242620869991e1};
    for (int i = 1; i < 8; i++) w8[0] -= 2.0 * w8[i];
    const double* w = w2;
    size_t s = 1;
    if (order == 4) { w = w4; s = 2; }
    else if (order == 6) { w = w6; s = 4; }
    else if (order == 8) { w = w8; s = 8; }
    for (size_t i = 0; i < s; i++) {
        c[i] = c[2 * s - 1 - i] = (i == 0) ? w[s - 1] / 2.0 : (w[s - i] + w[s - 1 - i]) / 2.0;
        d[i] = d[2 * s - 2 - i] = w[s - 1 - i];
    }
    return 2 * s;
}

// Classic fourth-order Runge-Kutta (Kutta 1901) in the reference's layout
// (CPU/omp_solver.cpp:772-797): a is the (s-1) x (s-1) row-major lower
// triangle of the Butcher tableau without its zero first row, b the weights.
// Order 4 is the only tableau; every order maps to it.
size_t runge_kutta(int /*order*/, double* a, double* b) {
    const double a4[9] = {0.5, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 1.0};
    const double b4[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    for (int i = 0; i < 9; i++) a[i] = a4[i];
    for (int i = 0; i < 4; i++) b[i] = b4[i];
    return 4;
}

}  // namespace nbh
