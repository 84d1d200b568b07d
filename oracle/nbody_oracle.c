/*
 * nbody_oracle.c -- CPU oracle for the softened all-pairs force + leapfrog path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The shipped solver (libnbody_hip.so) never calls into it.
 *
 * Plain C99 + OpenMP (parallel over i only, so every sum has a fixed order and
 * the results are deterministic for any thread count).  All arithmetic is
 * IEEE double unless a function says otherwise; build with -ffp-contract=off
 * so no fused multiply-add changes the rounding of the IC generators.
 *
 * Restated from the reference (citations relative to the reference tree):
 *   - pair force  F_kq = -G m_k m_q (x_k - x_q)/|x_k - x_q|^3
 *       CPU/force_calculation.cpp:78-82 (scalar), :597-611 (AVX2 blocked)
 *     generalised here with Plummer softening: |d|^3 -> (|d|^2 + eps^2)^(3/2);
 *     eps = 0 gives the reference formula exactly.
 *   - reduced (Newton's 3rd law) evaluation: each unordered pair once, +f on
 *     k and -f on q   CPU/force_calculation.cpp:477-690
 *   - DKD Yoshida leapfrog   CPU/omp_solver.cpp:305-421
 *   - explicit Euler and Runge-Kutta drivers   CPU/omp_solver.cpp:60-137, 512-606
 *   - force-error / deviation / inertia metrics of the accuracy experiments
 *       CPU/nbody_system.cpp:437-567, CPU/force_calculation.cpp:1911-1919
 *   - Yoshida c/d coefficients   CPU/omp_solver.cpp:677-726
 *   - kinetic / potential energy   CPU/force_calculation.cpp:1740-1748, 1880-1904
 *   - ring + centre initial condition and its analytic force
 *       CPU/nbody_system.cpp:235-300
 *   - rigid-rotation prediction   CPU/nbody_system.cpp:302-324
 *
 * Functions with no reference counterpart (softening closed forms, Plummer and
 * uniform-sphere generators, compensated sums, extrapolation features, the
 * FP64 MLP) are pinned by the tests against closed forms, long-double brute
 * force and distribution statistics; each says how in its comment.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------- */
/* Random numbers: splitmix64 seeding + xoshiro256** (Blackman & Vigna 2018). */
/* The product library implements the same published generator so that both  */
/* sides draw the same initial conditions from the same seed.                 */
/* ------------------------------------------------------------------------- */
typedef struct { uint64_t s[4]; } orc_rng;

static uint64_t orc_splitmix64(uint64_t *x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void orc_rng_seed(orc_rng *r, uint64_t seed)
{
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = orc_splitmix64(&x);
}

static uint64_t orc_rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t orc_rng_next(orc_rng *r)
{
    uint64_t *s = r->s;
    uint64_t out = orc_rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = orc_rotl(s[3], 45);
    return out;
}

/* uniform double in [0, 1) with 53 random bits */
static double orc_uniform(orc_rng *r) { return (double)(orc_rng_next(r) >> 11) * 0x1.0p-53; }

ORC_API void orc_rng_stream(uint64_t seed, size_t n, double *out)
{
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = orc_uniform(&r);
}

/* ------------------------------------------------------------------------- */
/* Initial conditions                                                        */
/* ------------------------------------------------------------------------- */

/* Uniform density in the unit sphere (rejection from the cube), m = 1/N, v = 0.
 * Cold-collapse set-up; no reference counterpart (the reference's random cube
 * CPU/nbody_system.cpp:154-190 is unseeded and never called). */
ORC_API void orc_ic_sphere(size_t n, uint64_t seed, double *m, double *x, double *y,
                           double *z, double *vx, double *vy, double *vz)
{
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (size_t i = 0; i < n; i++) {
        double px, py, pz;
        do {
            px = 2.0 * orc_uniform(&r) - 1.0;
            py = 2.0 * orc_uniform(&r) - 1.0;
            pz = 2.0 * orc_uniform(&r) - 1.0;
        } while (px * px + py * py + pz * pz >= 1.0);
        m[i] = 1.0 / (double)n;
        x[i] = px; y[i] = py; z[i] = pz;
        vx[i] = 0.0; vy[i] = 0.0; vz[i] = 0.0;
    }
}

/* Isotropic unit vector times len: z uniform in [-1,1], azimuth uniform. */
static void orc_iso(orc_rng *r, double len, double *ox, double *oy, double *oz)
{
    double cz = 2.0 * orc_uniform(r) - 1.0;
    double ph = 2.0 * M_PI * orc_uniform(r);
    double sz = sqrt(1.0 - cz * cz);
    *ox = len * sz * cos(ph);
    *oy = len * sz * sin(ph);
    *oz = len * cz;
}

/* Plummer sphere, G = M = 1, scale a, truncated at r < rmax (resampled).
 * Radii by inverting M(r) (Aarseth, Henon & Wielen 1974, A&A 37, 183);
 * speeds by von Neumann rejection on g(q) = q^2 (1-q^2)^(7/2), q = v/v_esc,
 * with the bound 0.1 > max g = 0.092; v_esc = sqrt(2) (r^2 + a^2)^(-1/4)
 * for G = M = 1.  Equal masses 1/N; the centre of mass and its velocity are
 * then moved to the origin (mass-weighted means subtracted). */
ORC_API void orc_ic_plummer(size_t n, uint64_t seed, double a, double rmax, double *m,
                            double *x, double *y, double *z, double *vx, double *vy,
                            double *vz)
{
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (size_t i = 0; i < n; i++) {
        double rad;
        for (;;) {
            double u = orc_uniform(&r);
            if (u <= 0.0) continue;
            rad = a / sqrt(pow(u, -2.0 / 3.0) - 1.0);
            if (rad < rmax) break;
        }
        orc_iso(&r, rad, &x[i], &y[i], &z[i]);
        double q, g;
        do {
            q = orc_uniform(&r);
            g = 0.1 * orc_uniform(&r);
        } while (g > q * q * pow(1.0 - q * q, 3.5));
        double vesc = sqrt(2.0) * pow(rad * rad + a * a, -0.25);
        orc_iso(&r, q * vesc, &vx[i], &vy[i], &vz[i]);
        m[i] = 1.0 / (double)n;
    }
    double cm[6] = {0, 0, 0, 0, 0, 0}, mt = 0.0;
    for (size_t i = 0; i < n; i++) {
        cm[0] += m[i] * x[i]; cm[1] += m[i] * y[i]; cm[2] += m[i] * z[i];
        cm[3] += m[i] * vx[i]; cm[4] += m[i] * vy[i]; cm[5] += m[i] * vz[i];
        mt += m[i];
    }
    for (int k = 0; k < 6; k++) cm[k] /= mt;
    for (size_t i = 0; i < n; i++) {
        x[i] -= cm[0]; y[i] -= cm[1]; z[i] -= cm[2];
        vx[i] -= cm[3]; vy[i] -= cm[4]; vz[i] -= cm[5];
    }
}

/* Ring of N-1 equal masses at radius R rotating rigidly at omega around a
 * central mass ratio*m (the last particle), with the exact centripetal force.
 * Restates CPU/nbody_system.cpp:235-300 (init_stable_orbiting_particles); the
 * mass normalisation sums 1/sqrt(2-2cos(theta_i)) over half the ring exactly as
 * the reference does (:241-248).  G = 1. */
ORC_API void orc_ic_ring(size_t n_total, double R, double omega, double ratio, double *m,
                         double *x, double *y, double *z, double *vx, double *vy,
                         double *vz, double *fx, double *fy, double *fz)
{
    const double G = 1.0;
    size_t n = n_total - 1;
    double total_mass = R * R * R * omega * omega / G;
    double sum = 0.0;
    for (size_t i = 1; i <= (n - 1) / 2; i++)
        sum += (2.0 + 2.0 * cos((M_PI * (double)(n - 2 * i)) / (double)n)) /
               pow(2.0 - 2.0 * cos((double)i * 2 * M_PI / (double)n), 3.0 / 2.0);
    if (n % 2 == 0) sum += 1.0 / 4.0;
    sum += ratio;
    total_mass /= sum;
    for (size_t i = 0; i < n; i++) {
        /* one sincos per angle: what the reference's compiled code evaluates
           (gcc merges its sin/cos pairs), so the ICs agree bit for bit */
        double th = (double)i * 2 * M_PI / (double)n, sn, cs;
        sincos(th, &sn, &cs);
        m[i] = total_mass;
        x[i] = R * cs; y[i] = R * sn; z[i] = 0.0;
        vx[i] = -R * omega * sn; vy[i] = R * omega * cs; vz[i] = 0.0;
        fx[i] = -G * m[i] * R * omega * omega * cs;
        fy[i] = -G * m[i] * R * omega * omega * sn;
        fz[i] = 0.0;
    }
    m[n] = ratio * total_mass;
    x[n] = y[n] = z[n] = 0.0;
    vx[n] = vy[n] = vz[n] = 0.0;
    fx[n] = fy[n] = fz[n] = 0.0;
}

/* Rigid-rotation prediction of the ring after fraction `phase` of a revolution
 * (CPU/nbody_system.cpp:302-324 with phase = curr_step/steps_per_rev). */
ORC_API void orc_ring_predict(size_t n_total, double R, double omega, double phase,
                              double *x, double *y, double *z, double *vx, double *vy,
                              double *vz)
{
    size_t n = n_total - 1;
    for (size_t i = 0; i < n; i++) {
        double th = ((double)i / (double)n + phase) * 2 * M_PI;
        x[i] = R * cos(th); y[i] = R * sin(th); z[i] = 0.0;
        vx[i] = -R * omega * sin(th); vy[i] = R * omega * cos(th); vz[i] = 0.0;
    }
    x[n] = y[n] = z[n] = 0.0;
    vx[n] = vy[n] = vz[n] = 0.0;
}

/* Analytic softened acceleration of ring particle k (G = 1): every |d|^3 of
 * the reference's closed form becomes (|d|^2 + eps^2)^(3/2).  The ring
 * particles pull particle k inwards with G m R (1-cos th)/(R^2(2-2cos th)+eps^2)^(3/2)
 * each and the centre with G M_c R/(R^2+eps^2)^(3/2); the centre feels zero.
 * Summed over all n-1 partners directly (no half-ring folding), so it is an
 * independent check of orc_ic_ring's normalisation at eps = 0. */
ORC_API void orc_ring_accel_exact(size_t n_total, double R, double mring, double mcentre,
                                  double eps, double *ax, double *ay, double *az)
{
    size_t n = n_total - 1;
    long double inward = 0.0L;
    for (size_t i = 1; i < n; i++) {
        long double th = (long double)i * 2.0L * 3.14159265358979323846264338327950288L / (long double)n;
        long double c = cosl(th);
        long double d2 = (long double)R * R * (2.0L - 2.0L * c) + (long double)eps * eps;
        inward += (long double)mring * R * (1.0L - c) / (d2 * sqrtl(d2));
    }
    long double dc = (long double)R * R + (long double)eps * eps;
    inward += (long double)mcentre * R / (dc * sqrtl(dc));
    for (size_t k = 0; k < n; k++) {
        long double th = (long double)k * 2.0L * 3.14159265358979323846264338327950288L / (long double)n;
        ax[k] = (double)(-inward * cosl(th));
        ay[k] = (double)(-inward * sinl(th));
        az[k] = 0.0;
    }
    ax[n] = ay[n] = az[n] = 0.0;
}

/* ------------------------------------------------------------------------- */
/* Compensated (Neumaier 1974) accumulation                                  */
/* ------------------------------------------------------------------------- */
typedef struct { double s, c; } orc_ksum;

static inline void orc_kadd(orc_ksum *k, double v)
{
    double t = k->s + v;
    if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v;
    else k->c += (v - t) + k->s;
    k->s = t;
}

static inline double orc_kval(const orc_ksum *k) { return k->s + k->c; }

/* ------------------------------------------------------------------------- */
/* Forces / accelerations                                                    */
/* ------------------------------------------------------------------------- */

/* Softened acceleration of every particle from all others (G = 1 scaled by G):
 *   a_i = G sum_{j != i} m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^(3/2)
 * Pairs at exactly zero separation contribute nothing (the self pair, and the
 * reference's eq_cond / invalid-particle mask FPGA/device/ring_lrbd_lf_no_sync_local.cl:157).
 * Neumaier-compensated per component; the divide uses sqrt like
 * CPU/force_calculation.cpp:78-82. */
ORC_API void orc_accel(size_t n, const double *m, const double *x, const double *y,
                       const double *z, double eps, double G, double *ax, double *ay,
                       double *az)
{
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < n; i++) {
        orc_ksum sx = {0, 0}, sy = {0, 0}, sz = {0, 0};
        for (size_t j = 0; j < n; j++) {
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            double s2 = r2 + e2;
            double inv = 1.0 / (s2 * sqrt(s2));
            double f = m[j] * inv;
            orc_kadd(&sx, f * dx);
            orc_kadd(&sy, f * dy);
            orc_kadd(&sz, f * dz);
        }
        ax[i] = G * orc_kval(&sx);
        ay[i] = G * orc_kval(&sy);
        az[i] = G * orc_kval(&sz);
    }
}

/* Same sum in long double (x87 80-bit), uncompensated: an independent
 * brute-force reference for small N. */
ORC_API void orc_accel_ld(size_t n, const double *m, const double *x, const double *y,
                          const double *z, double eps, double G, double *ax, double *ay,
                          double *az)
{
    const long double e2 = (long double)eps * eps;
    for (size_t i = 0; i < n; i++) {
        long double sx = 0, sy = 0, sz = 0;
        for (size_t j = 0; j < n; j++) {
            long double dx = (long double)x[j] - x[i], dy = (long double)y[j] - y[i],
                        dz = (long double)z[j] - z[i];
            long double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0L) continue;
            long double s2 = r2 + e2;
            long double f = (long double)m[j] / (s2 * sqrtl(s2));
            sx += f * dx; sy += f * dy; sz += f * dz;
        }
        ax[i] = (double)(G * sx);
        ay[i] = (double)(G * sy);
        az[i] = (double)(G * sz);
    }
}

/* Ring-share acceleration: ni i-particles against nj j-particles, both given
 * as interleaved (x, y, z, m) quadruples -- the FPGA particle_t layout
 * (FPGA/host/solver.cpp:721).  This is one received round of the reference's
 * hybrid ring (CPU/hybrid_solver.cpp:84-95) in full (non-reduced) form.
 * Zero-separation pairs contribute nothing, so i may be a subset of j. */
ORC_API void orc_accel_block(size_t ni, const double *pi, size_t nj, const double *pj,
                             double eps, double G, double *acc /* ni x 3 */)
{
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < ni; i++) {
        orc_ksum sx = {0, 0}, sy = {0, 0}, sz = {0, 0};
        const double xi = pi[4 * i], yi = pi[4 * i + 1], zi = pi[4 * i + 2];
        for (size_t j = 0; j < nj; j++) {
            double dx = pj[4 * j] - xi, dy = pj[4 * j + 1] - yi, dz = pj[4 * j + 2] - zi;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            double s2 = r2 + e2;
            double f = pj[4 * j + 3] / (s2 * sqrt(s2));
            orc_kadd(&sx, f * dx);
            orc_kadd(&sy, f * dy);
            orc_kadd(&sz, f * dz);
        }
        acc[3 * i] = G * orc_kval(&sx);
        acc[3 * i + 1] = G * orc_kval(&sy);
        acc[3 * i + 2] = G * orc_kval(&sz);
    }
}

/* Reduced (Newton's 3rd law) force evaluation with the reference's pairing:
 * for k < q the pair force f = -G m_k m_q (x_k - x_q)/(r^2+eps^2)^(3/2) is added
 * to F_k and subtracted from F_q, as in compute_forces_avx2_blocked
 * (CPU/force_calculation.cpp:530-531, 617-630).  Output is FORCE (not
 * acceleration), overwritten.  Serial: the pair order is k-major. */
ORC_API void orc_forces_reduced(size_t n, const double *m, const double *x, const double *y,
                                const double *z, double eps, double G, double *fx,
                                double *fy, double *fz)
{
    const double e2 = eps * eps;
    for (size_t i = 0; i < n; i++) fx[i] = fy[i] = fz[i] = 0.0;
    for (size_t k = 0; k < n; k++) {
        double pre = -G * m[k];
        for (size_t q = k + 1; q < n; q++) {
            double dx = x[k] - x[q], dy = y[k] - y[q], dz = z[k] - z[q];
            double s2 = dx * dx + dy * dy + dz * dz + e2;
            double dist = sqrt(s2);
            double sc = pre * m[q] / (dist * dist * dist);
            fx[k] += sc * dx; fy[k] += sc * dy; fz[k] += sc * dz;
            fx[q] -= sc * dx; fy[q] -= sc * dy; fz[q] -= sc * dz;
        }
    }
}

/* One ring round of the reference's reduced scheme, as hybrid_ring_solver
 * runs it per received block (CPU/hybrid_solver.cpp:81-95 calling
 * compute_forces_avx2_blocked, CPU/force_calculation.cpp:477-690): the local
 * block of n particles (positions lx..lz, masses mass[rank*n + k]) against
 * the received block (rx..rz, masses mass[partner*n + q], partner =
 * (rank + recv_round) % comm_sz, :531).  Pairs q >= k when rank < partner,
 * else q > k (:530); the pair force f = -G m_k m_q (x_k - x_q) /
 * (r^2+eps^2)^(3/2) is ADDED to the local force F_k and SUBTRACTED from the
 * received block's force R_q (:617-630, :681-688: both accumulate into the
 * caller's arrays).  Neumaier-compensated per component, starting from the
 * caller's values; pairs at exactly zero separation contribute nothing. */
ORC_API void orc_ring_round(size_t n, const double *mass, size_t rank, size_t comm_sz, size_t recv_round,
                            const double *lx, const double *ly, const double *lz, const double *rx,
                            const double *ry, const double *rz, double eps, double G, double *fx, double *fy,
                            double *fz, double *rfx, double *rfy, double *rfz)
{
    const double e2 = eps * eps;
    const size_t partner = (rank + recv_round) % comm_sz;
    const double *ml = mass + rank * n, *mr = mass + partner * n;
    double *out[6] = {fx, fy, fz, rfx, rfy, rfz};
    orc_ksum *acc = (orc_ksum *)malloc(6 * (n ? n : 1) * sizeof(orc_ksum));
    if (!acc) return;
    for (int c = 0; c < 6; c++)
        for (size_t i = 0; i < n; i++) acc[c * n + i] = (orc_ksum){out[c][i], 0.0};
    for (size_t k = 0; k < n; k++) {
        const double pre = -G * ml[k];
        for (size_t q = (rank < partner) ? k : k + 1; q < n; q++) {
            double d[3] = {lx[k] - rx[q], ly[k] - ry[q], lz[k] - rz[q]};
            double s2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + e2;
            if (!(s2 > 0.0)) continue;
            double dist = sqrt(s2);
            double sc = pre * mr[q] / (dist * dist * dist);
            for (int c = 0; c < 3; c++) {
                orc_kadd(&acc[c * n + k], sc * d[c]);
                orc_kadd(&acc[(3 + c) * n + q], -(sc * d[c]));
            }
        }
    }
    for (int c = 0; c < 6; c++)
        for (size_t i = 0; i < n; i++) out[c][i] = orc_kval(&acc[c * n + i]);
    free(acc);
}

/* ------------------------------------------------------------------------- */
/* Integrator                                                                */
/* ------------------------------------------------------------------------- */

/* Yoshida (1990) composition coefficients for DKD leapfrog of order 2/4/6/8,
 * built exactly as CPU/omp_solver.cpp:677-726: w from the paper's tables,
 * c[i] = c[2s-1-i] = (w[s-i] + w[s-1-i])/2 (c[0] = w[s-1]/2),
 * d[i] = d[2s-2-i] = w[s-1-i].  Unknown orders fall back to 2 (:703-707).
 * Returns lf_steps = 2s (number of c's; there are 2s-1 d's). */
ORC_API size_t orc_yoshida(int order, double *c, double *d)
{
    double w2[1] = {1.0};
    double w4[2] = {-cbrt(2.0) / (2.0 - cbrt(2.0)), 1.0 / (2.0 - cbrt(2.0))};
    double w6[4] = {1.0, -1.17767998417887, 0.235573213359357, 0.78451361047756};
    for (int i = 1; i < 4; i++) w6[0] -= 2.0 * w6[i];
    double w8[8] = {1.0, -0.161582374150097e1, -0.244699182370524e1, -0.716989419708120e-2,
                    0.244002732616735e1, 0.157739928123617, 0.182020630970714e1,
                    0.104242620869991e1};
    for (int i = 1; i < 8; i++) w8[0] -= 2.0 * w8[i];
    double *w;
    size_t s;
    switch (order) {
    case 4: w = w4; s = 2; break;
    case 6: w = w6; s = 4; break;
    case 8: w = w8; s = 8; break;
    default: w = w2; s = 1; break;
    }
    for (size_t i = 0; i < s; i++) {
        if (i == 0) c[i] = c[2 * s - 1 - i] = w[s - 1] / 2.0;
        else c[i] = c[2 * s - 1 - i] = (w[s - i] + w[s - 1 - i]) / 2.0;
        d[i] = d[2 * s - 2 - i] = w[s - 1 - i];
    }
    return 2 * s;
}

/* DKD leapfrog, restating reduced_solver_omp_lf (CPU/omp_solver.cpp:305-421):
 * per step, for each of the lf_steps-1 substeps: drift x += c[s] v dt, evaluate
 * accelerations, kick v += d[s] a dt; then the final drift c[lf_steps-1].
 * The reference kicks with F dt / m where F = m a (:394-396); here the
 * acceleration is used directly (equal up to one rounding per kick). */
ORC_API void orc_leapfrog(size_t n, const double *m, double *x, double *y, double *z,
                          double *vx, double *vy, double *vz, size_t steps, double dt,
                          const double *c, const double *d, size_t lf_steps, double eps,
                          double G)
{
    double *ax = (double *)malloc(3 * n * sizeof(double));
    double *ay = ax + n, *az = ax + 2 * n;
    for (size_t t = 0; t < steps; t++) {
        for (size_t s = 0; s + 1 < lf_steps; s++) {
            for (size_t i = 0; i < n; i++) {
                x[i] += c[s] * vx[i] * dt;
                y[i] += c[s] * vy[i] * dt;
                z[i] += c[s] * vz[i] * dt;
            }
            orc_accel(n, m, x, y, z, eps, G, ax, ay, az);
            for (size_t i = 0; i < n; i++) {
                vx[i] += d[s] * ax[i] * dt;
                vy[i] += d[s] * ay[i] * dt;
                vz[i] += d[s] * az[i] * dt;
            }
        }
        for (size_t i = 0; i < n; i++) {
            x[i] += c[lf_steps - 1] * vx[i] * dt;
            y[i] += c[lf_steps - 1] * vy[i] * dt;
            z[i] += c[lf_steps - 1] * vz[i] * dt;
        }
    }
    free(ax);
}

/* Explicit Euler, restating reduced_solver_omp (CPU/omp_solver.cpp:60-137):
 * per step, accelerations of the current positions, then x += v dt with the
 * velocity before the kick, then v += a dt (the reference: v += F dt / m). */
ORC_API void orc_euler(size_t n, const double *m, double *x, double *y, double *z,
                       double *vx, double *vy, double *vz, size_t steps, double dt,
                       double eps, double G)
{
    double *ax = (double *)malloc(3 * (n ? n : 1) * sizeof(double));
    double *ay = ax + n, *az = ax + 2 * n;
    for (size_t t = 0; t < steps; t++) {
        orc_accel(n, m, x, y, z, eps, G, ax, ay, az);
        for (size_t i = 0; i < n; i++) {
            x[i] += vx[i] * dt;
            y[i] += vy[i] * dt;
            z[i] += vz[i] * dt;
            vx[i] += ax[i] * dt;
            vy[i] += ay[i] * dt;
            vz[i] += az[i] * dt;
        }
    }
    free(ax);
}

/* Explicit Runge-Kutta for x' = v, v' = a(x), restating reduced_solver_omp_rk
 * (CPU/omp_solver.cpp:512-606): stage k records its velocity kv_k (kv_0 = v)
 * and the acceleration ka_k of its positions; stage k+1 starts at
 * x0 + (sum_{j<=k} a[k*(s-1)+j] kv_j) dt and v0 + (sum_j a[...] ka_j) dt; the
 * step ends at x0 + (sum_j b_j kv_j) dt, v0 + (sum_j b_j ka_j) dt.  Sums run in
 * stage order from zero, as the reference's.  s <= 8. */
ORC_API void orc_rk(size_t n, const double *m, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz, size_t steps, double dt,
                    const double *a, const double *b, size_t s, double eps, double G)
{
    const size_t nn = n ? n : 1;
    double *buf = (double *)malloc((12 + 6 * s) * nn * sizeof(double));
    double *x0 = buf;              /* step start: component d at x0 + d nn */
    double *u0 = x0 + 3 * nn;
    double *kv = u0 + 3 * nn;      /* stage k, component d: kv + (3k+d) nn */
    double *ka = kv + 3 * s * nn;
    double *sx = ka + 3 * s * nn;  /* stage positions */
    double *sv = sx + 3 * nn;      /* stage velocities */
    const double *X[3] = {x, y, z};
    const double *V[3] = {vx, vy, vz};
    for (size_t t = 0; t < steps; t++) {
        for (int d = 0; d < 3; d++) {
            memcpy(x0 + d * nn, X[d], n * sizeof(double));
            memcpy(u0 + d * nn, V[d], n * sizeof(double));
            memcpy(sx + d * nn, X[d], n * sizeof(double));
            memcpy(sv + d * nn, V[d], n * sizeof(double));
        }
        for (size_t k = 0; k < s; k++) {
            for (int d = 0; d < 3; d++) memcpy(kv + (3 * k + d) * nn, sv + d * nn, n * sizeof(double));
            orc_accel(n, m, sx, sx + nn, sx + 2 * nn, eps, G, ka + 3 * k * nn, ka + (3 * k + 1) * nn,
                      ka + (3 * k + 2) * nn);
            if (k + 1 == s) break;
            for (int d = 0; d < 3; d++)
                for (size_t i = 0; i < n; i++) {
                    double px = 0.0, pv = 0.0;
                    for (size_t j = 0; j <= k; j++) {
                        px += a[k * (s - 1) + j] * kv[(3 * j + d) * nn + i];
                        pv += a[k * (s - 1) + j] * ka[(3 * j + d) * nn + i];
                    }
                    sx[d * nn + i] = px * dt + x0[d * nn + i];
                    sv[d * nn + i] = pv * dt + u0[d * nn + i];
                }
        }
        double *Xo[3] = {x, y, z};
        double *Vo[3] = {vx, vy, vz};
        for (int d = 0; d < 3; d++)
            for (size_t i = 0; i < n; i++) {
                double px = 0.0, pv = 0.0;
                for (size_t j = 0; j < s; j++) {
                    px += b[j] * kv[(3 * j + d) * nn + i];
                    pv += b[j] * ka[(3 * j + d) * nn + i];
                }
                Xo[d][i] = px * dt + x0[d * nn + i];
                Vo[d][i] = pv * dt + u0[d * nn + i];
            }
    }
    free(buf);
}

/* ------------------------------------------------------------------------- */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------- */

/* Force-error metrics, restating calculate_force_error
 * (CPU/nbody_system.cpp:458-567, double-precision branch):
 *   out[0] mean |F-Fr|, out[1] max |F-Fr|,
 *   out[2] sum of |atan2(det, dot)| of the xy projections over particles
 *          with Fr_x != 0 and Fr_y != 0, / (2 pi N),
 *   out[3] sum ||F|-|Fr|| / (L0 N), out[4] max angle / (2 pi),
 *   out[5] max ||F|-|Fr|| / L0,
 * L0 = |Fr_0| computed in float from float-rounded components (the reference
 * reads its single-precision copy there, :566).  The reference's compiled
 * loop contracts products into fused multiply-adds (|v|^2 as
 * fma(z, z, fma(x, x, y*y)), dot as fma(fx, rx, fy*ry), det as
 * fma(fx, ry, -(fy*rx))); they are written out so the values agree bit for bit. */
static double orc_sq3(double x, double y, double z) { return fma(z, z, fma(x, x, y * y)); }

ORC_API void orc_force_error(size_t n, const double *fx, const double *fy, const double *fz,
                             const double *rx, const double *ry, const double *rz, double *out)
{
    double err = 0.0, lerr = 0.0, lmax = 0.0, aerr = 0.0, amax = 0.0, dmax = 0.0;
    for (size_t i = 0; i < n; i++) {
        double dl = sqrt(orc_sq3(fx[i] - rx[i], fy[i] - ry[i], fz[i] - rz[i]));
        double sl = sqrt(orc_sq3(fx[i], fy[i], fz[i]));
        double rl = sqrt(orc_sq3(rx[i], ry[i], rz[i]));
        double ld = fabs(sl - rl);
        double ang = fabs(atan2(fma(fx[i], ry[i], -(fy[i] * rx[i])), fma(fx[i], rx[i], fy[i] * ry[i])));
        int xy = rx[i] != 0.0 && ry[i] != 0.0;
        if (dl > dmax) dmax = dl;
        if (ld > lmax) lmax = ld;
        if (xy && ang > amax) amax = ang;
        err += dl;
        if (xy) aerr += ang;
        lerr += ld;
    }
    float gx = (float)rx[0], gy = (float)ry[0], gz = (float)rz[0];
    float l2 = fmaf(gz, gz, fmaf(gx, gx, gy * gy));
    double L0 = (double)sqrtf(l2);
    out[0] = err / (double)n;
    out[1] = dmax;
    out[2] = aerr / (2 * M_PI * (double)n);
    out[3] = lerr / (L0 * (double)n);
    out[4] = amax / (2 * M_PI);
    out[5] = lmax / L0;
}

/* calculate_deviation (CPU/nbody_system.cpp:437-456): out[0] mean |c - r|,
 * out[1] max ||c| - |r|| / |r| over |r| > 0 (|v|^2 contracted as above). */
ORC_API void orc_deviation(size_t n, const double *cx, const double *cy, const double *cz,
                           const double *rx, const double *ry, const double *rz, double *out)
{
    double mean = 0.0, worst = 0.0;
    for (size_t i = 0; i < n; i++) {
        mean += sqrt(orc_sq3(cx[i] - rx[i], cy[i] - ry[i], cz[i] - rz[i]));
        double lc = sqrt(orc_sq3(cx[i], cy[i], cz[i]));
        double lr = sqrt(orc_sq3(rx[i], ry[i], rz[i]));
        if (lr > 0.0 && fabs(lc - lr) / lr > worst) worst = fabs(lc - lr) / lr;
    }
    out[0] = mean / (double)n;
    out[1] = worst;
}

/* compute_inertia (CPU/force_calculation.cpp:1911-1919): sum m |x|^2 in order;
 * the compiled loop is vectorised with an in-order sum, so m |x|^2 (|x|^2
 * contracted as above) is rounded before it is added. */
ORC_API double orc_inertia(size_t n, const double *m, const double *x, const double *y, const double *z)
{
    double I = 0.0;
    for (size_t i = 0; i < n; i++) I += m[i] * orc_sq3(x[i], y[i], z[i]);
    return I;
}

/* T = 1/2 sum m v^2 (CPU/force_calculation.cpp:1740-1748) and
 * U = -G sum_{k<q} m_k m_q / sqrt(r^2 + eps^2) (:1880-1904 with softening, as
 * FPGA/host/accuracy_evaluation.cpp:33-35 applies EPS2).  Both sums are
 * compensated; U is accumulated per row k in parallel, rows combined in order. */
ORC_API void orc_energy(size_t n, const double *m, const double *x, const double *y,
                        const double *z, const double *vx, const double *vy,
                        const double *vz, double eps, double G, double *T, double *U)
{
    orc_ksum kt = {0, 0};
    for (size_t i = 0; i < n; i++)
        orc_kadd(&kt, m[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]));
    *T = 0.5 * orc_kval(&kt);
    double *row = (double *)malloc((n ? n : 1) * sizeof(double));
    const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t k = 0; k < n; k++) {
        orc_ksum s = {0, 0};
        for (size_t q = k + 1; q < n; q++) {
            double dx = x[k] - x[q], dy = y[k] - y[q], dz = z[k] - z[q];
            double s2 = dx * dx + dy * dy + dz * dz + e2;
            if (s2 == 0.0) continue;
            orc_kadd(&s, m[k] * m[q] / sqrt(s2));
        }
        row[k] = orc_kval(&s);
    }
    orc_ksum ku = {0, 0};
    for (size_t k = 0; k < n; k++) orc_kadd(&ku, row[k]);
    *U = -G * orc_kval(&ku);
    free(row);
}

/* max_i |a_i - a_i^ref| / |a_i^ref| over particles with nonzero reference
 * acceleration; also writes the mean relative error. */
ORC_API double orc_max_rel_err(size_t n, const double *ax, const double *ay, const double *az,
                               const double *rx, const double *ry, const double *rz,
                               double *mean_out)
{
    double mx = 0.0, mean = 0.0;
    size_t cnt = 0;
    for (size_t i = 0; i < n; i++) {
        double ref = sqrt(rx[i] * rx[i] + ry[i] * ry[i] + rz[i] * rz[i]);
        if (ref == 0.0) continue;
        double dx = ax[i] - rx[i], dy = ay[i] - ry[i], dz = az[i] - rz[i];
        double e = sqrt(dx * dx + dy * dy + dz * dz) / ref;
        if (e > mx) mx = e;
        mean += e;
        cnt++;
    }
    if (mean_out) *mean_out = cnt ? mean / (double)cnt : 0.0;
    return mx;
}
