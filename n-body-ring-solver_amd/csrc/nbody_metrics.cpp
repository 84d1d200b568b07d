// nbody_metrics.cpp -- host-side accuracy diagnostics of the reference's
// omp-accuracy / omp-ratio experiments: force-error metrics, state deviations,
// moment of inertia and the ring's rigid-rotation prediction.  They run on
// host planes that the GPU path produced (nb_get_state, nb_compute_forces).
//
// The reference's metric loops are compiled with contraction into fused
// multiply-adds (g++ -O3 on an FMA-capable target): x*x + y*y + z*z becomes
// fma(z, z, fma(x, x, y*y)), a*b + c*d becomes fma(a, b, c*d) and a*b - c*d
// becomes fma(a, b, -(c*d)).  The same operations are written out here with
// std::fma (and this file is compiled with -ffp-contract=off so the compiler
// adds no others), so the values match the reference's bit for bit.
#include <cmath>
#include <cstddef>
#include <string>

#include "nbody.h"
#include "nbody_host.h"

namespace {

bool planes_ok(const double* const p[3]) { return p && p[0] && p[1] && p[2]; }

double sq3(double x, double y, double z) { return std::fma(z, z, std::fma(x, x, y * y)); }

double len3(double x, double y, double z) { return std::sqrt(sq3(x, y, z)); }

}  // namespace

extern "C" {

nb_status nb_force_error(size_t n, const double* const calc[3], const double* const ref[3], double out[6]) {
    if (!out) return nbh::set_error(NB_ERR_ARG, "nb_force_error: null output");
    if (n && (!planes_ok(calc) || !planes_ok(ref))) return nbh::set_error(NB_ERR_ARG, "nb_force_error: null plane");
    if (n == 0) return nbh::set_error(NB_ERR_ARG, "nb_force_error: N must be > 0");
    double err = 0.0, len_err = 0.0, max_len = 0.0, ang_err = 0.0, max_ang = 0.0, max_diff = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double fx = calc[0][i], fy = calc[1][i], fz = calc[2][i];
        const double rx = ref[0][i], ry = ref[1][i], rz = ref[2][i];
        const double diff = len3(fx - rx, fy - ry, fz - rz);
        const double ldiff = std::fabs(len3(fx, fy, fz) - len3(rx, ry, rz));
        // signed angle between the xy projections (the reference's angle
        // ignores z, CPU/nbody_system.cpp:478-481)
        const double dot = std::fma(fx, rx, fy * ry);
        const double det = std::fma(fx, ry, -(fy * rx));
        const double ang = std::fabs(std::atan2(det, dot));
        const bool has_xy = rx != 0.0 && ry != 0.0;
        if (diff > max_diff) max_diff = diff;
        if (ldiff > max_len) max_len = ldiff;
        if (has_xy && ang > max_ang) max_ang = ang;
        err += diff;
        if (has_xy) ang_err += ang;
        len_err += ldiff;
    }
    // the normaliser from particle 0's single-precision reference force
    // (CPU/nbody_system.cpp:566 reads force_sp even in double precision)
    const float gx = float(ref[0][0]), gy = float(ref[1][0]), gz = float(ref[2][0]);
    const float l2 = std::fmaf(gz, gz, std::fmaf(gx, gx, gy * gy));
    const double L0 = double(std::sqrt(l2));
    const double N = double(n);
    out[0] = err / N;
    out[1] = max_diff;
    out[2] = ang_err / (2 * M_PI * N);
    out[3] = len_err / (L0 * N);
    out[4] = max_ang / (2 * M_PI);
    out[5] = max_len / L0;
    return NB_OK;
}

nb_status nb_deviation(size_t n, const double* const calc[3], const double* const ref[3], double out[2]) {
    if (!out) return nbh::set_error(NB_ERR_ARG, "nb_deviation: null output");
    if (n && (!planes_ok(calc) || !planes_ok(ref))) return nbh::set_error(NB_ERR_ARG, "nb_deviation: null plane");
    if (n == 0) return nbh::set_error(NB_ERR_ARG, "nb_deviation: N must be > 0");
    double mean = 0.0, worst = 0.0;
    for (size_t i = 0; i < n; i++) {
        mean += len3(calc[0][i] - ref[0][i], calc[1][i] - ref[1][i], calc[2][i] - ref[2][i]);
        const double lc = len3(calc[0][i], calc[1][i], calc[2][i]);
        const double lr = len3(ref[0][i], ref[1][i], ref[2][i]);
        if (lr > 0.0) {
            const double e = std::fabs(lc - lr) / lr;
            if (e > worst) worst = e;
        }
    }
    out[0] = mean / double(n);
    out[1] = worst;
    return NB_OK;
}

nb_status nb_inertia(size_t n, const double* mass, const double* const pos[3], double* I) {
    if (!I) return nbh::set_error(NB_ERR_ARG, "nb_inertia: null output");
    if (n && (!mass || !planes_ok(pos))) return nbh::set_error(NB_ERR_ARG, "nb_inertia: null input");
    double s = 0.0;
    // the reference's loop is vectorised with an in-order sum: m |x|^2 is
    // rounded before it is added
    for (size_t i = 0; i < n; i++) s += mass[i] * sq3(pos[0][i], pos[1][i], pos[2][i]);
    *I = s;
    return NB_OK;
}

nb_status nb_ring_predict(size_t n, double R, double omega, size_t step, size_t steps_per_rev,
                          double* const pos[3], double* const vel[3]) {
    if (n < 2) return nbh::set_error(NB_ERR_ARG, "nb_ring_predict: the ring needs N >= 2");
    if (steps_per_rev == 0) return nbh::set_error(NB_ERR_ARG, "nb_ring_predict: steps_per_rev must be > 0");
    if ((pos && !planes_ok(pos)) || (vel && !planes_ok(vel)))
        return nbh::set_error(NB_ERR_ARG, "nb_ring_predict: null plane");
    const size_t m = n - 1;
    const double offset = double(step) / double(steps_per_rev);
    for (size_t i = 0; i < m; i++) {
        const double th = (double(i) / double(m) + offset) * 2 * M_PI;
        double sn, cs;
        ::sincos(th, &sn, &cs);
        if (pos) {
            pos[0][i] = R * cs;
            pos[1][i] = R * sn;
            pos[2][i] = 0.0;
        }
        if (vel) {
            vel[0][i] = -R * omega * sn;
            vel[1][i] = R * omega * cs;
            vel[2][i] = 0.0;
        }
    }
    for (int k = 0; k < 3; k++) {
        if (pos) pos[k][m] = 0.0;
        if (vel) vel[k][m] = 0.0;
    }
    return NB_OK;
}

}  // extern "C"
