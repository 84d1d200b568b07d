// ref_shim.cpp -- extern "C" entry points over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY (the oracle's second leg).  Built by oracle/Makefile
// together with the unmodified reference files CPU/force_calculation.cpp,
// CPU/omp_solver.cpp and CPU/nbody_system.cpp, compiled where they lie; the
// result goes to oracle/_ref/libnbref.so.  Used by tests/golden/make_golden.py
// to write fixtures, by the tests that check the oracle against the
// reference, and by the cpu_baseline leg of bench.py.
//
// Each wrapper sets up an NBody_system and the per-thread scratch exactly as
// the reference's leapfrog driver does (CPU/omp_solver.cpp:259-300), calls the
// reference routine, and copies the results out.
#include <omp.h>

#include <array>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "force_calculation.h"
#include "nbody_system.h"
#include "omp_solver.h"

// Defined in CPU/omp_solver.cpp:257 with a Data_type argument; the header
// prototype at CPU/omp_solver.h:11 omits it, so declare the real signature.
void reduced_solver_omp_lf(NBody_system& nb_sys, size_t time_steps, double delta_t, double* c,
                           double* d, size_t lf_steps, bool blocked, Data_type data_type);

namespace {

struct Scratch {
    size_t T = 0, N = 0;
    double*** lf = nullptr;
    double*** lrf = nullptr;
    float*** lf_sp = nullptr;
    float*** lrf_sp = nullptr;
    double* recv_pos[DIM] = {nullptr, nullptr, nullptr};
    double* recv_force[DIM] = {nullptr, nullptr, nullptr};
    float* recv_pos_sp[DIM] = {nullptr, nullptr, nullptr};
    float* recv_force_sp[DIM] = {nullptr, nullptr, nullptr};

    static size_t padded(size_t n) { return ((n * sizeof(double) + 63) / 64) * 64; }

    Scratch(size_t n) : T(omp_get_max_threads()), N(n) {
        lf = (double***)malloc(T * sizeof(double**));
        lrf = (double***)malloc(T * sizeof(double**));
        lf_sp = (float***)malloc(T * sizeof(float**));
        lrf_sp = (float***)malloc(T * sizeof(float**));
        for (size_t t = 0; t < T; t++) {
            lf[t] = (double**)malloc(DIM * sizeof(double*));
            lrf[t] = (double**)malloc(DIM * sizeof(double*));
            lf_sp[t] = (float**)malloc(DIM * sizeof(float*));
            lrf_sp[t] = (float**)malloc(DIM * sizeof(float*));
            for (int d = 0; d < DIM; d++) {
                lf[t][d] = (double*)aligned_alloc(64, padded(n));
                lrf[t][d] = (double*)aligned_alloc(64, padded(n));
                lf_sp[t][d] = (float*)aligned_alloc(64, padded(n));
                lrf_sp[t][d] = (float*)aligned_alloc(64, padded(n));
            }
        }
        for (int d = 0; d < DIM; d++) {
            recv_pos[d] = (double*)aligned_alloc(64, padded(n));
            recv_force[d] = (double*)aligned_alloc(64, padded(n));
            recv_pos_sp[d] = (float*)aligned_alloc(64, padded(n));
            recv_force_sp[d] = (float*)aligned_alloc(64, padded(n));
        }
    }
    ~Scratch() {
        for (size_t t = 0; t < T; t++) {
            for (int d = 0; d < DIM; d++) {
                free(lf[t][d]); free(lrf[t][d]); free(lf_sp[t][d]); free(lrf_sp[t][d]);
            }
            free(lf[t]); free(lrf[t]); free(lf_sp[t]); free(lrf_sp[t]);
        }
        free(lf); free(lrf); free(lf_sp); free(lrf_sp);
        for (int d = 0; d < DIM; d++) {
            free(recv_pos[d]); free(recv_force[d]); free(recv_pos_sp[d]); free(recv_force_sp[d]);
        }
    }
};

void load_system(NBody_system& s, size_t n, const double* m, const double* const* pos,
                 const double* const* vel) {
    s.alloc_particles(n, n);
    for (size_t i = 0; i < n; i++) {
        s.mass[i] = m[i];
        s.mass_sp[i] = float(m[i]);
        for (int d = 0; d < DIM; d++) {
            s.pos[d][i] = pos[d][i];
            s.pos_sp[d][i] = float(pos[d][i]);
            s.vel[d][i] = vel ? vel[d][i] : 0.0;
            s.vel_sp[d][i] = float(s.vel[d][i]);
            s.force[d][i] = 0.0;
            s.force_sp[d][i] = 0.0f;
        }
    }
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) int ref_max_threads() { return omp_get_max_threads(); }

// One reduced force evaluation with the reference kernel `variant`:
//   0 compute_forces_avx2_blocked            (CPU/force_calculation.cpp:477)
//   1 compute_forces_avx512_rsqrt_4i_blocked (CPU/force_calculation.cpp:1137)
//   2 compute_forces_avx512_rsqrt_blocked    (CPU/force_calculation.cpp:894)
//   3 compute_forces_avx2                    (CPU/force_calculation.cpp:109)
//   4 compute_forces_avx2_blocked_sp         (CPU/force_calculation.cpp:692), FP32
// then F += recv_F as the leapfrog driver does (CPU/omp_solver.cpp:375-378).
// Writes FORCE (G m_k sum ...), not acceleration.
__attribute__((visibility("default"))) int ref_forces(int variant, size_t n, const double* m,
                                                      const double* const* pos, double* const* force) {
    NBody_system s;
    load_system(s, n, m, pos, nullptr);
    Scratch sc(n);
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            sc.recv_pos[d][i] = s.pos[d][i];
            sc.recv_force[d][i] = 0.0;
            sc.recv_pos_sp[d][i] = s.pos_sp[d][i];
            sc.recv_force_sp[d][i] = 0.0f;
        }
    switch (variant) {
    case 0: compute_forces_avx2_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T); break;
    case 1: compute_forces_avx512_rsqrt_4i_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T); break;
    case 2: compute_forces_avx512_rsqrt_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T); break;
    case 3: compute_forces_avx2(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T); break;
    case 4: compute_forces_avx2_blocked_sp(s, sc.recv_pos_sp, sc.recv_force_sp, sc.lf_sp, sc.lrf_sp, n, 0, 1, 0, sc.T); break;
    default: return 1;
    }
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            if (variant == 4) force[d][i] = double(s.force_sp[d][i] + sc.recv_force_sp[d][i]);
            else force[d][i] = s.force[d][i] + sc.recv_force[d][i];
        }
    return 0;
}

// One ring round of the reference's reduced scheme with variant 0
// (compute_forces_avx2_blocked) or 1 (compute_forces_avx512_rsqrt_4i_blocked),
// as hybrid_ring_solver calls it per round (CPU/hybrid_solver.cpp:81-95) but
// in one process: the system holds comm_sz * n masses and the local block's
// n positions, the received block is recv_pos; force and recv_force come in
// with the caller's values and the kernel accumulates into them.
__attribute__((visibility("default"))) int ref_ring_round(int variant, size_t n, size_t comm_sz, size_t rank,
                                                          size_t recv_round, const double* mass,
                                                          const double* const* local_pos,
                                                          const double* const* recv_pos, double* const* force,
                                                          double* const* recv_force) {
    if (comm_sz == 0 || rank >= comm_sz || n == 0) return 1;
    NBody_system s;
    s.alloc_particles(comm_sz * n, n);
    for (size_t i = 0; i < comm_sz * n; i++) {
        s.mass[i] = mass[i];
        s.mass_sp[i] = float(mass[i]);
    }
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < comm_sz * n; i++) s.pos[d][i] = i < n ? local_pos[d][i] : 0.0;
    Scratch sc(n);
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            s.force[d][i] = force[d][i];
            sc.recv_pos[d][i] = recv_pos[d][i];
            sc.recv_force[d][i] = recv_force[d][i];
        }
    switch (variant) {
    case 0:
        compute_forces_avx2_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, rank, comm_sz, recv_round, sc.T);
        break;
    case 1:
        compute_forces_avx512_rsqrt_4i_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, rank, comm_sz,
                                               recv_round, sc.T);
        break;
    default: return 1;
    }
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            force[d][i] = s.force[d][i];
            recv_force[d][i] = sc.recv_force[d][i];
        }
    return 0;
}

// Repeated force evaluations for timing (the cpu_baseline leg): `reps` calls of
// variant's kernel on the same positions; returns the best wall time in seconds.
__attribute__((visibility("default"))) double ref_time_forces(int variant, size_t n, const double* m,
                                                              const double* const* pos, int reps) {
    NBody_system s;
    load_system(s, n, m, pos, nullptr);
    Scratch sc(n);
    double best = 1e300;
    for (int r = 0; r < reps; r++) {
        for (int d = 0; d < DIM; d++)
            for (size_t i = 0; i < n; i++) {
                sc.recv_pos[d][i] = s.pos[d][i];
                sc.recv_force[d][i] = 0.0;
                s.force[d][i] = 0.0;
            }
        double t0 = omp_get_wtime();
        if (variant == 1)
            compute_forces_avx512_rsqrt_4i_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T);
        else
            compute_forces_avx2_blocked(s, sc.recv_pos, sc.recv_force, sc.lf, sc.lrf, n, 0, 1, 0, sc.T);
        double t1 = omp_get_wtime();
        if (t1 - t0 < best) best = t1 - t0;
    }
    return best;
}

// The reference's leapfrog driver (CPU/omp_solver.cpp:257) in double precision:
// `steps` steps of DKD with the given c[lf_steps], d[lf_steps-1]; state in place.
__attribute__((visibility("default"))) int ref_leapfrog(size_t n, const double* m, double* const* pos,
                                                        double* const* vel, size_t steps, double dt,
                                                        const double* c, const double* d, size_t lf_steps) {
    NBody_system s;
    load_system(s, n, m, (const double* const*)pos, (const double* const*)vel);
    double cc[16], dd[16];
    if (lf_steps > 16) return 1;
    memcpy(cc, c, lf_steps * sizeof(double));
    memcpy(dd, d, (lf_steps - 1) * sizeof(double));
    reduced_solver_omp_lf(s, steps, dt, cc, dd, lf_steps, true, Data_type::DOUBLE_PRECISION);
    for (int k = 0; k < DIM; k++)
        for (size_t i = 0; i < n; i++) {
            pos[k][i] = s.pos[k][i];
            vel[k][i] = s.vel[k][i];
        }
    return 0;
}

// Ring + centre initial condition with the analytic force
// (NBody_system::init_stable_orbiting_particles, CPU/nbody_system.cpp:235).
__attribute__((visibility("default"))) int ref_init_ring(size_t n, double R, double omega, double ratio,
                                                         double* m, double* const* pos, double* const* vel,
                                                         double* const* force) {
    NBody_system s;
    s.init_stable_orbiting_particles(n, R, omega, ratio);
    for (size_t i = 0; i < n; i++) {
        m[i] = s.mass[i];
        for (int d = 0; d < DIM; d++) {
            pos[d][i] = s.pos[d][i];
            vel[d][i] = s.vel[d][i];
            force[d][i] = s.force[d][i];
        }
    }
    return 0;
}

// Energies: compute_kinetic_energy, compute_potential_energy_avx2 and the
// scalar calculate_potential_energy (CPU/force_calculation.cpp:1740-1904).
__attribute__((visibility("default"))) int ref_energy(size_t n, const double* m, const double* const* pos,
                                                      const double* const* vel, double* T, double* U_avx2,
                                                      double* U_scalar) {
    NBody_system s;
    load_system(s, n, m, pos, vel);
    *T = compute_kinetic_energy(s);
    *U_avx2 = compute_potential_energy_avx2(s);
    *U_scalar = calculate_potential_energy(s);
    return 0;
}

// Snapshot frames written by NBody_system::save_particles
// (CPU/nbody_system.cpp:133-152) into NBody_<N>.dat in the working directory:
// frame k is written at time step k with the given dt, after scaling the
// velocities by (1 + k) so the frames differ.
__attribute__((visibility("default"))) int ref_save_particles(size_t n, const double* m, const double* const* pos,
                                                              const double* const* vel, const double* const* force,
                                                              double dt, size_t frames) {
    NBody_system s;
    load_system(s, n, m, pos, vel);
    for (size_t k = 0; k < frames; k++) {
        for (int d = 0; d < DIM; d++)
            for (size_t i = 0; i < n; i++) {
                s.vel[d][i] = vel[d][i] * double(1 + k);
                s.force[d][i] = force[d][i];
            }
        s.save_particles(k, dt);
    }
    s.output_file.close();
    return 0;
}

// The reference's explicit Euler driver reduced_solver_omp (CPU/omp_solver.cpp:16)
// in double precision with compute_forces_avx2_blocked; state in place.
__attribute__((visibility("default"))) int ref_euler(size_t n, const double* m, double* const* pos,
                                                     double* const* vel, size_t steps, double dt) {
    NBody_system s;
    load_system(s, n, m, (const double* const*)pos, (const double* const*)vel);
    reduced_solver_omp(s, steps, dt, true, Vectorization_type::AVX2, Data_type::DOUBLE_PRECISION);
    for (int k = 0; k < DIM; k++)
        for (size_t i = 0; i < n; i++) {
            pos[k][i] = s.pos[k][i];
            vel[k][i] = s.vel[k][i];
        }
    return 0;
}

// The reference's Runge-Kutta driver reduced_solver_omp_rk (CPU/omp_solver.cpp:458)
// with its a[(rk_steps-1)^2] / b[rk_steps]; state in place.  That routine
// pins OpenMP to one thread (:460); the previous count is restored after it.
__attribute__((visibility("default"))) int ref_rk(size_t n, const double* m, double* const* pos,
                                                  double* const* vel, size_t steps, double dt,
                                                  const double* a, const double* b, size_t rk_steps) {
    if (rk_steps < 1 || rk_steps > 8) return 1;
    NBody_system s;
    load_system(s, n, m, (const double* const*)pos, (const double* const*)vel);
    double aa[64] = {0}, bb[8] = {0};
    memcpy(aa, a, (rk_steps - 1) * (rk_steps - 1) * sizeof(double));
    memcpy(bb, b, rk_steps * sizeof(double));
    const int threads = omp_get_max_threads();
    reduced_solver_omp_rk(s, steps, dt, aa, bb, rk_steps, true);
    omp_set_num_threads(threads);
    for (int k = 0; k < DIM; k++)
        for (size_t i = 0; i < n; i++) {
            pos[k][i] = s.pos[k][i];
            vel[k][i] = s.vel[k][i];
        }
    return 0;
}

// calculate_force_error (CPU/nbody_system.cpp:458) in double precision:
// sys.force = calc, ref.force = ref and ref's single-precision copy rounded
// from it, as the ring initialisation leaves it (:262-275).
__attribute__((visibility("default"))) int ref_force_error(size_t n, const double* const* calc,
                                                           const double* const* ref, double* out) {
    NBody_system a, b;
    a.alloc_particles(n, n);
    b.alloc_particles(n, n);
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            a.force[d][i] = calc[d][i];
            a.force_sp[d][i] = float(calc[d][i]);
            b.force[d][i] = ref[d][i];
            b.force_sp[d][i] = float(ref[d][i]);
        }
    const std::array<double, 6> e = calculate_force_error(a, b, Data_type::DOUBLE_PRECISION, false);
    for (int k = 0; k < 6; k++) out[k] = e[k];
    return 0;
}

// calculate_deviation (CPU/nbody_system.cpp:437).
__attribute__((visibility("default"))) int ref_deviation(size_t n, double* const* calc, double* const* ref,
                                                         double* out) {
    double* c[DIM] = {calc[0], calc[1], calc[2]};
    double* r[DIM] = {ref[0], ref[1], ref[2]};
    const std::array<double, 2> e = calculate_deviation(c, r, n);
    out[0] = e[0];
    out[1] = e[1];
    return 0;
}

// compute_inertia (CPU/force_calculation.cpp:1911).
__attribute__((visibility("default"))) double ref_inertia(size_t n, const double* m, const double* const* pos) {
    NBody_system s;
    load_system(s, n, m, pos, nullptr);
    return compute_inertia(s);
}

// NBody_system::predict_stable_orbiting_particles (CPU/nbody_system.cpp:302)
// applied to a ring of n particles set up by init_stable_orbiting_particles.
__attribute__((visibility("default"))) int ref_ring_predict(size_t n, double R, double omega, size_t step,
                                                            size_t steps_per_rev, double* const* pos,
                                                            double* const* vel) {
    NBody_system s;
    s.init_stable_orbiting_particles(n, R, omega, 1.0);
    s.predict_stable_orbiting_particles(R, omega, step, steps_per_rev);
    for (int d = 0; d < DIM; d++)
        for (size_t i = 0; i < n; i++) {
            pos[d][i] = s.pos[d][i];
            vel[d][i] = s.vel[d][i];
        }
    return 0;
}

}  // extern "C"
