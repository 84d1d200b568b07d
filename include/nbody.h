/*
 * nbody.h -- C ABI of the MI355X (gfx950) direct N-body solver.
 *
 * The drop-in boundary for the reference's force-plus-leapfrog path.  Every
 * entry point takes plain pointers and sizes, returns an nb_status, never
 * throws, and keeps its error text for nb_last_error().  Host arrays are
 * owned by the caller; device memory is owned by the context.  Arrays use the
 * reference's struct-of-arrays layout (CPU/nbody_system.h:41-74): a position,
 * velocity or force argument is `double* const v[3]` (three length-N planes,
 * like NBody_system::pos[DIM]), masses are one length-N array.  G = 1 as in
 * the reference (CPU/nbody_system.h:12).
 *
 * Softening: every pair term uses (|d|^2 + eps^2)^(3/2) in place of |d|^3;
 * eps = 0 is the reference's unsoftened force exactly.  Pairs at exactly zero
 * separation contribute nothing (the reference's self-pair exclusion,
 * CPU/force_calculation.cpp:530, and the FPGA eq_cond / w==0 mask,
 * FPGA/device/ring_lrbd_lf_no_sync_local.cl:157).
 *
 * Threading: a context is used by one host thread at a time.  Work is ordered
 * on the context's own HIP stream; entry points that return host data
 * synchronise that stream before returning.  nb_step with NB_FORM_LEARNED
 * also runs the near field on a second stream of the context, forked from
 * and joined back to the first by events inside every substep, so the
 * first stream orders everything a caller sees.
 */
#ifndef NBODY_H
#define NBODY_H

#include <stddef.h>
#include <stdint.h>

#if defined(__GNUC__)
#define NB_API __attribute__((visibility("default")))
#else
#define NB_API
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nb_ctx nb_ctx;

typedef enum nb_status {
    NB_OK = 0,
    NB_ERR_ARG = 1,       /* bad argument: null pointer, size, unknown enum */
    NB_ERR_DEVICE = 2,    /* no usable gfx950 device / HIP runtime error */
    NB_ERR_ALLOC = 3,     /* device or host allocation failed */
    NB_ERR_STATE = 4,     /* call order: e.g. stepping before nb_set_state */
    NB_ERR_LAUNCH = 5     /* a kernel launch failed */
} nb_status;

/* Force-evaluation form (the reference's -p dp|sp switch, CPU/main.cpp:247-261,
 * becomes a precision form). */
typedef enum nb_form {
    NB_FORM_FP64 = 0,     /* FP64 differences, rsq + cubic correction, FP64 sums */
    NB_FORM_MIXED = 1,    /* FP64 differences, FP32 rsqrt (+1 Newton step) and
                             pair arithmetic, FP32 per-tile sums flushed to FP64 */
    NB_FORM_LEARNED = 2,  /* nb_step only: exact mixed sum every `period`-th
                             substep, learned extrapolation in between */
    NB_FORM_MIXED_RAW = 3, /* as NB_FORM_MIXED without the Newton step (v_rsq_f32
                             is already accurate to about 1 ulp) */
    NB_FORM_FP32 = 4      /* every step in FP32: positions rounded to FP32, FP32
                             differences, rsqrt + Newton step, FP32 sums -- the
                             reference's -p sp kernels (compute_forces_avx2_blocked_sp,
                             CPU/force_calculation.cpp:692-891); always on the
                             tile kernel; not for the ring round */
} nb_form;

/* Initial conditions (system set-up).  NB_IC_RING is the reference's ring of
 * N-1 masses around a central mass (CPU/nbody_system.cpp:235-300). */
typedef enum nb_ic {
    NB_IC_RING = 0,       /* p0 = R, p1 = omega, p2 = ratio (centre mass / ring mass) */
    NB_IC_SPHERE = 1,     /* uniform unit sphere, v = 0, m = 1/N (cold collapse) */
    NB_IC_PLUMMER = 2     /* p0 = scale a, p1 = truncation radius; m = 1/N, G = M = 1 */
} nb_ic;

/* Text of the last error on this thread ("" if none). */
NB_API const char* nb_last_error(void);

/* Library version string. */
NB_API const char* nb_version(void);

/* Create a context on HIP device `device`.  Replaces the reference's system
 * construction and per-thread scratch allocation (CPU/omp_solver.cpp:259-300). */
NB_API nb_status nb_create(nb_ctx** out, int device);
NB_API nb_status nb_destroy(nb_ctx* ctx);

/* Initial conditions on the host, deterministic for (kind, N, seed, params).
 * NB_IC_RING also writes the analytic force when `force` is non-null
 * (NBody_system::init_stable_orbiting_particles, CPU/nbody_system.cpp:235). */
NB_API nb_status nb_init_system(int kind, size_t n, uint64_t seed, const double* params, double* mass,
                         double* const pos[3], double* const vel[3], double* const force[3]);

/* Yoshida composition coefficients for DKD leapfrog of order 2/4/6/8
 * (CPU/omp_solver.cpp:677-726).  c has lf_steps entries, d lf_steps-1; both
 * buffers need room for 16.  Unknown orders fall back to 2, as the reference
 * does (:703-707). */
NB_API nb_status nb_yoshida(int order, double* c, double* d, size_t* lf_steps);

/* Upload the system (NBody_system fields mass/pos/vel, CPU/nbody_system.h:41-74).
 * vel may be null (zero velocities). */
NB_API nb_status nb_set_state(nb_ctx* ctx, size_t n, const double* mass, const double* const pos[3],
                       const double* const vel[3]);

/* Download positions and/or velocities (either may be null). */
NB_API nb_status nb_get_state(nb_ctx* ctx, double* const pos[3], double* const vel[3]);

/* Accelerations a_i = G sum_j m_j d_ij / (|d_ij|^2 + eps^2)^(3/2) of the
 * current state, written to acc[3][N].  Replaces one evaluation of
 * compute_forces_avx2_blocked / compute_forces_avx512_rsqrt_4i_blocked
 * (CPU/force_calculation.cpp:477, :1137) with all ordered pairs evaluated. */
NB_API nb_status nb_accel(nb_ctx* ctx, int form, double eps, double* const acc[3]);

/* Forces F_i = m_i a_i, the quantity the reference's kernels produce
 * (CPU/force_calculation.cpp:78-82 sums -G m_k m_q d/|d|^3).  Overwrites
 * force[3][N]; the reference accumulates (+=) into zeroed arrays. */
NB_API nb_status nb_compute_forces(nb_ctx* ctx, int form, double eps, double* const force[3]);

/* As nb_accel, but leaves the result in the context's device acceleration
 * buffer (no host copy): for callers that keep the state resident in HBM. */
NB_API nb_status nb_accel_device(nb_ctx* ctx, int form, double eps);

/* One ring round on one GPU: accelerations of ni i-particles from nj
 * j-particles, both interleaved (x, y, z, m) as the FPGA particle_t
 * (FPGA/host/solver.cpp:721).  Replaces the local-block x received-block
 * evaluation of hybrid_ring_solver (CPU/hybrid_solver.cpp:84-95), in full
 * rather than reduced form.  acc is ni x 3, row-major.  i may be a subset of
 * j: zero-separation pairs contribute nothing.  Host buffers; the j set is
 * streamed to the device in blocks of at most j_block particles (0 = all). */
NB_API nb_status nb_accel_block(nb_ctx* ctx, int form, double eps, size_t ni, const double* pi,
                         size_t nj, const double* pj, size_t j_block, double* acc);

/* As nb_accel_block with DEVICE pointers: d_pi (ni x 4), d_pj (nj x 4) and
 * d_acc4 (ni x 4 doubles: ax, ay, az, 0) in HBM, work queued on the context's
 * stream without synchronising.  For inputs already resident on the GPU.
 * Whether every j mass is equal (then the pair terms leave it out) is found
 * by two small kernels on the same stream, not on the host. */
NB_API nb_status nb_accel_block_device(nb_ctx* ctx, int form, double eps, size_t ni, const double* d_pi,
                                size_t nj, const double* d_pj, size_t j_block, double* d_acc4);

/* As nb_accel_block_device, allocating nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and summation order, so the results equal
 * nb_accel_block_device's bit for bit, but the j-split partial sums (and
 * the check whether every j mass is equal) go to a static device array of
 * 256 MiB + 256 bytes that the library holds on every device it runs on
 * until the process ends: calls whose partial sums need more (more than
 * the 131072 x 1048576 ring share in j-blocks of 131072) fail with NB_ERR_ARG
 * before anything is queued.  Every call on one device shares that array, so
 * calls that overlap on one device (two contexts, or two host threads) race;
 * calls queued on one context's stream are ordered.  Not timed by
 * nb_timing_*; nj = 0 writes zeros with a fill kernel. */
NB_API nb_status nb_accel_block_device_noalloc(nb_ctx* ctx, int form, double eps, size_t ni, const double* d_pi,
                                               size_t nj, const double* d_pj, size_t j_block, double* d_acc4);

/* One ring round in the reference's reduced form: replaces one call of
 * compute_forces_avx2_blocked / compute_forces_avx512_rsqrt_4i_blocked as
 * hybrid_ring_solver makes it per received block (CPU/force_calculation.h:7-13,
 * CPU/force_calculation.cpp:477-690, called at CPU/hybrid_solver.cpp:84-95),
 * with the same arguments minus the per-thread scratch the context replaces:
 * local_pos are the rank's local_n positions (nb_sys.pos), local_force its
 * forces (nb_sys.force), recv_pos / recv_force the received block's,
 * mass the global masses (comm_sz * local_n, nb_sys.mass).  With partner =
 * (rank + recv_round) % comm_sz, every pair of local particle k (mass
 * mass[rank * local_n + k]) and received particle q (mass
 * mass[partner * local_n + q]) with q >= k if rank < partner, else q > k
 * (CPU/force_calculation.cpp:530-531), is evaluated once, and its force
 * f = -G m_k m_q (x_k - x_q) / (|x_k - x_q|^2 + eps^2)^(3/2) is ADDED to
 * local_force[.][k] and SUBTRACTED from recv_force[.][q]: both arrays
 * accumulate, as the reference's do.  eps = 0 is the reference exactly
 * (coincident particles contribute nothing instead of NaN).  Host planes;
 * synchronises. */
NB_API nb_status nb_forces_ring_round(nb_ctx* ctx, int form, double eps, size_t local_n, const double* mass,
                                      const double* const local_pos[3], double* const local_force[3],
                                      const double* const recv_pos[3], double* const recv_force[3], size_t rank,
                                      size_t comm_sz, size_t recv_round);

/* As nb_forces_ring_round with DEVICE records: d_local and d_recv hold
 * local_n (x, y, z, m) records each (the masses the caller indexes as the
 * reference does), d_local_force4 and d_recv_force4 local_n (fx, fy, fz, 0)
 * records that accumulate.  Queued on the context's stream, no
 * synchronisation; the partial sums live in the context's scratch. */
NB_API nb_status nb_forces_ring_round_device(nb_ctx* ctx, int form, double eps, size_t local_n,
                                             const double* d_local, double* d_local_force4, const double* d_recv,
                                             double* d_recv_force4, size_t rank, size_t comm_sz, size_t recv_round);

/* Advance the resident state by time_steps DKD leapfrog steps of size dt with
 * coefficients c[lf_steps], d[lf_steps-1] (reduced_solver_omp_lf,
 * CPU/omp_solver.cpp:257-421): per step, for each substep s < lf_steps-1,
 * drift c[s], force, kick d[s]; then drift c[lf_steps-1].  Consecutive drifts
 * are fused into the kick kernel (as FPGA/device/ring_lrbd_lf_no_sync_local.cl:853).
 * form NB_FORM_LEARNED requires a trained extrapolator (nb_learned_*). */
NB_API nb_status nb_step(nb_ctx* ctx, size_t time_steps, double dt, const double* c, const double* d,
                  size_t lf_steps, int form, double eps);

/* Whole run in the FPGA host's form: replaces lrbd_solver_no_sync (and its
 * all-FP32 twin lrbd_solver_sp with form NB_FORM_FP32),
 * FPGA/host/solver.cpp:1112, as FPGA/host/main.cpp:495-531 calls it.
 * pos4, vel4, force4: n records of 4 doubles (x, y, z, unused: OpenCL's
 * cl_double3 with its 32-byte stride); the 4th word is neither read nor
 * written.  The pair force on particle i is K coeff_i coeff_j d/|d|^3
 * (softened with eps), the kick divides it by mass_i, so coeff are the
 * pair weights (NULL: coeff = mass, which is what the reference's host
 * passes) and K the scalar factor (G); K coeff_i / mass_i must be one value
 * for every particle (masses positive and finite), else NB_ERR_ARG before
 * anything changes.  Runs time_steps DKD steps of dt with c[lf_steps],
 * d[lf_steps-1] on the device (nb_step, the FPGA integrator's merged drifts,
 * FPGA/device/ring_lrbd_lf_no_sync_local.cl:820-903), then writes pos4 and
 * vel4 back and, if force4 is not NULL, force4 = the last force evaluation's
 * K coeff_i a_i (the FPGA's local minus remote force,
 * FPGA/host/solver.cpp:1563-1571).  *seconds (if not NULL) = the device time
 * of the steps, from events on the context's stream (the reference returns
 * its integrator kernel's time); host copies are outside it.  The run
 * replaces the context's resident state (as nb_set_state with the pair
 * weights as masses).  time_steps = 0 or n = 0 changes nothing.  The
 * reference's MPI scatter over FPGAs, topology and power sampling have no
 * counterpart on one GPU. */
NB_API nb_status nb_run_records(nb_ctx* ctx, int form, double eps, double K, const double* mass,
                                const double* coeff, double* pos4, double* vel4, double* force4, size_t n,
                                size_t time_steps, double dt, const double* c, const double* d, size_t lf_steps,
                                double* seconds);

/* Advance the resident state by time_steps explicit Euler steps of size dt
 * (reduced_solver_omp, CPU/omp_solver.cpp:16-137, the reference CLI's default
 * integrator): per step, a = force/m of the current state, x += v dt with the
 * velocity before the kick, then v += a dt.  form: NB_FORM_FP64 / NB_FORM_MIXED /
 * NB_FORM_MIXED_RAW. */
NB_API nb_status nb_step_euler(nb_ctx* ctx, size_t time_steps, double dt, int form, double eps);

/* Explicit Runge-Kutta tableau of the reference's -i rk driver
 * (CPU/omp_solver.cpp:772-797): order 4 (classic RK4) is the only one; any
 * other order falls back to it, as the reference does.  a receives
 * (rk_steps-1)^2 coefficients row-major (room for 49), b rk_steps (room for 8). */
NB_API nb_status nb_runge_kutta(int order, double* a, double* b, size_t* rk_steps);

/* Advance the resident state by time_steps explicit Runge-Kutta steps
 * (reduced_solver_omp_rk, CPU/omp_solver.cpp:458-606) for the first-order
 * system x' = v, v' = a(x): stage k has velocity kv_k and acceleration ka_k
 * (kv_0 = v, ka_k from the stage positions); stage k+1 starts at
 * x + dt sum_{j<=k} a[k*(rk_steps-1)+j] kv_j, v + dt sum_{j<=k} a[...] ka_j;
 * the step ends at x + dt sum_j b_j kv_j, v + dt sum_j b_j ka_j.  Only the
 * lower triangle of a is read.  1 <= rk_steps <= 8; the stage buffers take
 * 2 * rk_steps + 1 records of 32 bytes per particle in HBM. */
NB_API nb_status nb_step_rk(nb_ctx* ctx, size_t time_steps, double dt, const double* a, const double* b,
                            size_t rk_steps, int form, double eps);

/* Kinetic and potential energy of the resident state:
 * T = 1/2 sum m v^2 (CPU/force_calculation.cpp:1740), U = -G sum_{i<j} m_i m_j /
 * sqrt(r^2+eps^2) (:1880, softened as FPGA/host/accuracy_evaluation.cpp:35;
 * a coincident pair adds m_i m_j / eps, and is skipped when eps = 0).
 * FP64 with compensated, fixed-order sums (bitwise reproducible). */
NB_API nb_status nb_energy(nb_ctx* ctx, double eps, double* T, double* U);

/* Timing of force kernels on the context's stream, measured with HIP events
 * around every force launch since the last reset: total milliseconds and the
 * number of force evaluations.  Enabling adds two event records per launch. */
NB_API nb_status nb_timing_enable(nb_ctx* ctx, int enable);
NB_API nb_status nb_timing_read(nb_ctx* ctx, double* force_ms, uint64_t* force_evals, double* step_ms);
NB_API nb_status nb_timing_reset(nb_ctx* ctx);

/* Force-kernel selection for the resident system (nb_accel, nb_step):
 * NB_KERNEL_AUTO picks by size; NB_KERNEL_TILE evaluates every ordered pair
 * (i-particles in registers, j-tiles through LDS); NB_KERNEL_SYMMETRIC
 * evaluates each unordered pair once and applies it to both particles (the
 * reference's reduced scheme, CPU/force_calculation.cpp:530-531). */
typedef enum nb_kernel {
    NB_KERNEL_AUTO = 0,
    NB_KERNEL_TILE = 1,
    NB_KERNEL_SYMMETRIC = 2
} nb_kernel;

NB_API nb_status nb_set_kernel(nb_ctx* ctx, int kernel);

/* What the last force evaluation of the resident system ran (nb_accel,
 * nb_compute_forces, nb_accel_device, nb_step, nb_step_euler, nb_step_rk and
 * the learned form's exact substeps): for callers that label timings or
 * profiles with the kernel the library chose.  No reference counterpart. */
typedef struct nb_kernel_info {
    int kernel;              /* NB_KERNEL_TILE or NB_KERNEL_SYMMETRIC; NB_KERNEL_AUTO before any evaluation */
    int form;                /* NB_FORM_FP64, NB_FORM_MIXED or NB_FORM_MIXED_RAW */
    int per_lane;            /* particles per lane: symmetric K, tile i-particles */
    int block;               /* symmetric: particles per block; tile: j-splits */
    uint64_t items;          /* symmetric: work items; tile: workgroups */
    uint64_t scratch_bytes;  /* partial-sum bytes the evaluation wrote */
    char name[64];           /* kernel and its parameters, e.g. "k_accel_sym<FP64, K=8>" */
} nb_kernel_info;

NB_API nb_status nb_kernel_info_read(nb_ctx* ctx, nb_kernel_info* out);

/* Snapshots in the reference's binary layout (NBody_system::save_particles,
 * CPU/nbody_system.cpp:133-152): frames of doubles [t, m[N], x[N], y[N],
 * z[N], vx[N], vy[N], vz[N], fx[N], fy[N], fz[N]] one after another, host
 * byte order.  Host-only (no context): write the arrays of nb_get_state /
 * nb_compute_forces (force may be NULL: zeros), count the frames of a file
 * for a given N, read frame `frame` (any output may be NULL: skipped). */
NB_API nb_status nb_snapshot_write(const char* path, int append, double t, size_t n, const double* mass,
                                   const double* const pos[3], const double* const vel[3],
                                   const double* const force[3]);
NB_API nb_status nb_snapshot_frames(const char* path, size_t n, size_t* frames);
NB_API nb_status nb_snapshot_read(const char* path, size_t n, size_t frame, double* t, double* mass,
                                  double* const pos[3], double* const vel[3], double* const force[3]);

/* Wait for all work queued on the context's stream. */
NB_API nb_status nb_synchronize(nb_ctx* ctx);

/* ---------------------------------------------------------------------------
 * Accuracy diagnostics of the reference's omp-accuracy / omp-ratio
 * experiments (CPU/omp_solver.cpp:899-1243).  Host-only (no context), on
 * host planes such as those of nb_get_state / nb_compute_forces.
 * ------------------------------------------------------------------------- */

/* Force-error metrics of calculate_force_error (CPU/nbody_system.cpp:458-567,
 * double-precision branch) of calc against ref:
 *   out[0] mean |F - F_ref|          out[1] max |F - F_ref|
 *   out[2] sum of xy-plane angles between F and F_ref over particles whose
 *          reference x and y components are both non-zero, / (2 pi N)
 *   out[3] mean ||F| - |F_ref|| / L0  out[4] max such angle / (2 pi)
 *   out[5] max ||F| - |F_ref|| / L0
 * L0 is the length of particle 0's reference force as the reference reads
 * it, from its single-precision copy (components rounded to float, the
 * squares summed and rooted in float); L0 = 0 (particle 0 is the ring's
 * centre) gives infinities, as in the reference. */
NB_API nb_status nb_force_error(size_t n, const double* const calc[3], const double* const ref[3], double out[6]);

/* calculate_deviation (CPU/nbody_system.cpp:437-456): out[0] mean |calc - ref|,
 * out[1] max ||calc| - |ref|| / |ref| over particles with |ref| > 0. */
NB_API nb_status nb_deviation(size_t n, const double* const calc[3], const double* const ref[3], double out[2]);

/* Moment of inertia sum m |x|^2 in index order (compute_inertia,
 * CPU/force_calculation.cpp:1911-1919). */
NB_API nb_status nb_inertia(size_t n, const double* mass, const double* const pos[3], double* I);

/* Rigid-rotation state of the NB_IC_RING system after step/steps_per_rev of
 * a revolution (predict_stable_orbiting_particles, CPU/nbody_system.cpp:302-324):
 * ring particle i at angle (i/(N-1) + step/steps_per_rev) 2 pi, the centre
 * at rest in the origin.  Writes pos and vel (either may be NULL). */
NB_API nb_status nb_ring_predict(size_t n, double R, double omega, size_t step, size_t steps_per_rev,
                                 double* const pos[3], double* const vel[3]);

/* ---------------------------------------------------------------------------
 * Learned acceleration extrapolator (NB_FORM_LEARNED).  No counterpart in the
 * reference.  In nb_step with NB_FORM_LEARNED every `period`-th substep runs
 * the exact direct sum (the configured exact form); the others extrapolate
 * each particle's acceleration from its last four EXACT samples: the cubic
 * through them (Lagrange weights at their substeps) plus, on evenly spaced
 * histories, a 15 -> 128 -> 128 -> 128 -> 3 MLP correction (BF16 weights on
 * MFMA, FP32 accumulate) in units of the third difference.  Each exact substep
 * checks the extrapolation made for it; while the RMS relative error exceeds
 * `threshold` every substep is exact, and the network is used only in cycles
 * after it passed a check itself.  Particles with a rough history (|third
 * difference| > rho |a|) or that missed the last check by more than kappa
 * times the RMS acceleration get their exact rows of the direct sum.  With a
 * near-field radius, pairs closer than it are summed exactly on every substep
 * (sorted cell grid) and only the remainder is extrapolated.
 * ------------------------------------------------------------------------- */
typedef struct nb_learned_stats {
    uint64_t exact_evals;      /* direct sums run by learned-form steps */
    uint64_t predicted_evals;  /* substeps served by the extrapolator */
    uint64_t fallbacks;        /* exact checks that exceeded the threshold */
    uint64_t subset_rows;      /* rough or held particles summed exactly on extrapolated substeps */
    double last_check_err;     /* RMS relative acceleration error at the last check */
    double max_check_err;      /* largest such error since configure */
    uint64_t net_evals;        /* extrapolated substeps that applied the network */
    double last_cubic_err;     /* the cubic alone at the last check */
    double last_check_max;     /* largest per-particle relative error at the last check */
    uint64_t checks;           /* exact substeps that checked an extrapolation made for them */
    uint64_t net_checks;       /* of those, checks that also evaluated the network's extrapolation */
    double net_cubic_err_sum;  /* sum over net_checks of the cubic's RMS relative error */
    double net_err_sum;        /* sum over net_checks of the network's RMS relative error */
    uint64_t rough_exact;      /* substeps due for extrapolation after one whose exact rows exceeded the
                                  rough share of N: the direct sum ran instead (and on the following
                                  periods' substeps); counted in exact_evals */
} nb_learned_stats;

/* period >= 2 (exact every period-th substep), threshold > 0 (0 = always
 * exact), exact_form NB_FORM_FP64 / NB_FORM_MIXED / NB_FORM_MIXED_RAW. */
NB_API nb_status nb_learned_configure(nb_ctx* ctx, int period, double threshold, int exact_form);

/* rho >= 0: roughness ratio |D3| / |a_n| above which a particle is summed
 * exactly instead of extrapolated (default 1e-2; 0 = every particle exact). */
NB_API nb_status nb_learned_set_flag_ratio(nb_ctx* ctx, double rho);

/* kappa > 0: a particle whose extrapolation missed the exact acceleration at
 * a check by more than kappa times the RMS acceleration is summed exactly on
 * the extrapolated substeps until the next check (default 1e-3; INFINITY
 * never holds on this ground). */
NB_API nb_status nb_learned_set_hold_ratio(nb_ctx* ctx, double kappa);

/* share >= 0: when an extrapolated substep had more than share * N
 * particles to sum exactly (rough or held), such rows cost about what the
 * direct sum does: the next substep runs the direct sum instead, and so do
 * the substeps of the next 8 periods, twice as many each further time in a
 * row up to 64 periods (default 0.25; 1 or more: never). */
NB_API nb_status nb_learned_set_rough_share(nb_ctx* ctx, double share);

/* rc >= 0: near-field radius.  With rc > 0 pairs closer than rc (tapered
 * smoothly from rc/2) are summed exactly on every substep from a sorted cell
 * grid and only the remainder a - a_near is extrapolated.  0 (default): the
 * whole acceleration is extrapolated. */
NB_API nb_status nb_learned_set_near_radius(nb_ctx* ctx, double rc);

/* Near-field accelerations of the resident state, as used by the split:
 * acc_i = sum_{j : 0 < r_ij < rc} m_j S(r_ij) d_ij / (r_ij^2 + eps^2)^(3/2),
 * S = 1 for r <= rc/2, 0 for r >= rc, C2 quintic in r^2/rc^2 between.
 * form NB_FORM_FP64: FP64 pair arithmetic; a mixed form: FP32 pair arithmetic
 * on FP64 differences (what the learned form uses with that exact form). */
NB_API nb_status nb_near_accel(nb_ctx* ctx, int form, double rc, double eps, double* const acc[3]);

/* The same sums on the cell grid of this context's last near-field build
 * (nb_near_accel, or the learned form's exact substeps) when it was built
 * for the same particle count and radius, else on a new grid: the current
 * positions are gathered in the grid's order and the cells searched reach
 * farther by the largest displacement since the build, measured on the
 * device, so the same pairs are summed (in another order) however far the
 * particles moved.  What the learned form's extrapolated substeps use. */
NB_API nb_status nb_near_accel_kept(nb_ctx* ctx, int form, double rc, double eps, double* const acc[3]);

/* 0: cubic extrapolation corrected by the network (default once trained);
 * 1: cubic extrapolation alone (the method without the network). */
NB_API nb_status nb_learned_set_mode(nb_ctx* ctx, int mode);

/* Run `time_steps` exact steps (form) from the resident state, advancing it,
 * and record training pairs in device memory: for every substep from the
 * 4*period-th on, the features of the exact samples one to four periods back
 * and, as target, the residual of their cubic at the substep (remainder after
 * the near field when a radius is set) in units of |D3|, as components in the
 * particle's frame.  The pairs match the configured period and radius. */
NB_API nb_status nb_learned_collect(nb_ctx* ctx, size_t time_steps, double dt, const double* c,
                             const double* d, size_t lf_steps, int form, double eps);

/* Forget the network (re-initialise it from `seed`: hidden layers uniform in
 * +-1/sqrt(fan_in), output layer zero), its optimiser state and the collected
 * training pairs. */
NB_API nb_status nb_learned_reset(nb_ctx* ctx, uint64_t seed);

/* Train the network on the collected pairs with the library's own MFMA
 * forward/backward kernels and Adam: `iters` minibatches of `batch`. */
NB_API nb_status nb_learned_train(nb_ctx* ctx, int iters, int batch, double lr, uint64_t seed,
                           double* final_loss);

NB_API nb_status nb_learned_stats_read(nb_ctx* ctx, nb_learned_stats* out);

/* Network parameters as FP32 master copies, packed W1 b1 W2 b2 W3 b3 W4 b4
 * (Wk row-major [out][in]).  count receives the number of floats. */
NB_API nb_status nb_mlp_param_count(size_t* count);
NB_API nb_status nb_mlp_get_params(nb_ctx* ctx, float* params);
NB_API nb_status nb_mlp_set_params(nb_ctx* ctx, const float* params);

/* Test hooks for the network kernels on host buffers: forward y = f(x) for
 * `batch` rows of 15 inputs, and the gradient of the loss
 * L = sum_rows sum_k (y_k - target_k)^2 / (2 batch) with respect to every
 * parameter (packed as above). */
NB_API nb_status nb_mlp_forward(nb_ctx* ctx, size_t batch, const float* x, float* y);
NB_API nb_status nb_mlp_grad(nb_ctx* ctx, size_t batch, const float* x, const float* target,
                      float* grads, double* loss);

/* Test hook for the optimiser kernel: one Adam step (beta1 0.9, beta2 0.999,
 * epsilon 1e-8, bias-corrected) of the network's FP32 parameters with the
 * host gradient `grads` (packed as above) and learning rate lr; the moments
 * and the step count persist in the context until nb_mlp_set_params,
 * nb_learned_reset or training start them again. */
NB_API nb_status nb_mlp_adam_step(nb_ctx* ctx, const float* grads, double lr);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_H */
