// nbody_cli.cpp -- command-line driver of the GPU force + integrator path.
//
// Mirrors the reference's flags (CPU/main.cpp:11-289): -t -N -start -step -n
// -s -T -r -ratio-start -ratio-stop -ratio-steps -rev-steps -max-error -i
// -order -V -S -p -o -v, with the same defaults (Euler, order 1, N = 1000,
// s = 1000, T = 1 ...), the same "unknown value -> default, with a message"
// behaviour and dt = T/s (CPU/omp_solver.cpp:640).  Run types:
//   -t omp           one run of the chosen integrator (-i euler|lf|rk,
//                    CPU/omp_solver.cpp:635-797) on the GPU
//   -t omp-accuracy  force-accumulation error on permuted rings
//                    (omp_accuracy, CPU/omp_solver.cpp:899-990)
//   -t omp-ratio     step-by-step deviation from the ring's rigid rotation
//                    until it collapses (omp_ratio_accuracy, :995-1243)
// The reference's other run types (cache and scaling benchmarks, MPI and
// hybrid runs) are not provided and end with a message.  -V is accepted for
// compatibility (every vectorisation variant maps to the gfx950 kernels),
// -S full selects the kernel that evaluates every ordered pair, and -p sp
// selects the all-FP32 form (the reference's single-precision kernels,
// CPU/force_calculation.cpp:692-891) unless -form names another.  The experiments' permutations come from
// a seeded generator (-seed) instead of the reference's std::random_device,
// so their runs repeat.  Added options:
//   -eps <e>       softening length (default 0: the reference's force)
//   -ic ring|sphere|plummer, -seed <k>, -a <scale>, -rmax <r>
//   -form fp64|mixed|mixed-raw|fp32|learned
//   -device <d>
//   -period <k> -threshold <x> -rc <r> -train-steps <n> -train-iters <n>  (learned form)
//   -snapshot <file>   initial and final state (with FP64 forces) as frames in
//                      the reference's NBody_<N>.dat layout (CPU/nbody_system.cpp:133)
// Output: the reference's stdout lines plus energies; -t omp with -o appends
// one CSV row per run under a header written when the file is new (the FPGA
// accuracy format, FPGA/host/main.cpp:582, extended); the experiments write
// the reference's own CSV rows (default file output.csv, as the reference).
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "nbody.h"

namespace {

struct Args {
    size_t N = 1000;
    size_t start_N = 1000, step_N = 1000, num_steps = 1;
    size_t steps = 1000;
    double T = 1.0;
    double ratio = 100000.0;
    double ratio_start = 1e6, ratio_stop = 1e9;
    size_t ratio_steps = 3;
    size_t rev_steps = 1000;
    double max_error = 1.0;
    bool verbose = false;
    std::string out;
    bool out_given = false;
    std::string exec = "omp";
    std::string integ = "euler";
    int order = 1;
    std::string vect = "AVX2";
    std::string solver = "reduced";
    std::string prec = "dp";
    double eps = 0.0;
    std::string ic = "ring";
    uint64_t seed = 42;
    double a = 1.0, rmax = 20.0;
    std::string form;
    int device = 0;
    std::string snapshot;
    int period = 4;
    double threshold = 3e-3;
    double rc = 0.25;   // near-field radius of the learned form (0: no split)
    size_t train_steps = 200;
    int train_iters = 16000;   // then half as many at a lower rate
};

bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

void parse(int argc, char** argv, Args& a) {
    const std::map<std::string, int> execs = {{"omp", 0},        {"omp-cache", 1},     {"omp-weak", 2},
                                              {"omp-strong", 3}, {"omp-accuracy", 4},  {"omp-ratio", 5},
                                              {"mpi-bandwidth", 6}, {"hybrid", 7},     {"hybrid-weak", 8},
                                              {"hybrid-strong", 9}};
    const std::map<std::string, int> integs = {{"euler", 0}, {"lf", 1}, {"rk", 2}};
    const std::map<std::string, int> vects = {{"AVX2", 0}, {"AVX512", 1}, {"AVX2-RSQRT", 2}, {"AVX512-RSQRT", 3},
                                              {"AVX512-RSQRT-4I", 4}};
    const std::map<std::string, int> solvers = {{"full", 0}, {"reduced", 1}};
    const std::map<std::string, int> precs = {{"dp", 0}, {"sp", 1}};
    for (int i = 1; i < argc; i++) {
        const std::string f = argv[i];
        const bool has = i + 1 < argc;
        if (f == "-N" && has) a.N = strtoull(argv[++i], nullptr, 10);
        else if (f == "-start" && has) a.start_N = strtoull(argv[++i], nullptr, 10);
        else if (f == "-step" && has) a.step_N = strtoull(argv[++i], nullptr, 10);
        else if (f == "-n" && has) a.num_steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-ratio-start" && has) a.ratio_start = atof(argv[++i]);
        else if (f == "-ratio-stop" && has) a.ratio_stop = atof(argv[++i]);
        else if (f == "-ratio-steps" && has) a.ratio_steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-rev-steps" && has) a.rev_steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-max-error" && has) a.max_error = atof(argv[++i]);
        else if (f == "-s" && has) a.steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-T" && has) a.T = atof(argv[++i]);
        else if (f == "-r" && has) a.ratio = atof(argv[++i]);
        else if (f == "-v") a.verbose = true;
        else if (f == "-o" && has) {
            a.out = argv[++i];
            a.out_given = true;
        } else if (f == "-snapshot" && has) a.snapshot = argv[++i];
        else if (f == "-order" && has) a.order = int(strtoull(argv[++i], nullptr, 10));
        else if (f == "-t" && has) {
            a.exec = argv[++i];
            if (!execs.count(a.exec)) {
                printf("Unknown Execution type: %s\n Defaulting to omp\n", a.exec.c_str());
                a.exec = "omp";
            }
        } else if (f == "-i" && has) {
            a.integ = argv[++i];
            if (!integs.count(a.integ)) {
                printf("Unknown Integration kind: %s\n Defaulting to euler\n", a.integ.c_str());
                a.integ = "euler";
            }
        } else if (f == "-V" && has) {
            a.vect = argv[++i];
            if (!vects.count(a.vect)) {
                printf("Unknown Vectorization type: %s\n Defaulting to AVX2\n", a.vect.c_str());
                a.vect = "AVX2";
            }
        } else if (f == "-S" && has) {
            a.solver = argv[++i];
            if (!solvers.count(a.solver)) {
                printf("Unknown Solver type: %s\n Defaulting to Reduced Solver\n", a.solver.c_str());
                a.solver = "reduced";
            }
        } else if (f == "-p" && has) {
            a.prec = argv[++i];
            if (!precs.count(a.prec)) {
                printf("Unknown Data type: %s\n Defaulting to double precision\n", a.prec.c_str());
                a.prec = "dp";
            }
        } else if (f == "-eps" && has) a.eps = atof(argv[++i]);
        else if (f == "-ic" && has) a.ic = argv[++i];
        else if (f == "-seed" && has) a.seed = strtoull(argv[++i], nullptr, 10);
        else if (f == "-a" && has) a.a = atof(argv[++i]);
        else if (f == "-rmax" && has) a.rmax = atof(argv[++i]);
        else if (f == "-form" && has) a.form = argv[++i];
        else if (f == "-device" && has) a.device = atoi(argv[++i]);
        else if (f == "-period" && has) a.period = atoi(argv[++i]);
        else if (f == "-threshold" && has) a.threshold = atof(argv[++i]);
        else if (f == "-rc" && has) a.rc = atof(argv[++i]);
        else if (f == "-train-steps" && has) a.train_steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-train-iters" && has) a.train_iters = atoi(argv[++i]);
    }
    // the reference's sanity resets (CPU/main.cpp:271-288)
    if (a.N >= 100000000 || a.start_N >= 100000000 || a.start_N + a.step_N * a.num_steps >= 100000000) {
        printf("Input problem size might be too large!\nMake sure to stay below 10^8 particles.\nResetting to defaults");
        a.N = 1000;
        a.start_N = 1000;
        a.step_N = 1000;
        a.num_steps = 1;
    }
    if (a.T <= 0.0) {
        printf("T has to be >= 0!\n Resetting to 1.0\n");
        a.T = 1.0;
    }
    if (a.ratio < 0.0) {
        printf("ratio has to be >= 0!\n Resetting to 1e6\n");
        a.ratio = 1e6;
    }
    if (a.eps < 0.0) {
        printf("eps has to be >= 0!\n Resetting to 0\n");
        a.eps = 0.0;
    }
    if (a.form.empty()) a.form = (a.prec == "sp") ? "fp32" : "fp64";
    if (a.out.empty()) a.out = "output.csv";   // the reference's default output file
}

int form_id(const std::string& f) {
    if (f == "fp64") return NB_FORM_FP64;
    if (f == "mixed") return NB_FORM_MIXED;
    if (f == "mixed-raw") return NB_FORM_MIXED_RAW;
    if (f == "fp32") return NB_FORM_FP32;
    if (f == "learned") return NB_FORM_LEARNED;
    printf("Unknown form: %s\n Defaulting to fp64\n", f.c_str());
    return NB_FORM_FP64;
}

#define CHECK(call)                                                                  \
    do {                                                                             \
        nb_status s_ = (call);                                                       \
        if (s_ != NB_OK) {                                                           \
            fprintf(stderr, "error (%d) in %s: %s\n", int(s_), #call, nb_last_error()); \
            return 1;                                                                \
        }                                                                            \
    } while (0)

struct Planes {
    std::vector<double> buf;
    double* p[3];
    explicit Planes(size_t n) : buf(3 * n) { bind(n); }
    Planes(const Planes& o) : buf(o.buf) { bind(o.buf.size() / 3); }
    Planes& operator=(const Planes&) = delete;
    void bind(size_t n) {
        for (int k = 0; k < 3; k++) p[k] = buf.data() + k * n;
    }
};

// A system in the reference's struct-of-arrays layout.
struct System {
    size_t n;
    std::vector<double> mass;
    Planes pos, vel, force;
    explicit System(size_t n_) : n(n_), mass(n_), pos(n_), vel(n_), force(n_) {}
    // particle i takes the values of particle perm[i] (apply_permutation,
    // CPU/nbody_system.cpp:344-383)
    void permute(const std::vector<size_t>& perm) {
        std::vector<double> t(n);
        auto apply = [&](double* v) {
            for (size_t i = 0; i < n; i++) t[i] = v[perm[i]];
            std::copy(t.begin(), t.end(), v);
        };
        apply(mass.data());
        for (int k = 0; k < 3; k++) {
            apply(pos.p[k]);
            apply(vel.p[k]);
            apply(force.p[k]);
        }
    }
};

std::vector<size_t> seeded_permutation(size_t n, uint64_t seed) {
    std::vector<size_t> perm(n);
    std::iota(perm.begin(), perm.end(), size_t(0));
    std::mt19937_64 g(seed);
    std::shuffle(perm.begin(), perm.end(), g);
    return perm;
}

// Mean |a - b| over the particles; with verbose, every particle's values and
// difference as print_deviation writes them (CPU/nbody_system.cpp:416-435).
double mean_deviation(const Planes& a, const Planes& b, size_t n, const char* kind, bool verbose) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double dx = a.p[0][i] - b.p[0][i], dy = a.p[1][i] - b.p[1][i], dz = a.p[2][i] - b.p[2][i];
        s += std::sqrt(dx * dx + dy * dy + dz * dz);
        if (verbose) {
            printf("calc_%s[%zu]=[%f,%f,%f]\n", kind, i, a.p[0][i], a.p[1][i], a.p[2][i]);
            printf("ref_%s[%zu] =[%f,%f,%f]\n", kind, i, b.p[0][i], b.p[1][i], b.p[2][i]);
            printf("%s_diff[%zu] =[%f,%f,%f]\n\n", kind, i, dx, dy, dz);
        }
    }
    return n ? s / double(n) : 0.0;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void print_coefficients(const double* c, const double* d, size_t lf) {
    double sum = 0.0;
    for (size_t i = 0; i < lf; i++) {
        printf("c[%zu] = %f\n", i, c[i]);
        sum += c[i];
    }
    printf("sum(c) = %f\n", sum);
    sum = 0.0;
    for (size_t i = 0; i + 1 < lf; i++) {
        printf("d[%zu] = %f\n", i, d[i]);
        sum += d[i];
    }
    printf("sum(d) = %f\n", sum);
}

// -t omp: one run of the chosen integrator (omp_main, CPU/omp_solver.cpp:635-797).
int run_omp(nb_ctx* ctx, Args& a) {
    const size_t N = a.N;
    const double dt = a.T / double(a.steps);
    int form = form_id(a.form);
    if (form == NB_FORM_LEARNED && a.integ != "lf") {
        printf("The learned form drives the leapfrog only; using -i lf\n");
        a.integ = "lf";
    }
    System sys(N);
    int kind = NB_IC_RING;
    double params[3] = {1.0, 2.0 * M_PI, a.ratio};
    if (a.ic == "sphere") kind = NB_IC_SPHERE;
    else if (a.ic == "plummer") {
        kind = NB_IC_PLUMMER;
        params[0] = a.a;
        params[1] = a.rmax;
    } else if (a.ic != "ring") {
        printf("Unknown initial condition: %s\n Defaulting to ring\n", a.ic.c_str());
        a.ic = "ring";
    }
    CHECK(nb_init_system(kind, N, a.seed, params, sys.mass.data(), sys.pos.p, sys.vel.p, sys.force.p));
    const Planes pos0 = sys.pos, vel0 = sys.vel;
    printf("N = %zu, s = %zu\n", N, a.steps);

    double c[16], d[16], ra[49], rb[8];
    size_t lf = 0, rk = 0;
    int order = a.order;
    double evals_per_step = 1.0;
    if (a.integ == "lf") {
        if (order != 2 && order != 4 && order != 6 && order != 8) {
            printf("Order %d not supported! Defaulting to 2nd order\n", order);
            order = 2;
        }
        CHECK(nb_yoshida(order, c, d, &lf));
        print_coefficients(c, d, lf);
        evals_per_step = double(lf - 1);
    } else if (a.integ == "rk") {
        if (order != 4) {
            printf("Order %d not supported! Defaulting to 4th order\n", order);
            order = 4;
        }
        CHECK(nb_runge_kutta(order, ra, rb, &rk));
        evals_per_step = double(rk);
    } else {
        order = 1;
    }

    CHECK(nb_set_state(ctx, N, sys.mass.data(), sys.pos.p, sys.vel.p));
    double T0, U0;
    CHECK(nb_energy(ctx, a.eps, &T0, &U0));
    auto snapshot = [&](double t, bool append) -> nb_status {
        Planes p(N), v(N), f(N);
        nb_status s = nb_get_state(ctx, p.p, v.p);
        if (s == NB_OK) s = nb_compute_forces(ctx, NB_FORM_FP64, a.eps, f.p);
        if (s == NB_OK) s = nb_snapshot_write(a.snapshot.c_str(), append ? 1 : 0, t, N, sys.mass.data(), p.p, v.p, f.p);
        return s;
    };
    if (!a.snapshot.empty()) CHECK(snapshot(0.0, false));

    if (form == NB_FORM_LEARNED) {
        CHECK(nb_learned_configure(ctx, a.period, a.threshold, NB_FORM_MIXED));
        CHECK(nb_learned_set_near_radius(ctx, a.rc));
        CHECK(nb_learned_set_flag_ratio(ctx, 0.1));
        CHECK(nb_learned_set_hold_ratio(ctx, 3e-3));
        printf("Training the extrapolator on %zu exact steps ...\n", a.train_steps);
        double loss = 0.0;
        CHECK(nb_learned_collect(ctx, a.train_steps, dt, c, d, lf, NB_FORM_MIXED, a.eps));
        // the library's measured schedule: train_iters minibatches at 2e-3, then
        // train_iters / 2 at 3e-4
        CHECK(nb_learned_train(ctx, a.train_iters, 4096, 2e-3, a.seed, &loss));
        if (a.train_iters / 2 > 0) CHECK(nb_learned_train(ctx, a.train_iters / 2, 4096, 3e-4, a.seed + 1, &loss));
        printf("final training loss %.4e\n", loss);
        CHECK(nb_set_state(ctx, N, sys.mass.data(), sys.pos.p, sys.vel.p));   // restart from the initial state
    }

    CHECK(nb_synchronize(ctx));
    const auto t0 = std::chrono::steady_clock::now();
    if (a.integ == "lf") CHECK(nb_step(ctx, a.steps, dt, c, d, lf, form, a.eps));
    else if (a.integ == "rk") CHECK(nb_step_rk(ctx, a.steps, dt, ra, rb, rk, form, a.eps));
    else CHECK(nb_step_euler(ctx, a.steps, dt, form, a.eps));
    CHECK(nb_synchronize(ctx));
    const double secs = seconds_since(t0);

    const double pairs = evals_per_step * double(a.steps) * (double(N) * double(N) - double(N));
    if (a.integ == "lf") printf("Reduced Symplectic Solver took %fs\n", secs);
    else if (a.integ == "rk") printf("Reduced Runge-Kutta Solver took %fs\n", secs);
    else printf("Reduced Euler blocked solver took %fs\n", secs);
    printf("Performance: %fMPairs/s\n", pairs / (1e6 * secs));
    printf("Simulated steps/s: %f\n", double(a.steps) / secs);

    Planes pos(N), vel(N);
    CHECK(nb_get_state(ctx, pos.p, vel.p));
    const double pos_dev = mean_deviation(pos, pos0, N, "position", a.verbose);
    const double vel_dev = mean_deviation(vel, vel0, N, "velocity", a.verbose);
    printf("Average position deviation from reference: %.13f\n", pos_dev);
    printf("Average velocity deviation from reference: %.13f\n", vel_dev);
    double T1, U1;
    CHECK(nb_energy(ctx, a.eps, &T1, &U1));
    const double E0 = T0 + U0, E1 = T1 + U1;
    const double edev = std::fabs((E0 - E1) / E0);   // CPU/omp_solver.cpp:1205
    printf("E0 = %.15e, E = %.15e, |dE/E| = %.6e\n", E0, E1, edev);
    if (!a.snapshot.empty()) {
        CHECK(snapshot(dt * double(a.steps), true));
        printf("Snapshot frames (t = 0 and t = %g) written to %s\n", dt * double(a.steps), a.snapshot.c_str());
    }

    double max_rel = 0.0;
    if (form != NB_FORM_FP64 && N > 1) {
        Planes a64(N), af(N);
        CHECK(nb_accel(ctx, NB_FORM_FP64, a.eps, a64.p));
        CHECK(nb_accel(ctx, form == NB_FORM_LEARNED ? NB_FORM_MIXED : form, a.eps, af.p));
        for (size_t i = 0; i < N; i++) {
            const double r = std::sqrt(a64.p[0][i] * a64.p[0][i] + a64.p[1][i] * a64.p[1][i] + a64.p[2][i] * a64.p[2][i]);
            if (r == 0.0) continue;
            const double dx = af.p[0][i] - a64.p[0][i], dy = af.p[1][i] - a64.p[1][i], dz = af.p[2][i] - a64.p[2][i];
            max_rel = std::fmax(max_rel, std::sqrt(dx * dx + dy * dy + dz * dz) / r);
        }
        printf("Max relative force error vs FP64: %.6e\n", max_rel);
    }

    if (a.out_given) {
        const bool exists = file_exists(a.out);
        FILE* f = fopen(a.out.c_str(), exists ? "a" : "w");
        if (f) {
            if (!exists)
                fprintf(f, "N, s, T, dt, integrator, order, eps, ic, seed, form, time_s, GPairs_s, steps_s, T_kin, U, E, "
                           "E_dev, I'', pos_dev, vel_dev, max_rel_force_err\n");
            fprintf(f, "%zu, %zu, %f, %e, %s, %d, %e, %s, %llu, %s, %f, %f, %f, %.15e, %.15e, %.15e, %e, %e, %e, %e, %e\n",
                    N, a.steps, a.T, dt, a.integ.c_str(), order, a.eps, a.ic.c_str(), (unsigned long long)a.seed,
                    a.form.c_str(), secs, pairs / (1e9 * secs), double(a.steps) / secs, T1, U1, E1, edev,
                    4 * T1 + 2 * U1, pos_dev, vel_dev, max_rel);
            fclose(f);
        } else {
            fprintf(stderr, "cannot open %s\n", a.out.c_str());
        }
    }
    return 0;
}

FILE* open_table(const std::string& path, const char* header) {
    const bool exists = file_exists(path);
    FILE* f = fopen(path.c_str(), exists ? "a" : "w");
    if (f && !exists) fputs(header, f);
    return f;
}

// Exact-step form of the experiments: learned (a stepping form) is refused.
bool experiment_form(const Args& a, int& form) {
    form = form_id(a.form);
    if (form == NB_FORM_LEARNED) {
        fprintf(stderr, "the accuracy experiments take fp64, mixed, mixed-raw or fp32 forms\n");
        return false;
    }
    return true;
}

// -t omp-accuracy (omp_accuracy, CPU/omp_solver.cpp:899-990): for each mass
// ratio (geometric from ratio_start to ratio_stop in ratio_steps steps) and
// N = start_N * step_N^i, the ring is permuted, its forces are evaluated once
// and compared with the analytic ones.
int run_accuracy(nb_ctx* ctx, const Args& a) {
    int form;
    if (!experiment_form(a, form)) return 1;
    if (a.eps != 0.0) printf("-eps is ignored here: the analytic ring force is unsoftened\n");
    const double R = 1.0, omega = 2.0 * M_PI;
    FILE* out = open_table(a.out, "N, precision, ratio, average_force_length_error[%], average_force_angle_error[%]\n");
    if (!out) {
        fprintf(stderr, "cannot open %s\n", a.out.c_str());
        return 1;
    }
    const char* prec = form == NB_FORM_FP64 ? "dp" : "sp";
    const double root = std::pow(a.ratio_stop / a.ratio_start, 1.0 / double(a.ratio_steps));
    printf("ratio_root: %f, ratio_start: %f\n", root, a.ratio_start);
    uint64_t run = 0;
    for (size_t rs = 0; rs <= a.ratio_steps; rs++) {
        const double ratio = a.ratio_start * std::pow(root, double(rs));
        for (size_t i = 0; i < a.num_steps; i++) {
            const size_t n = a.start_N * size_t(std::round(std::pow(double(a.step_N), double(i))));
            System sys(n);
            const double params[3] = {R, omega, ratio};
            CHECK(nb_init_system(NB_IC_RING, n, 0, params, sys.mass.data(), sys.pos.p, sys.vel.p, sys.force.p));
            sys.permute(seeded_permutation(n, a.seed + run++));
            printf("################### N = %zu, s = %zu ###################\n", n, size_t(1));
            printf("curr_ratio = %.3e\n", ratio);
            CHECK(nb_set_state(ctx, n, sys.mass.data(), sys.pos.p, sys.vel.p));
            Planes f(n);
            const auto t0 = std::chrono::steady_clock::now();
            CHECK(nb_compute_forces(ctx, form, 0.0, f.p));
            printf("Solver took: %fs\n", seconds_since(t0));
            double e[6];
            CHECK(nb_force_error(n, f.p, sys.force.p, e));
            printf("Average force error: %.10e\n", e[0]);
            printf("Largest force error: %.10e\n", e[1]);
            printf("Largest to Average ratio: %.10e\n", e[1] / e[0]);
            printf("Average normalized angle error: %.10e\n", e[2]);
            printf("Average normalized length error: %.10e\n", e[3]);
            printf("Largest normalized angle error: %.10e\n", e[4]);
            printf("Largest normalized length error: %.10e\n\n", e[5]);
            fprintf(out, "%zu, %s, %.5e, %.5e, %.6e\n", n, prec, ratio, e[3] * 100.0, e[2] * 100.0);
        }
    }
    fclose(out);
    return 0;
}

// -t omp-ratio (omp_ratio_accuracy, CPU/omp_solver.cpp:995-1243): second-order
// leapfrog steps of 1/rev_steps revolutions on a permuted ring, each compared
// with the rigid rotation, until the energy error reaches 1 or the largest
// relative position or velocity deviation reaches max_error.
int run_ratio(nb_ctx* ctx, const Args& a) {
    int form;
    if (!experiment_form(a, form)) return 1;
    FILE* out = open_table(a.out, "N, precision, s, s_collapse, ratio, max_error, last_E_dev, last_avg_pos_dev, "
                                  "last_avg_vel_dev, max_pos_dev, max_vel_dev\n");
    if (!out) {
        fprintf(stderr, "cannot open %s\n", a.out.c_str());
        return 1;
    }
    const char* prec = form == NB_FORM_FP64 ? "dp" : "sp";
    double c[16], d[16];
    size_t lf = 0;
    CHECK(nb_yoshida(2, c, d, &lf));
    print_coefficients(c, d, lf);
    const double R = 1.0, omega = 2.0 * M_PI;
    const double dt = 1.0 / double(a.rev_steps);
    uint64_t run = 0;
    for (size_t k = 0; k < a.num_steps; k++) {
        const size_t n = a.start_N + k * a.step_N;
        const double root = std::pow(a.ratio_stop / a.ratio_start, 1.0 / double(a.ratio_steps));
        printf("ratio_root: %f, ratio_start: %f\n", root, a.ratio_start);
        for (size_t rs = 0; rs <= a.ratio_steps; rs++) {
            const double ratio = a.ratio_start * std::pow(root, double(rs));
            System sys(n);
            const double params[3] = {R, omega, ratio};
            CHECK(nb_init_system(NB_IC_RING, n, 0, params, sys.mass.data(), sys.pos.p, sys.vel.p, sys.force.p));
            const std::vector<size_t> perm = seeded_permutation(n, a.seed + run++);
            sys.permute(perm);
            CHECK(nb_set_state(ctx, n, sys.mass.data(), sys.pos.p, sys.vel.p));
            double T0, U0, I0;
            CHECK(nb_energy(ctx, a.eps, &T0, &U0));
            CHECK(nb_inertia(n, sys.mass.data(), sys.pos.p, &I0));
            const double E0 = T0 + U0;
            printf("########Initial system condition########\n");
            printf("System size N = %zu\n", n);
            printf("Mass ratio: %f\n", ratio);
            printf("T0 = %f\n", T0);
            printf("U0 = %f\n", U0);
            printf("E0 = %f\n", E0);
            printf("I0 = %f\n", I0);
            printf("I0'' = %f\n", 4 * T0 + 2 * U0);
            printf("Number of system revolutions to simulate: %f\n", 1.0);
            double e_sol = 0.0, pos_sol = 0.0, vel_sol = 0.0, pmax_sol = 0.0, vmax_sol = 0.0;
            size_t collapse = a.steps;
            System ref(n), pred(n);
            Planes pos(n), vel(n);
            for (size_t t = 1; t <= a.steps; t++) {
                if (a.verbose) printf("curr: %zu\n", t);
                CHECK(nb_step(ctx, 1, dt, c, d, lf, form, a.eps));
                CHECK(nb_ring_predict(n, R, omega, t, a.rev_steps, pred.pos.p, pred.vel.p));
                for (int q = 0; q < 3; q++)
                    for (size_t i = 0; i < n; i++) {
                        ref.pos.p[q][i] = pred.pos.p[q][perm[i]];
                        ref.vel.p[q][i] = pred.vel.p[q][perm[i]];
                    }
                double T, U, I;
                CHECK(nb_energy(ctx, a.eps, &T, &U));
                CHECK(nb_get_state(ctx, pos.p, vel.p));
                CHECK(nb_inertia(n, sys.mass.data(), pos.p, &I));
                const double E = T + U;
                double pd[2], vd[2];
                CHECK(nb_deviation(n, pos.p, ref.pos.p, pd));
                CHECK(nb_deviation(n, vel.p, ref.vel.p, vd));
                const double e_dev = std::fabs((E0 - E) / E0);
                if (a.verbose) {
                    printf("\n########System condition after %zu timesteps########\n", t);
                    printf("T = %f\n", T);
                    printf("U = %f\n", U);
                    printf("I = %f\n", I);
                    printf("I'' = %f\n", 4 * T + 2 * U);
                    printf("E = %f\n", E);
                    printf("E_dev = %f\n", e_dev);
                    printf("Average pos deviation from reference: %.13f\n", pd[0]);
                    printf("Average vel deviation from reference: %.13f\n", vd[0]);
                    printf("Maximum pos length deviation from reference: %.13f\n", pd[1]);
                    printf("Maximum vel length deviation from reference: %.13f\n", vd[1]);
                }
                if (e_dev >= 1.0 || pd[1] >= a.max_error || vd[1] >= a.max_error) {
                    collapse = t;
                    break;
                }
                e_sol = e_dev;
                pos_sol = pd[0];
                vel_sol = vd[0];
                pmax_sol = pd[1];
                vmax_sol = vd[1];
            }
            printf("true solution = %zu\n", collapse);
            printf("last E_dev = %.10e, last max pos dev = %.10e, last max vel dev = %.10e\n", e_sol, pmax_sol, vmax_sol);
            fprintf(out, "%zu, %s, %zu, %zu, %f, %f, %f, %f, %f, %f, %f\n", n, prec, a.steps, collapse, ratio,
                    a.max_error, e_sol, pos_sol, vel_sol, pmax_sol, vmax_sol);
            fflush(out);
        }
        printf("\n\n--------------------------------------------\n--------------------------------------------\n\n\n");
    }
    fclose(out);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    parse(argc, argv, a);
    if (a.exec != "omp" && a.exec != "omp-accuracy" && a.exec != "omp-ratio") {
        printf("Execution type %s is not provided by the GPU path (omp, omp-accuracy, omp-ratio are)\n",
               a.exec.c_str());
        return 2;
    }
    if (a.vect != "AVX2") printf("Vectorization type %s maps to the gfx950 force kernel\n", a.vect.c_str());
    nb_ctx* ctx = nullptr;
    CHECK(nb_create(&ctx, a.device));
    if (a.solver == "full") CHECK(nb_set_kernel(ctx, NB_KERNEL_TILE));
    int rc;
    if (a.exec == "omp-accuracy") rc = run_accuracy(ctx, a);
    else if (a.exec == "omp-ratio") rc = run_ratio(ctx, a);
    else rc = run_omp(ctx, a);
    nb_destroy(ctx);
    return rc;
}
