// nbody_cli.cpp -- command-line driver of the GPU force + leapfrog path.
//
// Mirrors the reference's flags (CPU/main.cpp:11-289): -t -N -s -T -i -order
// -V -p -r -o -v, with dt = T/s (CPU/omp_solver.cpp:640), the same defaults
// and the same "unknown value -> default, with a message" behaviour.  The
// GPU path is the leapfrog driver (-t omp -i lf); -V is accepted for
// compatibility (every vectorisation variant maps to the gfx950 kernel) and
// -p sp selects the mixed-precision form.  Added options:
//   -eps <e>       softening length (default 0: the reference's force)
//   -ic ring|sphere|plummer, -seed <k>, -a <scale>, -rmax <r>
//   -form fp64|mixed|mixed-raw|learned
//   -device <d>
//   -period <k> -threshold <x> -rc <r> -train-steps <n> -train-iters <n>  (learned form)
//   -snapshot <file>   initial and final state (with FP64 forces) as frames in
//                      the reference's NBody_<N>.dat layout (CPU/nbody_system.cpp:133)
// Output: the reference's stdout lines (coefficients, run time, MPairs/s,
// deviations from the initial state) plus energies; with -o, one CSV row per
// run appended under a header written when the file is new (the FPGA
// accuracy format, FPGA/host/main.cpp:582, extended).
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "nbody.h"

namespace {

struct Args {
    size_t N = 1000;
    size_t steps = 1000;
    double T = 1.0;
    double ratio = 100000.0;
    bool verbose = false;
    std::string out;
    std::string exec = "omp";
    std::string integ = "lf";
    int order = 1;
    std::string vect = "AVX2";
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
    int train_iters = 4000;
};

bool file_exists(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0;
}

void parse(int argc, char** argv, Args& a) {
    const std::map<std::string, int> execs = {{"omp", 0}};
    const std::map<std::string, int> integs = {{"euler", 0}, {"lf", 1}, {"rk", 2}};
    const std::map<std::string, int> vects = {{"AVX2", 0}, {"AVX512", 1}, {"AVX2-RSQRT", 2}, {"AVX512-RSQRT", 3},
                                              {"AVX512-RSQRT-4I", 4}};
    const std::map<std::string, int> precs = {{"dp", 0}, {"sp", 1}};
    for (int i = 1; i < argc; i++) {
        const std::string f = argv[i];
        const bool has = i + 1 < argc;
        if (f == "-N" && has) a.N = strtoull(argv[++i], nullptr, 10);
        else if (f == "-s" && has) a.steps = strtoull(argv[++i], nullptr, 10);
        else if (f == "-T" && has) a.T = atof(argv[++i]);
        else if (f == "-r" && has) a.ratio = atof(argv[++i]);
        else if (f == "-v") a.verbose = true;
        else if (f == "-o" && has) a.out = argv[++i];
        else if (f == "-snapshot" && has) a.snapshot = argv[++i];
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
                printf("Unknown Integration kind: %s\n Defaulting to lf\n", a.integ.c_str());
                a.integ = "lf";
            }
        } else if (f == "-V" && has) {
            a.vect = argv[++i];
            if (!vects.count(a.vect)) {
                printf("Unknown Vectorization type: %s\n Defaulting to AVX2\n", a.vect.c_str());
                a.vect = "AVX2";
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
    if (a.N >= 100000000) {
        printf("Input problem size might be too large!\nMake sure to stay below 10^8 particles.\nResetting to defaults");
        a.N = 1000;
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
    if (a.form.empty()) a.form = (a.prec == "sp") ? "mixed" : "fp64";
}

int form_id(const std::string& f) {
    if (f == "fp64") return NB_FORM_FP64;
    if (f == "mixed") return NB_FORM_MIXED;
    if (f == "mixed-raw") return NB_FORM_MIXED_RAW;
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
    explicit Planes(size_t n) : buf(3 * n) {
        for (int k = 0; k < 3; k++) p[k] = buf.data() + k * n;
    }
};

double mean_deviation(const Planes& a, const Planes& b, size_t n) {
    double s = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double dx = a.p[0][i] - b.p[0][i], dy = a.p[1][i] - b.p[1][i], dz = a.p[2][i] - b.p[2][i];
        s += std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    return n ? s / double(n) : 0.0;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    parse(argc, argv, a);
    if (a.integ != "lf") {
        printf("Integration kind %s is not provided by the GPU path; using lf\n", a.integ.c_str());
        a.integ = "lf";
    }
    if (a.vect != "AVX2") printf("Vectorization type %s maps to the gfx950 force kernel\n", a.vect.c_str());
    const size_t N = a.N;
    const double dt = a.T / double(a.steps);
    const int form = form_id(a.form);

    nb_ctx* ctx = nullptr;
    CHECK(nb_create(&ctx, a.device));
    std::vector<double> mass(N);
    Planes pos(N), vel(N), force(N);
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
    CHECK(nb_init_system(kind, N, a.seed, params, mass.data(), pos.p, vel.p, force.p));
    Planes pos0 = pos, vel0 = vel;
    for (int k = 0; k < 3; k++) {
        pos0.p[k] = pos0.buf.data() + k * N;
        vel0.p[k] = vel0.buf.data() + k * N;
    }
    printf("N = %zu, s = %zu\n", N, a.steps);

    double c[16], d[16];
    size_t lf = 0;
    if (a.order != 2 && a.order != 4 && a.order != 6 && a.order != 8)
        printf("Order %d not supported! Defaulting to 2nd order\n", a.order);
    CHECK(nb_yoshida(a.order, c, d, &lf));
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

    CHECK(nb_set_state(ctx, N, mass.data(), pos.p, vel.p));
    double T0, U0;
    CHECK(nb_energy(ctx, a.eps, &T0, &U0));
    auto snapshot = [&](double t, bool append) -> nb_status {
        Planes p(N), v(N), f(N);
        nb_status s = nb_get_state(ctx, p.p, v.p);
        if (s == NB_OK) s = nb_compute_forces(ctx, NB_FORM_FP64, a.eps, f.p);
        if (s == NB_OK) s = nb_snapshot_write(a.snapshot.c_str(), append ? 1 : 0, t, N, mass.data(), p.p, v.p, f.p);
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
        CHECK(nb_learned_train(ctx, a.train_iters, 4096, 1e-3, a.seed, &loss));
        printf("final training loss %.4e\n", loss);
        CHECK(nb_set_state(ctx, N, mass.data(), pos.p, vel.p));   // restart from the initial state
    }

    CHECK(nb_synchronize(ctx));
    const auto t0 = std::chrono::steady_clock::now();
    CHECK(nb_step(ctx, a.steps, dt, c, d, lf, form, a.eps));
    CHECK(nb_synchronize(ctx));
    const auto t1 = std::chrono::steady_clock::now();
    const double secs = std::chrono::duration<double>(t1 - t0).count();

    const double pairs = double(lf - 1) * double(a.steps) * (double(N) * double(N) - double(N));
    printf("Reduced Symplectic Solver took %fs\n", secs);
    printf("Performance: %fMPairs/s\n", pairs / (1e6 * secs));
    printf("Simulated steps/s: %f\n", double(a.steps) / secs);

    CHECK(nb_get_state(ctx, pos.p, vel.p));
    const double pos_dev = mean_deviation(pos, pos0, N), vel_dev = mean_deviation(vel, vel0, N);
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

    if (!a.out.empty()) {
        const bool exists = file_exists(a.out);
        FILE* f = fopen(a.out.c_str(), exists ? "a" : "w");
        if (f) {
            if (!exists)
                fprintf(f, "N, s, T, dt, order, eps, ic, seed, form, time_s, GPairs_s, steps_s, T_kin, U, E, E_dev, I'', "
                           "pos_dev, vel_dev, max_rel_force_err\n");
            fprintf(f, "%zu, %zu, %f, %e, %zu, %e, %s, %llu, %s, %f, %f, %f, %.15e, %.15e, %.15e, %e, %e, %e, %e, %e\n",
                    N, a.steps, a.T, dt, lf, a.eps, a.ic.c_str(), (unsigned long long)a.seed, a.form.c_str(), secs,
                    pairs / (1e9 * secs), double(a.steps) / secs, T1, U1, E1, edev, 4 * T1 + 2 * U1, pos_dev, vel_dev,
                    max_rel);
            fclose(f);
        } else {
            fprintf(stderr, "cannot open %s\n", a.out.c_str());
        }
    }
    nb_destroy(ctx);
    return 0;
}
