/* sanitize_check.c -- TEST INFRASTRUCTURE ONLY.  Drives every function of
 * the CPU oracle (nbody_oracle.c, nbody_mlp_oracle.c) on small inputs,
 * including the edge cases the parity tests use (empty systems, one
 * particle, coincident particles at eps = 0, ragged sizes), so that a build
 * with -fsanitize=address,undefined (oracle/Makefile, target `sanitize`)
 * reports any out-of-bounds access, leak or undefined behaviour in the
 * oracle itself.  It checks no numbers: the pytest suite does that.
 * Exit status 0 when every call returned. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_rng_stream(uint64_t seed, size_t n, double *out);
void orc_ic_sphere(size_t n, uint64_t seed, double *m, double *x, double *y, double *z, double *vx, double *vy,
                   double *vz);
void orc_ic_plummer(size_t n, uint64_t seed, double a, double rmax, double *m, double *x, double *y, double *z,
                    double *vx, double *vy, double *vz);
void orc_ic_ring(size_t n_total, double R, double omega, double ratio, double *m, double *x, double *y, double *z,
                 double *vx, double *vy, double *vz, double *fx, double *fy, double *fz);
void orc_ring_predict(size_t n_total, double R, double omega, double phase, double *x, double *y, double *z,
                      double *vx, double *vy, double *vz);
void orc_ring_accel_exact(size_t n_total, double R, double mring, double mcentre, double eps, double *ax,
                          double *ay, double *az);
void orc_accel(size_t n, const double *m, const double *x, const double *y, const double *z, double eps, double G,
               double *ax, double *ay, double *az);
void orc_accel_ld(size_t n, const double *m, const double *x, const double *y, const double *z, double eps,
                  double G, double *ax, double *ay, double *az);
void orc_accel_block(size_t ni, const double *pi, size_t nj, const double *pj, double eps, double G, double *acc);
void orc_forces_reduced(size_t n, const double *m, const double *x, const double *y, const double *z, double eps,
                        double G, double *fx, double *fy, double *fz);
size_t orc_yoshida(int order, double *c, double *d);
void orc_leapfrog(size_t n, const double *m, double *x, double *y, double *z, double *vx, double *vy, double *vz,
                  size_t steps, double dt, const double *c, const double *d, size_t lf_steps, double eps, double G);
void orc_euler(size_t n, const double *m, double *x, double *y, double *z, double *vx, double *vy, double *vz,
               size_t steps, double dt, double eps, double G);
void orc_rk(size_t n, const double *m, double *x, double *y, double *z, double *vx, double *vy, double *vz,
            size_t steps, double dt, const double *a, const double *b, size_t s, double eps, double G);
void orc_force_error(size_t n, const double *fx, const double *fy, const double *fz, const double *rx,
                     const double *ry, const double *rz, double *out);
void orc_deviation(size_t n, const double *cx, const double *cy, const double *cz, const double *rx,
                   const double *ry, const double *rz, double *out);
double orc_inertia(size_t n, const double *m, const double *x, const double *y, const double *z);
void orc_energy(size_t n, const double *m, const double *x, const double *y, const double *z, const double *vx,
                const double *vy, const double *vz, double eps, double G, double *T, double *U);
double orc_max_rel_err(size_t n, const double *ax, const double *ay, const double *az, const double *rx,
                       const double *ry, const double *rz, double *mean_out);
size_t orc_mlp_param_count(void);
double orc_bf16_round(double v);
void orc_mlp_forward(const double *params, size_t rows, const double *x, int bf16, double *y);
double orc_mlp_grad(const double *params, size_t rows, const double *x, const double *t, int bf16, double *grad);
void orc_adam(size_t n, double *p, const double *g, double *m, double *v, double lr, double b1, double b2,
              double eps, int t);
void orc_poly_extrap(size_t n3, const double *a0, const double *a1, const double *a2, const double *a3,
                     double *out);
void orc_lagrange4(const double *t, double x, double *w);
void orc_features(size_t n, const double *a0, const double *a1, const double *a2, const double *a3,
                  const double *v, double *x, double *scale, double *cubic, double *frame);
void orc_near_accel(size_t n, const double *m, const double *x, const double *y, const double *z, double rc,
                    double eps, double *ax, double *ay, double *az);
double orc_near_taper(double r, double rc);

/* Exactly-sized heap arrays, so that any overrun reaches a redzone. */
static double *arr(size_t n) {
    double *p = (double *)calloc(n ? n : 1, sizeof(double));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

/* One system of n particles: the force, integrator, energy and metric calls. */
static void system_calls(size_t n, int kind) {
    double *m = arr(n), *x = arr(n), *y = arr(n), *z = arr(n), *vx = arr(n), *vy = arr(n), *vz = arr(n);
    double *ax = arr(n), *ay = arr(n), *az = arr(n), *bx = arr(n), *by = arr(n), *bz = arr(n);
    if (kind == 0) orc_ic_sphere(n, 3, m, x, y, z, vx, vy, vz);
    else orc_ic_plummer(n, 3, 1.0, 20.0, m, x, y, z, vx, vy, vz);
    if (n > 3) {   /* coincident particles, as the eps = 0 masking tests use */
        x[1] = x[0]; y[1] = y[0]; z[1] = z[0];
    }
    for (int e = 0; e < 2; e++) {
        const double eps = e ? 0.01 : 0.0;
        orc_accel(n, m, x, y, z, eps, 1.0, ax, ay, az);
        orc_accel_ld(n, m, x, y, z, eps, 1.0, bx, by, bz);
        orc_forces_reduced(n, m, x, y, z, eps, 1.0, bx, by, bz);
        double T, U, mean, out[6];
        orc_energy(n, m, x, y, z, vx, vy, vz, eps, 1.0, &T, &U);
        (void)orc_max_rel_err(n, ax, ay, az, bx, by, bz, &mean);
        orc_force_error(n, ax, ay, az, bx, by, bz, out);
        orc_deviation(n, ax, ay, az, bx, by, bz, out);
        (void)orc_near_accel;
        orc_near_accel(n, m, x, y, z, 0.3, eps, bx, by, bz);
    }
    (void)orc_inertia(n, m, x, y, z);
    /* ring-share layout: (x, y, z, m) records, i a prefix of j */
    const size_t ni = n / 2 + 1 > n ? n : n / 2 + 1;
    double *pj = arr(4 * n), *acc = arr(3 * (ni ? ni : 1));
    for (size_t i = 0; i < n; i++) {
        pj[4 * i] = x[i]; pj[4 * i + 1] = y[i]; pj[4 * i + 2] = z[i]; pj[4 * i + 3] = m[i];
    }
    orc_accel_block(n ? ni : 0, pj, n, pj, 0.01, 1.0, acc);
    for (int order = 2; order <= 8; order += 2) {
        double c[16], d[16];
        const size_t lf = orc_yoshida(order, c, d);
        orc_leapfrog(n, m, x, y, z, vx, vy, vz, 2, 1e-3, c, d, lf, 0.01, 1.0);
    }
    orc_euler(n, m, x, y, z, vx, vy, vz, 2, 1e-3, 0.01, 1.0);
    {   /* classic RK4 in the (s-1)^2 row-major tableau layout */
        const double a[9] = {0.5, 0, 0, 0, 0.5, 0, 0, 0, 1.0};
        const double b[4] = {1.0 / 6, 1.0 / 3, 1.0 / 3, 1.0 / 6};
        orc_rk(n, m, x, y, z, vx, vy, vz, 2, 1e-3, a, b, 4, 0.01, 1.0);
        const double a1[1] = {0.0};
        const double b1[1] = {1.0};
        orc_rk(n, m, x, y, z, vx, vy, vz, 1, 1e-3, a1, b1, 1, 0.01, 1.0);
    }
    free(pj); free(acc);
    free(m); free(x); free(y); free(z); free(vx); free(vy); free(vz);
    free(ax); free(ay); free(az); free(bx); free(by); free(bz);
}

static void ring_calls(size_t n_total) {
    double *m = arr(n_total), *x = arr(n_total), *y = arr(n_total), *z = arr(n_total);
    double *vx = arr(n_total), *vy = arr(n_total), *vz = arr(n_total);
    double *fx = arr(n_total), *fy = arr(n_total), *fz = arr(n_total);
    orc_ic_ring(n_total, 1.0, 0.5, 1e5, m, x, y, z, vx, vy, vz, fx, fy, fz);
    orc_ring_predict(n_total, 1.0, 0.5, 0.1, x, y, z, vx, vy, vz);
    orc_ring_accel_exact(n_total, 1.0, 1e-3, 1.0, 0.0, fx, fy, fz);
    orc_ring_accel_exact(n_total, 1.0, 1e-3, 1.0, 0.01, fx, fy, fz);
    free(m); free(x); free(y); free(z); free(vx); free(vy); free(vz); free(fx); free(fy); free(fz);
}

static void learned_calls(size_t rows) {
    const size_t np = orc_mlp_param_count();
    double *p = arr(np), *g = arr(np), *mm = arr(np), *vv = arr(np);
    double *xin = arr(15 * rows), *t = arr(3 * rows), *out = arr(3 * rows);
    orc_rng_stream(5, np, p);
    for (size_t i = 0; i < np; i++) p[i] = 0.1 * (p[i] - 0.5);
    orc_rng_stream(6, 15 * rows, xin);
    orc_rng_stream(7, 3 * rows, t);
    for (int bf = 0; bf < 2; bf++) {
        orc_mlp_forward(p, rows, xin, bf, out);
        (void)orc_mlp_grad(p, rows, xin, t, bf, g);
    }
    orc_adam(np, p, g, mm, vv, 1e-3, 0.9, 0.999, 1e-8, 1);
    (void)orc_bf16_round(1.0 / 3.0);
    double *a[4], *v = arr(3 * rows), *feat = arr(15 * rows), *scale = arr(rows), *cubic = arr(3 * rows),
                  *frame = arr(9 * rows), *ex = arr(3 * rows);
    for (int k = 0; k < 4; k++) {
        a[k] = arr(3 * rows);
        orc_rng_stream(10 + k, 3 * rows, a[k]);
    }
    orc_rng_stream(20, 3 * rows, v);
    if (rows > 1) {   /* degenerate histories: all zero, and collinear */
        memset(a[0], 0, 3 * sizeof(double)); memset(a[1], 0, 3 * sizeof(double));
        memset(a[2], 0, 3 * sizeof(double)); memset(a[3], 0, 3 * sizeof(double));
        for (int k = 0; k < 4; k++) { a[k][3] = k; a[k][4] = 0.0; a[k][5] = 0.0; }
    }
    orc_features(rows, a[0], a[1], a[2], a[3], v, feat, scale, cubic, frame);
    orc_poly_extrap(3 * rows, a[0], a[1], a[2], a[3], ex);
    const double tk[4] = {0.0, -1.0, -2.5, -3.0};
    double w[4];
    orc_lagrange4(tk, 1.0, w);
    for (double r = 0.0; r <= 0.5; r += 0.05) (void)orc_near_taper(r, 0.3);
    for (int k = 0; k < 4; k++) free(a[k]);
    free(v); free(feat); free(scale); free(cubic); free(frame); free(ex);
    free(p); free(g); free(mm); free(vv); free(xin); free(t); free(out);
}

int main(void) {
    const size_t sizes[] = {0, 1, 2, 3, 17, 64, 257};
    for (size_t s = 0; s < sizeof sizes / sizeof sizes[0]; s++)
        for (int kind = 0; kind < 2; kind++) system_calls(sizes[s], kind);
    ring_calls(2);
    ring_calls(65);
    learned_calls(1);
    learned_calls(33);
    printf("oracle sanitizer run: ok\n");
    return 0;
}
