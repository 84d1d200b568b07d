/*
 * nbody_mlp_oracle.c -- CPU oracle for the learned acceleration extrapolator.
 *
 * TEST INFRASTRUCTURE ONLY (see nbody_oracle.c).  No reference counterpart:
 * the reference has no learned component.  Pinned by the tests against
 * closed forms (the cubic extrapolation is exact on cubics; the features are
 * invariant under scaling), and the backward pass against central finite
 * differences of the FP64 loss.
 *
 * Network: 15 -> 128 -> 128 -> 128 -> 3, SiLU hidden activations, linear
 * output.  Parameters are packed W1 b1 W2 b2 W3 b3 W4 b4 with W row-major
 * [out][in].  With `bf16` set, every operand of a matrix product (weights,
 * inputs, hidden activations, back-propagated deltas) is first rounded to
 * bfloat16 (round to nearest even), which is what the GPU's MFMA kernels
 * consume; sums stay in FP64.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

enum { NIN = 15, NH = 128, NOUT = 3 };

static const size_t layer_in[4] = {NIN, NH, NH, NH};
static const size_t layer_out[4] = {NH, NH, NH, NOUT};

ORC_API size_t orc_mlp_param_count(void)
{
    size_t n = 0;
    for (int l = 0; l < 4; l++) n += layer_out[l] * layer_in[l] + layer_out[l];
    return n;
}

/* float -> bfloat16 -> float, round to nearest even (NaN kept quiet). */
static double bf16r(double v)
{
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        if (u & 0x007fffffu) u |= 0x00400000u;
        u &= 0xffff0000u;
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
        u &= 0xffff0000u;
    }
    memcpy(&f, &u, 4);
    return (double)f;
}

ORC_API double orc_bf16_round(double v) { return bf16r(v); }

static double silu(double z) { return z / (1.0 + exp(-z)); }

static double silu_d(double z)
{
    double s = 1.0 / (1.0 + exp(-z));
    return s * (1.0 + z * (1.0 - s));
}

static const double *layer_w(const double *p, int l)
{
    for (int k = 0; k < l; k++) p += layer_out[k] * layer_in[k] + layer_out[k];
    return p;
}

/* Forward pass for one row; keeps pre-activations z[l][] and inputs h[l][]. */
static void fwd_row(const double *p, const double *x, int bf16, double h[4][NH], double z[4][NH],
                    double *y)
{
    for (size_t k = 0; k < NIN; k++) h[0][k] = bf16 ? bf16r(x[k]) : x[k];
    for (int l = 0; l < 4; l++) {
        const double *W = layer_w(p, l);
        const double *b = W + layer_out[l] * layer_in[l];
        for (size_t o = 0; o < layer_out[l]; o++) {
            double s = 0.0;
            for (size_t i = 0; i < layer_in[l]; i++) {
                double w = bf16 ? bf16r(W[o * layer_in[l] + i]) : W[o * layer_in[l] + i];
                s += w * h[l][i];
            }
            s += b[o];
            z[l][o] = s;
            if (l < 3) {
                double a = silu(s);
                h[l + 1][o] = bf16 ? bf16r(a) : a;
            } else {
                y[o] = s;
            }
        }
    }
}

ORC_API void orc_mlp_forward(const double *params, size_t rows, const double *x, int bf16, double *y)
{
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < rows; r++) {
        double h[4][NH], z[4][NH];
        fwd_row(params, x + r * NIN, bf16, h, z, y + r * NOUT);
    }
}

/* Loss L = sum_rows sum_k (y_k - t_k)^2 / (2 rows) and its gradient with
 * respect to every parameter (packed like the parameters).  Returns L. */
ORC_API double orc_mlp_grad(const double *params, size_t rows, const double *x, const double *t,
                            int bf16, double *grad)
{
    const size_t np = orc_mlp_param_count();
    memset(grad, 0, np * sizeof(double));
    double loss = 0.0;
    for (size_t r = 0; r < rows; r++) {
        double h[4][NH], z[4][NH], y[NOUT];
        fwd_row(params, x + r * NIN, bf16, h, z, y);
        double delta[NH], prev[NH];
        for (size_t k = 0; k < NOUT; k++) {
            double e = y[k] - t[r * NOUT + k];
            loss += 0.5 * e * e / (double)rows;
            delta[k] = e / (double)rows;
        }
        for (int l = 3; l >= 0; l--) {
            const double *W = layer_w(params, l);
            double *gW = grad + (layer_w(params, l) - params);
            double *gb = gW + layer_out[l] * layer_in[l];
            double dl[NH];
            for (size_t o = 0; o < layer_out[l]; o++) dl[o] = bf16 ? bf16r(delta[o]) : delta[o];
            for (size_t o = 0; o < layer_out[l]; o++) {
                gb[o] += delta[o];
                for (size_t i = 0; i < layer_in[l]; i++) gW[o * layer_in[l] + i] += dl[o] * h[l][i];
            }
            if (l == 0) break;
            for (size_t i = 0; i < layer_in[l]; i++) {
                double s = 0.0;
                for (size_t o = 0; o < layer_out[l]; o++) {
                    double w = bf16 ? bf16r(W[o * layer_in[l] + i]) : W[o * layer_in[l] + i];
                    s += w * dl[o];
                }
                prev[i] = s * silu_d(z[l - 1][i]);
            }
            memcpy(delta, prev, layer_in[l] * sizeof(double));
        }
    }
    return loss;
}

/* One Adam step (Kingma & Ba 2015) with bias correction, step count t >= 1. */
ORC_API void orc_adam(size_t n, double *p, const double *g, double *m, double *v, double lr,
                      double b1, double b2, double eps, int t)
{
    double c1 = 1.0 - pow(b1, t), c2 = 1.0 - pow(b2, t);
    for (size_t i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / c1) / (sqrt(v[i] / c2) + eps);
    }
}

/* Cubic extrapolation from the last four equally spaced samples:
 * a_{n+1} = 4 a_n - 6 a_{n-1} + 4 a_{n-2} - a_{n-3} (exact for cubics). */
ORC_API void orc_poly_extrap(size_t n3, const double *a0, const double *a1, const double *a2,
                             const double *a3, double *out)
{
    for (size_t i = 0; i < n3; i++) out[i] = 4.0 * a0[i] - 6.0 * a1[i] + 4.0 * a2[i] - a3[i];
}

/* Lagrange weights of the cubic through samples at times t[0..3] (distinct),
 * evaluated at x: value = sum_k w[k] a(t[k]).  For t = (0, -1, -2, -3) and
 * x = 1 they are (4, -6, 4, -1), the equally spaced case above. */
ORC_API void orc_lagrange4(const double *t, double x, double *w)
{
    for (int k = 0; k < 4; k++) {
        double num = 1.0, den = 1.0;
        for (int j = 0; j < 4; j++)
            if (j != k) {
                num *= x - t[j];
                den *= t[k] - t[j];
            }
        w[k] = num / den;
    }
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Local frame of the network's output: e1 = D3/|D3|; e2 = the unit part of D2
 * orthogonal to e1 (of D1 when D2 is parallel to within 1e-6, else of the
 * coordinate axis least aligned with e1); e3 = e1 x e2.  All zero when
 * D3 = 0. */
static void frame_of(const double *d1, const double *d2, const double *d3, double e[9])
{
    memset(e, 0, 9 * sizeof(double));
    double n3 = sqrt(dot3(d3, d3));
    if (!(n3 > 0.0)) return;
    double e1[3] = {d3[0] / n3, d3[1] / n3, d3[2] / n3}, p[3];
    const double *cand[2] = {d2, d1};
    int found = 0;
    for (int c = 0; c < 2 && !found; c++) {
        double pr = dot3(cand[c], e1), nc = sqrt(dot3(cand[c], cand[c]));
        for (int k = 0; k < 3; k++) p[k] = cand[c][k] - pr * e1[k];
        double np_ = sqrt(dot3(p, p));
        if (np_ > 1e-6 * nc && np_ > 0.0) {
            for (int k = 0; k < 3; k++) p[k] /= np_;
            found = 1;
        }
    }
    if (!found) {
        int ax = 0;
        for (int k = 1; k < 3; k++)
            if (fabs(e1[k]) < fabs(e1[ax])) ax = k;
        double u[3] = {0.0, 0.0, 0.0};
        u[ax] = 1.0;
        double pr = dot3(u, e1);
        for (int k = 0; k < 3; k++) p[k] = u[k] - pr * e1[k];
        double np_ = sqrt(dot3(p, p));
        for (int k = 0; k < 3; k++) p[k] /= np_;
    }
    e[0] = e1[0]; e[1] = e1[1]; e[2] = e1[2];
    e[3] = p[0]; e[4] = p[1]; e[5] = p[2];
    e[6] = e1[1] * p[2] - e1[2] * p[1];
    e[7] = e1[2] * p[0] - e1[0] * p[2];
    e[8] = e1[0] * p[1] - e1[1] * p[0];
}

/* Extrapolation features of particle i from its last four (exact) samples
 * a0 = newest .. a3 (each 3 doubles) and velocity v:
 *   D1 = a0 - a1, D2 = a0 - 2a1 + a2, D3 = a0 - 3a1 + 3a2 - a3 (backward
 *   differences; for equally spaced samples the cubic one spacing ahead is
 *   a0 + D1 + D2 + D3), S = |D1| + |D2| + |D3|,
 *   u0 = a0/|a0|, u1 = D1/S, u2 = D2/S, u3 = D3/S, u4 = v/|v| (0 for 0/0),
 *   x = (|u1|, |u2|, |u3|, u0.u1, u0.u2, u0.u3, u0.u4, u1.u2, u1.u3, u1.u4,
 *        u2.u3, u2.u4, u3.u4, clamp(log10(S/|a0|), -8, 2)/4 (-2 if S or |a0|
 *        is 0), 10 u1.(u2 x u3))
 * -- 15 numbers invariant under rotations and under scaling of the
 * accelerations and of v; the output scale |D3|, the local frame (frame_of)
 * in which the network's 3 outputs are components, and the equally spaced
 * cubic.  The network predicts the residual (a_next - cubic) / |D3| in that
 * frame, so its correction turns with the particle's history. */
ORC_API void orc_features(size_t n, const double *a0, const double *a1, const double *a2,
                          const double *a3, const double *v, double *x, double *scale,
                          double *cubic, double *frame)
{
    for (size_t i = 0; i < n; i++) {
        double d1[3], d2[3], d3[3];
        const double *A0 = a0 + 3 * i, *V = v + 3 * i;
        for (int k = 0; k < 3; k++) {
            double p0 = a0[3 * i + k], p1 = a1[3 * i + k], p2 = a2[3 * i + k], p3 = a3[3 * i + k];
            d1[k] = p0 - p1;
            d2[k] = p0 - 2.0 * p1 + p2;
            d3[k] = p0 - 3.0 * p1 + 3.0 * p2 - p3;
            cubic[3 * i + k] = p0 + d1[k] + d2[k] + d3[k];
        }
        double n0 = sqrt(dot3(A0, A0)), n1 = sqrt(dot3(d1, d1)), n2 = sqrt(dot3(d2, d2)), n3 = sqrt(dot3(d3, d3));
        double nv = sqrt(dot3(V, V)), S = n1 + n2 + n3;
        double i0 = n0 > 0 ? 1.0 / n0 : 0.0, iS = S > 0 ? 1.0 / S : 0.0, iv = nv > 0 ? 1.0 / nv : 0.0;
        double u[5][3];
        for (int k = 0; k < 3; k++) {
            u[0][k] = A0[k] * i0;
            u[1][k] = d1[k] * iS;
            u[2][k] = d2[k] * iS;
            u[3][k] = d3[k] * iS;
            u[4][k] = V[k] * iv;
        }
        double *xr = x + 15 * i;
        xr[0] = n1 * iS;
        xr[1] = n2 * iS;
        xr[2] = n3 * iS;
        xr[3] = dot3(u[0], u[1]);
        xr[4] = dot3(u[0], u[2]);
        xr[5] = dot3(u[0], u[3]);
        xr[6] = dot3(u[0], u[4]);
        xr[7] = dot3(u[1], u[2]);
        xr[8] = dot3(u[1], u[3]);
        xr[9] = dot3(u[1], u[4]);
        xr[10] = dot3(u[2], u[3]);
        xr[11] = dot3(u[2], u[4]);
        xr[12] = dot3(u[3], u[4]);
        if (S > 0 && n0 > 0) {
            double l = log10(S / n0);
            xr[13] = (l < -8.0 ? -8.0 : (l > 2.0 ? 2.0 : l)) / 4.0;
        } else {
            xr[13] = -2.0;
        }
        double c23[3] = {u[2][1] * u[3][2] - u[2][2] * u[3][1], u[2][2] * u[3][0] - u[2][0] * u[3][2],
                         u[2][0] * u[3][1] - u[2][1] * u[3][0]};
        xr[14] = 10.0 * dot3(u[1], c23);
        scale[i] = n3;
        frame_of(d1, d2, d3, frame + 9 * i);
    }
}

/* Taper of the learned form's near field as a function of q = r^2 / rc^2:
 * S = 1 for q <= 1/4 (r <= rc/2), 0 for q >= 1 (r >= rc), and in between
 * S = 1 - u^3 (10 - 15u + 6u^2), u = (q - 1/4) / (3/4) (the C2 quintic
 * "smootherstep"; smooth in r as q is). */
static double taper_q(double q)
{
    if (q >= 1.0) return 0.0;
    if (q <= 0.25) return 1.0;
    double u = (q - 0.25) * (4.0 / 3.0);
    return 1.0 - u * u * u * (10.0 - 15.0 * u + 6.0 * u * u);
}

/* Near-field acceleration of the learned form's force split (brute force):
 * a_i = sum_{j : 0 < r_ij < rc} m_j S(r_ij) d_ij / (r_ij^2 + eps^2)^(3/2).
 * Pinned by the tests: S's joins (value and C2 order), the two-body case, and
 * for rc beyond the system size the near field equals the full softened sum. */
ORC_API void orc_near_accel(size_t n, const double *m, const double *x, const double *y, const double *z,
                            double rc, double eps, double *ax, double *ay, double *az)
{
    const double e2 = eps * eps, rc2 = rc * rc;
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < n; i++) {
        double sx = 0, sy = 0, sz = 0;
        for (size_t j = 0; j < n; j++) {
            double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0 || r2 >= rc2) continue;
            double s2 = r2 + e2;
            double f = m[j] * taper_q(r2 / rc2) / (s2 * sqrt(s2));
            sx += f * dx; sy += f * dy; sz += f * dz;
        }
        ax[i] = sx; ay[i] = sy; az[i] = sz;
    }
}

/* The taper S(r) on its own (for the tests). */
ORC_API double orc_near_taper(double r, double rc)
{
    return taper_q((r * r) / (rc * rc));
}
