// nbody_learned.hip -- the learned acceleration extrapolator (NB_FORM_LEARNED).
//
// No counterpart in the reference.  Every force evaluation of the leapfrog is
// a "substep".  Every `period`-th substep is exact (the configured direct sum);
// the others are extrapolated from the last four EXACT samples a_0..a_3 of
// each particle (newest first, taken at substeps t_0 > t_1 > t_2 > t_3):
//     a(t) ~ C(t) + |D3| tau w(t) (f1 e1 + f2 e2 + f3 e3),   C = sum_k l_k(t) a_k,
// where l_k are the Lagrange weights of the cubic through the four samples at
// their substeps, D1..D3 the backward differences of the samples, f = f(x) the
// MLP of mlp_kernels.hip on 15 rotation- and scale-invariant numbers x (norms,
// dot products and a triple product of a_0, D1..D3 and v, normalised; the
// log of S / |a_0|, S = |D1|+|D2|+|D3|), and (e1, e2, e3) the particle's own
// frame (e1 along D3, e2 in the D3-D2 plane): the correction turns with the
// particle's history, as the physics does, instead of being learned per
// lab-frame axis.  The network is trained on the residual of
// the cubic one period after the newest sample of a history spaced `period`
// substeps apart, in units of |D3| tau; at an earlier substep t = t_0 + k that
// residual is taken in proportion to the nodal polynomial,
// w = prod_k (t - t_k) / prod_k (t_0 + period - t_k), the shape of the
// interpolation error.  The network is applied only to such evenly spaced
// histories (after a fallback the spacing is uneven for a few periods and
// the cubic runs alone).
// Only exact samples enter the history, so extrapolation errors never feed
// back into later extrapolations: the scheme is stable however long it runs.
// (Extrapolating from extrapolated samples evaluates one polynomial further
// and further ahead; the history errors it leaves grow geometrically from
// cycle to cycle.)
// On an exact substep the extrapolation for the same substep is compared with
// the exact result.  While the RMS relative error sqrt(sum|da|^2 / sum|a|^2)
// over the checked particles exceeds `threshold` (or is NaN) every substep is
// exact (fallback); the first extrapolation is used only after a passed check.
// The check is made for the cubic alone and, when the network applies, with
// it; the network is used in the next cycle only if it passed, so a network
// producing garbage never reaches the integrator (the cubic carries on).  Particles
// whose history is rough, |D3| > rho |a_0| (close encounters), or whose
// extrapolation missed at the last check by more than kappa a_rms are not
// extrapolated: their rows of the direct sum are computed exactly on every
// extrapolated substep, and they are left out of the check.
// With a near-field radius rc > 0 the force is split: pairs closer than rc
// (tapered from rc/2, nbody_near.hip) are summed exactly on every substep
// from a sorted cell grid, and only the smooth remainder a - a_near is
// extrapolated (the history holds it), which removes the close-encounter
// noise the extrapolation cannot follow.
// When an extrapolated substep had to sum more than a share of the rows
// exactly (rough_share, a quarter by default: such rows cost about what the
// direct sum does), the next substeps run the direct sum for 8 periods,
// doubling for each further time in a row (up to 64).
// Mode 1 drops the network (f = 0): the method without it.
// No substep waits for its own work on the host: the flagged rows' count
// stays on the device (their kernels size themselves from it), a check is
// decided on the device, and the host takes both from pinned copies one
// substep later, after queueing that substep's near field.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mlp_kernels.h"
#include "nbody_host.h"
#include "nbody_learned.h"
#include "nbody_near.h"

using nbk::P4;
using nbh::DevBuf;

namespace nbl {

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Local frame of the network's output: e1 = D3/|D3|; e2 = the unit part of D2
// orthogonal to e1 (of D1 when D2 is parallel to within 1e-6, else of the
// coordinate axis least aligned with e1); e3 = e1 x e2.  Zero when D3 = 0.
__device__ void frame_of(const double* d1, const double* d2, const double* d3, double* e) {
    for (int k = 0; k < 9; k++) e[k] = 0.0;
    const double n3 = sqrt(dot3(d3, d3));
    if (!(n3 > 0.0)) return;
    const double e1[3] = {d3[0] / n3, d3[1] / n3, d3[2] / n3};
    double p[3];
    bool found = false;
    for (int c = 0; c < 2 && !found; c++) {
        const double* q = c == 0 ? d2 : d1;
        const double pr = dot3(q, e1), nq = sqrt(dot3(q, q));
        for (int k = 0; k < 3; k++) p[k] = q[k] - pr * e1[k];
        const double np_ = sqrt(dot3(p, p));
        if (np_ > 1e-6 * nq && np_ > 0.0) {
            for (int k = 0; k < 3; k++) p[k] /= np_;
            found = true;
        }
    }
    if (!found) {
        int ax = 0;
        for (int k = 1; k < 3; k++)
            if (fabs(e1[k]) < fabs(e1[ax])) ax = k;
        double u[3] = {0.0, 0.0, 0.0};
        u[ax] = 1.0;
        const double pr = dot3(u, e1);
        for (int k = 0; k < 3; k++) p[k] = u[k] - pr * e1[k];
        const double np_ = sqrt(dot3(p, p));
        for (int k = 0; k < 3; k++) p[k] /= np_;
    }
    e[0] = e1[0]; e[1] = e1[1]; e[2] = e1[2];
    e[3] = p[0]; e[4] = p[1]; e[5] = p[2];
    e[6] = e1[1] * p[2] - e1[2] * p[1];
    e[7] = e1[2] * p[0] - e1[0] * p[2];
    e[8] = e1[0] * p[1] - e1[1] * p[0];
}

// Features of particle i from the samples in slots s0..s3 (newest first):
// backward differences D1..D3, S = |D1|+|D2|+|D3|, unit vectors u0 = a0/|a0|,
// u1..u3 = D1..D3 / S, u4 = v/|v|, and the 15 rotation- and scale-invariant
// numbers (|u1|, |u2|, |u3|, u0.u1, u0.u2, u0.u3, u0.u4, u1.u2, u1.u3, u1.u4,
// u2.u3, u2.u4, u3.u4, clamp(log10(S/|a0|), -8, 2)/4 (-2 for 0), 10 u1.(u2 x u3));
// also the cubic with weights w, the scale |D3|, |a0| and the output frame.
__device__ __forceinline__ void features(const P4* __restrict__ hist, int4 slot, double4 w, long n, long i,
                                         const P4& v, float* x, double* cubic, double& scale, double& anorm,
                                         double* e) {
    const P4& a0 = hist[(long)slot.x * n + i];
    const P4& a1 = hist[(long)slot.y * n + i];
    const P4& a2 = hist[(long)slot.z * n + i];
    const P4& a3 = hist[(long)slot.w * n + i];
    const double A0[3] = {a0.x, a0.y, a0.z}, A1[3] = {a1.x, a1.y, a1.z}, A2[3] = {a2.x, a2.y, a2.z},
                 A3[3] = {a3.x, a3.y, a3.z}, V[3] = {v.x, v.y, v.z};
    double d1[3], d2[3], d3[3];
    for (int k = 0; k < 3; k++) {
        d1[k] = A0[k] - A1[k];
        d2[k] = A0[k] - 2.0 * A1[k] + A2[k];
        d3[k] = A0[k] - 3.0 * A1[k] + 3.0 * A2[k] - A3[k];
        cubic[k] = w.x * A0[k] + w.y * A1[k] + w.z * A2[k] + w.w * A3[k];
    }
    const double n0 = sqrt(dot3(A0, A0)), n1 = sqrt(dot3(d1, d1)), n2 = sqrt(dot3(d2, d2)), n3 = sqrt(dot3(d3, d3));
    const double nv = sqrt(dot3(V, V)), S = n1 + n2 + n3;
    const double i0 = n0 > 0 ? 1.0 / n0 : 0.0, iS = S > 0 ? 1.0 / S : 0.0, iv = nv > 0 ? 1.0 / nv : 0.0;
    double u[5][3];
    for (int k = 0; k < 3; k++) {
        u[0][k] = A0[k] * i0;
        u[1][k] = d1[k] * iS;
        u[2][k] = d2[k] * iS;
        u[3][k] = d3[k] * iS;
        u[4][k] = V[k] * iv;
    }
    x[0] = (float)(n1 * iS);
    x[1] = (float)(n2 * iS);
    x[2] = (float)(n3 * iS);
    x[3] = (float)dot3(u[0], u[1]);
    x[4] = (float)dot3(u[0], u[2]);
    x[5] = (float)dot3(u[0], u[3]);
    x[6] = (float)dot3(u[0], u[4]);
    x[7] = (float)dot3(u[1], u[2]);
    x[8] = (float)dot3(u[1], u[3]);
    x[9] = (float)dot3(u[1], u[4]);
    x[10] = (float)dot3(u[2], u[3]);
    x[11] = (float)dot3(u[2], u[4]);
    x[12] = (float)dot3(u[3], u[4]);
    x[13] = (S > 0 && n0 > 0) ? (float)(fmin(fmax(log10(S / n0), -8.0), 2.0) / 4.0) : -2.0f;
    const double c23[3] = {u[2][1] * u[3][2] - u[2][2] * u[3][1], u[2][2] * u[3][0] - u[2][0] * u[3][2],
                           u[2][0] * u[3][1] - u[2][1] * u[3][0]};
    x[14] = (float)(10.0 * dot3(u[1], c23));
    x[15] = 0.f;
    scale = n3;
    anorm = n0;
    frame_of(d1, d2, d3, e);
}

// X[i] = features, cub[i] = (cubic, |D3|); with `mask`, particles whose
// history is rough (|D3| > rho |a_0|, or not finite) or that are held after a
// missed check are marked and listed.
__global__ __launch_bounds__(BLK) void k_features(const P4* __restrict__ hist, int4 slot, double4 w, long n,
                                                  const P4* __restrict__ vel, float* __restrict__ X,
                                                  P4* __restrict__ cub, float4* __restrict__ frame, double rho,
                                                  unsigned char* __restrict__ mask, int* __restrict__ flags,
                                                  int* __restrict__ nflags, const unsigned char* __restrict__ held) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    float x[16];
    double c[3], s, a0, e[9];
    features(hist, slot, w, n, i, vel[i], x, c, s, a0, e);
    float4* xo = (float4*)(X + i * 16);
    for (int q = 0; q < 4; q++) xo[q] = make_float4(x[4 * q], x[4 * q + 1], x[4 * q + 2], x[4 * q + 3]);
    cub[i] = P4{c[0], c[1], c[2], s};
    for (int k = 0; k < 3; k++)
        frame[3 * i + k] = make_float4((float)e[3 * k], (float)e[3 * k + 1], (float)e[3 * k + 2], 0.f);
    if (mask) {
        const bool rough = !(s <= rho * a0) || held[i] != 0;
        mask[i] = rough ? 1 : 0;
        if (rough) flags[atomicAdd(nflags, 1)] = (int)i;
    }
}

// out = cubic + the network's correction |D3| tau (y1 e1 + y2 e2 + y3 e3)
// with its component along the velocity removed (use_net), or the cubic: a
// correction orthogonal to v does no work, so the network cannot make the
// energy drift (a learned correction with a bias along v would heat or cool
// the system however small its RMS error).
__global__ __launch_bounds__(BLK) void k_combine(const P4* __restrict__ cub, const float4* __restrict__ frame,
                                                 const float* __restrict__ Y, const P4* __restrict__ vel, float tau,
                                                 int use_net, long n, P4* __restrict__ out) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    P4 c = cub[i];
    if (use_net) {
        const float4 y = *(const float4*)(Y + i * 4);
        const float4 e1 = frame[3 * i], e2 = frame[3 * i + 1], e3 = frame[3 * i + 2];
        const double f = c.m * (double)tau;
        double dx = f * ((double)y.x * e1.x + (double)y.y * e2.x + (double)y.z * e3.x);
        double dy = f * ((double)y.x * e1.y + (double)y.y * e2.y + (double)y.z * e3.y);
        double dz = f * ((double)y.x * e1.z + (double)y.y * e2.z + (double)y.z * e3.z);
        const P4 v = vel[i];
        const double v2 = v.x * v.x + v.y * v.y + v.z * v.z;
        if (v2 > 0.0) {
            const double pr = (dx * v.x + dy * v.y + dz * v.z) / v2;
            dx -= pr * v.x;
            dy -= pr * v.y;
            dz -= pr * v.z;
        }
        c.x += dx;
        c.y += dy;
        c.z += dz;
    }
    out[i] = P4{c.x, c.y, c.z, 0.0};
}

// Training pair: T[row] = components of (a_exact - a_near - cubic) / |D3| in
// the particle's frame (0 where |D3| = 0)
__global__ __launch_bounds__(BLK) void k_targets(const P4* __restrict__ cub, const float4* __restrict__ frame,
                                                 const P4* __restrict__ acc, const P4* __restrict__ near, long n,
                                                 float* __restrict__ T) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const P4 c = cub[i];
    P4 a = acc[i];
    if (near) {
        const P4 q = near[i];
        a.x -= q.x; a.y -= q.y; a.z -= q.z;
    }
    const double inv = c.m > 0 ? 1.0 / c.m : 0.0;
    const double rx = (a.x - c.x) * inv, ry = (a.y - c.y) * inv, rz = (a.z - c.z) * inv;
    float t[3];
    for (int k = 0; k < 3; k++) {
        const float4 e = frame[3 * i + k];
        t[k] = (float)(rx * e.x + ry * e.y + rz * e.z);
    }
    *(float4*)(T + i * 4) = make_float4(t[0], t[1], t[2], 0.f);
}

// Per block, over the particles not marked rough: sum |pred - (acc - near)|^2,
// sum |acc|^2 and the max relative error (pred extrapolates acc - near).
__global__ __launch_bounds__(BLK) void k_check(const P4* __restrict__ pred, const P4* __restrict__ acc,
                                               const P4* __restrict__ near, const unsigned char* __restrict__ mask,
                                               long n, double* __restrict__ out) {
    __shared__ double s0[BLK], s1[BLK], s2[BLK];
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    double e2 = 0.0, a2 = 0.0, rel = 0.0;
    if (i < n && !mask[i]) {
        const P4 p = pred[i], a = acc[i];
        const P4 q = near ? near[i] : P4{0.0, 0.0, 0.0, 0.0};
        const double dx = p.x + q.x - a.x, dy = p.y + q.y - a.y, dz = p.z + q.z - a.z;
        e2 = dx * dx + dy * dy + dz * dz;
        a2 = a.x * a.x + a.y * a.y + a.z * a.z;
        rel = a2 > 0 ? sqrt(e2 / a2) : 0.0;
        if (!(rel == rel)) rel = INFINITY;
    }
    s0[threadIdx.x] = e2;
    s1[threadIdx.x] = a2;
    s2[threadIdx.x] = rel;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = s0[0];
        out[3 * blockIdx.x + 1] = s1[0];
        out[3 * blockIdx.x + 2] = s2[0];
    }
}

// The outcome of a check, decided on the device from k_check's block sums:
// the hold kernel reads it there, the host from a pinned copy when it next
// needs it (after queueing the next substep's near field).
struct Decision {
    double e2[2], a2[2], mx[2];   // squared error, squared exact value, largest relative error: cubic, network
    double thr2;                  // squared miss above which a particle is held
    int net_ok;                   // the network's extrapolation passed (it is the one judged and held against)
    int ok;                       // the judged extrapolation passed: extrapolation may go on
};

// One wave: per set, the block sums of k_check added in a fixed order (lane
// l takes blocks l, l + 64, ..., then a fixed tree), then the verdict:
// the network passes when it applied and its error is within the threshold
// (NaN fails), the check passes when the judged extrapolation's does.
__global__ __launch_bounds__(64) void k_decide(const double* __restrict__ chk, long nb, int sets, double threshold,
                                               double kappa, long n, Decision* __restrict__ d) {
    const int lane = threadIdx.x;
    Decision r{};
    for (int k = 0; k < 2; k++) {
        double e2 = 0.0, a2 = 0.0, mx = 0.0;
        if (k < sets) {
            const double* h = chk + 3 * nb * k;
            for (long b = lane; b < nb; b += 64) {
                e2 += h[3 * b];
                a2 += h[3 * b + 1];
                mx = fmax(mx, h[3 * b + 2]);
            }
            for (int o = 32; o > 0; o >>= 1) {
                e2 += __shfl_xor(e2, o, 64);
                a2 += __shfl_xor(a2, o, 64);
                mx = fmax(mx, __shfl_xor(mx, o, 64));
            }
        }
        r.e2[k] = e2;
        r.a2[k] = a2;
        r.mx[k] = mx;
    }
    if (lane != 0) return;
    auto rel = [](double e2, double a2) { return a2 > 0 ? sqrt(e2 / a2) : INFINITY; };
    const double err_c = rel(r.e2[0], r.a2[0]);
    const double err_n = sets > 1 ? rel(r.e2[1], r.a2[1]) : INFINITY;
    r.net_ok = (sets > 1 && err_n <= threshold) ? 1 : 0;
    const double err = r.net_ok ? err_n : err_c;
    r.ok = (err <= threshold) ? 1 : 0;
    r.thr2 = kappa * kappa * r.a2[0] / (double)(n > 1 ? n : 1);
    *d = r;
}

// After a check: held[i] = whether particle i's judged extrapolation (the
// network's when it passed, else the cubic) missed the exact remainder by
// more than sqrt(d->thr2) (NaN counts as a miss).
__global__ __launch_bounds__(BLK) void k_hold(const P4* __restrict__ cub, const P4* __restrict__ net,
                                              const P4* __restrict__ acc, const P4* __restrict__ near, long n,
                                              const Decision* __restrict__ d, unsigned char* __restrict__ held) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const double thr2 = d->thr2;
    const P4 p = d->net_ok ? net[i] : cub[i], a = acc[i];
    const P4 q = near ? near[i] : P4{0.0, 0.0, 0.0, 0.0};
    const double dx = a.x - q.x - p.x, dy = a.y - q.y - p.y, dz = a.z - q.z - p.z;
    held[i] = (dx * dx + dy * dy + dz * dz <= thr2) ? 0 : 1;
}

// dst = a - b (b may be null: dst = a)
__global__ __launch_bounds__(BLK) void k_sub(const P4* __restrict__ a, const P4* __restrict__ b, long n,
                                             P4* __restrict__ dst) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    P4 v = a[i];
    if (b) {
        const P4 q = b[i];
        v.x -= q.x; v.y -= q.y; v.z -= q.z;
    }
    dst[i] = P4{v.x, v.y, v.z, 0.0};
}

// dst = a + b
__global__ __launch_bounds__(BLK) void k_add(const P4* __restrict__ a, const P4* __restrict__ b, long n,
                                             P4* __restrict__ dst) {
    const long i = (long)blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const P4 u = a[i], v = b[i];
    dst[i] = P4{u.x + v.x, u.y + v.y, u.z + v.z, 0.0};
}

// The flagged rows, m = min(*count, cap) read on the device (grids sized
// for cap): gathered into a dense set, and their exact sums scattered back.
__global__ __launch_bounds__(BLK) void k_gather(const P4* __restrict__ src, const int* __restrict__ idx,
                                                const int* __restrict__ count, int cap, P4* __restrict__ dst) {
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k < min(*count, cap)) dst[k] = src[idx[k]];
}

__global__ __launch_bounds__(BLK) void k_scatter(const P4* __restrict__ src, const int* __restrict__ idx,
                                                 const int* __restrict__ count, int cap, P4* __restrict__ dst) {
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k < min(*count, cap)) dst[idx[k]] = src[k];
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Minibatch: rows drawn with a counter-based hash of (seed, iteration, slot);
// targets scaled by 1/tau and clipped to +-clip.
__global__ __launch_bounds__(BLK) void k_batch(const float* __restrict__ dsX, const float* __restrict__ dsT, long rows,
                                               unsigned long long seed, int it, int batch, float inv_tau, float clip,
                                               float* __restrict__ bX, float* __restrict__ bT) {
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k >= batch) return;
    const unsigned long long h = mix64(seed * 0x9E3779B97F4A7C15ULL + ((unsigned long long)it << 32) + (unsigned)k + 1);
    const long r = (long)(h % (unsigned long long)rows);
    const float4* xs = (const float4*)(dsX + r * 16);
    float4* xd = (float4*)(bX + (long)k * 16);
    for (int q = 0; q < 4; q++) xd[q] = xs[q];
    const float4 t = *(const float4*)(dsT + r * 4);
    auto c = [&](float v) { return fminf(clip, fmaxf(-clip, v * inv_tau)); };
    *(float4*)(bT + (long)k * 4) = make_float4(c(t.x), c(t.y), c(t.z), 0.f);
}

// sum of squares of the dataset targets, per block (for tau)
__global__ __launch_bounds__(BLK) void k_sumsq(const float* __restrict__ dsT, long rows, double* __restrict__ out) {
    __shared__ double s[BLK];
    double acc = 0.0;
    for (long r = (long)blockIdx.x * BLK + threadIdx.x; r < rows; r += (long)gridDim.x * BLK) {
        const float4 t = *(const float4*)(dsT + r * 4);
        acc += (double)t.x * t.x + (double)t.y * t.y + (double)t.z * t.z;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

inline dim3 blocks(long n) { return dim3((unsigned)((n + BLK - 1) / BLK)); }

// After an extrapolated substep whose exact rows exceed rough_share N the
// direct sum runs instead, and on the substeps of the next kRoughHold
// periods, twice as many for each further time in a row (up to 8 times).
constexpr int kRoughHold = 8;
constexpr int kRoughDoublings = 3;
// The near-field grid is sorted anew after at most this many substeps.
constexpr int kMaxGridAge = 4;

}  // namespace

struct Learned {
    // configuration
    int period = 4;
    double threshold = 1e-3;
    double rho = 1e-2;   // roughness ratio |D3| / |a_n| above which a particle is summed exactly
    double kappa = 1e-3; // miss at a check, in units of the RMS acceleration, above which a particle is held
    double rc = 0.0;     // near-field radius: the history holds a - a_near (0: no split)
    double rough_share = 0.25;   // exact rows, as a share of N, above which the direct sum runs instead
    nbn::Near near;
    int exact_kform = nbk::FORM_MIXED_NR;
    int mode = 0;
    // network: fp32 master parameters, Adam moments, gradient, device copies
    DevBuf params, am, av, grad, netmem, lossbuf;
    mlp::Net net{};
    bool initialised = false, trained = false;
    int adam_t = 0;
    float tau = 1.0f;
    // training set
    DevBuf dsX, dsT;
    long ds_rows = 0, ds_cap = 0;
    // training scratch
    DevBuf bX, bT, bY, bZ, bH, bD, dbpart, losspart;
    // per-particle state
    size_t n = 0;
    DevBuf hist, X, Y, cub, frame, pred, chk, mask, held, flags, nflags, sub_p, sub_part, sub_acc, sub_shape, nearacc;
    int rows_grid = 0;         // resident blocks of the flagged rows' force kernel
    int head = 3, count = 0;   // ring of the last four exact samples (head newest)
    double htime[4] = {0, 0, 0, 0};   // their times, in steps
    long sub = 0;              // substeps since the state was set
    double clock = 0.0;        // time of the resident positions, in steps, since the state was set
    int substeps = 1;          // force evaluations per step of the running integrator (lf - 1)
    double dt = 0.0;           // the step the history's times are counted in
    long since_exact = 0;
    bool fallback = true;
    bool pred_net = false;     // the last extrapolation applied the network
    bool net_ok = false;       // the network passed the last check: use it until the next
    bool clear_held = true;    // held to be cleared before the next substep
    long prev_rough = -1;            // flagged rows of the previous substep when it was extrapolated, else -1
    int* nflag_h = nullptr;          // pinned host copy of the flag count
    hipEvent_t nflag_ev = nullptr;   // recorded after that copy
    bool nflag_wait = false;         // a copy not yet taken by take_flag_count()
    hipStream_t side = nullptr;      // the near field runs here, beside the main stream's work
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    int grid_age = 0;                // substeps since the near-field grid was last rebuilt
    // the last check's decision: on the device, its pinned copy, the event
    // after the copy; pending until the host has taken it (check_finish)
    DevBuf dec;
    Decision* dec_h = nullptr;
    hipEvent_t dec_ev = nullptr;
    bool dec_pending = false;
    bool dec_net = false;            // the pending check evaluated the network
    long rough_hold = 0;             // substeps still exact after the exact rows outgrew their share of N
    int rough_streak = 0;            // times in a row the exact rows outgrew it
    nb_learned_stats st{};
};

namespace {

nb_status hip_err(hipError_t e, const char* what) {
    return nbh::set_error(e == hipErrorOutOfMemory ? NB_ERR_ALLOC : NB_ERR_DEVICE,
                          std::string(what) + ": " + hipGetErrorString(e));
}

#define LH(call, what)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return hip_err(e_, what);  \
    } while (0)

Learned* get(nb_ctx* c) {
    Learned*& l = nbh::ctx_learned(c);
    if (!l) l = new (std::nothrow) Learned;
    return l;
}

// Device copies of the network live in one allocation.
nb_status net_alloc(Learned& L) {
    if (L.netmem.p) return NB_OK;
    LH(L.netmem.ensure(mlp::NET_BYTES + 1024), "allocate network");
    LH(L.params.ensure(mlp::NPARAM * sizeof(float)), "allocate parameters");
    LH(L.am.ensure(mlp::NPARAM * sizeof(float)), "allocate Adam moments");
    LH(L.av.ensure(mlp::NPARAM * sizeof(float)), "allocate Adam moments");
    LH(L.grad.ensure(mlp::NPARAM * sizeof(float)), "allocate gradient");
    LH(L.lossbuf.ensure(64), "allocate loss");
    char* p = (char*)L.netmem.p;
    auto take = [&](long bytes) {
        char* q = p;
        p += (bytes + 63) & ~63L;
        return q;
    };
    L.net.W1 = (const unsigned short*)take(128 * 16 * 2);
    L.net.W2 = (const unsigned short*)take(128 * 128 * 2);
    L.net.W3 = (const unsigned short*)take(128 * 128 * 2);
    L.net.W4 = (const unsigned short*)take(16 * 128 * 2);
    L.net.W2T = (const unsigned short*)take(128 * 128 * 2);
    L.net.W3T = (const unsigned short*)take(128 * 128 * 2);
    L.net.W4T = (const unsigned short*)take(128 * 16 * 2);
    L.net.b1 = (const float*)take(128 * 4);
    L.net.b2 = (const float*)take(128 * 4);
    L.net.b3 = (const float*)take(128 * 4);
    L.net.b4 = (const float*)take(16 * 4);
    L.net.Wtile = (const unsigned short*)take(mlp::TILES * 64 * 4 * 2);
    return NB_OK;
}

// Upload host parameters (fp32) and refresh the device copies.
nb_status net_set(Learned& L, const float* host, hipStream_t st) {
    if (nb_status s = net_alloc(L)) return s;
    LH(hipMemcpyAsync(L.params.p, host, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice, st), "upload parameters");
    LH(hipMemsetAsync(L.am.p, 0, mlp::NPARAM * sizeof(float), st), "clear Adam moments");
    LH(hipMemsetAsync(L.av.p, 0, mlp::NPARAM * sizeof(float), st), "clear Adam moments");
    LH(mlp::publish((const float*)L.params.p, L.net, st), "publish kernel");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    L.adam_t = 0;
    L.initialised = true;
    return NB_OK;
}

// Seeded initialisation: W ~ U(-1/sqrt(in), 1/sqrt(in)), b = 0; the output
// layer starts at zero so the untrained network leaves the cubic unchanged.
nb_status net_init(Learned& L, uint64_t seed, hipStream_t st) {
    std::vector<float> h(mlp::NPARAM, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](long off, int out, int in) {
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt((float)in), 1.0f / std::sqrt((float)in));
        for (long k = 0; k < (long)out * in; k++) h[off + k] = u(rng);
    };
    fill(mlp::OFF_W1, 128, 15);
    fill(mlp::OFF_W2, 128, 128);
    fill(mlp::OFF_W3, 128, 128);
    return net_set(L, h.data(), st);
}

nb_status ensure_particles(Learned& L, size_t n) {
    if (L.n == n && L.hist.p) return NB_OK;
    const size_t m = std::max<size_t>(n, 1);
    LH(L.hist.ensure(4 * m * sizeof(P4)), "allocate history");
    LH(L.X.ensure(m * 16 * sizeof(float)), "allocate features");
    LH(L.Y.ensure(m * 4 * sizeof(float)), "allocate network output");
    LH(L.cub.ensure(m * sizeof(P4)), "allocate extrapolation");
    LH(L.frame.ensure(m * 3 * sizeof(float4)), "allocate output frames");
    LH(L.pred.ensure(m * sizeof(P4)), "allocate prediction");
    LH(L.chk.ensure(((m + BLK - 1) / BLK) * 6 * sizeof(double)), "allocate check sums");
    LH(L.mask.ensure(m), "allocate roughness mask");
    LH(L.held.ensure(m), "allocate holds");
    L.clear_held = true;
    LH(L.nearacc.ensure(m * sizeof(P4)), "allocate near-field accelerations");
    LH(L.flags.ensure(m * sizeof(int)), "allocate flags");
    LH(L.nflags.ensure(sizeof(int)), "allocate flag counter");
    if (!L.nflag_h) LH(hipHostMalloc((void**)&L.nflag_h, sizeof(int), hipHostMallocDefault), "allocate pinned count");
    if (!L.nflag_ev) LH(hipEventCreateWithFlags(&L.nflag_ev, hipEventDisableTiming), "create event");
    if (!L.side) LH(hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking), "create stream");
    if (!L.ev_fork) LH(hipEventCreateWithFlags(&L.ev_fork, hipEventDisableTiming), "create event");
    if (!L.ev_join) LH(hipEventCreateWithFlags(&L.ev_join, hipEventDisableTiming), "create event");
    LH(L.dec.ensure(sizeof(Decision)), "allocate check decision");
    if (!L.dec_h) LH(hipHostMalloc((void**)&L.dec_h, sizeof(Decision), hipHostMallocDefault), "allocate pinned decision");
    if (!L.dec_ev) LH(hipEventCreateWithFlags(&L.dec_ev, hipEventDisableTiming), "create event");
    L.n = n;
    L.count = 0;
    L.head = 3;
    return NB_OK;
}

// Lagrange weights at time t of the cubic through the history samples
// (newest first, at their times), and the network weight: the nodal
// polynomial at t over its value one period after the newest sample, when the
// samples are spaced exactly one period apart (period substeps of the mean
// substep interval, 1/substeps steps; else 0: the network is not applied).
// Times are the positions' times, in steps: substeps of a higher-order
// Yoshida step are not evenly spaced.
struct Weights {
    double l[4];
    double wnet;
};

Weights weights(const Learned& L, double t) {
    Weights W{};
    double tk[4];
    for (int k = 0; k < 4; k++) tk[k] = L.htime[(L.head - k + 4) & 3];
    const double x = t;
    for (int k = 0; k < 4; k++) {
        double num = 1.0, den = 1.0;
        for (int j = 0; j < 4; j++)
            if (j != k) {
                num *= x - tk[j];
                den *= tk[k] - tk[j];
            }
        W.l[k] = num / den;
    }
    const double P = (double)L.period / (double)std::max(1, L.substeps);
    bool even = true;
    for (int k = 0; k < 3; k++) even = even && std::fabs(tk[k] - tk[k + 1] - P) <= 1e-9 * P;
    if (even) {
        double num = 1.0, den = 1.0;
        for (int k = 0; k < 4; k++) {
            num *= x - tk[k];
            den *= tk[0] + P - tk[k];
        }
        W.wnet = num / den;
    }
    return W;
}

inline int4 ring_slots(const Learned& L) {
    return make_int4(L.head, (L.head + 3) & 3, (L.head + 2) & 3, (L.head + 1) & 3);
}

// Extrapolated accelerations (remainders, with the split) of every particle
// for the current substep, in two parts so that work needing only the flags
// can run between them.  predict_features: the features, the cubic (in
// L.cub), the particles to be summed exactly (marked and listed); with
// `count`, queues the copy of how many are marked to pinned host memory
// (take_flag_count() takes it at the next substep).  predict_finish: the
// network's correction when `allow_net`, the network is on and trained and
// the history is evenly spaced (noted in L.pred_net), combined into `out`.
nb_status predict_features(nb_ctx* c, Learned& L, bool count) {
    hipStream_t st = nbh::ctx_stream(c);
    const long n = (long)L.n;
    const Weights W = weights(L, L.clock);
    LH(hipMemsetAsync(L.nflags.p, 0, sizeof(int), st), "clear flag counter");
    hipLaunchKernelGGL(k_features, blocks(n), dim3(BLK), 0, st, (const P4*)L.hist.p, ring_slots(L),
                       make_double4(W.l[0], W.l[1], W.l[2], W.l[3]), n, (const P4*)nbh::ctx_vel(c), (float*)L.X.p,
                       (P4*)L.cub.p, (float4*)L.frame.p, L.rho, (unsigned char*)L.mask.p, (int*)L.flags.p,
                       (int*)L.nflags.p, (const unsigned char*)L.held.p);
    LH(hipGetLastError(), "features kernel");
    if (count) {
        LH(hipMemcpyAsync(L.nflag_h, L.nflags.p, sizeof(int), hipMemcpyDeviceToHost, st), "download flag count");
        LH(hipEventRecord(L.nflag_ev, st), "record event");
        L.nflag_wait = true;
    }
    return NB_OK;
}

nb_status predict_finish(nb_ctx* c, Learned& L, P4* out, bool allow_net) {
    hipStream_t st = nbh::ctx_stream(c);
    const long n = (long)L.n;
    const Weights W = weights(L, L.clock);
    const bool use_net = allow_net && L.mode == 0 && L.trained && W.wnet != 0.0;
    L.pred_net = use_net;
    if (use_net) LH(mlp::forward(L.net, (const float*)L.X.p, (int)n, (float*)L.Y.p, false, nullptr, nullptr, 0, st), "network forward");
    hipLaunchKernelGGL(k_combine, blocks(n), dim3(BLK), 0, st, (const P4*)L.cub.p, (const float4*)L.frame.p,
                       (const float*)L.Y.p, (const P4*)nbh::ctx_vel(c), (float)(L.tau * W.wnet), use_net ? 1 : 0, n,
                       out);
    LH(hipGetLastError(), "combine kernel");
    return NB_OK;
}

// The flag count of the last predict_features(count = true), if not yet taken: into
// L.prev_rough and the stats.  Taken at the top of the next substep, when the
// device has long passed the copy (the work queued behind it keeps it busy).
nb_status take_flag_count(Learned& L) {
    if (!L.nflag_wait) return NB_OK;
    LH(hipEventSynchronize(L.nflag_ev), "hipEventSynchronize");
    L.nflag_wait = false;
    L.prev_rough = std::min<long>(std::max(*L.nflag_h, 0), (long)L.n);
    L.st.subset_rows += (uint64_t)L.prev_rough;
    return NB_OK;
}

// Exact sums of the flagged particles into L.sub_acc; exact_subset_scatter()
// writes them into acc.  Their count stays on the device: the kernels read
// it and size the work themselves (grids for the worst case, every row).
nb_status exact_subset_compute(nb_ctx* c, Learned& L, double eps, int kform) {
    hipStream_t st = nbh::ctx_stream(c);
    const int n = (int)L.n;
    if (L.rows_grid <= 0) {
        int dev = 0, ncu = 0;
        LH(hipGetDevice(&dev), "hipGetDevice");
        LH(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "device attribute");
        L.rows_grid = std::max(1, 4 * ncu);
    }
    LH(L.sub_p.ensure((size_t)n * sizeof(P4)), "allocate subset");
    LH(L.sub_part.ensure(nbk::accel_rows_part_records(L.rows_grid, n) * sizeof(P4)), "allocate subset partials");
    LH(L.sub_acc.ensure((size_t)n * sizeof(P4)), "allocate subset output");
    LH(L.sub_shape.ensure(sizeof(int4)), "allocate subset shape");
    hipLaunchKernelGGL(k_gather, blocks(n), dim3(BLK), 0, st, (const P4*)nbh::ctx_pm(c), (const int*)L.flags.p,
                       (const int*)L.nflags.p, n, (P4*)L.sub_p.p);
    LH(hipGetLastError(), "gather kernel");
    nbk::RowsLaunch A;
    A.form = kform;
    A.pi = (const P4*)L.sub_p.p;
    A.count = (const int*)L.nflags.p;
    A.ni_max = n;
    A.pj = (const P4*)nbh::ctx_pm(c);
    A.n = n;
    A.eps2 = eps * eps;
    A.G = 1.0;
    A.grid = L.rows_grid;
    A.part = (P4*)L.sub_part.p;
    A.shape = (int4*)L.sub_shape.p;
    A.acc = (P4*)L.sub_acc.p;
    LH(nbk::launch_accel_rows(A, st), "subset force kernels");
    return NB_OK;
}

nb_status exact_subset_scatter(nb_ctx* c, Learned& L) {
    const int n = (int)L.n;
    hipLaunchKernelGGL(k_scatter, blocks(n), dim3(BLK), 0, nbh::ctx_stream(c), (const P4*)L.sub_acc.p,
                       (const int*)L.flags.p, (const int*)L.nflags.p, n, nbh::ctx_acc(c));
    LH(hipGetLastError(), "scatter kernel");
    return NB_OK;
}

// Compare the extrapolations made for this substep (the cubic alone in
// L.cub, and with the network in L.pred when it was applied) with the exact
// acc over the particles not marked, decide on the device (k_decide) and hold
// the particles that missed (k_hold reads the decision there); the decision
// travels to pinned host memory behind an event.  Nothing waits here:
// check_finish() takes the decision when the schedule needs it.
nb_status check_enqueue(nb_ctx* c, Learned& L) {
    hipStream_t st = nbh::ctx_stream(c);
    const long n = (long)L.n;
    const long nb = (n + BLK - 1) / BLK;
    const int sets = L.pred_net ? 2 : 1;
    const P4* pset[2] = {(const P4*)L.cub.p, (const P4*)L.pred.p};
    const P4* near = L.rc > 0 ? (const P4*)L.nearacc.p : nullptr;
    for (int k = 0; k < sets; k++) {
        hipLaunchKernelGGL(k_check, dim3((unsigned)nb), dim3(BLK), 0, st, pset[k], (const P4*)nbh::ctx_acc(c), near,
                           (const unsigned char*)L.mask.p, n, (double*)L.chk.p + 3 * nb * k);
        LH(hipGetLastError(), "check kernel");
    }
    hipLaunchKernelGGL(k_decide, dim3(1), dim3(64), 0, st, (const double*)L.chk.p, nb, sets, L.threshold, L.kappa, n,
                       (Decision*)L.dec.p);
    LH(hipGetLastError(), "decision kernel");
    hipLaunchKernelGGL(k_hold, blocks(n), dim3(BLK), 0, st, (const P4*)L.cub.p, (const P4*)L.pred.p,
                       (const P4*)nbh::ctx_acc(c), near, n, (const Decision*)L.dec.p, (unsigned char*)L.held.p);
    LH(hipGetLastError(), "hold kernel");
    LH(hipMemcpyAsync(L.dec_h, L.dec.p, sizeof(Decision), hipMemcpyDeviceToHost, st), "download decision");
    LH(hipEventRecord(L.dec_ev, st), "record event");
    L.dec_pending = true;
    L.dec_net = L.pred_net;
    return NB_OK;
}

// Take the pending check's decision (waiting for its copy): the fallback
// state, whether the network may be used until the next check, the stats.
nb_status check_finish(Learned& L) {
    if (!L.dec_pending) return NB_OK;
    LH(hipEventSynchronize(L.dec_ev), "hipEventSynchronize");
    L.dec_pending = false;
    const Decision& r = *L.dec_h;
    auto rel = [](double e2, double a2) { return a2 > 0 ? std::sqrt(e2 / a2) : INFINITY; };
    const double err_c = rel(r.e2[0], r.a2[0]);
    const double err_n = L.dec_net ? rel(r.e2[1], r.a2[1]) : INFINITY;
    L.net_ok = r.net_ok != 0;
    L.st.checks++;
    if (L.dec_net) {
        L.st.net_checks++;
        L.st.net_cubic_err_sum += err_c;
        L.st.net_err_sum += err_n;
    }
    const double err = L.net_ok ? err_n : err_c;
    L.st.last_check_err = err;
    L.st.last_check_max = L.net_ok ? r.mx[1] : r.mx[0];
    L.st.last_cubic_err = err_c;
    L.st.max_check_err = std::max(L.st.max_check_err, std::isnan(err) ? INFINITY : err);
    if (!r.ok) L.st.fallbacks++;
    L.fallback = !r.ok;
    return NB_OK;
}

// Push the exact acc (minus the near field) into the history ring.
nb_status push_history(nb_ctx* c, Learned& L) {
    L.head = (L.head + 1) & 3;
    L.htime[L.head] = L.clock;
    const long n = (long)L.n;
    hipLaunchKernelGGL(k_sub, blocks(n), dim3(BLK), 0, nbh::ctx_stream(c), (const P4*)nbh::ctx_acc(c),
                       L.rc > 0 ? (const P4*)L.nearacc.p : nullptr, n, (P4*)L.hist.p + (long)L.head * n);
    LH(hipGetLastError(), "history kernel");
    L.count = std::min(L.count + 1, 4);
    return NB_OK;
}

// Near-field accelerations of the current positions (when the split is on).
nb_status near_field(nb_ctx* c, Learned& L, double eps) {
    if (!(L.rc > 0)) return NB_OK;
    LH(L.near.near_accel(nbh::ctx_pm(c), (int)L.n, L.rc, eps, (P4*)L.nearacc.p, nbh::ctx_stream(c),
                         L.exact_kform != nbk::FORM_FP64),
       "near-field kernels");
    return NB_OK;
}

// The near field of the current positions on the side stream, after all work
// queued so far on the main stream; near_field_join() makes the main stream
// wait for it.  On an exact substep it runs beside the all-pairs kernel and
// fills the CUs that kernel's waves leave idle; on an extrapolated one beside
// the extrapolation and the exact sums of the flagged particles.  The grid
// is rebuilt every min(period, kMaxGridAge) substeps (with period 4 on the
// exact ones of the regular cycle); in between the kept grid is searched
// wider by the displacement since its build (the same pairs, without the
// sort; the wider search outgrows the sort it saves after ~4 substeps at
// rc = 0.35: period 8 3.40 ms per step rebuilding every 8th, 3.28 every 4th,
// 3.30 every 2nd).
nb_status near_field_fork(nb_ctx* c, Learned& L, double eps) {
    if (!(L.rc > 0)) return NB_OK;
    const bool rebuild = L.grid_age + 1 >= std::min(L.period, kMaxGridAge);
    LH(hipEventRecord(L.ev_fork, nbh::ctx_stream(c)), "record event");
    LH(hipStreamWaitEvent(L.side, L.ev_fork, 0), "stream wait");
    LH(L.near.near_accel(nbh::ctx_pm(c), (int)L.n, L.rc, eps, (P4*)L.nearacc.p, L.side,
                         L.exact_kform != nbk::FORM_FP64, rebuild),
       "near-field kernels");
    L.grid_age = rebuild ? 0 : L.grid_age + 1;
    LH(hipEventRecord(L.ev_join, L.side), "record event");
    return NB_OK;
}

nb_status near_field_join(nb_ctx* c, Learned& L) {
    if (!(L.rc > 0)) return NB_OK;
    LH(hipStreamWaitEvent(nbh::ctx_stream(c), L.ev_join, 0), "stream wait");
    return NB_OK;
}

// Append this substep's (features, target) pairs: features from the exact
// samples one to four periods back in the collection ring (R slots, sample t
// in slot t mod R), target the exact remainder now.
nb_status record(nb_ctx* c, Learned& L, const DevBuf& ring, long R, long t) {
    hipStream_t st = nbh::ctx_stream(c);
    const long n = (long)L.n, P = L.period;
    if (L.ds_rows + n > L.ds_cap) return NB_OK;
    float* X = (float*)L.dsX.p + L.ds_rows * 16;
    float* T = (float*)L.dsT.p + L.ds_rows * 4;
    const int4 slot = make_int4((int)((t - P) % R), (int)((t - 2 * P) % R), (int)((t - 3 * P) % R), (int)((t - 4 * P) % R));
    hipLaunchKernelGGL(k_features, blocks(n), dim3(BLK), 0, st, (const P4*)ring.p, slot, make_double4(4.0, -6.0, 4.0, -1.0),
                       n, (const P4*)nbh::ctx_vel(c), X, (P4*)L.cub.p, (float4*)L.frame.p, 0.0, nullptr, nullptr,
                       nullptr, nullptr);
    LH(hipGetLastError(), "features kernel");
    hipLaunchKernelGGL(k_targets, blocks(n), dim3(BLK), 0, st, (const P4*)L.cub.p, (const float4*)L.frame.p,
                       (const P4*)nbh::ctx_acc(c), L.rc > 0 ? (const P4*)L.nearacc.p : nullptr, n, T);
    LH(hipGetLastError(), "targets kernel");
    L.ds_rows += n;
    return NB_OK;
}

struct Schedule {   // drift/kick coefficients of substep s in step t
    double kdt, c1, c2;
};

Schedule schedule(size_t t, size_t steps, size_t s, double dt, const double* cc, const double* dd, size_t lf) {
    Schedule q{dd[s] * dt, 0.0, 0.0};
    if (s + 2 < lf) {
        q.c1 = cc[s + 1] * dt;
    } else {
        q.c1 = cc[lf - 1] * dt;
        if (t + 1 < steps) q.c2 = cc[0] * dt;
    }
    return q;
}

nb_status exact_force(nb_ctx* c, int kform, double eps) {
    nbh::ForceOut out;
    LH(nbh::force_eval(c, kform, eps, out), "force kernel");
    LH(nbh::finish_accel(c, out), "reduce kernel");
    return NB_OK;
}

// The direct sum with its finish fused with the kick and drift of schedule
// q, as the exact forms step (acc is written too).  The near field, which
// reads the positions on the side stream, is joined before the finish moves
// them.
nb_status exact_force_kick(nb_ctx* c, Learned& L, int kform, double eps, const Schedule& q) {
    nbh::ForceOut out;
    LH(nbh::force_eval(c, kform, eps, out), "force kernel");
    if (nb_status e = near_field_join(c, L)) return e;
    LH(nbh::finish_kick_drift(c, out, q.kdt, q.c1, q.c2), "kick/drift kernel");
    return NB_OK;
}

nb_status kick_drift(nb_ctx* c, const Schedule& q) {
    nbh::ForceOut self;
    self.sym = false;
    self.part = nbh::ctx_acc(c);
    self.S = 1;
    LH(nbh::finish_kick_drift(c, self, q.kdt, q.c1, q.c2), "kick/drift kernel");
    return NB_OK;
}

}  // namespace

void destroy(Learned* l) {
    if (!l) return;
    if (l->side) (void)hipStreamSynchronize(l->side);   // its near-field work uses the buffers below
    l->near.release();
    if (l->nflag_ev) (void)hipEventDestroy(l->nflag_ev);
    if (l->ev_fork) (void)hipEventDestroy(l->ev_fork);
    if (l->ev_join) (void)hipEventDestroy(l->ev_join);
    if (l->side) (void)hipStreamDestroy(l->side);
    if (l->nflag_h) (void)hipHostFree(l->nflag_h);
    if (l->dec_ev) (void)hipEventDestroy(l->dec_ev);
    if (l->dec_h) (void)hipHostFree(l->dec_h);
    for (DevBuf* b : {&l->params, &l->am, &l->av, &l->grad, &l->netmem, &l->lossbuf, &l->dsX, &l->dsT, &l->bX, &l->bT,
                      &l->bY, &l->bZ, &l->bH, &l->bD, &l->dbpart, &l->losspart, &l->hist, &l->X, &l->Y, &l->cub, &l->frame,
                      &l->pred, &l->chk, &l->mask, &l->held, &l->flags, &l->nflags, &l->sub_p, &l->sub_part, &l->sub_acc,
                      &l->sub_shape, &l->nearacc, &l->dec})
        b->release();
    delete l;
}

void on_new_state(Learned* l) {
    if (!l) return;
    (void)check_finish(*l);   // a check or flag count still pending counts in the stats
    (void)take_flag_count(*l);
    l->count = 0;
    l->head = 3;
    l->since_exact = 0;
    l->sub = 0;
    l->clock = 0.0;
    l->fallback = true;
    l->pred_net = l->net_ok = false;
    l->clear_held = true;
    l->prev_rough = -1;
    l->rough_hold = 0;
    l->rough_streak = 0;
    l->grid_age = 1 << 20;   // the next near field sorts a new grid
}

nb_status step_learned(nb_ctx* c, size_t time_steps, double dt, const double* cc, const double* dd, size_t lf,
                       double eps) {
    Learned* Lp = nbh::ctx_learned(c);
    if (!Lp) return nbh::set_error(NB_ERR_STATE, "nb_step: configure the learned form first (nb_learned_configure)");
    Learned& L = *Lp;
    const size_t n = nbh::ctx_n(c);
    if (n == 0 || time_steps == 0) return NB_OK;
    if (nb_status s = ensure_particles(L, n)) return s;
    if (dt != L.dt) {   // history times are in steps of dt: a new dt starts a new history
        on_new_state(&L);
        L.dt = dt;
    }
    hipStream_t st = nbh::ctx_stream(c);
    if (L.clear_held) {
        LH(hipMemsetAsync(L.held.p, 0, n, st), "clear holds");
        L.clear_held = false;
    }
    LH(nbk::launch_drift((int)n, nbh::ctx_pm(c), nbh::ctx_vel(c), cc[0] * dt, st), "drift kernel");
    L.clock += cc[0];
    L.substeps = (int)(lf - 1);
    for (size_t t = 0; t < time_steps; t++) {
        for (size_t s = 0; s + 1 < lf; s++) {
            const Schedule q = schedule(t, time_steps, s, dt, cc, dd, lf);
            const bool full = (L.count == 4);
            const bool held_exact = L.rough_hold > 0;
            if (held_exact) L.rough_hold--;
            // a held substep with four more after it neither needs the near
            // field nor enters the history: those four replace its sample
            // before anything is extrapolated
            const bool quiet = held_exact && L.rough_hold >= 4;
            // exact on the schedule's own terms, known without the pending check
            const bool sched = !full || held_exact || (L.since_exact + 1 >= L.period);
            // this substep's near field is queued first: the GPU works on it
            // while the host takes the last check's decision and the last
            // flag count (copies the device passed earlier)
            if (!quiet) {
                if (nb_status e = near_field_fork(c, L, eps)) return e;
            }
            if (nb_status e = check_finish(L)) return e;
            if (nb_status e = take_flag_count(L)) return e;
            const long prev_rough = L.prev_rough;   // the previous substep's flagged rows, if it extrapolated
            L.prev_rough = -1;
            const bool rough = prev_rough >= 0 && (double)prev_rough > L.rough_share * (double)n;
            if (prev_rough >= 0 && !rough) L.rough_streak = 0;
            const bool exact = sched || L.fallback;
            if (exact) {
                const bool checking = full && !held_exact;
                if (checking) {
                    if (nb_status e = predict_features(c, L, false)) return e;
                    if (nb_status e = predict_finish(c, L, (P4*)L.pred.p, true)) return e;
                }
                if (quiet) L.grid_age = 1 << 20;   // the next near field sorts a new grid
                if (nb_status e = exact_force_kick(c, L, L.exact_kform, eps, q)) return e;
                L.st.exact_evals++;
                // extrapolation starts only after a check has passed, so a
                // network that produces garbage never reaches the integrator
                if (checking) {
                    if (nb_status e = check_enqueue(c, L)) return e;
                }
                if (!quiet) {
                    if (nb_status e = push_history(c, L)) return e;
                }
                L.since_exact = 0;
            } else if (rough) {
                // the previous substep had so many exact rows (each a full
                // row of N, the direct sum's two) that they cost about what
                // the direct sum does: it runs instead, here and on the
                // substeps of the next kRoughHold periods (doubled for each
                // further time in a row)
                if (nb_status e = exact_force_kick(c, L, L.exact_kform, eps, q)) return e;
                if (nb_status e = push_history(c, L)) return e;
                L.st.exact_evals++;
                L.st.rough_exact++;
                L.since_exact = 0;
                L.rough_hold = ((long)kRoughHold * L.period) << std::min(L.rough_streak, kRoughDoublings);
                L.rough_streak++;
            } else {
                // the near field runs on the side stream while the main
                // stream sums the flagged particles exactly (their count
                // stays on the device) and extrapolates the others -- the
                // exact rows first: they need only the flags, and so run
                // beside the near field rather than after the network
                if (nb_status e = predict_features(c, L, true)) return e;
                if (nb_status e = exact_subset_compute(c, L, eps, L.exact_kform)) return e;
                if (nb_status e = predict_finish(c, L, L.rc > 0 ? (P4*)L.pred.p : nbh::ctx_acc(c), L.net_ok)) return e;
                if (L.rc > 0) {
                    if (nb_status e = near_field_join(c, L)) return e;
                    hipLaunchKernelGGL(k_add, blocks((long)n), dim3(BLK), 0, nbh::ctx_stream(c), (const P4*)L.pred.p,
                                       (const P4*)L.nearacc.p, (long)n, nbh::ctx_acc(c));
                    LH(hipGetLastError(), "add kernel");
                }
                if (nb_status e = exact_subset_scatter(c, L)) return e;
                if (nb_status e = kick_drift(c, q)) return e;
                L.st.predicted_evals++;
                if (L.pred_net) L.st.net_evals++;
                L.since_exact++;
            }
            L.sub++;
            L.clock += (q.c1 + q.c2) / dt;
        }
    }
    return NB_OK;
}

// nb_near_accel and nb_near_accel_kept: the near field of the resident state
// into host planes, on a new grid (rebuild) or the kept one.
nb_status near_accel_host(nb_ctx* c, int form, double rc, double eps, double* const acc[3], bool rebuild,
                          const char* name) {
    if (!c || !acc || !acc[0] || !acc[1] || !acc[2]) return nbh::set_error(NB_ERR_ARG, std::string(name) + ": null argument");
    const int kf = nbh::kernel_form(form);
    if (kf < 0) return nbh::set_error(NB_ERR_ARG, std::string(name) + ": form must be FP64 or a mixed form");
    if (!(rc > 0.0) || !std::isfinite(rc) || !(eps >= 0.0))
        return nbh::set_error(NB_ERR_ARG, std::string(name) + ": need rc > 0, eps >= 0");
    const size_t n = nbh::ctx_n(c);
    if (n == 0) return NB_OK;
    Learned* L = get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, std::string(name) + ": out of host memory");
    hipStream_t st = nbh::ctx_stream(c);
    nbh::DevBuf out;   // released on every path below
    std::vector<P4> h(n);
    const char* what = "allocate output";
    hipError_t e = out.ensure(n * sizeof(P4));
    if (e == hipSuccess) {
        what = "near-field kernels";
        e = L->near.near_accel(nbh::ctx_pm(c), (int)n, rc, eps, (P4*)out.p, st, kf != nbk::FORM_FP64, rebuild);
    }
    if (e == hipSuccess) {
        what = "download";
        e = hipMemcpyAsync(h.data(), out.p, n * sizeof(P4), hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) {
        what = "hipStreamSynchronize";
        e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess) (void)hipStreamSynchronize(st);   // nothing queued may still use the buffer
    out.release();
    if (e != hipSuccess) return hip_err(e, what);
    for (size_t i = 0; i < n; i++) {
        acc[0][i] = h[i].x;
        acc[1][i] = h[i].y;
        acc[2][i] = h[i].z;
    }
    return NB_OK;
}

}  // namespace nbl

using nbl::Learned;
using nbl::hip_err;

extern "C" {

nb_status nb_learned_configure(nb_ctx* c, int period, double threshold, int exact_form) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_configure: null context");
    if (period < 2) return nbh::set_error(NB_ERR_ARG, "nb_learned_configure: period must be >= 2");
    if (!(threshold >= 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_learned_configure: threshold must be >= 0");
    const int kf = nbh::kernel_form(exact_form);
    if (kf < 0) return nbh::set_error(NB_ERR_ARG, "nb_learned_configure: exact_form must be FP64 or a mixed form");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_configure: out of host memory");
    nbl::on_new_state(L);   // takes a pending check first, so the stats below start from zero
    L->period = period;
    L->threshold = threshold;
    L->exact_kform = kf;
    L->st = nb_learned_stats{};
    return NB_OK;
}

nb_status nb_learned_set_flag_ratio(nb_ctx* c, double rho) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_flag_ratio: null context");
    if (!(rho >= 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_flag_ratio: ratio must be >= 0");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_set_flag_ratio: out of host memory");
    L->rho = rho;
    return NB_OK;
}

nb_status nb_learned_set_rough_share(nb_ctx* c, double share) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_rough_share: null context");
    if (!(share >= 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_rough_share: share must be >= 0");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_set_rough_share: out of host memory");
    L->rough_share = share;
    return NB_OK;
}

nb_status nb_learned_set_hold_ratio(nb_ctx* c, double kappa) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_hold_ratio: null context");
    if (!(kappa > 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_hold_ratio: ratio must be > 0 (inf: never hold)");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_set_hold_ratio: out of host memory");
    L->kappa = kappa;
    return NB_OK;
}

nb_status nb_learned_set_near_radius(nb_ctx* c, double rc) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_near_radius: null context");
    if (!(rc >= 0.0) || !std::isfinite(rc)) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_near_radius: radius must be finite and >= 0");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_set_near_radius: out of host memory");
    L->rc = rc;
    nbl::on_new_state(L);
    return NB_OK;
}

nb_status nb_near_accel(nb_ctx* c, int form, double rc, double eps, double* const acc[3]) {
    return nbl::near_accel_host(c, form, rc, eps, acc, true, "nb_near_accel");
}

nb_status nb_near_accel_kept(nb_ctx* c, int form, double rc, double eps, double* const acc[3]) {
    return nbl::near_accel_host(c, form, rc, eps, acc, false, "nb_near_accel_kept");
}

nb_status nb_learned_set_mode(nb_ctx* c, int mode) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_mode: null context");
    if (mode != 0 && mode != 1) return nbh::set_error(NB_ERR_ARG, "nb_learned_set_mode: mode must be 0 or 1");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_set_mode: out of host memory");
    L->mode = mode;
    return NB_OK;
}

nb_status nb_learned_collect(nb_ctx* c, size_t time_steps, double dt, const double* cc, const double* dd, size_t lf,
                             int form, double eps) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_collect: null context");
    if (!cc || !dd || lf < 2 || lf > 16 || !std::isfinite(dt))
        return nbh::set_error(NB_ERR_ARG, "nb_learned_collect: bad coefficients or dt");
    if (!(eps >= 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_learned_collect: eps must be >= 0");
    const int kf = nbh::kernel_form(form);
    if (kf < 0) return nbh::set_error(NB_ERR_ARG, "nb_learned_collect: form must be FP64 or a mixed form");
    Learned* Lp = nbl::get(c);
    if (!Lp) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_collect: out of host memory");
    Learned& L = *Lp;
    const size_t n = nbh::ctx_n(c);
    if (n == 0 || time_steps == 0) return NB_OK;
    if (nb_status s = nbl::ensure_particles(L, n)) return s;
    nbl::on_new_state(&L);
    const long P = L.period, R = 4 * P;
    if (R > 64) return nbh::set_error(NB_ERR_ARG, "nb_learned_collect: period must be <= 16");
    const long subs = (long)time_steps * (long)(lf - 1);
    const long want = L.ds_rows + std::max(0L, subs - R) * (long)n;
    if (want > L.ds_cap) {   // grow, keeping what was collected before
        DevBuf x2, t2;
        LH(x2.ensure(want * 16 * sizeof(float)), "allocate training features");
        LH(t2.ensure(want * 4 * sizeof(float)), "allocate training targets");
        hipStream_t st = nbh::ctx_stream(c);
        if (L.ds_rows) {
            LH(hipMemcpyAsync(x2.p, L.dsX.p, L.ds_rows * 16 * sizeof(float), hipMemcpyDeviceToDevice, st), "copy");
            LH(hipMemcpyAsync(t2.p, L.dsT.p, L.ds_rows * 4 * sizeof(float), hipMemcpyDeviceToDevice, st), "copy");
            LH(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
        L.dsX.release();
        L.dsT.release();
        L.dsX = x2;
        L.dsT = t2;
        x2.p = t2.p = nullptr;
        L.ds_cap = want;
    }
    hipStream_t st = nbh::ctx_stream(c);
    DevBuf ring;   // the last R exact remainders, sample t in slot t mod R
    LH(ring.ensure((size_t)R * n * sizeof(P4)), "allocate collection ring");
    LH(nbk::launch_drift((int)n, nbh::ctx_pm(c), nbh::ctx_vel(c), cc[0] * dt, st), "drift kernel");
    long sub = 0;
    for (size_t t = 0; t < time_steps; t++)
        for (size_t s = 0; s + 1 < lf; s++, sub++) {
            if (nb_status e = nbl::exact_force(c, kf, eps)) return e;
            if (nb_status e = nbl::near_field(c, L, eps)) return e;
            if (sub >= R)
                if (nb_status e = nbl::record(c, L, ring, R, sub)) return e;
            hipLaunchKernelGGL(nbl::k_sub, nbl::blocks((long)n), dim3(nbl::BLK), 0, st, (const P4*)nbh::ctx_acc(c),
                               L.rc > 0 ? (const P4*)L.nearacc.p : nullptr, (long)n,
                               (P4*)ring.p + (sub % R) * (long)n);
            LH(hipGetLastError(), "history kernel");
            if (nb_status e = nbl::kick_drift(c, nbl::schedule(t, time_steps, s, dt, cc, dd, lf))) return e;
        }
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    ring.release();
    nbl::on_new_state(&L);
    return NB_OK;
}

nb_status nb_learned_train(nb_ctx* c, int iters, int batch, double lr, uint64_t seed, double* final_loss) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_train: null context");
    if (iters < 0 || batch < 16 || batch > (1 << 20) || !(lr > 0.0))
        return nbh::set_error(NB_ERR_ARG, "nb_learned_train: need iters >= 0, 16 <= batch <= 2^20, lr > 0");
    Learned* Lp = nbh::ctx_learned(c);
    if (!Lp || Lp->ds_rows == 0) return nbh::set_error(NB_ERR_STATE, "nb_learned_train: no training data collected");
    Learned& L = *Lp;
    hipStream_t st = nbh::ctx_stream(c);
    if (!L.initialised)
        if (nb_status s = nbl::net_init(L, seed, st)) return s;
    const long b = ((long)batch + 15) & ~15L;
    LH(L.bX.ensure(b * 16 * sizeof(float)), "allocate batch");
    LH(L.bT.ensure(b * 4 * sizeof(float)), "allocate batch");
    LH(L.bY.ensure(b * 4 * sizeof(float)), "allocate batch");
    LH(L.bZ.ensure(3 * b * 128 * sizeof(float)), "allocate activations");
    LH(L.bH.ensure((16 + 3 * 128) * b * sizeof(unsigned short)), "allocate activations");
    LH(L.bD.ensure((16 + 3 * 128) * b * sizeof(unsigned short)), "allocate deltas");
    LH(L.dbpart.ensure((b / 16) * (16 + 3 * 128) * sizeof(float)), "allocate bias partials");
    LH(L.losspart.ensure((b / 16) * sizeof(float)), "allocate loss partials");
    // target scale: RMS of the collected residuals
    {
        const int nbk_ = 256;
        nbh::DevBuf ss;   // released on every path below
        std::vector<double> h(nbk_);
        const char* what = "allocate";
        hipError_t e = ss.ensure(nbk_ * sizeof(double));
        if (e == hipSuccess) {
            what = "sum-of-squares kernel";
            hipLaunchKernelGGL(nbl::k_sumsq, dim3(nbk_), dim3(nbl::BLK), 0, st, (const float*)L.dsT.p, L.ds_rows,
                               (double*)ss.p);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            what = "download";
            e = hipMemcpyAsync(h.data(), ss.p, nbk_ * sizeof(double), hipMemcpyDeviceToHost, st);
        }
        if (e == hipSuccess) {
            what = "hipStreamSynchronize";
            e = hipStreamSynchronize(st);
        }
        if (e != hipSuccess) (void)hipStreamSynchronize(st);   // nothing queued may still use the buffer
        ss.release();
        if (e != hipSuccess) return hip_err(e, what);
        double s = 0.0;
        for (double v : h) s += v;
        const double rms = std::sqrt(s / (3.0 * (double)L.ds_rows));
        L.tau = (rms > 0 && std::isfinite(rms)) ? (float)rms : 1.0f;
    }
    const float clip = 8.0f;
    for (int it = 0; it < iters; it++) {
        hipLaunchKernelGGL(nbl::k_batch, dim3((unsigned)((b + nbl::BLK - 1) / nbl::BLK)), dim3(nbl::BLK), 0, st,
                           (const float*)L.dsX.p, (const float*)L.dsT.p, L.ds_rows, (unsigned long long)seed, it,
                           (int)b, 1.0f / L.tau, clip, (float*)L.bX.p, (float*)L.bT.p);
        LH(hipGetLastError(), "batch kernel");
        LH(mlp::forward(L.net, (const float*)L.bX.p, (int)b, (float*)L.bY.p, true, (float*)L.bZ.p,
                        (unsigned short*)L.bH.p, b, st), "network forward");
        LH(mlp::backward(L.net, (const float*)L.bY.p, (const float*)L.bT.p, (int)b, (const float*)L.bZ.p,
                         (const unsigned short*)L.bH.p, (unsigned short*)L.bD.p, b, (float*)L.dbpart.p,
                         (float*)L.losspart.p, (float*)L.grad.p, (float*)L.lossbuf.p, st), "network backward");
        LH(mlp::adam((float*)L.params.p, (const float*)L.grad.p, (float*)L.am.p, (float*)L.av.p, (int)mlp::NPARAM,
                     (float)lr, ++L.adam_t, st), "Adam kernel");
        LH(mlp::publish((const float*)L.params.p, L.net, st), "publish kernel");
    }
    float loss = 0.f;
    LH(hipMemcpyAsync(&loss, L.lossbuf.p, sizeof(float), hipMemcpyDeviceToHost, st), "download loss");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (final_loss) *final_loss = 0.5 * (double)loss / (double)b;   // last minibatch, before its update
    L.trained = true;
    return NB_OK;
}

nb_status nb_learned_reset(nb_ctx* c, uint64_t seed) {
    if (!c) return nbh::set_error(NB_ERR_ARG, "nb_learned_reset: null context");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_learned_reset: out of host memory");
    if (nb_status s = nbl::net_init(*L, seed, nbh::ctx_stream(c))) return s;
    L->trained = false;
    L->ds_rows = 0;
    L->tau = 1.0f;
    nbl::on_new_state(L);
    return NB_OK;
}

nb_status nb_learned_stats_read(nb_ctx* c, nb_learned_stats* out) {
    if (!c || !out) return nbh::set_error(NB_ERR_ARG, "nb_learned_stats_read: null argument");
    Learned* L = nbh::ctx_learned(c);
    if (L) {   // a check or flag count still pending counts
        if (nb_status e = nbl::check_finish(*L)) return e;
        if (nb_status e = nbl::take_flag_count(*L)) return e;
    }
    *out = L ? L->st : nb_learned_stats{};
    return NB_OK;
}

nb_status nb_mlp_param_count(size_t* count) {
    if (!count) return nbh::set_error(NB_ERR_ARG, "nb_mlp_param_count: null argument");
    *count = (size_t)mlp::NPARAM;
    return NB_OK;
}

nb_status nb_mlp_get_params(nb_ctx* c, float* params) {
    if (!c || !params) return nbh::set_error(NB_ERR_ARG, "nb_mlp_get_params: null argument");
    Learned* L = nbh::ctx_learned(c);
    if (!L || !L->initialised) return nbh::set_error(NB_ERR_STATE, "nb_mlp_get_params: no network yet");
    hipStream_t st = nbh::ctx_stream(c);
    LH(hipMemcpyAsync(params, L->params.p, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, st), "download parameters");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    return NB_OK;
}

nb_status nb_mlp_set_params(nb_ctx* c, const float* params) {
    if (!c || !params) return nbh::set_error(NB_ERR_ARG, "nb_mlp_set_params: null argument");
    Learned* L = nbl::get(c);
    if (!L) return nbh::set_error(NB_ERR_ALLOC, "nb_mlp_set_params: out of host memory");
    if (nb_status s = nbl::net_set(*L, params, nbh::ctx_stream(c))) return s;
    L->trained = true;
    return NB_OK;
}

static nb_status mlp_prepare(nb_ctx* c, size_t batch, const float* x, Learned*& L, long& b, const char* fn) {
    if (!c || !x || batch == 0 || batch > (1u << 22)) return nbh::set_error(NB_ERR_ARG, std::string(fn) + ": bad arguments");
    L = nbh::ctx_learned(c);
    if (!L || !L->initialised) return nbh::set_error(NB_ERR_STATE, std::string(fn) + ": set or train the network first");
    b = ((long)batch + 15) & ~15L;
    hipStream_t st = nbh::ctx_stream(c);
    LH(L->bX.ensure(b * 16 * sizeof(float)), "allocate batch");
    LH(L->bT.ensure(b * 4 * sizeof(float)), "allocate batch");
    LH(L->bY.ensure(b * 4 * sizeof(float)), "allocate batch");
    std::vector<float> xp((size_t)batch * 16, 0.f);
    for (size_t r = 0; r < batch; r++) std::memcpy(&xp[r * 16], x + r * 15, 15 * sizeof(float));
    LH(hipMemcpyAsync(L->bX.p, xp.data(), xp.size() * sizeof(float), hipMemcpyHostToDevice, st), "upload inputs");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    return NB_OK;
}

nb_status nb_mlp_forward(nb_ctx* c, size_t batch, const float* x, float* y) {
    Learned* L;
    long b;
    if (!y) return nbh::set_error(NB_ERR_ARG, "nb_mlp_forward: null output");
    if (nb_status s = mlp_prepare(c, batch, x, L, b, "nb_mlp_forward")) return s;
    hipStream_t st = nbh::ctx_stream(c);
    LH(mlp::forward(L->net, (const float*)L->bX.p, (int)batch, (float*)L->bY.p, false, nullptr, nullptr, 0, st), "network forward");
    std::vector<float> yp((size_t)batch * 4);
    LH(hipMemcpyAsync(yp.data(), L->bY.p, yp.size() * sizeof(float), hipMemcpyDeviceToHost, st), "download outputs");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (size_t r = 0; r < batch; r++) std::memcpy(y + r * 3, &yp[r * 4], 3 * sizeof(float));
    return NB_OK;
}

nb_status nb_mlp_grad(nb_ctx* c, size_t batch, const float* x, const float* target, float* grads, double* loss) {
    Learned* L;
    long b;
    if (!target || !grads) return nbh::set_error(NB_ERR_ARG, "nb_mlp_grad: null argument");
    if (nb_status s = mlp_prepare(c, batch, x, L, b, "nb_mlp_grad")) return s;
    hipStream_t st = nbh::ctx_stream(c);
    LH(L->bZ.ensure(3 * b * 128 * sizeof(float)), "allocate activations");
    LH(L->bH.ensure((16 + 3 * 128) * b * sizeof(unsigned short)), "allocate activations");
    LH(L->bD.ensure((16 + 3 * 128) * b * sizeof(unsigned short)), "allocate deltas");
    LH(L->dbpart.ensure((b / 16) * (16 + 3 * 128) * sizeof(float)), "allocate bias partials");
    LH(L->losspart.ensure((b / 16) * sizeof(float)), "allocate loss partials");
    std::vector<float> tp((size_t)b * 4, 0.f);
    for (size_t r = 0; r < batch; r++) std::memcpy(&tp[r * 4], target + r * 3, 3 * sizeof(float));
    LH(hipMemcpyAsync(L->bT.p, tp.data(), tp.size() * sizeof(float), hipMemcpyHostToDevice, st), "upload targets");
    LH(hipMemsetAsync(L->grad.p, 0, mlp::NPARAM * sizeof(float), st), "clear gradient");
    LH(mlp::forward(L->net, (const float*)L->bX.p, (int)batch, (float*)L->bY.p, true, (float*)L->bZ.p,
                    (unsigned short*)L->bH.p, b, st), "network forward");
    LH(mlp::backward(L->net, (const float*)L->bY.p, (const float*)L->bT.p, (int)batch, (const float*)L->bZ.p,
                     (const unsigned short*)L->bH.p, (unsigned short*)L->bD.p, b, (float*)L->dbpart.p,
                     (float*)L->losspart.p, (float*)L->grad.p, (float*)L->lossbuf.p, st), "network backward");
    float lsum = 0.f;
    LH(hipMemcpyAsync(grads, L->grad.p, mlp::NPARAM * sizeof(float), hipMemcpyDeviceToHost, st), "download gradient");
    LH(hipMemcpyAsync(&lsum, L->lossbuf.p, sizeof(float), hipMemcpyDeviceToHost, st), "download loss");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (loss) *loss = 0.5 * (double)lsum / (double)batch;
    return NB_OK;
}

nb_status nb_mlp_adam_step(nb_ctx* c, const float* grads, double lr) {
    if (!c || !grads || !(lr > 0.0)) return nbh::set_error(NB_ERR_ARG, "nb_mlp_adam_step: bad arguments");
    Learned* L = nbh::ctx_learned(c);
    if (!L || !L->initialised) return nbh::set_error(NB_ERR_STATE, "nb_mlp_adam_step: set or train the network first");
    hipStream_t st = nbh::ctx_stream(c);
    LH(hipMemcpyAsync(L->grad.p, grads, mlp::NPARAM * sizeof(float), hipMemcpyHostToDevice, st), "upload gradient");
    LH(mlp::adam((float*)L->params.p, (const float*)L->grad.p, (float*)L->am.p, (float*)L->av.p, (int)mlp::NPARAM,
                 (float)lr, ++L->adam_t, st), "Adam kernel");
    LH(mlp::publish((const float*)L->params.p, L->net, st), "publish kernel");
    LH(hipStreamSynchronize(st), "hipStreamSynchronize");
    return NB_OK;
}

}  // extern "C"
