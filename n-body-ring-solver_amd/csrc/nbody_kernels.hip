// nbody_kernels.hip -- gfx950 kernels for the softened all-pairs force,
// the fused leapfrog kick/drift and the energy diagnostic.
//
// Data layout in HBM: one 32-byte record per particle, (x, y, z, m) for
// position+mass (the FPGA particle_t = double4, FPGA/host/solver.cpp:721),
// (vx, vy, vz, 0) for velocity and (ax, ay, az, 0) for acceleration.
//
// Force kernel scheme: a 256-thread workgroup owns 256*IPT i-particles held in
// registers (IPT per lane).  It streams its j-range through LDS in tiles of
// 256 records (one coalesced 32-byte load per lane, double-buffered, one
// barrier per tile); every lane of a wave then reads the same LDS record
// (broadcast, no bank conflicts) and evaluates IPT pairs.  For small N the
// j-range is split over blockIdx.y so the grid still fills 256 CUs; the
// partial sums are added in a fixed order by the kick/drift or reduce kernel,
// so results are bitwise reproducible.
//
// Pair term (reference CPU/force_calculation.cpp:78-82, softened):
//   a_i += m_j (x_j - x_i) / (|x_j - x_i|^2 + eps^2)^(3/2)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "nbody_kernels.h"

namespace nbk {

// LDS record of the mixed form: FP64 position (differences stay FP64), FP32 mass.
struct __attribute__((aligned(32))) P4m {
    double x, y, z;
    float m, pad;
};

__device__ __forceinline__ P4 load_p4(const P4* __restrict__ p, int j, int jend) {
    if (j < jend) return p[j];
    return P4{0.0, 0.0, 0.0, 0.0};
}

// ---------------------------------------------------------------------------
// Pair kernels
// ---------------------------------------------------------------------------

// FP64: v_rsq_f64 gives y ~ r2^(-1/2) to about 2^-22 relative; with
// e = 1 - r2 y^2 the exact y_true^3 = y^3 (1-e)^(-3/2) = y^3 (1 + 3/2 e + 15/8 e^2 + O(e^3)),
// so two FMAs restore full FP64 accuracy (|e|^3 ~ 2^-66): the polynomial is
// evaluated as (s + cshift)^2 + coff, s = r2 y^2 (kCubicShift, kCubicOffset),
// and mj carries its factor 15/8 (staged with kCubicMass).  The constants
// arrive as kernel arguments so they stay in registers (as literals the
// compiler rebuilds them with two moves per pair).
template <bool EPS0, bool MUL = true>
__device__ __forceinline__ void pair_fp64(double xi, double yi, double zi, double xj, double yj,
                                          double zj, double mj, double eps2, double cshift, double coff,
                                          double& ax, double& ay, double& az) {
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const double y = __builtin_amdgcn_rsq(r2);
    const double t = y * y;
    const double u = __builtin_fma(r2, t, cshift);
    const double y3 = t * y;
    const double p = __builtin_fma(u, u, coff);
    double s = (MUL ? mj * y3 : y3) * p;   // MUL false: every mass equal, multiplied in afterwards
    if (EPS0) s = (r2 > 0.0) ? s : 0.0;   // zero separation (self pair) contributes 0
    ax = __builtin_fma(s, dx, ax);
    ay = __builtin_fma(s, dy, ay);
    az = __builtin_fma(s, dz, az);
}

// Mixed: FP64 differences rounded to FP32, FP32 rsqrt (1 ulp) plus one
// Newton-Raphson step applied to the cube, y^3 (1 + 3/2 (1 - r2 y^2))
// = c y^3 (F - r2 y^2) (F = kNewtonOffset; mj carries c = kNewtonMass),
// FP32 pair arithmetic.
template <bool EPS0, bool NR, bool MUL = true>
__device__ __forceinline__ void pair_mixed(double xi, double yi, double zi, double xj, double yj,
                                           double zj, float mj, float eps2, float& ax, float& ay,
                                           float& az) {
    const float dx = (float)(xj - xi), dy = (float)(yj - yi), dz = (float)(zj - zi);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    const float y = __builtin_amdgcn_rsqf(r2);
    const float y2 = y * y;
    float s = (MUL ? mj * y : y) * y2;
    if (NR) s = s * __builtin_fmaf(-r2, y2, kNewtonOffset);
    if (EPS0) s = (r2 > 0.0f) ? s : 0.0f;
    ax = __builtin_fmaf(s, dx, ax);
    ay = __builtin_fmaf(s, dy, ay);
    az = __builtin_fmaf(s, dz, az);
}

// FP32 throughout (FORM_FP32): FP32 positions and differences, the mixed
// form's rsqrt and Newton step, FP32 sums.
template <bool MUL = true>
__device__ __forceinline__ void pair_fp32(float xi, float yi, float zi, float xj, float yj, float zj, float mj,
                                          float eps2, bool eps0, float& ax, float& ay, float& az) {
    const float dx = xj - xi, dy = yj - yi, dz = zj - zi;
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    const float y = __builtin_amdgcn_rsqf(r2);
    const float y2 = y * y;
    float s = (MUL ? mj * y : y) * y2 * __builtin_fmaf(-r2, y2, kNewtonOffset);
    if (eps0) s = (r2 > 0.0f) ? s : 0.0f;
    ax = __builtin_fmaf(s, dx, ax);
    ay = __builtin_fmaf(s, dy, ay);
    az = __builtin_fmaf(s, dz, az);
}

// LDS record of the FP32 form.
struct __attribute__((aligned(16))) P4f {
    float x, y, z, m;
};

// ---------------------------------------------------------------------------
// Force kernel (asymmetric: all ordered pairs of i-set x j-range)
// ---------------------------------------------------------------------------
// One work item: rows [ibase, ibase + BLOCK IPT) of pi (those below ni)
// against pj[jbeg, jend), written to o[i].  Every thread of the block takes
// part (the j-tiles are staged through LDS with barriers).  eq > 0 (every j
// mass equals eq): a staged mass is 1 for a particle and 0 for padding, the
// full tiles take the pair terms without it, and eq times the form's mass
// factor joins G in the final scale.
template <int FORM, bool EPS0, int IPT>
__device__ __forceinline__ void accel_item(const P4* __restrict__ pi, int ni, const P4* __restrict__ pj, int jbeg,
                                           int jend, double eps2, double G, double cshift, double coff,
                                           P4* __restrict__ o, int ibase, double eq = 0.0) {
    constexpr bool F32 = (FORM == FORM_FP32);
    constexpr bool MIXED = (FORM != FORM_FP64);
    constexpr bool NR = (FORM == FORM_MIXED_NR) || F32;
    // the factor of the correction polynomial carried by the staged masses
    constexpr double MSCALE = MIXED ? (NR ? kNewtonMass : 1.0) : kCubicMass;
    const bool eqm = eq > 0.0;
    G = eqm ? G * (eq * MSCALE) : G;
    using Rec = typename std::conditional<F32, P4f, typename std::conditional<MIXED, P4m, P4>::type>::type;
    __shared__ Rec tile[2][BLOCK];

    const int t = threadIdx.x;

    double xi[IPT], yi[IPT], zi[IPT];
    double ax[IPT], ay[IPT], az[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int i = ibase + k * BLOCK + t;
        const P4 p = (i < ni) ? pi[i] : P4{0.0, 0.0, 0.0, 0.0};
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z;
        ax[k] = ay[k] = az[k] = 0.0;
    }

    const int ntiles = (jend > jbeg) ? (jend - jbeg + BLOCK - 1) / BLOCK : 0;
    P4 nxt = load_p4(pj, jbeg + t, jend);
    const float eps2f = (float)eps2;
    // the FP32 form: positions rounded once, sums over the whole j-range in FP32
    float xs[IPT], ys[IPT], zs[IPT], gx[IPT], gy[IPT], gz[IPT];
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        xs[k] = (float)xi[k]; ys[k] = (float)yi[k]; zs[k] = (float)zi[k];
        gx[k] = gy[k] = gz[k] = 0.0f;
    }

    // the pair loop over one staged tile: MUL false leaves the masses out
    auto sweep = [&](const Rec* tl, int cnt8, auto mul) {
        constexpr bool MUL = decltype(mul)::value;
        if constexpr (F32) {
            for (int jj = 0; jj < cnt8; jj += 8) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const P4f q = tl[jj + u];
#pragma unroll
                    for (int k = 0; k < IPT; k++)
                        pair_fp32<MUL>(xs[k], ys[k], zs[k], q.x, q.y, q.z, q.m, eps2f, EPS0, gx[k], gy[k], gz[k]);
                }
            }
        } else if constexpr (MIXED) {
            float fx[IPT], fy[IPT], fz[IPT];
#pragma unroll
            for (int k = 0; k < IPT; k++) fx[k] = fy[k] = fz[k] = 0.0f;
            for (int jj = 0; jj < cnt8; jj += 8) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const P4m q = tl[jj + u];
#pragma unroll
                    for (int k = 0; k < IPT; k++)
                        pair_mixed<EPS0, NR, MUL>(xi[k], yi[k], zi[k], q.x, q.y, q.z, q.m, eps2f,
                                                  fx[k], fy[k], fz[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < IPT; k++) {   // flush the FP32 tile sums into FP64
                ax[k] += (double)fx[k];
                ay[k] += (double)fy[k];
                az[k] += (double)fz[k];
            }
        } else {
            for (int jj = 0; jj < cnt8; jj += 8) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const P4 q = tl[jj + u];
#pragma unroll
                    for (int k = 0; k < IPT; k++)
                        pair_fp64<EPS0, MUL>(xi[k], yi[k], zi[k], q.x, q.y, q.z, q.m, eps2, cshift, coff,
                                             ax[k], ay[k], az[k]);
                }
            }
        }
    };

    for (int tt = 0; tt < ntiles; tt++) {
        const int buf = tt & 1;
        const double mj = eqm ? (nxt.m != 0.0 ? 1.0 : 0.0) : nxt.m * MSCALE;
        if constexpr (F32) {
            tile[buf][t] = P4f{(float)nxt.x, (float)nxt.y, (float)nxt.z, (float)mj};
        } else if constexpr (MIXED) {
            tile[buf][t] = P4m{nxt.x, nxt.y, nxt.z, (float)mj, 0.0f};
        } else {
            tile[buf][t] = P4{nxt.x, nxt.y, nxt.z, mj};
        }
        __syncthreads();
        const int j0 = jbeg + tt * BLOCK;
        if (tt + 1 < ntiles) nxt = load_p4(pj, j0 + BLOCK + t, jend);
        const int cnt = min(BLOCK, jend - j0);
        const int cnt8 = (cnt + 7) & ~7;   // entries past jend are zero-mass records
        if (eqm && cnt == BLOCK) sweep(tile[buf], cnt8, std::integral_constant<bool, false>{});
        else sweep(tile[buf], cnt8, std::integral_constant<bool, true>{});
    }

    if constexpr (F32) {
#pragma unroll
        for (int k = 0; k < IPT; k++) {
            ax[k] = (double)gx[k]; ay[k] = (double)gy[k]; az[k] = (double)gz[k];
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; k++) {
        const int i = ibase + k * BLOCK + t;
        if (i < ni) o[i] = P4{G * ax[k], G * ay[k], G * az[k], 0.0};
    }
}

// eq_dev (if set) says on the device whether every j mass is equal
// (k_mass_check); else eq_host does (0: not known equal).
template <int FORM, bool EPS0, int IPT>
__global__ __launch_bounds__(BLOCK) void k_accel(const P4* __restrict__ pi, int ni,
                                                 const P4* __restrict__ pj, int jbeg_all,
                                                 int jend_all, int jchunk, double eps2, double G,
                                                 double cshift, double coff, P4* __restrict__ out,
                                                 long out_stride, int split_base, double eq_host,
                                                 const EqState* __restrict__ eq_dev) {
    const int split = blockIdx.y;
    const int jbeg = jbeg_all + split * jchunk;
    const int jend = min(jend_all, jbeg + jchunk);
    const double eq = eq_dev ? (eq_dev->ok ? eq_dev->m0 : 0.0) : eq_host;
    accel_item<FORM, EPS0, IPT>(pi, ni, pj, jbeg, jend, eps2, G, cshift, coff,
                                out + (long)(split_base + split) * out_stride, blockIdx.x * BLOCK * IPT, eq);
}

// Whether every mass of pj[0, nj) equals the first (> 0, finite): the
// first kernel takes it as the candidate, the second clears `ok` at any
// other mass (plain stores of one value; no atomics needed).
__global__ void k_mass_check_init(const P4* __restrict__ pj, EqState* __restrict__ st) {
    const double m0 = pj[0].m;
    st->m0 = m0;
    st->ok = (m0 > 0.0 && m0 < INFINITY) ? 1 : 0;
}

__global__ __launch_bounds__(BLOCK) void k_mass_check(const P4* __restrict__ pj, int nj, EqState* __restrict__ st) {
    const double m0 = st->m0;
    bool same = true;
    for (int j = blockIdx.x * BLOCK + threadIdx.x; j < nj; j += gridDim.x * BLOCK) same = same && pj[j].m == m0;
    if (!same) st->ok = 0;
}

hipError_t launch_mass_check(const P4* pj, int nj, EqState* st, hipStream_t stream) {
    if (nj <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mass_check_init, dim3(1), dim3(1), 0, stream, pj, st);
    hipLaunchKernelGGL(k_mass_check, dim3(std::min(1024, (nj + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, stream, pj, nj,
                       st);
    return hipGetLastError();
}

// Rows [0, m) of pi, m = min(*count, ni_max) read on the device, against all
// of pj[0, n): a resident grid takes the items (row block, j-split) in turn.
// The split count adapts to m -- enough splits that the items fill the grid,
// each split a multiple of 64 particles and at least 256 -- and is written
// with m to *shape (x = m, y = splits) for the reduction; partial sums go to
// part[split * m + row].
template <int FORM, bool EPS0>
__global__ __launch_bounds__(BLOCK) void k_accel_rows(const P4* __restrict__ pi, const int* __restrict__ count,
                                                      int ni_max, const P4* __restrict__ pj, int n, double eps2,
                                                      double G, double cshift, double coff, P4* __restrict__ part,
                                                      int4* __restrict__ shape) {
    const int m = min(max(*count, 0), ni_max);
    const int ng = (m + BLOCK - 1) / BLOCK;
    const int smax = max(1, n / 256);
    const int seff = ng > 0 ? min(max((int)gridDim.x / ng, 1), smax) : 1;
    const int jchunk = ((n + seff - 1) / seff + 63) & ~63;
    const int splits = (n + jchunk - 1) / jchunk;
    if (blockIdx.x == 0 && threadIdx.x == 0) *shape = make_int4(m, splits, 0, 0);
    for (int it = blockIdx.x; it < ng * splits; it += gridDim.x) {   // uniform per block
        const int sp = it % splits;
        const int jbeg = sp * jchunk;
        accel_item<FORM, EPS0, 1>(pi, m, pj, jbeg, min(n, jbeg + jchunk), eps2, G, cshift, coff,
                                  part + (long)sp * m, (it / splits) * BLOCK);
        __syncthreads();   // the next item restages the LDS tiles
    }
}

// ---------------------------------------------------------------------------
// Partial-sum reduction, kick and drift
// ---------------------------------------------------------------------------

// acc_i = sum_s part[s][i] in order s = 0..S-1.
__device__ __forceinline__ void sum_parts(const P4* __restrict__ part, int S, long stride, int i,
                                          double& ax, double& ay, double& az) {
    ax = ay = az = 0.0;
    for (int s = 0; s < S; s++) {
        const P4 p = part[(long)s * stride + i];
        ax += p.x; ay += p.y; az += p.z;
    }
}

__global__ __launch_bounds__(BLOCK) void k_reduce(const P4* __restrict__ part, int S, long stride,
                                                  int n, P4* __restrict__ acc) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double ax, ay, az;
    sum_parts(part, S, stride, i, ax, ay, az);
    acc[i] = P4{ax, ay, az, 0.0};
}

// v += kdt a; x += c1dt v; x += c2dt v  (the kick d[s] dt followed by the drift
// of the next substep; c2dt carries the first drift of the next step when the
// last drift of this step is merged with it, FPGA integrator :853).
__global__ __launch_bounds__(BLOCK) void k_kick_drift(const P4* __restrict__ part, int S, long stride,
                                                      int n, P4* __restrict__ pm, P4* __restrict__ vel,
                                                      P4* __restrict__ acc, double kdt, double c1dt,
                                                      double c2dt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double ax, ay, az;
    sum_parts(part, S, stride, i, ax, ay, az);
    if (acc != part) acc[i] = P4{ax, ay, az, 0.0};
    P4 v = vel[i];
    v.x = __builtin_fma(kdt, ax, v.x);
    v.y = __builtin_fma(kdt, ay, v.y);
    v.z = __builtin_fma(kdt, az, v.z);
    vel[i] = v;
    P4 p = pm[i];
    p.x = __builtin_fma(c1dt, v.x, p.x);
    p.y = __builtin_fma(c1dt, v.y, p.y);
    p.z = __builtin_fma(c1dt, v.z, p.z);
    p.x = __builtin_fma(c2dt, v.x, p.x);
    p.y = __builtin_fma(c2dt, v.y, p.y);
    p.z = __builtin_fma(c2dt, v.z, p.z);
    pm[i] = p;
}

// The same two operations for many splits and few particles (the sequential
// sum of S partials per thread is latency-bound there): 16 waves per 64
// particles, wave w adds splits w, w + 16, ... in order, and the 16 sums are
// added in wave order -- a fixed order, so results stay bitwise reproducible.
constexpr int kWideWaves = 16;

template <bool KICK>
__global__ __launch_bounds__(64 * kWideWaves) void k_parts_wide(const P4* __restrict__ part, int S, long stride,
                                                                int n, P4* __restrict__ pm, P4* __restrict__ vel,
                                                                P4* __restrict__ acc, double kdt, double c1dt,
                                                                double c2dt) {
    __shared__ double sh[kWideWaves][3][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (i < n)
        for (int q = w; q < S; q += kWideWaves) {
            const P4 p = part[(long)q * stride + i];
            ax += p.x; ay += p.y; az += p.z;
        }
    sh[w][0][lane] = ax;
    sh[w][1][lane] = ay;
    sh[w][2][lane] = az;
    __syncthreads();
    if (w != 0 || i >= n) return;
    ax = ay = az = 0.0;
    for (int q = 0; q < kWideWaves; q++) {
        ax += sh[q][0][lane];
        ay += sh[q][1][lane];
        az += sh[q][2][lane];
    }
    if (!KICK) {
        acc[i] = P4{ax, ay, az, 0.0};
        return;
    }
    if (acc != part) acc[i] = P4{ax, ay, az, 0.0};
    P4 v = vel[i];
    v.x = __builtin_fma(kdt, ax, v.x);
    v.y = __builtin_fma(kdt, ay, v.y);
    v.z = __builtin_fma(kdt, az, v.z);
    vel[i] = v;
    P4 p = pm[i];
    p.x = __builtin_fma(c1dt, v.x, p.x);
    p.y = __builtin_fma(c1dt, v.y, p.y);
    p.z = __builtin_fma(c1dt, v.z, p.z);
    p.x = __builtin_fma(c2dt, v.x, p.x);
    p.y = __builtin_fma(c2dt, v.y, p.y);
    p.z = __builtin_fma(c2dt, v.z, p.z);
    pm[i] = p;
}

// x += cdt v
__global__ __launch_bounds__(BLOCK) void k_drift(int n, P4* __restrict__ pm, const P4* __restrict__ vel,
                                                 double cdt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const P4 v = vel[i];
    P4 p = pm[i];
    p.x = __builtin_fma(cdt, v.x, p.x);
    p.y = __builtin_fma(cdt, v.y, p.y);
    p.z = __builtin_fma(cdt, v.z, p.z);
    pm[i] = p;
}

// Explicit Euler update of the reference's first-order driver
// (CPU/omp_solver.cpp:124-137): the position moves with the velocity it had
// before the kick, then v += a dt.
__global__ __launch_bounds__(BLOCK) void k_euler(int n, P4* __restrict__ pm, P4* __restrict__ vel,
                                                 const P4* __restrict__ acc, double dt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    P4 p = pm[i];
    P4 v = vel[i];
    const P4 a = acc[i];
    p.x = __builtin_fma(v.x, dt, p.x);
    p.y = __builtin_fma(v.y, dt, p.y);
    p.z = __builtin_fma(v.z, dt, p.z);
    v.x = __builtin_fma(a.x, dt, v.x);
    v.y = __builtin_fma(a.y, dt, v.y);
    v.z = __builtin_fma(a.z, dt, v.z);
    pm[i] = p;
    vel[i] = v;
}

// One Runge-Kutta combination (CPU/omp_solver.cpp:544-572, 576-605):
//   x = x0 + dt * sum_j w_j kv_j,   v = kv_0 + dt * sum_j w_j ka_j
// over stages j = 0..count-1, kv_j / ka_j being stage j's velocity and
// acceleration (records j*stride + i; kv_0 is the step's starting velocity).
// The sums run in stage order from zero, as the reference's do; the mass
// travels with x0.
__global__ __launch_bounds__(BLOCK) void k_rk_combine(int n, const P4* __restrict__ x0, const P4* __restrict__ kv,
                                                      const P4* __restrict__ ka, long stride, RKTerms T, double dt,
                                                      P4* __restrict__ x_out, P4* __restrict__ v_out) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, tx = 0.0, ty = 0.0, tz = 0.0;
    for (int j = 0; j < T.count; j++) {
        const double w = T.w[j];
        const P4 u = kv[j * stride + i];
        const P4 a = ka[j * stride + i];
        sx = __builtin_fma(w, u.x, sx);
        sy = __builtin_fma(w, u.y, sy);
        sz = __builtin_fma(w, u.z, sz);
        tx = __builtin_fma(w, a.x, tx);
        ty = __builtin_fma(w, a.y, ty);
        tz = __builtin_fma(w, a.z, tz);
    }
    const P4 p = x0[i];
    const P4 v = kv[i];
    x_out[i] = P4{__builtin_fma(sx, dt, p.x), __builtin_fma(sy, dt, p.y), __builtin_fma(sz, dt, p.z), p.m};
    v_out[i] = P4{__builtin_fma(tx, dt, v.x), __builtin_fma(ty, dt, v.y), __builtin_fma(tz, dt, v.z), 0.0};
}

// ---------------------------------------------------------------------------
// Energy diagnostic
// ---------------------------------------------------------------------------

// Error-free sum (Knuth TwoSum): s + c == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& c) {
    s = a + b;
    const double bb = s - a;
    c = (a - (s - bb)) + (b - bb);
}

struct __attribute__((aligned(16))) DD { double s, c; };

__device__ __forceinline__ void dd_add(DD& acc, double v) {
    double s, e;
    two_sum(acc.s, v, s, e);
    acc.s = s;
    acc.c += e;
}

__device__ __forceinline__ void dd_merge(DD& a, const DD& b) {
    double s, e;
    two_sum(a.s, b.s, s, e);
    a.s = s;
    a.c += e + b.c;
}

// m_j / sqrt(d2 + eps2) with the FP64 rsq and its quadratic correction
// (1-e)^(-1/2) = 1 + e/2 + 3/8 e^2.  A pair with d2 + eps2 == 0 (coincident,
// unsoftened) gives 0; with eps > 0 coincident particles give m_j / eps, as
// the softened potential of FPGA/host/accuracy_evaluation.cpp:35 does.
__device__ __forceinline__ double pot_term(const P4& me, const P4& q, double eps2) {
    const double dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
    const double s2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const double y = __builtin_amdgcn_rsq(s2);
    const double e = __builtin_fma(-s2, y * y, 1.0);
    const double inv = y * __builtin_fma(e, __builtin_fma(e, 0.375, 0.5), 1.0);
    return (s2 > 0.0) ? q.m * inv : 0.0;
}

// Upper-triangle potential rows: row_i = sum_{j > i} m_j / sqrt(|x_j - x_i|^2
// + eps^2) (coincident pairs skipped only when eps = 0), so that
// U = -sum_i m_i row_i counts every pair once.  Tiles of
// BLOCK particles; workgroup (g, s) owns the rows of tiles I1 = g and
// I2 = nT-1-g, whose tile lists J = I..nT-1 together hold nT+1 tiles (every
// workgroup row has the same work), and the s-th of `splits` chunks of that
// concatenated list.  It writes its partial of each of its particles' rows to
// out[s*n + i], zero where its chunk holds no tile of that row, so every
// (s, i) is written once.  Terms are summed in FP64 per 64 j (all positive:
// relative error <= 63 u) and the 64-term sums added with TwoSum.
__global__ __launch_bounds__(BLOCK) void k_potential(const P4* __restrict__ pm, int n, int nT, int splits,
                                                     double eps2, DD* __restrict__ out) {
    __shared__ P4 tile[BLOCK];
    const int t = threadIdx.x;
    const int g = blockIdx.x, s = blockIdx.y;
    const int I1 = g, I2 = nT - 1 - g;
    const int len1 = nT - I1;
    const int len2 = (I2 != I1) ? nT - I2 : 0;
    const int total = len1 + len2;
    const int per = (total + splits - 1) / splits;
    const int c0 = min(total, s * per), c1 = min(total, c0 + per);
    for (int seg = 0; seg < 2; seg++) {
        if (seg == 1 && len2 == 0) break;
        const int I = seg ? I2 : I1;
        const int base = seg ? len1 : 0;
        const int len = seg ? len2 : len1;
        const int ka = max(c0, base) - base, kb = min(c1, base + len) - base;   // tiles I+ka .. I+kb-1
        const int i = I * BLOCK + t;
        const P4 me = (i < n) ? pm[i] : P4{0.0, 0.0, 0.0, 0.0};
        DD acc{0.0, 0.0};
        P4 nxt = (ka < kb) ? load_p4(pm, (I + ka) * BLOCK + t, n) : P4{0.0, 0.0, 0.0, 0.0};
        for (int k = ka; k < kb; k++) {
            tile[t] = nxt;
            __syncthreads();
            if (k + 1 < kb) nxt = load_p4(pm, (I + k + 1) * BLOCK + t, n);
            if (k == 0) {   // the diagonal tile: j > i only
                for (int j0 = 0; j0 < BLOCK; j0 += 64) {
                    double part = 0.0;
                    for (int jj = j0; jj < j0 + 64; jj++) {
                        const double v = pot_term(me, tile[jj], eps2);
                        part += (jj > t) ? v : 0.0;
                    }
                    dd_add(acc, part);
                }
            } else {
                for (int j0 = 0; j0 < BLOCK; j0 += 64) {
                    double part = 0.0;
#pragma unroll 8
                    for (int jj = j0; jj < j0 + 64; jj++) part += pot_term(me, tile[jj], eps2);
                    dd_add(acc, part);
                }
            }
            __syncthreads();
        }
        if (i < n) out[(long)s * n + i] = acc;
    }
}

// Per block: U_b = sum_i m_i row_i (row over all splits), T_b = sum_i m_i |v_i|^2,
// compensated, fixed tree order.  out[b] = {U.s, U.c, T.s, T.c}.
__global__ __launch_bounds__(BLOCK) void k_energy_reduce(const DD* __restrict__ rows, int S, int n,
                                                         const P4* __restrict__ pm,
                                                         const P4* __restrict__ vel,
                                                         double* __restrict__ out) {
    __shared__ DD su[BLOCK], st[BLOCK];
    const int t = threadIdx.x;
    const int i = blockIdx.x * BLOCK + t;
    DD u{0.0, 0.0}, k{0.0, 0.0};
    if (i < n) {
        DD row{0.0, 0.0};
        for (int s = 0; s < S; s++) dd_merge(row, rows[(long)s * n + i]);
        const double m = pm[i].m;
        const double r = row.s + row.c;
        u.s = m * r;
        u.c = __builtin_fma(m, r, -u.s);   // exact product error
        const P4 v = vel[i];
        const double v2 = __builtin_fma(v.z, v.z, __builtin_fma(v.y, v.y, v.x * v.x));
        k.s = m * v2;
        k.c = __builtin_fma(m, v2, -k.s);
    }
    su[t] = u;
    st[t] = k;
    __syncthreads();
    for (int w = BLOCK / 2; w > 0; w >>= 1) {
        if (t < w) {
            dd_merge(su[t], su[t + w]);
            dd_merge(st[t], st[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[4 * blockIdx.x + 0] = su[0].s;
        out[4 * blockIdx.x + 1] = su[0].c;
        out[4 * blockIdx.x + 2] = st[0].s;
        out[4 * blockIdx.x + 3] = st[0].c;
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
template <int FORM, bool EPS0, int IPT>
static hipError_t launch_accel_t(const AccelLaunch& L, hipStream_t st) {
    const dim3 grid((L.ni + BLOCK * IPT - 1) / (BLOCK * IPT), L.splits);
    hipLaunchKernelGGL((k_accel<FORM, EPS0, IPT>), grid, dim3(BLOCK), 0, st, L.pi, L.ni, L.pj, L.jbeg, L.jend,
                       L.jchunk, L.eps2, L.G, kCubicShift, kCubicOffset, L.out, L.out_stride, L.split_base, L.eq_mass,
                       L.eq_dev);
    return hipGetLastError();
}

template <int FORM, bool EPS0>
static hipError_t launch_accel_ipt(const AccelLaunch& L, hipStream_t st) {
    switch (L.ipt) {
    case 4: return launch_accel_t<FORM, EPS0, 4>(L, st);
    case 1: return launch_accel_t<FORM, EPS0, 1>(L, st);
    default: return hipErrorInvalidValue;
    }
}

template <int FORM>
static hipError_t launch_accel_eps(const AccelLaunch& L, hipStream_t st) {
    return (L.eps2 == 0.0) ? launch_accel_ipt<FORM, true>(L, st) : launch_accel_ipt<FORM, false>(L, st);
}

hipError_t launch_accel(const AccelLaunch& L, hipStream_t st) {
    if (L.ni <= 0 || L.splits <= 0) return hipSuccess;
    switch (L.form) {
    case FORM_FP64: return launch_accel_eps<FORM_FP64>(L, st);
    case FORM_MIXED: return launch_accel_eps<FORM_MIXED>(L, st);
    case FORM_MIXED_NR: return launch_accel_eps<FORM_MIXED_NR>(L, st);
    case FORM_FP32: return launch_accel_eps<FORM_FP32>(L, st);
    default: return hipErrorInvalidValue;
    }
}

// acc[i] = sum over s of part[s * m + i] in order s = 0..S-1 for i < m,
// (m, S) = shape->x, shape->y from k_accel_rows: 16 waves per block take
// every 16th split of a 64-row block, then wave 0 adds the 16 sums in order;
// the blocks take the row blocks in turn.
__global__ __launch_bounds__(64 * kWideWaves) void k_parts_rows(const P4* __restrict__ part,
                                                                const int4* __restrict__ shape,
                                                                P4* __restrict__ acc) {
    __shared__ double sh[kWideWaves][3][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = shape->x, S = shape->y;
    for (int rb = blockIdx.x; rb * 64 < m; rb += gridDim.x) {   // uniform per block
        const int i = rb * 64 + lane;
        double ax = 0.0, ay = 0.0, az = 0.0;
        if (i < m)
            for (int q = w; q < S; q += kWideWaves) {
                const P4 p = part[(long)q * m + i];
                ax += p.x; ay += p.y; az += p.z;
            }
        sh[w][0][lane] = ax;
        sh[w][1][lane] = ay;
        sh[w][2][lane] = az;
        __syncthreads();
        if (w == 0 && i < m) {
            ax = ay = az = 0.0;
            for (int q = 0; q < kWideWaves; q++) {
                ax += sh[q][0][lane];
                ay += sh[q][1][lane];
                az += sh[q][2][lane];
            }
            acc[i] = P4{ax, ay, az, 0.0};
        }
        __syncthreads();   // sh is rewritten for the next row block
    }
}

template <int FORM>
static hipError_t launch_accel_rows_t(const RowsLaunch& L, hipStream_t st) {
    auto k = (L.eps2 == 0.0) ? k_accel_rows<FORM, true> : k_accel_rows<FORM, false>;
    hipLaunchKernelGGL(k, dim3(L.grid), dim3(BLOCK), 0, st, L.pi, L.count, L.ni_max, L.pj, L.n, L.eps2, L.G,
                       kCubicShift, kCubicOffset, L.part, L.shape);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int rblocks = std::max(1, std::min(L.grid, (L.ni_max + 63) / 64));
    hipLaunchKernelGGL(k_parts_rows, dim3(rblocks), dim3(64 * kWideWaves), 0, st, (const P4*)L.part,
                       (const int4*)L.shape, L.acc);
    return hipGetLastError();
}

hipError_t launch_accel_rows(const RowsLaunch& L, hipStream_t st) {
    if (L.ni_max <= 0 || L.n <= 0 || L.grid <= 0) return hipSuccess;
    switch (L.form) {
    case FORM_FP64: return launch_accel_rows_t<FORM_FP64>(L, st);
    case FORM_MIXED: return launch_accel_rows_t<FORM_MIXED>(L, st);
    case FORM_MIXED_NR: return launch_accel_rows_t<FORM_MIXED_NR>(L, st);
    case FORM_FP32: return launch_accel_rows_t<FORM_FP32>(L, st);
    default: return hipErrorInvalidValue;
    }
}

size_t accel_rows_part_records(int grid, int ni_max) {
    return std::max((size_t)grid * BLOCK, (size_t)std::max(ni_max, 1));
}

hipError_t launch_reduce(const P4* part, int S, long stride, int n, P4* acc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (S >= kWideWaves)
        hipLaunchKernelGGL(k_parts_wide<false>, dim3((n + 63) / 64), dim3(64 * kWideWaves), 0, st, part, S, stride, n,
                           nullptr, nullptr, acc, 0.0, 0.0, 0.0);
    else
        hipLaunchKernelGGL(k_reduce, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, part, S, stride, n, acc);
    return hipGetLastError();
}

hipError_t launch_kick_drift(const P4* part, int S, long stride, int n, P4* pm, P4* vel, P4* acc,
                             double kdt, double c1dt, double c2dt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (S >= kWideWaves)
        hipLaunchKernelGGL(k_parts_wide<true>, dim3((n + 63) / 64), dim3(64 * kWideWaves), 0, st, part, S, stride, n,
                           pm, vel, acc, kdt, c1dt, c2dt);
    else
        hipLaunchKernelGGL(k_kick_drift, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, part, S, stride,
                           n, pm, vel, acc, kdt, c1dt, c2dt);
    return hipGetLastError();
}

hipError_t launch_drift(int n, P4* pm, const P4* vel, double cdt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_drift, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, n, pm, vel, cdt);
    return hipGetLastError();
}

hipError_t launch_euler(int n, P4* pm, P4* vel, const P4* acc, double dt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_euler, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, n, pm, vel, acc, dt);
    return hipGetLastError();
}

hipError_t launch_rk_combine(int n, const P4* x0, const P4* kv, const P4* ka, long stride, const RKTerms& T,
                             double dt, P4* x_out, P4* v_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (T.count < 1 || T.count > kMaxRK) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rk_combine, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, n, x0, kv, ka, stride, T, dt,
                       x_out, v_out);
    return hipGetLastError();
}

hipError_t launch_energy(const P4* pm, const P4* vel, int n, int splits, double eps2, void* rows,
                         double* blocks_out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int nT = (n + BLOCK - 1) / BLOCK;
    if (splits < 1 || splits > nT + 1) return hipErrorInvalidValue;
    const dim3 grid((nT + 1) / 2, splits);
    hipLaunchKernelGGL(k_potential, grid, dim3(BLOCK), 0, st, pm, n, nT, splits, eps2, (DD*)rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_energy_reduce, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, st, (const DD*)rows,
                       splits, n, pm, vel, blocks_out);
    return hipGetLastError();
}

}  // namespace nbk
