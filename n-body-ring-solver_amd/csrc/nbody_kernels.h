// nbody_kernels.h -- internal interface between the C ABI layer and the
// gfx950 kernels in nbody_kernels.hip.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

namespace nbk {

constexpr int BLOCK = 256;   // threads per workgroup (4 wave64s)

// One particle record in HBM: (x, y, z, m) / (vx, vy, vz, 0) / (ax, ay, az, 0).
struct __attribute__((aligned(32))) P4 {
    double x, y, z, m;
};

// FORM_FP32: every step in FP32 (positions rounded to FP32, FP32 differences,
// rsqrt + Newton step, FP32 sums): the reference's -p sp kernels
// (CPU/force_calculation.cpp:692-891); tile kernel only.
enum Form { FORM_FP64 = 0, FORM_MIXED = 1, FORM_MIXED_NR = 2, FORM_FP32 = 3 };

// Correction of the hardware reciprocal square root, shared by the pair
// functions.  With y ~ r2^(-1/2) and s = r2 y^2 = 1 - e, the exact
// r2^(-3/2) = y^3 (1 - e)^(-3/2).
//  FP64, to second order: 1 + 3/2 e + 15/8 e^2 = 15/8 ((s - 7/5)^2 + 28/75),
//    so w = y^3 ((s + kCubicShift)^2 + kCubicOffset) takes two FMAs and the
//    factor 15/8 rides on the masses (kCubicMass), which the pair multiplies
//    in anyway.
//  FP32, to first order (one Newton-Raphson step on the cube):
//    1 + 3/2 e = 3/2 (5/3 - s).  With F = fl32(5/3) the factor
//    c = 1/(F - 1) (kNewtonMass, on the masses) makes c (F - s) exactly 1 at
//    s = 1, so w = y^3 (F - s) takes one FMA; c differs from 3/2 by 6e-8
//    relative, a change of the first-order term far below FP32 rounding.
constexpr double kCubicShift = -1.4;
constexpr double kCubicOffset = 28.0 / 75.0;
constexpr double kCubicMass = 1.875;
constexpr float kNewtonOffset = 5.0f / 3.0f;
constexpr double kNewtonMass = 1.0 / ((double)kNewtonOffset - 1.0);

// Launch description of the force kernel: accelerations of pi[0..ni) from
// pj[jbeg..jend), the j-range cut into `splits` chunks of `jchunk` (one per
// blockIdx.y); chunk s writes out[(split_base + s) * out_stride + i].
// Whether every mass of a j set is equal (ok) and the value (m0), found on
// the device by launch_mass_check.
struct EqState {
    double m0;
    int ok, pad;
};
hipError_t launch_mass_check(const P4* pj, int nj, EqState* st, hipStream_t stream);

struct AccelLaunch {
    int form = FORM_FP64;
    int ipt = 1;              // i-particles per lane (1 or 4)
    const P4* pi = nullptr;
    int ni = 0;
    const P4* pj = nullptr;
    int jbeg = 0, jend = 0, jchunk = 0, splits = 1;
    double eps2 = 0.0, G = 1.0;
    P4* out = nullptr;
    long out_stride = 0;
    int split_base = 0;
    // every j mass equal: the pair terms leave it out and the final scale
    // takes it; known on the host (eq_mass > 0) or decided on the device
    // (eq_dev, from launch_mass_check; takes precedence)
    double eq_mass = 0.0;
    const EqState* eq_dev = nullptr;
};

hipError_t launch_accel(const AccelLaunch& L, hipStream_t st);

// Rows [0, min(*count, ni_max)) of pi against pj[0, n), the row count read on
// the device (no host round trip): a resident grid of `grid` blocks, splits
// over j chosen on the device, partial sums in `part` (at least
// accel_rows_part_records(grid, ni_max) records), their in-order sums in acc.
struct RowsLaunch {
    int form = FORM_FP64;
    const P4* pi = nullptr;
    const int* count = nullptr;
    int ni_max = 0;
    const P4* pj = nullptr;
    int n = 0;
    double eps2 = 0.0, G = 1.0;
    int grid = 0;
    P4* part = nullptr;
    int4* shape = nullptr;   // (m, splits) for the reduction
    P4* acc = nullptr;
};

hipError_t launch_accel_rows(const RowsLaunch& L, hipStream_t st);
size_t accel_rows_part_records(int grid, int ni_max);

// Symmetric kernel (nbody_sym.hip): blocks of S = 64*K particles, work items
// (A, b0, b1, action slot) = one wave sweeping j-blocks [b0, b1) against
// i-block A (slot layout described in nbody_sym.hip).
struct SymLaunch {
    int form = FORM_FP64;
    int K = 8;
    const P4* pm = nullptr;
    int n = 0;
    const int4* items = nullptr;
    int nitems = 0;
    int nB = 0;
    double eps2 = 0.0;
    P4* part = nullptr;          // action slots: nB*W slots of S records, W = (nB - 1)/L + 1
    P4* react64 = nullptr;       // FP64 form: reaction slots, nB(nB-1)/2 slots of S records
    float4* react32 = nullptr;   // mixed forms: FP32 reaction slots, nB(nB-1)/2 slots of S float4
    // two-set form (one ring round): j-chunks from pj, n records like pm;
    // pairs q > k (q >= k with incl); nB more reaction slots for the diagonal
    const P4* pj = nullptr;
    bool incl = false;
    // every mass equal (one form only): the sums leave out m times the
    // form's mass factor, which the finish applies (SymSums::scale)
    bool eqm = false;
};

hipError_t launch_accel_sym(const SymLaunch& L, hipStream_t st);

// The partial sums one symmetric evaluation leaves (exactly one of react64 /
// react32 set) and the item length L they were cut with.
struct SymSums {
    const P4* part = nullptr;
    const P4* react64 = nullptr;
    const float4* react32 = nullptr;
    int nB = 0, S = 0, L = 1;
    double scale = 1.0;   // factor of the sums: the mass factored out of an equal-mass evaluation, else 1
};
hipError_t launch_sym_reduce(const SymSums& P, int n, P4* acc, hipStream_t st);
hipError_t launch_sym_kick_drift(const SymSums& P, int n, P4* pm, P4* vel, P4* acc, double kdt, double c1dt,
                                 double c2dt, hipStream_t st);
// Grouped symmetric kernel (nbody_sym.hip): workgroups of kGroupWaves waves
// over W consecutive blocks, each walking a contiguous range of the chunk
// sequence; ranges[wg] = (g, B * K + c, chunks, first action slot); action
// slots of kGroupWaves blocks, each three component planes of S doubles;
// reaction slots (g, B), B >= g W, of three planes of S values (doubles, or
// floats for the mixed forms).
// four waves: three workgroups a CU at three waves per SIMD; groups of two
// measured slower (N = 262144 FP64 33.4 against 24.4 ms)
constexpr int kGroupWaves = 4;
struct SymgLaunch {
    int form = FORM_FP64;
    int K = 8;
    const P4* pm = nullptr;
    const P4* pj = nullptr;   // two-set form: the received set
    bool incl = false;
    int n = 0;
    const int4* ranges = nullptr;
    int nwg = 0;
    int nB = 0;
    double eps2 = 0.0;
    double* act = nullptr;       // action slots: W blocks of 3 planes (x, y, z) of S doubles
    double* react64 = nullptr;   // FP64 form: reaction slots of 3 planes of S doubles
    float* react32 = nullptr;    // mixed forms: reaction slots of 3 planes of S floats
    bool eqm = false;   // every mass equal: as SymLaunch::eqm
};
hipError_t launch_accel_symg(const SymgLaunch& L, hipStream_t st);

// The grouped kernel's partial sums: rowseg[g] = (first action slot of row g,
// number of its segments).
struct SymgSums {
    const double* act = nullptr;
    const int2* rowseg = nullptr;
    const double* react64 = nullptr;
    const float* react32 = nullptr;
    int nB = 0, S = 0;
    double scale = 1.0;   // as SymSums::scale
};
// acc = the sums (kick: then v += kdt a, x += c1dt v + c2dt v as the kick/drift kernels)
hipError_t launch_symg_finish(const SymgSums& P, int n, P4* pm, P4* vel, P4* acc, bool kick, double kdt, double c1dt,
                              double c2dt, hipStream_t st);
hipError_t launch_symg_ring_finish(const SymgSums& P, int n, const P4* pl, const P4* pr, P4* fl, P4* fr,
                                   hipStream_t st);

// Two-set finish: fl[i] += m_l a_l (actions), fr[i] += m_r a_r (reactions
// incl. the diagonal slots), masses from the records pl / pr.
hipError_t launch_ring_finish(const SymSums& P, int n, const P4* pl, const P4* pr, P4* fl, P4* fr, hipStream_t st);
hipError_t launch_reduce(const P4* part, int S, long stride, int n, P4* acc, hipStream_t st);
hipError_t launch_kick_drift(const P4* part, int S, long stride, int n, P4* pm, P4* vel, P4* acc,
                             double kdt, double c1dt, double c2dt, hipStream_t st);
hipError_t launch_drift(int n, P4* pm, const P4* vel, double cdt, hipStream_t st);
// Euler update: x += v dt (old v), then v += a dt.
hipError_t launch_euler(int n, P4* pm, P4* vel, const P4* acc, double dt, hipStream_t st);

// Runge-Kutta stage combination: x_out = x0 + dt sum_j w_j kv_j (mass from
// x0), v_out = kv_0 + dt sum_j w_j ka_j, j < count; stage j's records sit at
// j * stride.
constexpr int kMaxRK = 8;
struct RKTerms {
    int count = 0;
    double w[kMaxRK] = {};
};
hipError_t launch_rk_combine(int n, const P4* x0, const P4* kv, const P4* ka, long stride, const RKTerms& T,
                             double dt, P4* x_out, P4* v_out, hipStream_t st);
// Potential rows over the upper triangle (every pair once) in `splits`
// chunks (1 <= splits <= ceil(n/BLOCK) + 1) and the per-block sums of
// m_i row_i and m_i |v_i|^2.  rows: splits * n * 16 bytes; blocks_out: 4
// doubles per 256-particle block.
hipError_t launch_energy(const P4* pm, const P4* vel, int n, int splits, double eps2, void* rows,
                         double* blocks_out, hipStream_t st);

}  // namespace nbk
