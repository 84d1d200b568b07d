// nbody_near.hip -- short-range ("near-field") accelerations on a sorted
// uniform grid, for the learned form's force split.
//
// a_near,i = sum_{j : 0 < r_ij < rc} m_j S(q_ij) d_ij / (r_ij^2 + eps^2)^(3/2),
// q = r^2 / rc^2, S = 1 for q <= 1/4 (r <= rc/2) falling to 0 at q = 1 with
// the C2 quintic 1 - u^3 (10 - 15u + 6u^2), u = (q - 1/4) / (3/4); the
// remainder a - a_near (pairs farther than rc/2, tapered) varies smoothly in
// time and is what the learned form extrapolates.  A polynomial in r^2 needs
// no square root per pair.
//
// Layout: columns of cross-section (rc/2)^2, cut along z into slabs of
// height rc/8 (kZFine = 4 slabs per rc/2).  A cell (x, y, z) has the key
// x << 32 | y << 21 | z (x, y biased by 1024 and clamped to [0, 2048), z by
// 2^20 and clamped to [0, 2^21)), so the 2^22 (x, y) columns are the keys'
// top bits and each column's cells are contiguous in key order.  Particles
// are radix-sorted by key (stable: inside a cell, index order); a table gives
// every column's run [start, end) of the sorted array.  The thin z slabs
// order a dense column's particles by height, so its 64-particle chunks are
// compact (about 10 % fewer examined pairs than index order inside rc/2
// cells).  Clamping only merges far cells, and two particles closer than rc
// stay within two columns and eight slabs of each other after it, so results
// stay exact (only slower for far-flung systems).
//
// Work: a chunk is a run of consecutive sorted particles, one per lane: up to
// 64 particles of one dense column (kDense or more particles), or the
// particles of consecutive sparser columns of one group of kGroup columns
// (same x, y / kGroup equal), merged greedily up to 64, so sparse regions do
// not cost one launch slot per column.  The blocks the device holds at once
// take chunks from a work counter (no chunk count travels to the host); a
// block of five waves takes a chunk and the cells within rc of its bounding
// box (after a build: five columns across x, the chunk's y-range +- 2
// columns, its slabs +- 8): the columns are dealt to the waves in turn, each
// restricted by binary search to the slab range, streaming those runs
// through LDS; the five partial sums are added in wave order (bitwise
// reproducible).  Of
// each 64-particle batch of a run only the particles within rc of the chunk's
// bounding box are kept (ballot + prefix count, order preserved), so far
// corners of the cell box cost one distance test instead of a pair each.
// Pairs beyond rc get S = 0, so the pair code is branch-free; the FP32
// variant stages the masses already rounded to FP32.
//
// Reuse: between rebuilds (the learned form rebuilds on its exact substeps)
// the grid is kept; each evaluation gathers the current positions in the
// grid's order and measures the largest displacement d since the build, and
// k_near searches the build's cells within rc + d of each chunk's current
// bounding box: a partner within rc now was within rc + d of it at the
// build, so the result stays exact however far the particles moved (a state
// that moved too far for the cells to help is scanned whole).  The chunks'
// bounding boxes and every pair use the current positions.
//
// FP64 throughout, or (fp32) FP64 differences with FP32 pair arithmetic and
// FP32 sums over each 64-particle batch, flushed to FP64 -- the precision of
// the mixed force form.
//
// Bound: VALU.  About 26 operations per examined pair (FP32 variant).  With
// the bounding-box cull about 27 % of the examined pairs lie within rc (a CPU
// model of this chunking on a Plummer sphere of 262144 at rc = 0.25): 64
// particles at that density fill a box whose rc-neighbourhood is ~3.4 times
// the cut-off sphere, whatever its shape.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "nbody_near.h"

namespace nbn {

namespace {

constexpr int BLK = 256;
constexpr int kWaves = 5;                 // waves per chunk
constexpr int kColSide = 1 << 11;         // columns per side of the (x, y) table
constexpr int kColBias = kColSide / 2;
constexpr int kZBias = 1 << 20;
constexpr int kZMax = (1 << 21) - 1;
constexpr int kZFine = 4;                 // z slabs per rc/2: slab height rc/8
constexpr int kDense = 32;                // particles from which a column gets chunks of its own
constexpr int kGroup = 16;                // sparser columns are merged within groups of 16 along y

__device__ __forceinline__ int cell_xy(double x, double inv_h) {
    const double c = floor(x * inv_h) + (double)kColBias;
    return (int)fmin(fmax(c, 0.0), (double)(kColSide - 1));
}

__device__ __forceinline__ int cell_z(double z, double inv_h) {
    const double c = floor(z * (inv_h * kZFine)) + (double)kZBias;
    return (int)fmin(fmax(c, 0.0), (double)kZMax);
}

__device__ __forceinline__ unsigned long long pack(int ix, int iy, int iz) {
    return ((unsigned long long)ix << 32) | ((unsigned long long)iy << 21) | (unsigned long long)iz;
}

__device__ __forceinline__ int col_of(unsigned long long key) { return (int)(key >> 21); }

__global__ __launch_bounds__(BLK) void k_keys(const nbk::P4* __restrict__ pm, int n, double inv_h,
                                              unsigned long long* __restrict__ keys, int* __restrict__ idx) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const nbk::P4 p = pm[i];
    keys[i] = pack(cell_xy(p.x, inv_h), cell_xy(p.y, inv_h), cell_z(p.z, inv_h));
    idx[i] = i;
}

__global__ __launch_bounds__(BLK) void k_gather_sorted(const nbk::P4* __restrict__ pm, const int* __restrict__ order,
                                                       int n, nbk::P4* __restrict__ sorted) {
    const int p = blockIdx.x * BLK + threadIdx.x;
    if (p < n) sorted[p] = pm[order[p]];
}

__global__ __launch_bounds__(BLK) void k_clear_cols(int2* __restrict__ cols, int count) {
    const int c = blockIdx.x * BLK + threadIdx.x;
    if (c < count) cols[c] = make_int2(0, 0);
}

// Column runs of the sorted keys.
__global__ __launch_bounds__(BLK) void k_col_runs(const unsigned long long* __restrict__ keys, int n,
                                                  int2* __restrict__ cols) {
    const int p = blockIdx.x * BLK + threadIdx.x;
    if (p >= n) return;
    const int c = col_of(keys[p]);
    if (p == 0 || col_of(keys[p - 1]) != c) cols[c].x = p;
    if (p == n - 1 || col_of(keys[p + 1]) != c) cols[c].y = p + 1;
}

// Chunk starts (flag[p] = 1; flag was cleared), one thread per group of
// kGroup columns walking them in order: a dense column starts a chunk at its
// first particle and every 64th after it; sparser columns join the open chunk
// while it stays within 64 particles, else start a new one.
__global__ __launch_bounds__(BLK) void k_chunk_flags(const int2* __restrict__ cols, int ngroups,
                                                     int* __restrict__ flag) {
    const int g = blockIdx.x * BLK + threadIdx.x;
    if (g >= ngroups) return;
    int open = 64;   // particles in the open chunk (64: none open)
    for (int c = g * kGroup; c < (g + 1) * kGroup; c++) {
        const int2 run = cols[c];
        const int k = run.y - run.x;
        if (k <= 0) continue;
        if (k >= kDense) {
            for (int q = run.x; q < run.y; q += 64) flag[q] = 1;
            open = 64;
        } else if (open + k > 64) {
            flag[run.x] = 1;
            open = k;
        } else {
            open += k;
        }
    }
}

// The chunk list, its length and the work counter k_near takes chunks from
// (reset here, so no fill is needed; the displacement since the build is 0).
__global__ __launch_bounds__(BLK) void k_chunk_list(const int* __restrict__ flag, const int* __restrict__ pos, int n,
                                                    int* __restrict__ chunks, int* __restrict__ nchunks,
                                                    int* __restrict__ next, unsigned int* __restrict__ dmax) {
    const int p = blockIdx.x * BLK + threadIdx.x;
    if (p >= n) return;
    if (flag[p]) chunks[pos[p]] = p;
    if (p == n - 1) {
        *nchunks = pos[p] + flag[p];
        *next = 0;
        *dmax = 0u;
        chunks[pos[p] + flag[p]] = n;   // sentinel: the end of the last chunk
    }
}

// Before an evaluation on a kept grid: the work counter and the displacement
// maximum start at zero.
__global__ void k_reuse_reset(int* __restrict__ next, unsigned int* __restrict__ dmax) {
    *next = 0;
    *dmax = 0u;
}

// The current positions in the grid's order and the largest displacement
// since the build (an FP32 upper bound, as bits: positive floats order like
// their bit patterns).
__global__ __launch_bounds__(BLK) void k_restage(const nbk::P4* __restrict__ pm, const int* __restrict__ order,
                                                 const nbk::P4* __restrict__ built, int n, nbk::P4* __restrict__ cur,
                                                 unsigned int* __restrict__ dmax) {
    __shared__ float red[BLK / 64];
    const int p = blockIdx.x * BLK + threadIdx.x;
    float d = 0.0f;
    if (p < n) {
        const nbk::P4 q = pm[order[p]];
        cur[p] = q;
        const nbk::P4 b = built[p];
        const double dx = q.x - b.x, dy = q.y - b.y, dz = q.z - b.z;
        d = __double2float_ru(sqrt(dx * dx + dy * dy + dz * dz));
        if (!(d <= 3.0e38f)) d = 3.0e38f;   // a non-finite position: the widest reach
    }
    for (int o = 32; o > 0; o >>= 1) d = fmaxf(d, __shfl_xor(d, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
    __syncthreads();
    if (threadIdx.x == 0) {
        float m = red[0];
        for (int w = 1; w < BLK / 64; w++) m = fmaxf(m, red[w]);
        atomicMax(dmax, __float_as_uint(m));
    }
}

__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// first position in [lo, hi) with keys[pos] >= k (hi if none)
__device__ __forceinline__ int lower_bound(const unsigned long long* __restrict__ keys, int lo, int hi,
                                           unsigned long long k) {
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Order this lane's LDS accesses against the other lanes' of the same wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Taper argument: with s2 = r^2 + eps^2, u = (q - 1/4) (4/3) with
// q = r^2 / rc^2 is s2 k1 + k0, k1 = 4 / (3 rc^2), k0 = -(1/3 + k1 eps^2), so
// the pair needs r^2 + eps^2 only (one fused chain seeded with eps^2).
struct Taper {
    double k1, k0;
};

__host__ __device__ inline Taper make_taper(double rc, double eps) {
    const double k1 = 4.0 / (3.0 * rc * rc);
    return Taper{k1, -(1.0 / 3.0 + k1 * eps * eps)};
}

// One ordered pair: a += m S(q) w d, w = (r^2 + eps^2)^(-3/2) from v_rsq_f64
// and the correction (1-e)^(-3/2) ~ 1 + 3/2 e + 15/8 e^2 (e = 1 - s2 y^2),
// evaluated as (s2 y^2 + kCubicShift)^2 + kCubicOffset with the factor 15/8
// in the staged mass (stage64).
// For eps > 0 a zero separation needs no test: w stays finite and d = 0
// makes the term vanish.  With eps = 0 (EPS0) such pairs are selected out.
template <bool EPS0>
__device__ __forceinline__ void near_pair(const nbk::P4& me, const nbk::P4& o, double k1, double k0, double eps2,
                                          double& ax, double& ay, double& az) {
    const double dx = o.x - me.x, dy = o.y - me.y, dz = o.z - me.z;
    const double s2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const double u = fmin(fmax(__builtin_fma(s2, k1, k0), 0.0), 1.0);
    const double S = __builtin_fma(-u * u * u, __builtin_fma(u, __builtin_fma(6.0, u, -15.0), 10.0), 1.0);
    const double y = __builtin_amdgcn_rsq(s2);
    const double t = y * y;
    const double v = __builtin_fma(s2, t, nbk::kCubicShift);
    const double w = t * y * __builtin_fma(v, v, nbk::kCubicOffset);
    double f = o.m * S * w;
    if (EPS0) f = (s2 > 0.0) ? f : 0.0;
    ax = __builtin_fma(f, dx, ax);
    ay = __builtin_fma(f, dy, ay);
    az = __builtin_fma(f, dz, az);
}

// Staged record of the FP64 pair: the mass times the correction's factor 15/8.
__device__ __forceinline__ nbk::P4 stage64(const nbk::P4& p) {
    return nbk::P4{p.x, p.y, p.z, p.m * nbk::kCubicMass};
}

// Staged record of the FP32 pair: FP64 position, the mass times the Newton
// step's factor kNewtonMass, rounded to FP32 and kept in the low word of the
// fourth slot.
__device__ __forceinline__ nbk::P4 stage32(const nbk::P4& p) {
    return nbk::P4{p.x, p.y, p.z,
                   __builtin_bit_cast(double, make_int2(__builtin_bit_cast(int, (float)(p.m * nbk::kNewtonMass)), 0))};
}

__device__ __forceinline__ float staged_mass(const nbk::P4& o) {
    return __builtin_bit_cast(float, __builtin_bit_cast(int2, o.m).x);
}

// The same pair with FP64 differences rounded to FP32 and FP32 arithmetic:
// FP32 taper, v_rsq_f32 with the first-order correction of the cube,
// w = y^3 (1 + 3/2 (1 - s2 y^2)) = c y^3 (kNewtonOffset - s2 y^2) with c in
// the staged mass; o is a stage32 record.
template <bool EPS0>
__device__ __forceinline__ void near_pair32(const nbk::P4& me, const nbk::P4& o, float k1, float k0, float eps2,
                                            float& ax, float& ay, float& az) {
    const float dx = (float)(o.x - me.x), dy = (float)(o.y - me.y), dz = (float)(o.z - me.z);
    const float s2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    const float u = fminf(fmaxf(__builtin_fmaf(s2, k1, k0), 0.0f), 1.0f);
    const float S = __builtin_fmaf(-u * u * u, __builtin_fmaf(u, __builtin_fmaf(6.0f, u, -15.0f), 10.0f), 1.0f);
    const float y = __builtin_amdgcn_rsqf(s2);
    const float t = y * y;
    const float w = t * y * __builtin_fmaf(-s2, t, nbk::kNewtonOffset);
    float f = staged_mass(o) * S * w;
    if (EPS0) f = (s2 > 0.0f) ? f : 0.0f;
    ax = __builtin_fmaf(f, dx, ax);
    ay = __builtin_fmaf(f, dy, ay);
    az = __builtin_fmaf(f, dz, az);
}

// A resident grid of blocks takes the chunks in list order from the work
// counter `next` until the count `*nchunks` is reached (every block leaves
// then), so the launch needs no count from the host; a chunk's sums do not
// depend on which block takes it.
//
// src: positions in the grid's order (the build's, or current ones gathered
// through it); keys: the build's cells.  A pair within rc now is within
// rc + d of the chunk's current bounding box at the build (d: the largest
// displacement since, *dmax), so the cells from cell(lo - reach) to
// cell(hi + reach), reach = rc (1 + 2^-20) + d, hold every partner (the cell
// maps are monotone; the margin covers the rounding of lo - reach).  The
// columns of that (x, y) range are dealt to the waves in turn (column c to
// wave c mod kWaves); a range wider than `brute` columns (a far-flung or
// non-finite state) is scanned as the whole sorted array instead, split
// evenly over the waves, so the work stays bounded.
template <bool F32, bool EPS0>
__global__ __launch_bounds__(64 * kWaves) void k_near(const nbk::P4* __restrict__ src,
                                                      const unsigned long long* __restrict__ keys,
                                                      const int* __restrict__ order, const int2* __restrict__ cols,
                                                      const int* __restrict__ chunks,
                                                      const int* __restrict__ nchunks, int* __restrict__ next,
                                                      const unsigned int* __restrict__ dmax, int n, double inv_h,
                                                      double rc, double rc2, Taper tp, double eps2, int brute,
                                                      nbk::P4* __restrict__ out) {
    __shared__ nbk::P4 stage[kWaves][64];
    __shared__ nbk::P4 part[kWaves][64];
    __shared__ int taken;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int nch = *nchunks;
    const double reach = rc * (1.0 + 0x1p-20) + (double)__uint_as_float(*dmax);
    for (;;) {
        if (threadIdx.x == 0) taken = atomicAdd(next, 1);
        __syncthreads();
        const int ch = taken;   // the same for the whole block
        if (ch >= nch) break;
        const int begin = chunks[ch];
        const int cnt = chunks[ch + 1] - begin;   // 1..64
        const nbk::P4 me = src[begin + (lane < cnt ? lane : 0)];
        double ax = 0.0, ay = 0.0, az = 0.0;
        // bounding box of the chunk's particles (lanes past the chunk repeat its first)
        const double3 lo = make_double3(wave_min(me.x), wave_min(me.y), wave_min(me.z));
        const double3 hi = make_double3(wave_max(me.x), wave_max(me.y), wave_max(me.z));
        nbk::P4* L = stage[w];
        const int xa = cell_xy(lo.x - reach, inv_h), xb = cell_xy(hi.x + reach, inv_h);
        const int ya = cell_xy(lo.y - reach, inv_h), yb = cell_xy(hi.y + reach, inv_h);
        const int za = cell_z(lo.z - reach, inv_h), zb = cell_z(hi.z + reach, inv_h);
        const int ny = yb - ya + 1, ncol = (xb - xa + 1) * ny;
        const bool all = ncol > brute;   // wave-uniform
        // column rounds: 64 columns of this wave's share per round; the whole
        // array: one round whose single "column" is this wave's fifth of it
        const int rounds = all ? 1 : (ncol - w + 64 * kWaves - 1) / (64 * kWaves);
        for (int r = 0; r < rounds; r++) {
            // lanes locate the z-runs of their columns c = w + kWaves (64 r + lane)
            int a = 0, b = 0;
            if (all) {
                a = (int)((long long)n * w / kWaves);
                b = (int)((long long)n * (w + 1) / kWaves);
            } else {
                const int c = w + kWaves * (64 * r + lane);
                if (c < ncol) {
                    const int X = xa + c / ny, Y = ya + c % ny;
                    const int2 run = cols[X * kColSide + Y];
                    a = lower_bound(keys, run.x, run.y, pack(X, Y, za));
                    b = lower_bound(keys, a, run.y, pack(X, Y, zb) + 1);
                }
            }
            const int kmax = all ? 1 : min(64, (ncol - w + kWaves - 1) / kWaves - 64 * r);
            for (int k = 0; k < kmax; k++) {
            const int ak = __shfl(a, k, 64), bk = __shfl(b, k, 64);
            for (int j0 = ak; j0 < bk; j0 += 64) {
                // keep only the particles within rc of the chunk's bounding box,
                // packed in order to the front of the stage
                const int jj = j0 + lane;
                nbk::P4 J = nbk::P4{0.0, 0.0, 0.0, 0.0};
                bool use = false;
                if (jj < bk) {
                    J = src[jj];
                    const double ex = fmax(fmax(lo.x - J.x, J.x - hi.x), 0.0);
                    const double ey = fmax(fmax(lo.y - J.y, J.y - hi.y), 0.0);
                    const double ez = fmax(fmax(lo.z - J.z, J.z - hi.z), 0.0);
                    use = ex * ex + ey * ey + ez * ez < rc2;
                }
                const unsigned long long keep = __ballot(use);
                const int nkeep = __popcll(keep);
                if (nkeep == 0) continue;   // wave-uniform
                const int m4 = (nkeep + 3) & ~3;   // fillers pad to a multiple of 4
                if (use) L[__popcll(keep & ((1ull << lane) - 1ull))] = F32 ? stage32(J) : stage64(J);
                // filler: this lane's position with zero mass, so it adds nothing
                // to any lane (d = 0 for this one, m = 0 for the others)
                if (lane >= nkeep && lane < m4) L[lane] = nbk::P4{me.x, me.y, me.z, 0.0};
                wave_sync();
                if constexpr (F32) {
                    const float k1 = (float)tp.k1, k0 = (float)tp.k0, e2 = (float)eps2;
                    float fx = 0.f, fy = 0.f, fz = 0.f;
                    for (int t = 0; t < m4; t += 4) {
                        nbk::P4 o[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) o[q] = L[t + q];
#pragma unroll
                        for (int q = 0; q < 4; q++) near_pair32<EPS0>(me, o[q], k1, k0, e2, fx, fy, fz);
                    }
                    ax += fx;
                    ay += fy;
                    az += fz;
                } else {
                    for (int t = 0; t < m4; t += 4) {
                        nbk::P4 o[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) o[q] = L[t + q];
#pragma unroll
                        for (int q = 0; q < 4; q++) near_pair<EPS0>(me, o[q], tp.k1, tp.k0, eps2, ax, ay, az);
                    }
                }
                wave_sync();
            }
            }
        }
        part[w][lane] = nbk::P4{ax, ay, az, 0.0};
        __syncthreads();
        if (w == 0 && lane < cnt) {
            double sx = 0.0, sy = 0.0, sz = 0.0;
            for (int q = 0; q < kWaves; q++) {
                sx += part[q][lane].x;
                sy += part[q][lane].y;
                sz += part[q][lane].z;
            }
            out[order[begin + lane]] = nbk::P4{sx, sy, sz, 0.0};
        }
        // the next chunk's part[] and `taken` are written only after every wave
        // has passed the barrier above and wave 0 has read part[]
    }
}

}  // namespace

hipError_t Near::near_accel(const nbk::P4* pm, int n, double rc, double eps, nbk::P4* out, hipStream_t st,
                            bool fp32, bool rebuild) {
    if (n <= 0) return hipSuccess;
    const double inv_h = 2.0 / rc;
    const int ncols = kColSide * kColSide;
    const dim3 grid((n + BLK - 1) / BLK);
    hipError_t e;
    const bool keep = !rebuild && built_n == n && built_rc == rc;
    if (!keep) built_n = 0;   // invalid until this build is queued
    // chunk buffer: n flags, then n + 1 scan results, the list (n + 1), the
    // count, the work counter, the displacement maximum
    auto slots = [&](int*& flag, int*& pos, int*& list, int*& nchunks, int*& next, unsigned int*& dmax) {
        flag = (int*)chunks.p;
        pos = flag + n;
        list = pos + n + 1;
        nchunks = list + n + 1;
        next = nchunks + 1;
        dmax = (unsigned int*)(next + 1);
    };
    auto launch_near = [&](const nbk::P4* src) -> hipError_t {
        int *flag, *pos, *list, *nchunks, *next;
        unsigned int* dmax;
        slots(flag, pos, list, nchunks, next, dmax);
        const int variant = (fp32 ? 2 : 0) + (eps > 0.0 ? 0 : 1);
        auto kernel = fp32 ? (eps > 0.0 ? k_near<true, false> : k_near<true, true>)
                           : (eps > 0.0 ? k_near<false, false> : k_near<false, true>);
        if (resident[variant] <= 0) {   // blocks the device holds at once: the persistent grid
            int dev = 0, ncu = 0, per_cu = 0;
            hipError_t q;
            if ((q = hipGetDevice(&dev)) != hipSuccess) return q;
            if ((q = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return q;
            if ((q = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kWaves, 0)) != hipSuccess)
                return q;
            resident[variant] = std::max(1, ncu * std::max(per_cu, 1));
        }
        const int blocks = std::min(resident[variant], n);   // never more chunks than particles
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(64 * kWaves), 0, st, src, (const unsigned long long*)keys_out.p,
                           (const int*)idx_out.p, (const int2*)cols.p, (const int*)list, (const int*)nchunks, next,
                           (const unsigned int*)dmax, n, inv_h, rc, rc * rc, make_taper(rc, eps), eps * eps,
                           std::max(n / 32, 256), out);
        return hipGetLastError();
    };
    if (keep) {
        if ((e = cur.ensure((size_t)n * sizeof(nbk::P4))) != hipSuccess) return e;
        int *flag, *pos, *list, *nchunks, *next;
        unsigned int* dmax;
        slots(flag, pos, list, nchunks, next, dmax);
        hipLaunchKernelGGL(k_reuse_reset, dim3(1), dim3(1), 0, st, next, dmax);
        hipLaunchKernelGGL(k_restage, grid, dim3(BLK), 0, st, pm, (const int*)idx_out.p, (const nbk::P4*)sorted.p, n,
                           (nbk::P4*)cur.p, dmax);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        return launch_near((const nbk::P4*)cur.p);
    }
    if ((e = keys_in.ensure((size_t)n * 8)) != hipSuccess) return e;
    if ((e = keys_out.ensure((size_t)n * 8)) != hipSuccess) return e;
    if ((e = idx_in.ensure((size_t)n * 4)) != hipSuccess) return e;
    if ((e = idx_out.ensure((size_t)n * 4)) != hipSuccess) return e;
    if ((e = sorted.ensure((size_t)n * sizeof(nbk::P4))) != hipSuccess) return e;
    if ((e = cols.ensure((size_t)ncols * sizeof(int2))) != hipSuccess) return e;
    if ((e = chunks.ensure((size_t)(4 * n + 3) * 4 + 64)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_keys, grid, dim3(BLK), 0, st, pm, n, inv_h, (unsigned long long*)keys_in.p, (int*)idx_in.p);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = 0;
    e = rocprim::radix_sort_pairs(nullptr, bytes, (const unsigned long long*)keys_in.p, (unsigned long long*)keys_out.p,
                                  (const int*)idx_in.p, (int*)idx_out.p, (unsigned int)n, 0, 43, st);
    if (e != hipSuccess) return e;
    if ((e = temp.ensure(bytes + 16)) != hipSuccess) return e;
    e = rocprim::radix_sort_pairs(temp.p, bytes, (const unsigned long long*)keys_in.p, (unsigned long long*)keys_out.p,
                                  (const int*)idx_in.p, (int*)idx_out.p, (unsigned int)n, 0, 43, st);
    if (e != hipSuccess) return e;
    const unsigned long long* keys = (const unsigned long long*)keys_out.p;
    int *flag, *pos, *list, *nchunks, *next;
    unsigned int* dmax;
    slots(flag, pos, list, nchunks, next, dmax);
    hipLaunchKernelGGL(k_gather_sorted, grid, dim3(BLK), 0, st, pm, (const int*)idx_out.p, n, (nbk::P4*)sorted.p);
    hipLaunchKernelGGL(k_clear_cols, dim3((ncols + BLK - 1) / BLK), dim3(BLK), 0, st, (int2*)cols.p, ncols);
    hipLaunchKernelGGL(k_col_runs, grid, dim3(BLK), 0, st, keys, n, (int2*)cols.p);
    if ((e = hipMemsetAsync(flag, 0, (size_t)n * sizeof(int), st)) != hipSuccess) return e;
    const int ngroups = ncols / kGroup;
    hipLaunchKernelGGL(k_chunk_flags, dim3((ngroups + BLK - 1) / BLK), dim3(BLK), 0, st, (const int2*)cols.p, ngroups,
                       flag);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t sbytes = 0;
    e = rocprim::exclusive_scan(nullptr, sbytes, (const int*)flag, pos, 0, (size_t)n, rocprim::plus<int>(), st);
    if (e != hipSuccess) return e;
    if ((e = scan_temp.ensure(sbytes + 16)) != hipSuccess) return e;
    e = rocprim::exclusive_scan(scan_temp.p, sbytes, (const int*)flag, pos, 0, (size_t)n, rocprim::plus<int>(), st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_chunk_list, grid, dim3(BLK), 0, st, (const int*)flag, (const int*)pos, n, list, nchunks,
                       next, dmax);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    built_n = n;
    built_rc = rc;
    return launch_near((const nbk::P4*)sorted.p);
}

void Near::release() {
    for (nbh::DevBuf* b : {&keys_in, &keys_out, &idx_in, &idx_out, &sorted, &temp, &scan_temp, &cols, &chunks, &cur})
        b->release();
    built_n = 0;
}

}  // namespace nbn
