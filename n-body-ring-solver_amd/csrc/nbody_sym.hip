// nbody_sym.hip -- symmetric (Newton's third law) force kernel for gfx950.
//
// Every unordered pair is evaluated once and its force applied to both
// particles, the GPU analogue of the reference's reduced solver
// (CPU/force_calculation.cpp:530-531, 617-630) and of the FPGA PE_2x2
// (FPGA/device/ring_lrbd_lf_no_sync_local.cl:161-184).
//
// Blocks of S = 64*K particles.  A work item (A, b0, b1) is one wave64 that
// holds the K particles of block A it owns per lane in registers (lane l owns
// A*S + k*64 + l) with their action accumulators, and sweeps the j-blocks
// B in [b0, b1), B >= A.  Each j-block is taken in K chunks of 64: the chunk
// is staged (twice, so indices never wrap) in the wave's LDS slot, and over 64
// steps lane l pairs its K particles with chunk particle (l + t) mod 64 read
// from LDS while the chunk's reaction accumulators travel with their particle
// through the lanes: after every step they rotate by one lane (DPP wave_rol:1,
// lane l receives lane l+1), so after 64 steps each sits in its home lane
// again.  Diagonal tiles (B == A) mask pairs with i >= j.
//
// Outputs are deterministic partial sums, no atomics.  Reactions of block B
// from block A (A < B) go to the triangular slot B(B-1)/2 + A of the reaction
// region (FP64 records in react64, or the FP32 sums of the mixed forms as
// float4 in react32); the actions of item (A, b0, ..) go to action slot
// A*W + (b0 - A)/L of part, W = (nB - 1)/L + 1 slots per block (the item
// carries that index); diagonal reactions fold into the actions.  Each slot
// holds S records, one per particle of the block.
// k_sym_kick_drift / k_sym_reduce add a block's slots in a fixed order.
//
// Two-set form (TWO): one round of the reference's ring (hybrid_ring_solver,
// CPU/hybrid_solver.cpp:81-95 over compute_forces_avx2_blocked,
// CPU/force_calculation.cpp:477-690): i-blocks come from the local set pm,
// j-chunks from the received set pj (same count n), every pair of the local
// particle k and the received particle q with q > k -- q >= k when `incl` --
// evaluated once.  The diagonal tile's reactions belong to the received set's
// block A and go to their own slot nB(nB-1)/2 + A instead of folding into the
// actions; k_ring_finish adds the local actions and the received reactions
// and accumulates m a into the two force arrays.
#include <hip/hip_runtime.h>

#include "nbody_kernels.h"

namespace nbk {

namespace {

constexpr int kWavesPerGroup = 4;

// Lane l receives the value held by lane l+1 (mod 64): DPP wave_rol:1, every
// lane enabled, so no lane keeps an old value.
constexpr int kDppWaveRol1 = 0x134;

__device__ __forceinline__ double rotate_down(double v) {
    const int2 b = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_mov_dpp(b.x, kDppWaveRol1, 0xF, 0xF, false);
    r.y = __builtin_amdgcn_mov_dpp(b.y, kDppWaveRol1, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}

__device__ __forceinline__ float rotate_down(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), kDppWaveRol1, 0xF, 0xF,
                                                              false));
}

// One unordered pair, FP64.  d = x_j - x_i; w = 1/(|d|^2+eps^2)^(3/2) from
// v_rsq_f64 and the second-order correction (1-e)^(-3/2) ~ 1 + 3/2 e + 15/8 e^2
// written as ((s + cshift)^2 + coff), s = r2 y^2 (kCubicShift, kCubicOffset),
// with the 15/8 already in both masses; a_i += m_j w d, b_j -= m_i w d.
// `keep` zeroes masked pairs.  MUL false (every mass equal, factored out of
// the sums): a_i += w d, b_j -= w d, the masses unused.
template <bool EPS0, bool MASK, bool MUL = true>
__device__ __forceinline__ void sym64(double xi, double yi, double zi, double mi, double xj, double yj,
                                      double zj, double mj, double eps2, double cshift, double coff, bool keep,
                                      double& ax, double& ay, double& az, double& bx, double& by,
                                      double& bz) {
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, __builtin_fma(dx, dx, eps2)));
    const double y = __builtin_amdgcn_rsq(r2);
    const double t = y * y;
    const double u = __builtin_fma(r2, t, cshift);
    const double p = __builtin_fma(u, u, coff);
    double w = (t * y) * p;
    if (EPS0) w = (r2 > 0.0) ? w : 0.0;
    if (MASK) w = keep ? w : 0.0;
    const double si = MUL ? mj * w : w, sj = MUL ? mi * w : w;
    ax = __builtin_fma(si, dx, ax);
    ay = __builtin_fma(si, dy, ay);
    az = __builtin_fma(si, dz, az);
    bx = __builtin_fma(-sj, dx, bx);
    by = __builtin_fma(-sj, dy, by);
    bz = __builtin_fma(-sj, dz, bz);
}

// Mixed: FP64 difference rounded to FP32, FP32 rsqrt and pair arithmetic,
// FP32 sums (flushed to FP64 per 64-particle chunk).  With NR the Newton-Raphson
// step is applied to the cube: y = rsq(r2) (1 ulp), e = 1 - r2 y^2, and
// r^-3 = y^3 (1 - e)^(-3/2) = y^3 (1 + 3/2 e) + O(e^2) -- the same first-order
// gain as refining y by y (3 - r2 y^2)/2 and cubing, for two instructions
// fewer: 1 + 3/2 e = c (F - r2 y^2) with F = kNewtonOffset and c = kNewtonMass
// already in both masses.
template <bool EPS0, bool MASK, bool NR, bool MUL = true>
__device__ __forceinline__ void sym32(double xi, double yi, double zi, float mi, double xj, double yj,
                                      double zj, float mj, float eps2, bool keep, float& ax, float& ay,
                                      float& az, float& bx, float& by, float& bz) {
    const float dx = (float)(xj - xi), dy = (float)(yj - yi), dz = (float)(zj - zi);
    const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, eps2)));
    const float y = __builtin_amdgcn_rsqf(r2);
    const float t = y * y;
    float w = t * y;
    if (NR) w = w * __builtin_fmaf(-r2, t, kNewtonOffset);
    if (EPS0) w = (r2 > 0.0f) ? w : 0.0f;
    if (MASK) w = keep ? w : 0.0f;
    const float si = MUL ? mj * w : w, sj = MUL ? mi * w : w;
    ax = __builtin_fmaf(si, dx, ax);
    ay = __builtin_fmaf(si, dy, ay);
    az = __builtin_fmaf(si, dz, az);
    bx = __builtin_fmaf(-sj, dx, bx);
    by = __builtin_fmaf(-sj, dy, by);
    bz = __builtin_fmaf(-sj, dz, bz);
}

// Order this lane's LDS accesses against the other lanes' of the same wave
// (LDS instructions of one wave execute in order; the fences stop the compiler
// from moving loads across the point where the chunk is (re)written).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ P4 load_or_zero(const P4* __restrict__ p, int j, int n) {
    return (j < n) ? p[j] : P4{0.0, 0.0, 0.0, 0.0};
}

}  // namespace

// One chunk (64 records staged twice in lds) against the lane's K particles;
// the chunk's reaction sums end in their home lanes.  DIAG: the chunk is the
// c-th of the wave's own block (pairs i < j only; i <= j with incl).  MUL
// false: the pair terms without masses (equal masses, padding-free chunk).
template <int FORM, bool EPS0, int K, bool DIAG, bool MUL = true>
__device__ __forceinline__ void sweep_chunk(const P4* __restrict__ lds, int lane, int c, int tmin,
                                            const double (&xi)[K], const double (&yi)[K], const double (&zi)[K],
                                            const double (&mi)[K], const float (&mif)[K], double eps2,
                                            double cshift, double coff, double (&ax)[K], double (&ay)[K],
                                            double (&az)[K], double& bx, double& by, double& bz) {
    constexpr bool MIXED = (FORM != FORM_FP64);
    constexpr bool NR = (FORM == FORM_MIXED_NR);
    if constexpr (!MIXED) {
        bx = by = bz = 0.0;
#pragma unroll 2
        for (int t = 0; t < 64; t++) {
            const P4 J = lds[lane + t];
            const bool same_ok = (t >= tmin) && (lane + t < 64);
#pragma unroll
            for (int k = 0; k < K; k++)
                sym64<EPS0, DIAG, MUL>(xi[k], yi[k], zi[k], mi[k], J.x, J.y, J.z, J.m, eps2, cshift, coff,
                                       (k < c) || (k == c && same_ok), ax[k], ay[k], az[k], bx, by, bz);
            bx = rotate_down(bx); by = rotate_down(by); bz = rotate_down(bz);
        }
    } else {
        const float eps2f = (float)eps2;
        float fbx = 0.f, fby = 0.f, fbz = 0.f;
        float fax[K], fay[K], faz[K];
#pragma unroll
        for (int k = 0; k < K; k++) fax[k] = fay[k] = faz[k] = 0.f;
#pragma unroll 1
        for (int t = 0; t < 64; t++) {
            const P4 J = lds[lane + t];
            const float mj = (float)J.m;
            const bool same_ok = (t >= tmin) && (lane + t < 64);
#pragma unroll
            for (int k = 0; k < K; k++)
                sym32<EPS0, DIAG, NR, MUL>(xi[k], yi[k], zi[k], mif[k], J.x, J.y, J.z, mj, eps2f,
                                           (k < c) || (k == c && same_ok), fax[k], fay[k], faz[k], fbx, fby, fbz);
            fbx = rotate_down(fbx); fby = rotate_down(fby); fbz = rotate_down(fbz);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            ax[k] += (double)fax[k];
            ay[k] += (double)fay[k];
            az[k] += (double)faz[k];
        }
        bx = (double)fbx; by = (double)fby; bz = (double)fbz;
    }
}

template <int FORM, bool EPS0, int K, bool TWO, bool EQM>
__global__ __launch_bounds__(64 * kWavesPerGroup) void k_accel_sym(const P4* __restrict__ pm,
                                                                   const P4* __restrict__ pj, int n,
                                                                   const int4* __restrict__ items, int nitems,
                                                                   int nB, double eps2, double cshift,
                                                                   double coff, int incl, P4* __restrict__ part,
                                                                   P4* __restrict__ react64,
                                                                   float4* __restrict__ react32) {
    constexpr bool MIXED = (FORM != FORM_FP64);
    constexpr bool NR = (FORM == FORM_MIXED_NR);
    constexpr int S = 64 * K;
    // the factor of the correction polynomial carried by the masses; EQM
    // (every mass equal, m > 0): the sums leave the factor m MSCALE to the
    // finish, a "mass" is 1 for a particle and 0 for padding, and the chunks
    // with no padding on either side take the pair terms without it
    constexpr double MSCALE = MIXED ? (NR ? kNewtonMass : 1.0) : kCubicMass;
    auto mass_of = [](double m) { return EQM ? (m != 0.0 ? 1.0 : 0.0) : m * MSCALE; };
    __shared__ P4 jt[kWavesPerGroup][128];   // each wave's chunk, stored twice

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int item = blockIdx.x * kWavesPerGroup + wave;
    if (item >= nitems) return;   // the whole wave leaves; no workgroup barrier is used below
    const int4 it = items[item];
    const int A = it.x, b0 = it.y, b1 = it.z, act = it.w;
    P4* __restrict__ lds = jt[wave];

    double xi[K], yi[K], zi[K], mi[K], ax[K], ay[K], az[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const P4 p = load_or_zero(pm, A * S + k * 64 + lane, n);
        xi[k] = p.x; yi[k] = p.y; zi[k] = p.z; mi[k] = mass_of(p.m);
        ax[k] = ay[k] = az[k] = 0.0;
    }
    float mif[K];
#pragma unroll
    for (int k = 0; k < K; k++) mif[k] = (float)mi[k];

    for (int B = b0; B < b1; B++) {
        const bool diag = (B == A);
        for (int c = 0; c < K; c++) {
            P4 q = load_or_zero(TWO ? pj : pm, B * S + c * 64 + lane, n);
            q.m = mass_of(q.m);
            lds[lane] = q;
            lds[lane + 64] = q;
            wave_sync();
            double bx = 0.0, by = 0.0, bz = 0.0;
            // padding: in the wave's own block (the last) or in this chunk
            const bool pad = (A + 1) * S > n || B * S + (c + 1) * 64 > n;
            if (diag)
                sweep_chunk<FORM, EPS0, K, true>(lds, lane, c, TWO ? 1 - incl : 1, xi, yi, zi, mi, mif, eps2, cshift,
                                                 coff, ax, ay, az, bx, by, bz);
            else if (EQM && !pad)
                sweep_chunk<FORM, EPS0, K, false, false>(lds, lane, c, 1, xi, yi, zi, mi, mif, eps2, cshift, coff, ax,
                                                         ay, az, bx, by, bz);
            else
                sweep_chunk<FORM, EPS0, K, false>(lds, lane, c, 1, xi, yi, zi, mi, mif, eps2, cshift, coff, ax, ay,
                                                  az, bx, by, bz);
            if (diag && !TWO) {
                // the chunk is this wave's own particles k == c: fold the reactions in
#pragma unroll
                for (int k = 0; k < K; k++)
                    if (k == c) {
                        ax[k] += bx; ay[k] += by; az[k] += bz;
                    }
            } else {
                // reactions of the mixed forms are FP32 sums: stored as such
                // (16 B), the FP64 form's as records (32 B); the two-set
                // form's diagonal reactions after the triangle
                const long tri = diag ? (long)nB * (nB - 1) / 2 + A : (long)B * (B - 1) / 2 + A;
                const long slot = tri * S + c * 64 + lane;
                if constexpr (MIXED) react32[slot] = make_float4((float)bx, (float)by, (float)bz, 0.0f);
                else react64[slot] = P4{bx, by, bz, 0.0};
            }
            wave_sync();
        }
    }
#pragma unroll
    for (int k = 0; k < K; k++) part[(long)act * S + k * 64 + lane] = P4{ax[k], ay[k], az[k], 0.0};
}

// ---------------------------------------------------------------------------
// Grouped, persistent form.  A workgroup of W waves holds W consecutive
// blocks (a group g: blocks gW .. gW+W-1, wave w block A = gW + w) and walks a
// contiguous range of the chunk sequence (g, B, c) -- rows g in order, j-blocks
// B = gW .. nB-1, chunks c = 0 .. K-1 -- that the host cut into equal parts,
// one per resident workgroup.  Per chunk: the workgroup stages the 64 records
// once (its 256 doubles spread over the threads, prefetched a chunk ahead), every wave sweeps
// them against its own block (B < A: nothing; B == A: the diagonal mask), the
// waves' reaction sums meet in LDS and wave 0 adds them in wave order (FP64)
// into reaction slot (g, B), g nB - W g(g-1)/2 + (B - gW): three component
// planes of S values (x, y, z; doubles, floats for the mixed forms), the
// chunk's 64 at c*64.  The actions stay in registers along the row and are
// written once per row segment, to action slot seg (W blocks, each three
// planes of S doubles); the segments of a row are consecutive.  No padding
// word is stored: 24 (12) bytes per partial sum.  Sums are in a fixed order:
// results are bitwise reproducible.
// ---------------------------------------------------------------------------

template <int FORM, bool EPS0, int K, int W, bool TWO, bool EQM>
__global__ __launch_bounds__(64 * W) void k_accel_symg(const P4* __restrict__ pm, const P4* __restrict__ pj, int n,
                                                       const int4* __restrict__ ranges, int nB, double eps2,
                                                       double cshift, double coff, int incl,
                                                       double* __restrict__ act, double* __restrict__ react64,
                                                       float* __restrict__ react32) {
    constexpr bool MIXED = (FORM != FORM_FP64);
    constexpr bool NR = (FORM == FORM_MIXED_NR);
    constexpr int S = 64 * K;
    constexpr double MSCALE = MIXED ? (NR ? kNewtonMass : 1.0) : kCubicMass;
    // EQM (every mass equal, m > 0): the sums leave the factor m MSCALE to the
    // finish; a staged "mass" is 1 for a particle and 0 for padding, and the
    // chunks with no padding on either side take the pair terms without it
    auto mass_of = [](double m) { return EQM ? (m != 0.0 ? 1.0 : 0.0) : m * MSCALE; };
    __shared__ P4 lds[128];             // the chunk, stored twice so lane + t never wraps
    __shared__ double red[W][3][64];    // the waves' reaction sums of the chunk

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x;
    const int4 rg = ranges[blockIdx.x];   // (g, B * K + c, chunks, first action slot)
    int g = rg.x, B = rg.y / K, c = rg.y - (rg.y / K) * K, slot = rg.w;
    const int count = rg.z;
    if (count <= 0) return;   // the whole workgroup leaves together
    const P4* __restrict__ src = TWO ? pj : pm;
    const int tmin = TWO ? 1 - incl : 1;

    double xi[K], yi[K], zi[K], mi[K], ax[K], ay[K], az[K];
    float mif[K];
    auto load_block = [&](int A) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const P4 p = load_or_zero(pm, A * S + k * 64 + lane, n);
            xi[k] = p.x; yi[k] = p.y; zi[k] = p.z; mi[k] = mass_of(p.m); mif[k] = (float)mi[k];
            ax[k] = ay[k] = az[k] = 0.0;
        }
        // settle these loads here (once per row): otherwise the sweeps, which
        // read them, wait at their first pair for every load in flight --
        // the next chunk's prefetch included (s_waitcnt vmcnt(0): vmcnt 0,
        // expcnt and lgkmcnt at their maxima)
        __builtin_amdgcn_s_waitcnt(0x0F70);
    };
    auto flush = [&]() {
        double* __restrict__ o = act + ((long)slot * W + wave) * (3 * S) + lane;
#pragma unroll
        for (int k = 0; k < K; k++) {
            o[k * 64] = ax[k];
            o[S + k * 64] = ay[k];
            o[2 * S + k * 64] = az[k];
        }
    };
    // the chunk's 256 doubles spread over the threads (NPT each; indexes past
    // 255 stage nothing); the mass takes its factor when staged, so nothing
    // waits for the load before the sweep it overlaps
    constexpr int NPT = (256 + 64 * W - 1) / (64 * W);
    auto fetch = [&](int b, int ch, double (&v)[NPT]) {
#pragma unroll
        for (int d = 0; d < NPT; d++) {
            const int idx = tid + d * 64 * W;
            const int j = b * S + ch * 64 + (idx >> 2);
            v[d] = (idx < 256 && j < n) ? ((const double*)(src + j))[idx & 3] : 0.0;
        }
    };
    // wave 0's summed reactions of the previous chunk: stored after the next
    // barrier, so the store is in flight during a sweep instead of being
    // waited for at a barrier.  FP64: lane l holds chunk particle l's three
    // sums (three 512-byte plane stores).  Mixed forms: lanes 0-47 hold four
    // consecutive particles' sums of one component (lane l: component l / 16,
    // particles 4 (l mod 16) ..), one 16-byte store each, so every 64 bytes
    // of a plane come from one quarter-wave (wave-wide 4-byte plane stores,
    // waited for at the next chunk, ran the kernel 9 % slower).
    long pend_at = -1;
    double px = 0.0, py = 0.0, pz = 0.0;
    float4 pq = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto store_pending = [&]() {
        if (pend_at < 0) return;
        if constexpr (MIXED) {
            *reinterpret_cast<float4*>(react32 + pend_at) = pq;
        } else {
            react64[pend_at] = px;
            react64[pend_at + S] = py;
            react64[pend_at + 2 * S] = pz;
        }
    };
    int A = g * W + wave;
    load_block(A);
    double nxt[NPT];
    fetch(B, c, nxt);
    for (int step = 0; step < count; step++) {
#pragma unroll
        for (int d = 0; d < NPT; d++) {
            const int idx = tid + d * 64 * W;
            if (idx < 256) {
                double* l = (double*)lds;
                const double v = (idx & 3) == 3 ? mass_of(nxt[d]) : nxt[d];
                l[idx] = v;
                l[idx + 256] = v;
            }
        }
        __syncthreads();   // the chunk is staged; the last chunk's reaction sums have been read
        // prefetch the next chunk (past the last row it lies beyond n: zeros,
        // nothing loaded), then store the last chunk's reactions
        int g2 = g, B2 = B, c2 = c + 1;
        if (c2 == K) {
            c2 = 0;
            if (++B2 == nB) { ++g2; B2 = g2 * W; }
        }
        fetch(B2, c2, nxt);
        if (wave == 0) store_pending();
        double bx = 0.0, by = 0.0, bz = 0.0;
        // padding: in the wave's own block (the last) or in this chunk
        const bool pad = (A + 1) * S > n || B * S + (c + 1) * 64 > n;
        if (B > A && EQM && !pad)
            sweep_chunk<FORM, EPS0, K, false, false>(lds, lane, c, tmin, xi, yi, zi, mi, mif, eps2, cshift, coff, ax,
                                                     ay, az, bx, by, bz);
        else if (B > A)
            sweep_chunk<FORM, EPS0, K, false>(lds, lane, c, tmin, xi, yi, zi, mi, mif, eps2, cshift, coff, ax, ay, az,
                                              bx, by, bz);
        else if (B == A) sweep_chunk<FORM, EPS0, K, true>(lds, lane, c, tmin, xi, yi, zi, mi, mif, eps2, cshift, coff,
                                                          ax, ay, az, bx, by, bz);
        red[wave][0][lane] = bx;
        red[wave][1][lane] = by;
        red[wave][2][lane] = bz;
        __syncthreads();   // every wave's reactions of this chunk are in LDS; the chunk is free
        if (wave == 0) {
            const long rslot = (long)g * nB - (long)W * g * (g - 1) / 2 + (B - g * W);
            if constexpr (MIXED) {
                if (lane < 48) {
                    const int comp = lane >> 4, p0 = (lane & 15) * 4;
                    double q[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                    for (int w = 0; w < W; w++) {
#pragma unroll
                        for (int u = 0; u < 4; u++) q[u] += red[w][comp][p0 + u];
                    }
                    pq = make_float4((float)q[0], (float)q[1], (float)q[2], (float)q[3]);
                    pend_at = rslot * (3 * S) + comp * S + c * 64 + p0;
                }
            } else {
                px = py = pz = 0.0;
#pragma unroll
                for (int w = 0; w < W; w++) {
                    px += red[w][0][lane]; py += red[w][1][lane]; pz += red[w][2][lane];
                }
                pend_at = rslot * (3 * S) + c * 64 + lane;
            }
        }
        if (g2 != g) {   // the row ends: its actions go to their slot, the next row's block is loaded
            flush();
            ++slot;
            A = g2 * W + wave;
            load_block(A);
        }
        g = g2; B = B2; c = c2;
    }
    if (wave == 0) store_pending();
    if (c != 0 || B != g * W) flush();   // the range ended inside a row
}

// Sums of the grouped form for particle i of block X (group gX = X / W):
// reactions from slots (g, X), g = 0 .. gX; actions from the segments of row
// gX (slots seg0 .. seg0 + nseg - 1, the wave X - gX W block of each).  Each
// slot is three planes of S values; eight slots' values are loaded together
// (in flight at once) and added one by one in slot order.
template <typename T, int U>
__device__ __forceinline__ void add_planes(const T* __restrict__ base, long stride, int count, int S, double& x,
                                           double& y, double& z) {
    int q = 0;
    for (; q + U <= count; q += U) {
        T vx[U], vy[U], vz[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const T* p = base + (long)(q + u) * stride;
            vx[u] = p[0];
            vy[u] = p[S];
            vz[u] = p[2 * S];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            x += (double)vx[u]; y += (double)vy[u]; z += (double)vz[u];
        }
    }
    for (; q < count; q++) {
        const T* p = base + (long)q * stride;
        x += (double)p[0]; y += (double)p[S]; z += (double)p[2 * S];
    }
}

template <bool F32R, int W>
__device__ __forceinline__ void symg_react(const double* __restrict__ react64, const float* __restrict__ react32,
                                           int nB, int S, int i, double& bx, double& by, double& bz) {
    constexpr int U = 8;
    const int X = i / S, r = i - X * S, gX = X / W;
    bx = by = bz = 0.0;
    // slot (g, X) = g nB - W g(g-1)/2 + (X - g W): consecutive g are nB - W g
    // slots apart, so the slots are visited by a running index
    long slot = X;
    int g = 0;
    for (; g + U <= gX + 1; g += U) {
        long at[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            at[u] = slot * (3L * S) + r;
            slot += nB - (long)W * (g + u) - W;
        }
        if constexpr (F32R) {
            float vx[U], vy[U], vz[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                vx[u] = react32[at[u]]; vy[u] = react32[at[u] + S]; vz[u] = react32[at[u] + 2 * S];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                bx += (double)vx[u]; by += (double)vy[u]; bz += (double)vz[u];
            }
        } else {
            double vx[U], vy[U], vz[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                vx[u] = react64[at[u]]; vy[u] = react64[at[u] + S]; vz[u] = react64[at[u] + 2 * S];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                bx += vx[u]; by += vy[u]; bz += vz[u];
            }
        }
    }
    for (; g <= gX; g++) {
        const long at = slot * (3L * S) + r;
        slot += nB - (long)W * g - W;
        if constexpr (F32R) {
            bx += (double)react32[at]; by += (double)react32[at + S]; bz += (double)react32[at + 2 * S];
        } else {
            bx += react64[at]; by += react64[at + S]; bz += react64[at + 2 * S];
        }
    }
}

template <int W>
__device__ __forceinline__ void symg_act(const double* __restrict__ act, const int2* __restrict__ rowseg, int S,
                                         int i, double& ax, double& ay, double& az) {
    const int X = i / S, r = i - X * S, gX = X / W, w = X - gX * W;
    const int2 sg = rowseg[gX];
    ax = ay = az = 0.0;
    add_planes<double, 8>(act + ((long)sg.x * W + w) * (3L * S) + r, (long)W * 3 * S, sg.y, S, ax, ay, az);
}

// The finish of particle i: its sums, the kick and the drift.
template <bool F32R, int W, bool KICK>
__device__ __forceinline__ void symg_finish_one(const double* __restrict__ act, const int2* __restrict__ rowseg,
                                                const double* __restrict__ react64,
                                                const float* __restrict__ react32, int nB, int S, int i,
                                                double scale, P4* __restrict__ pm, P4* __restrict__ vel,
                                                P4* __restrict__ acc, double kdt, double c1dt, double c2dt) {
    double bx, by, bz, ax, ay, az;
    symg_react<F32R, W>(react64, react32, nB, S, i, bx, by, bz);
    symg_act<W>(act, rowseg, S, i, ax, ay, az);
    ax = scale * (ax + bx);   // scale: the factored-out mass (1 when the sums hold it)
    ay = scale * (ay + by);
    az = scale * (az + bz);
    acc[i] = P4{ax, ay, az, 0.0};
    if constexpr (KICK) {
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
}

// With `pair`, thread t finishes particles t and n-1-t: a particle of block
// X adds X / W + 1 reaction slots and its row's action segments, which grow
// from the first blocks' ~95 to the last blocks' ~170 slots (N = 262144,
// mixed), so one particle from each end gives every thread about the same
// work (every wave of this kernel is resident at once, so it lasts as long
// as its heaviest wave).  Else thread t finishes particle t.  Each
// particle's sums keep their order.
template <bool F32R, int W, bool KICK>
__global__ __launch_bounds__(BLOCK) void k_symg_finish(const double* __restrict__ act,
                                                       const int2* __restrict__ rowseg,
                                                       const double* __restrict__ react64,
                                                       const float* __restrict__ react32, int nB, int S, int n,
                                                       int pair, double scale, P4* __restrict__ pm,
                                                       P4* __restrict__ vel, P4* __restrict__ acc, double kdt,
                                                       double c1dt, double c2dt) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (pair ? (n + 1) / 2 : n)) return;
    symg_finish_one<F32R, W, KICK>(act, rowseg, react64, react32, nB, S, t, scale, pm, vel, acc, kdt, c1dt, c2dt);
    if (pair && n - 1 - t != t)
        symg_finish_one<F32R, W, KICK>(act, rowseg, react64, react32, nB, S, n - 1 - t, scale, pm, vel, acc, kdt,
                                       c1dt, c2dt);
}

// Two-set finish of the grouped form: local forces from the actions, the
// received set's from the reactions, times the particles' masses, added;
// with `pair` thread t takes particles t and n-1-t, as k_symg_finish.
template <bool F32R, int W>
__device__ __forceinline__ void symg_ring_finish_one(const double* __restrict__ act,
                                                     const int2* __restrict__ rowseg,
                                                     const double* __restrict__ react64,
                                                     const float* __restrict__ react32, int nB, int S, int i,
                                                     const P4* __restrict__ pl, const P4* __restrict__ pr,
                                                     P4* __restrict__ fl, P4* __restrict__ fr) {
    double bx, by, bz, ax, ay, az;
    symg_react<F32R, W>(react64, react32, nB, S, i, bx, by, bz);
    symg_act<W>(act, rowseg, S, i, ax, ay, az);
    const double ml = pl[i].m, mr = pr[i].m;
    P4 f = fl[i];
    f.x += ml * ax; f.y += ml * ay; f.z += ml * az;
    fl[i] = f;
    P4 h = fr[i];
    h.x += mr * bx; h.y += mr * by; h.z += mr * bz;
    fr[i] = h;
}

template <bool F32R, int W>
__global__ __launch_bounds__(BLOCK) void k_symg_ring_finish(const double* __restrict__ act,
                                                            const int2* __restrict__ rowseg,
                                                            const double* __restrict__ react64,
                                                            const float* __restrict__ react32, int nB, int S, int n,
                                                            int pair, const P4* __restrict__ pl,
                                                            const P4* __restrict__ pr, P4* __restrict__ fl,
                                                            P4* __restrict__ fr) {
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if (t >= (pair ? (n + 1) / 2 : n)) return;
    symg_ring_finish_one<F32R, W>(act, rowseg, react64, react32, nB, S, t, pl, pr, fl, fr);
    if (pair && n - 1 - t != t)
        symg_ring_finish_one<F32R, W>(act, rowseg, react64, react32, nB, S, n - 1 - t, pl, pr, fl, fr);
}

// a_i for i in block X = sum of the reactions from blocks Y = 0 .. X-1, then
// of the actions of the items (X, b0 = X, X+L, ..) of row X, in that order.
// F32R: the reactions are the FP32 records of react32 (the mixed forms), else
// the P4 records of react64.  Slots are loaded eight at a time (the loads in
// flight together) and added one by one in slot order.
template <bool F32R>
__device__ __forceinline__ void sym_sum(const P4* __restrict__ part, const P4* __restrict__ react64,
                                        const float4* __restrict__ react32, int nB, int S, int L, int i,
                                        double& ax, double& ay, double& az) {
    constexpr int U = 8;
    const int X = i / S, r = i - X * S;
    ax = ay = az = 0.0;
    const long rrow = (long)X * (X - 1) / 2 * S + r;
    int Y = 0;
    if constexpr (F32R) {
        for (; Y + U <= X; Y += U) {
            float4 q[U];
#pragma unroll
            for (int u = 0; u < U; u++) q[u] = react32[rrow + (long)(Y + u) * S];
#pragma unroll
            for (int u = 0; u < U; u++) {
                ax += (double)q[u].x; ay += (double)q[u].y; az += (double)q[u].z;
            }
        }
        for (; Y < X; Y++) {
            const float4 q = react32[rrow + (long)Y * S];
            ax += (double)q.x; ay += (double)q.y; az += (double)q.z;
        }
    } else {
        for (; Y + U <= X; Y += U) {
            P4 q[U];
#pragma unroll
            for (int u = 0; u < U; u++) q[u] = react64[rrow + (long)(Y + u) * S];
#pragma unroll
            for (int u = 0; u < U; u++) {
                ax += q[u].x; ay += q[u].y; az += q[u].z;
            }
        }
        for (; Y < X; Y++) {
            const P4 p = react64[rrow + (long)Y * S];
            ax += p.x; ay += p.y; az += p.z;
        }
    }
    const int W = (nB - 1) / L + 1;
    const int cnt = (nB - 1 - X) / L + 1;   // items of row X
    const long arow = (long)X * W * S + r;
    int j = 0;
    for (; j + U <= cnt; j += U) {
        P4 q[U];
#pragma unroll
        for (int u = 0; u < U; u++) q[u] = part[arow + (long)(j + u) * S];
#pragma unroll
        for (int u = 0; u < U; u++) {
            ax += q[u].x; ay += q[u].y; az += q[u].z;
        }
    }
    for (; j < cnt; j++) {
        const P4 p = part[arow + (long)j * S];
        ax += p.x; ay += p.y; az += p.z;
    }
}

// Two-set finish: the local particle i's actions (row X) and the received
// particle i's reactions (triangular slots Y < X, then the diagonal slot),
// each in slot order, times the particle's mass, added to the force records.
template <bool F32R>
__device__ __forceinline__ void ring_react_sum(const P4* __restrict__ react64, const float4* __restrict__ react32,
                                               int nB, int S, int i, double& bx, double& by, double& bz) {
    const int X = i / S, r = i - X * S;
    bx = by = bz = 0.0;
    const long rrow = (long)X * (X - 1) / 2 * S + r;
    const long dslot = ((long)nB * (nB - 1) / 2 + X) * S + r;
    for (int Y = 0; Y <= X; Y++) {
        const long at = (Y < X) ? rrow + (long)Y * S : dslot;
        if constexpr (F32R) {
            const float4 q = react32[at];
            bx += (double)q.x; by += (double)q.y; bz += (double)q.z;
        } else {
            const P4 q = react64[at];
            bx += q.x; by += q.y; bz += q.z;
        }
    }
}

template <bool F32R>
__global__ __launch_bounds__(BLOCK) void k_ring_finish(const P4* __restrict__ part, const P4* __restrict__ react64,
                                                       const float4* __restrict__ react32, int nB, int S, int L,
                                                       int n, const P4* __restrict__ pl, const P4* __restrict__ pr,
                                                       P4* __restrict__ fl, P4* __restrict__ fr) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const int X = i / S, r = i - X * S;
    const int W = (nB - 1) / L + 1;
    const int cnt = (nB - 1 - X) / L + 1;   // items of row X
    const long arow = (long)X * W * S + r;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j = 0; j < cnt; j++) {
        const P4 p = part[arow + (long)j * S];
        ax += p.x; ay += p.y; az += p.z;
    }
    double bx, by, bz;
    ring_react_sum<F32R>(react64, react32, nB, S, i, bx, by, bz);
    const double ml = pl[i].m, mr = pr[i].m;
    P4 f = fl[i];
    f.x += ml * ax; f.y += ml * ay; f.z += ml * az;
    fl[i] = f;
    P4 g = fr[i];
    g.x += mr * bx; g.y += mr * by; g.z += mr * bz;
    fr[i] = g;
}

template <bool F32R>
__global__ __launch_bounds__(BLOCK) void k_sym_reduce(const P4* __restrict__ part, const P4* __restrict__ react64,
                                                      const float4* __restrict__ react32, int nB, int S, int L,
                                                      int n, double scale, P4* __restrict__ acc) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double ax, ay, az;
    sym_sum<F32R>(part, react64, react32, nB, S, L, i, ax, ay, az);
    acc[i] = P4{scale * ax, scale * ay, scale * az, 0.0};   // scale: the factored-out mass (else 1)
}

template <bool F32R>
__global__ __launch_bounds__(BLOCK) void k_sym_kick_drift(const P4* __restrict__ part,
                                                          const P4* __restrict__ react64,
                                                          const float4* __restrict__ react32, int nB, int S, int L,
                                                          int n, double scale, P4* __restrict__ pm,
                                                          P4* __restrict__ vel, P4* __restrict__ acc, double kdt,
                                                          double c1dt, double c2dt) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double ax, ay, az;
    sym_sum<F32R>(part, react64, react32, nB, S, L, i, ax, ay, az);
    ax *= scale;   // the factored-out mass (else 1)
    ay *= scale;
    az *= scale;
    acc[i] = P4{ax, ay, az, 0.0};
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

template <int FORM, bool EPS0, int K>
static hipError_t launch_sym_t(const SymLaunch& L, hipStream_t st) {
    const int groups = (L.nitems + kWavesPerGroup - 1) / kWavesPerGroup;
    if (L.pj)
        hipLaunchKernelGGL((k_accel_sym<FORM, EPS0, K, true, false>), dim3(groups), dim3(64 * kWavesPerGroup), 0, st,
                           L.pm, L.pj, L.n, L.items, L.nitems, L.nB, L.eps2, kCubicShift, kCubicOffset,
                           L.incl ? 1 : 0, L.part, L.react64, L.react32);
    else if (L.eqm)
        hipLaunchKernelGGL((k_accel_sym<FORM, EPS0, K, false, true>), dim3(groups), dim3(64 * kWavesPerGroup), 0, st,
                           L.pm, L.pm, L.n, L.items, L.nitems, L.nB, L.eps2, kCubicShift, kCubicOffset, 0, L.part,
                           L.react64, L.react32);
    else
        hipLaunchKernelGGL((k_accel_sym<FORM, EPS0, K, false, false>), dim3(groups), dim3(64 * kWavesPerGroup), 0, st,
                           L.pm, L.pm, L.n, L.items, L.nitems, L.nB, L.eps2, kCubicShift, kCubicOffset, 0, L.part,
                           L.react64, L.react32);
    return hipGetLastError();
}

template <int FORM, bool EPS0>
static hipError_t launch_sym_k(const SymLaunch& L, hipStream_t st) {
    switch (L.K) {
    case 8: return launch_sym_t<FORM, EPS0, 8>(L, st);
    case 6: return launch_sym_t<FORM, EPS0, 6>(L, st);
    case 4: return launch_sym_t<FORM, EPS0, 4>(L, st);
    default: return hipErrorInvalidValue;
    }
}

template <int FORM>
static hipError_t launch_sym_eps(const SymLaunch& L, hipStream_t st) {
    return (L.eps2 == 0.0) ? launch_sym_k<FORM, true>(L, st) : launch_sym_k<FORM, false>(L, st);
}

hipError_t launch_accel_sym(const SymLaunch& L, hipStream_t st) {
    if (L.nitems <= 0) return hipSuccess;
    // FP64 reactions need their record slots, the mixed forms' their FP32 slots
    if (L.form == FORM_FP64 ? !L.react64 : !L.react32) return hipErrorInvalidValue;
    switch (L.form) {
    case FORM_FP64: return launch_sym_eps<FORM_FP64>(L, st);
    case FORM_MIXED: return launch_sym_eps<FORM_MIXED>(L, st);
    case FORM_MIXED_NR: return launch_sym_eps<FORM_MIXED_NR>(L, st);
    default: return hipErrorInvalidValue;
    }
}

template <int FORM, bool EPS0, int K>
static hipError_t launch_symg_k(const SymgLaunch& L, hipStream_t st) {
    constexpr int W = kGroupWaves;
    if (L.pj)
        hipLaunchKernelGGL((k_accel_symg<FORM, EPS0, K, W, true, false>), dim3(L.nwg), dim3(64 * W), 0, st, L.pm, L.pj,
                           L.n, L.ranges, L.nB, L.eps2, kCubicShift, kCubicOffset, L.incl ? 1 : 0, L.act, L.react64,
                           L.react32);
    else if (L.eqm)
        hipLaunchKernelGGL((k_accel_symg<FORM, EPS0, K, W, false, true>), dim3(L.nwg), dim3(64 * W), 0, st, L.pm, L.pm,
                           L.n, L.ranges, L.nB, L.eps2, kCubicShift, kCubicOffset, 0, L.act, L.react64, L.react32);
    else
        hipLaunchKernelGGL((k_accel_symg<FORM, EPS0, K, W, false, false>), dim3(L.nwg), dim3(64 * W), 0, st, L.pm,
                           L.pm, L.n, L.ranges, L.nB, L.eps2, kCubicShift, kCubicOffset, 0, L.act, L.react64,
                           L.react32);
    return hipGetLastError();
}

template <int FORM, bool EPS0>
static hipError_t launch_symg_eps(const SymgLaunch& L, hipStream_t st) {
    switch (L.K) {
    case 8: return launch_symg_k<FORM, EPS0, 8>(L, st);
    case 6: return launch_symg_k<FORM, EPS0, 6>(L, st);
    case 4: return launch_symg_k<FORM, EPS0, 4>(L, st);
    default: return hipErrorInvalidValue;
    }
}

template <int FORM>
static hipError_t launch_symg_form(const SymgLaunch& L, hipStream_t st) {
    return (L.eps2 == 0.0) ? launch_symg_eps<FORM, true>(L, st) : launch_symg_eps<FORM, false>(L, st);
}

hipError_t launch_accel_symg(const SymgLaunch& L, hipStream_t st) {
    if (L.nwg <= 0) return hipSuccess;
    if (L.form == FORM_FP64 ? !L.react64 : !L.react32) return hipErrorInvalidValue;
    switch (L.form) {
    case FORM_FP64: return launch_symg_form<FORM_FP64>(L, st);
    case FORM_MIXED: return launch_symg_form<FORM_MIXED>(L, st);
    case FORM_MIXED_NR: return launch_symg_form<FORM_MIXED_NR>(L, st);
    default: return hipErrorInvalidValue;
    }
}

// Pairing the two ends in one thread halves the finish's waves: it pays from
// N = 262144 (FP64 172.6 -> 160.2 us, mixed 120.4 -> 115.4 us), where 2048
// waves remain, and costs at N = 131072 (83.0 -> 86.1 us, 53.8 -> 57.4 us).
static bool finish_pairs(int n) { return n >= 262144; }

hipError_t launch_symg_finish(const SymgSums& P, int n, P4* pm, P4* vel, P4* acc, bool kick, double kdt, double c1dt,
                              double c2dt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!P.react64 == !P.react32) return hipErrorInvalidValue;
    constexpr int W = kGroupWaves;
    const bool pair = finish_pairs(n);
    const dim3 grid(((pair ? (n + 1) / 2 : n) + BLOCK - 1) / BLOCK);
#define NB_SYMG_FINISH(F32R, KICK)                                                                                 \
    hipLaunchKernelGGL((k_symg_finish<F32R, W, KICK>), grid, dim3(BLOCK), 0, st, P.act, P.rowseg, P.react64,        \
                       P.react32, P.nB, P.S, n, pair ? 1 : 0, P.scale, pm, vel, acc, kdt, c1dt, c2dt)
    if (P.react32) {
        if (kick) NB_SYMG_FINISH(true, true);
        else NB_SYMG_FINISH(true, false);
    } else {
        if (kick) NB_SYMG_FINISH(false, true);
        else NB_SYMG_FINISH(false, false);
    }
#undef NB_SYMG_FINISH
    return hipGetLastError();
}

hipError_t launch_symg_ring_finish(const SymgSums& P, int n, const P4* pl, const P4* pr, P4* fl, P4* fr,
                                   hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!P.react64 == !P.react32) return hipErrorInvalidValue;
    constexpr int W = kGroupWaves;
    const bool pair = finish_pairs(n);
    const dim3 grid(((pair ? (n + 1) / 2 : n) + BLOCK - 1) / BLOCK);
    if (P.react32)
        hipLaunchKernelGGL((k_symg_ring_finish<true, W>), grid, dim3(BLOCK), 0, st, P.act, P.rowseg, P.react64,
                           P.react32, P.nB, P.S, n, pair ? 1 : 0, pl, pr, fl, fr);
    else
        hipLaunchKernelGGL((k_symg_ring_finish<false, W>), grid, dim3(BLOCK), 0, st, P.act, P.rowseg, P.react64,
                           P.react32, P.nB, P.S, n, pair ? 1 : 0, pl, pr, fl, fr);
    return hipGetLastError();
}

hipError_t launch_sym_reduce(const SymSums& P, int n, P4* acc, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!P.react64 == !P.react32) return hipErrorInvalidValue;   // exactly one reaction region
    const dim3 grid((n + BLOCK - 1) / BLOCK);
    if (P.react32)
        hipLaunchKernelGGL(k_sym_reduce<true>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB, P.S,
                           P.L, n, P.scale, acc);
    else
        hipLaunchKernelGGL(k_sym_reduce<false>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB, P.S,
                           P.L, n, P.scale, acc);
    return hipGetLastError();
}

hipError_t launch_ring_finish(const SymSums& P, int n, const P4* pl, const P4* pr, P4* fl, P4* fr, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!P.react64 == !P.react32) return hipErrorInvalidValue;
    const dim3 grid((n + BLOCK - 1) / BLOCK);
    if (P.react32)
        hipLaunchKernelGGL(k_ring_finish<true>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB, P.S,
                           P.L, n, pl, pr, fl, fr);
    else
        hipLaunchKernelGGL(k_ring_finish<false>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB, P.S,
                           P.L, n, pl, pr, fl, fr);
    return hipGetLastError();
}

hipError_t launch_sym_kick_drift(const SymSums& P, int n, P4* pm, P4* vel, P4* acc, double kdt, double c1dt,
                                 double c2dt, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    if (!P.react64 == !P.react32) return hipErrorInvalidValue;
    const dim3 grid((n + BLOCK - 1) / BLOCK);
    if (P.react32)
        hipLaunchKernelGGL(k_sym_kick_drift<true>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB,
                           P.S, P.L, n, P.scale, pm, vel, acc, kdt, c1dt, c2dt);
    else
        hipLaunchKernelGGL(k_sym_kick_drift<false>, grid, dim3(BLOCK), 0, st, P.part, P.react64, P.react32, P.nB,
                           P.S, P.L, n, P.scale, pm, vel, acc, kdt, c1dt, c2dt);
    return hipGetLastError();
}

}  // namespace nbk
