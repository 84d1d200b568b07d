// nbody_capi.hip -- the C ABI (include/nbody.h) over the gfx950 kernels.
//
// One context = one device + one HIP stream (default priority) + the
// device-resident system: positions+masses, velocities and accelerations as
// 32-byte records, plus scratch for split partial sums and the energy rows.
// All work is stream-ordered; calls returning host data synchronise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nbody.h"
#include "nbody_host.h"
#include "nbody_kernels.h"
#include "nbody_learned.h"

using nbk::P4;

namespace {

nb_status fail(nb_status s, const std::string& msg) { return nbh::set_error(s, msg); }

nb_status hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? NB_ERR_ALLOC : NB_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define NB_HIP(call, what)                          \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

constexpr size_t kMaxN = 100000000;   // the reference's bound (CPU/main.cpp:271)

// Scratch of nb_accel_block_device_noalloc: a fixed array in device memory
// that every GPU the library runs on holds until the process ends -- a
// 256-byte head (the j masses' check) and 256 MiB of partial sums (the
// 131072 x 1048576 ring share in j-blocks of 131072 needs exactly that).
// Host code never passes the array's name to a kernel or a copy: it uses the
// address hipGetSymbolAddress returns for the device (nb_ctx::scratch_ptr,
// taken in nb_create with the device current).
constexpr size_t kStaticScratchBytes = size_t(256) << 20;   // partial sums
constexpr size_t kScratchHead = 256;
static __device__ __align__(256) P4 scratch[(kScratchHead + kStaticScratchBytes) / sizeof(P4)];

__global__ void k_zero_records(P4* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = P4{0.0, 0.0, 0.0, 0.0};
}

}  // namespace

namespace nbh {

hipError_t DevBuf::ensure(size_t b) {
    if (b <= bytes) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    else p = nullptr;
    return e;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

}  // namespace nbh

using nbh::DevBuf;

struct nb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    size_t n = 0;
    std::vector<double> mass_h;
    double eq_mass = 0.0;   // the common mass when every particle has it (> 0, finite), else 0
    DevBuf pm, vel, acc, part, rows, eblk;
    DevBuf bi, bj, bout, bout2, eqstate;   // ring-share staging; eqstate: the j masses' check
    DevBuf rk;   // Runge-Kutta stage records: x0, kv[s], ka[s]
    // timing
    bool timing = false;
    std::vector<hipEvent_t> events;   // pairs (start, stop) per force launch
    size_t ev_used = 0;
    double force_ms = 0.0;
    uint64_t force_evals = 0;
    double step_ms = 0.0;
    hipEvent_t step_ev[2] = {nullptr, nullptr};
    nbl::Learned* learned = nullptr;
    int kernel = NB_KERNEL_AUTO;
    // symmetric-kernel plans for the resident size (one per K in use: the
    // FP64 and mixed forms block differently)
    struct SymPlan {
        size_t n = 0;
        int K = 0, nB = 0, L = 0, nitems = 0;
        DevBuf items;
    } sym[2];
    int sym_next = 0;
    // grouped symmetric plans: per-workgroup chunk ranges and the action
    // segments of each group row
    struct SymgPlan {
        size_t n = 0;
        int K = 0, nB = 0, nG = 0, nwg = 0, nseg = 0;
        long nslots = 0;
        DevBuf ranges, rowseg;
    } symg[2];
    int symg_next = 0;
    nb_kernel_info kinfo{};   // the last force evaluation of the resident system
    P4* scratch_ptr = nullptr;          // this device's address of the static scratch array
    hipError_t scratch_status = hipSuccess;
};

namespace nbh {

// i-particles per lane: four when the grid still fills the chip with them
// (ring share 131072 x 1048576 on 256 CUs: FP64 87.9 -> 82.5 ms against two).
int choose_ipt(const nb_ctx* c, long ni) { return ni >= long(c->num_cus) * 512 ? 4 : 1; }

// Number of j-chunks so that the grid holds about 8 waves per CU.
int choose_splits(const nb_ctx* c, long ni, long nj, int ipt) {
    const long waves = (ni + 64 * ipt - 1) / (64 * ipt);
    const long target = long(c->num_cus) * 16;   // four waves per SIMD to hide LDS and FMA latency
    long S = 1;
    if (waves > 0 && waves < target) S = (target + waves - 1) / waves;
    S = std::min(S, std::max(1L, nj / 64));
    S = std::min(S, 1024L);
    return int(S);
}

// Timing: record an event before a force launch; returns the index or -1.
static long timing_begin(nb_ctx* c) {
    if (!c->timing) return -1;
    if (c->ev_used + 2 > c->events.size()) {
        if (c->events.size() >= 16384) {   // fold what has completed so far
            (void)hipStreamSynchronize(c->stream);
            for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, c->events[k], c->events[k + 1]) == hipSuccess)
                    c->force_ms += ms;
            }
            c->ev_used = 0;
        } else {
            for (int k = 0; k < 2; k++) {
                hipEvent_t e;
                if (hipEventCreate(&e) != hipSuccess) return -1;
                c->events.push_back(e);
            }
        }
    }
    const long idx = long(c->ev_used);
    (void)hipEventRecord(c->events[idx], c->stream);
    c->ev_used += 2;
    return idx;
}

static void timing_end(nb_ctx* c, long idx) {
    if (idx < 0) return;
    (void)hipEventRecord(c->events[idx + 1], c->stream);
    c->force_evals++;
}

// Symmetric plan: blocks of S = 64*K, items of L tiles along each row.
// Tiles per work item of the symmetric kernel: L >= 2, ~128 items per CU
// where there are enough tiles.  Short items keep the last round of the
// dispatch from idling most of the chip, while single-tile items cost more
// in per-item loads, action slots and their reduction than they save.
// Measured: N = 262144 (L = 32 / 4) 27.6 / 24.4 ms FP64, 20.9 / 18.0 ms
// mixed; N = 65536 (L = 8 / 2 / 1) 1.91 / 1.75 / 1.77 ms FP64; N = 32768
// steps 450 us at L = 2 against 495 us at L = 1.
static long sym_item_tiles(const nb_ctx* c, long nB) {
    const long tiles = nB * (nB + 1) / 2;
    const long target = long(c->num_cus) * 128;
    return std::max(2L, tiles / target);
}

static hipError_t sym_plan(nb_ctx* c, size_t n, int K, const nb_ctx::SymPlan*& plan) {
    for (const auto& Q : c->sym)
        if (Q.n == n && Q.K == K) {
            plan = &Q;
            return hipSuccess;
        }
    auto& P = c->sym[c->sym_next];
    c->sym_next ^= 1;
    P.n = 0;   // invalid until built
    const long S = 64L * K;
    const long nB = (long(n) + S - 1) / S;
    const long L = sym_item_tiles(c, nB);
    const long W = (nB - 1) / L + 1;
    std::vector<int4> items;
    for (long A = 0; A < nB; A++)
        for (long b0 = A; b0 < nB; b0 += L)
            items.push_back(make_int4(int(A), int(b0), int(std::min(nB, b0 + L)), int(A * W + (b0 - A) / L)));
    // longest items first so the short row tails fill in at the end
    std::stable_sort(items.begin(), items.end(), [](const int4& a, const int4& b) { return (a.z - a.y) > (b.z - b.y); });
    hipError_t e = P.items.ensure(items.size() * sizeof(int4));
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(P.items.p, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return e;
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    P.n = n;
    P.K = K;
    P.nB = int(nB);
    P.L = int(L);
    P.nitems = int(items.size());
    plan = &P;
    return hipSuccess;
}

struct SymRegions {
    size_t act_bytes, react_bytes;
};

// Grouped plan: the chunk sequence (rows g of W blocks, j-blocks B >= g W,
// chunks c < K) cut into equal ranges, one per resident workgroup (12 waves
// per CU at 3 per SIMD); each workgroup's action slots (one per row it
// touches) follow the previous workgroup's, so a row's segments are
// consecutive slots.
// 24 ranges per CU: filling the chip exactly once (3 workgroups per CU) left
// it under-occupied (N = 262144: mixed 18.6 against 16.0 ms), many short
// ranges let the dispatcher balance the load at the price of more action
// segments (N = 262144, FP64 / mixed kernel ms at 3 / 12 / 24 / 48 / 96
// ranges per CU: 24.2 / 23.6 / 23.0 / 23.5 / 23.8 and 18.6 / 16.6 / 16.0 /
// 15.9 / 15.8, with the partial sums growing 0.60 / 0.75 / 0.95 / 1.35 /
// 2.16 GB FP64 in 32-byte records; three quarters of that in the component
// planes used now).
static int symg_workgroups(const nb_ctx* c, long units) {
    return int(std::max(1L, std::min(long(c->num_cus) * 24, units)));
}

static long symg_slots(long nB, long W) {
    const long nG = (nB + W - 1) / W;
    return nG * nB - W * nG * (nG - 1) / 2;
}

static hipError_t symg_plan(nb_ctx* c, size_t n, int K, const nb_ctx::SymgPlan*& plan) {
    for (const auto& Q : c->symg)
        if (Q.n == n && Q.K == K) {
            plan = &Q;
            return hipSuccess;
        }
    auto& P = c->symg[c->symg_next];
    c->symg_next ^= 1;
    P.n = 0;
    const long W = nbk::kGroupWaves, S = 64L * K;
    const long nB = (long(n) + S - 1) / S, nG = (nB + W - 1) / W;
    std::vector<long> rowstart(nG + 1, 0);
    for (long g = 0; g < nG; g++) rowstart[g + 1] = rowstart[g] + (nB - g * W) * K;
    const long U = rowstart[nG];
    const int nwg = symg_workgroups(c, U);
    std::vector<int4> ranges(nwg);
    std::vector<int2> rowseg(nG, make_int2(0, 0));
    long slot = 0, g = 0;
    for (int wg = 0; wg < nwg; wg++) {
        const long u0 = U * wg / nwg, u1 = U * (wg + 1) / nwg;
        while (rowstart[g + 1] <= u0) g++;
        long gl = g;
        while (rowstart[gl + 1] < u1) gl++;   // row of the last chunk u1 - 1
        const long off = u0 - rowstart[g];
        ranges[wg] = make_int4(int(g), int(g * W * K + off), int(u1 - u0), int(slot));
        for (long r = g; r <= gl; r++) {
            if (rowseg[r].y == 0) rowseg[r].x = int(slot + (r - g));
            rowseg[r].y++;
        }
        slot += gl - g + 1;
    }
    hipError_t e = P.ranges.ensure(ranges.size() * sizeof(int4));
    if (e == hipSuccess) e = P.rowseg.ensure(rowseg.size() * sizeof(int2));
    if (e == hipSuccess)
        e = hipMemcpyAsync(P.ranges.p, ranges.data(), ranges.size() * sizeof(int4), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(P.rowseg.p, rowseg.data(), rowseg.size() * sizeof(int2), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return e;
    P.n = n;
    P.K = K;
    P.nB = int(nB);
    P.nG = int(nG);
    P.nwg = nwg;
    P.nseg = int(slot);
    P.nslots = symg_slots(nB, W);
    plan = &P;
    return hipSuccess;
}

// Partial sums of the grouped kernel: action slots (nseg of W blocks of S
// particles, 3 doubles each: 24 B) then reaction slots (nslots of S
// particles: 3 doubles FP64, 3 floats for the mixed forms), every slot as
// three component planes.  `nseg` at most workgroups + rows.
static SymRegions symg_regions(long nB, long S, long nseg, int kform) {
    const long W = nbk::kGroupWaves;
    const size_t rec = 3 * ((kform == nbk::FORM_FP64) ? sizeof(double) : sizeof(float));
    return {size_t(nseg) * size_t(W) * size_t(S) * 3 * sizeof(double), size_t(symg_slots(nB, W)) * size_t(S) * rec};
}

// Partial sums of the symmetric kernel: the action slots (nB*W slots of S
// records of 32 B, W = (nB - 1)/L + 1) and after them the reaction slots
// (nB(nB-1)/2 slots of S records: 32 B for FP64, 16 B FP32 for the mixed
// forms).  N^2/32 + 32 N W bytes for FP64 at K = 8 (34 GB at N = 1048576),
// N^2/48 + 32 N W for the mixed forms at K = 6.
static SymRegions sym_regions(long nB, long S, long L, int kform) {
    const long W = (nB - 1) / L + 1;
    const size_t rec = (kform == nbk::FORM_FP64) ? sizeof(P4) : sizeof(float4);
    return {size_t(nB) * size_t(W) * size_t(S) * sizeof(P4), size_t(nB) * size_t(nB - 1) / 2 * size_t(S) * rec};
}

// Which kernel evaluates the resident system: the symmetric one from 16384
// particles up (it does half the pair work) while its partial-sum buffer
// fits: always up to 16 GiB, beyond that while the growth takes at most 80 %
// of the free device memory (N = 1M needs 35 GB FP64 / 24 GB mixed of the
// 288 GB of an MI355X, so one device runs it to N ~ 2.6M FP64); the tile
// kernel otherwise.
// The grouped symmetric kernel from 131072 particles on (where the blocks hold
// 8 / 6 particles per lane), the item kernel below it: with 4 per lane and
// fewer chunks per range the item kernel measured faster (N = 32768 mixed
// 283 against 294 us, N = 65536 1.00 against 1.04 ms per evaluation).
static bool grouped_size(long n) { return n >= 131072; }

static bool use_symmetric(const nb_ctx* c, int kform, int& K) {
    const long n = long(c->n);
    if (kform == nbk::FORM_FP32) return false;   // the FP32 form is a tile-kernel form
    // More particles per lane at large N (fewer LDS reads and rotations per
    // pair): 8 for FP64, 6 for the mixed forms, whose t-loop is left unrolled
    // once to fit the VGPRs (measured at N = 262144, mixed: K = 5 / 6 / 7 / 8
    // 17.9 / 17.2 / 17.6 / 17.8 ms; FP64 K = 6 / 8 24.8 / 24.75 ms).
    K = (n >= 131072) ? (kform == nbk::FORM_FP64 ? 8 : 6) : 4;
    const long S = 64L * K;
    const long nB = (n + S - 1) / S;
    SymRegions R;
    if (grouped_size(n)) {
        const long W = nbk::kGroupWaves, nG = (nB + W - 1) / W;
        R = symg_regions(nB, S, symg_workgroups(c, symg_slots(nB, W) * K) + nG, kform);
    } else {
        R = sym_regions(nB, S, sym_item_tiles(c, nB), kform);
    }
    const double need = double(R.act_bytes) + double(R.react_bytes);
    bool fits = need <= 16.0 * (1L << 30) || need <= double(c->part.bytes);
    if (!fits) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess)
            fits = need - double(c->part.bytes) <= 0.8 * double(free_b);
    }
    if (c->kernel == NB_KERNEL_TILE) return false;
    if (c->kernel == NB_KERNEL_SYMMETRIC) return fits && n >= 2;
    return fits && n >= 16384;   // measured crossover: N = 8192 tile 57 us / symmetric 74 us, 16384 196 / 164 us
}

static const char* form_label(int kform) {
    switch (kform) {
    case nbk::FORM_FP64: return "FP64";
    case nbk::FORM_MIXED_NR: return "MIXED_NR";
    case nbk::FORM_FP32: return "FP32";
    default: return "MIXED";
    }
}

static int public_form(int kform) {
    switch (kform) {
    case nbk::FORM_FP64: return NB_FORM_FP64;
    case nbk::FORM_MIXED_NR: return NB_FORM_MIXED;
    case nbk::FORM_FP32: return NB_FORM_FP32;
    default: return NB_FORM_MIXED_RAW;
    }
}

// Record what a force evaluation of the resident system ran (nb_kernel_info_read).
static void note_kernel(nb_ctx* c, int kernel, int kform, int per_lane, int block, uint64_t items,
                        uint64_t scratch_bytes, bool grouped = false, bool eqm = false) {
    nb_kernel_info& k = c->kinfo;
    k.kernel = kernel;
    k.form = public_form(kform);
    k.per_lane = per_lane;
    k.block = block;
    k.items = items;
    k.scratch_bytes = scratch_bytes;
    const char* eq = eqm ? ", equal masses" : "";
    if (kernel == NB_KERNEL_SYMMETRIC && grouped)
        std::snprintf(k.name, sizeof k.name, "k_accel_symg<%s, K=%d, W=%d%s>", form_label(kform), per_lane,
                      nbk::kGroupWaves, eq);
    else if (kernel == NB_KERNEL_SYMMETRIC)
        std::snprintf(k.name, sizeof k.name, "k_accel_sym<%s, K=%d%s>", form_label(kform), per_lane, eq);
    else
        std::snprintf(k.name, sizeof k.name, "k_accel<%s, IPT=%d>", form_label(kform), per_lane);
}

// One force evaluation of the resident state; `out` says where the partial
// sums are and how to add them.
// The mass an equal-mass evaluation leaves out of its sums: the common mass
// times the form's factor on the staged masses (0: the masses stay in).
static double eq_scale(const nb_ctx* c, int kform) {
    if (c->eq_mass == 0.0) return 0.0;
    switch (kform) {
    case nbk::FORM_FP64: return c->eq_mass * nbk::kCubicMass;
    case nbk::FORM_MIXED_NR: return c->eq_mass * nbk::kNewtonMass;
    case nbk::FORM_MIXED: return c->eq_mass;
    default: return 0.0;
    }
}

hipError_t force_eval(nb_ctx* c, int kform, double eps, ForceOut& out) {
    const long n = long(c->n);
    const double scale = eq_scale(c, kform);
    int K;
    if (use_symmetric(c, kform, K) && grouped_size(n)) {
        const nb_ctx::SymgPlan* plan = nullptr;
        hipError_t e = symg_plan(c, c->n, K, plan);
        if (e != hipSuccess) return e;
        const long S = 64L * K;
        const SymRegions R = symg_regions(plan->nB, S, plan->nseg, kform);
        e = c->part.ensure(R.act_bytes + R.react_bytes);
        if (e != hipSuccess) return e;
        char* const react = (char*)c->part.p + R.act_bytes;
        nbk::SymgLaunch L;
        L.form = kform;
        L.K = K;
        L.pm = (const P4*)c->pm.p;
        L.n = int(n);
        L.ranges = (const int4*)plan->ranges.p;
        L.nwg = plan->nwg;
        L.nB = plan->nB;
        L.eps2 = eps * eps;
        L.act = (double*)c->part.p;
        if (kform == nbk::FORM_FP64) L.react64 = (double*)react;
        else L.react32 = (float*)react;
        L.eqm = scale != 0.0;
        out.sym = true;
        out.grouped = true;
        out.gsums.act = L.act;
        out.gsums.rowseg = (const int2*)plan->rowseg.p;
        out.gsums.react64 = L.react64;
        out.gsums.react32 = L.react32;
        out.gsums.nB = plan->nB;
        out.gsums.S = int(S);
        out.gsums.scale = L.eqm ? scale : 1.0;
        note_kernel(c, NB_KERNEL_SYMMETRIC, kform, K, int(S), uint64_t(plan->nwg), R.act_bytes + R.react_bytes, true,
                    L.eqm);
        const long ev = timing_begin(c);
        e = nbk::launch_accel_symg(L, c->stream);
        timing_end(c, ev);
        return e;
    }
    if (use_symmetric(c, kform, K)) {
        const nb_ctx::SymPlan* plan = nullptr;
        hipError_t e = sym_plan(c, c->n, K, plan);
        if (e != hipSuccess) return e;
        const auto& P = *plan;
        const long S = 64L * K;
        const SymRegions R = sym_regions(P.nB, S, P.L, kform);
        e = c->part.ensure(R.act_bytes + R.react_bytes);
        if (e != hipSuccess) return e;
        char* const react = (char*)c->part.p + R.act_bytes;
        nbk::SymLaunch L;
        L.form = kform;
        L.K = K;
        L.pm = (const P4*)c->pm.p;
        L.n = int(n);
        L.items = (const int4*)P.items.p;
        L.nitems = P.nitems;
        L.nB = P.nB;
        L.eps2 = eps * eps;
        L.part = (P4*)c->part.p;
        if (kform == nbk::FORM_FP64) L.react64 = (P4*)react;
        else L.react32 = (float4*)react;
        L.eqm = scale != 0.0;
        out.sym = true;
        out.part = L.part;
        out.react64 = L.react64;
        out.react32 = L.react32;
        out.nB = P.nB;
        out.S = 64 * K;
        out.L = P.L;
        out.scale = L.eqm ? scale : 1.0;
        note_kernel(c, NB_KERNEL_SYMMETRIC, kform, K, int(S), uint64_t(P.nitems), R.act_bytes + R.react_bytes, false,
                    L.eqm);
        const long ev = timing_begin(c);
        e = nbk::launch_accel_sym(L, c->stream);
        timing_end(c, ev);
        return e;
    }
    nbk::AccelLaunch L;
    L.form = kform;
    L.ipt = choose_ipt(c, n);
    L.eq_mass = c->eq_mass;
    int S = choose_splits(c, n, n, L.ipt);
    long jchunk = (n + S - 1) / S;
    jchunk = (jchunk + 63) & ~63L;
    S = int((n + jchunk - 1) / jchunk);
    P4* part;
    if (S <= 1) {
        S = 1;
        jchunk = n;
        part = (P4*)c->acc.p;
    } else {
        hipError_t e = c->part.ensure(size_t(S) * n * sizeof(P4));
        if (e != hipSuccess) return e;
        part = (P4*)c->part.p;
    }
    L.pi = (const P4*)c->pm.p;
    L.ni = int(n);
    L.pj = (const P4*)c->pm.p;
    L.jbeg = 0;
    L.jend = int(n);
    L.jchunk = int(jchunk);
    L.splits = S;
    L.eps2 = eps * eps;
    L.G = 1.0;
    L.out = part;
    L.out_stride = n;
    out.sym = false;
    out.part = part;
    out.S = S;
    note_kernel(c, NB_KERNEL_TILE, kform, L.ipt, S, uint64_t((n + 256L * L.ipt - 1) / (256L * L.ipt)) * uint64_t(S),
                S > 1 ? uint64_t(S) * uint64_t(n) * sizeof(P4) : 0);
    const long ev = timing_begin(c);
    hipError_t e = nbk::launch_accel(L, c->stream);
    timing_end(c, ev);
    return e;
}

static nbk::SymSums sym_sums(const ForceOut& out) {
    nbk::SymSums P;
    P.part = out.part;
    P.react64 = out.react64;
    P.react32 = out.react32;
    P.nB = out.nB;
    P.S = out.S;
    P.L = out.L;
    P.scale = out.scale;
    return P;
}

// acc = sum of the partials of `out`.
hipError_t finish_accel(nb_ctx* c, const ForceOut& out) {
    const int n = int(c->n);
    if (out.sym && out.grouped)
        return nbk::launch_symg_finish(out.gsums, n, nullptr, nullptr, (P4*)c->acc.p, false, 0.0, 0.0, 0.0,
                                       c->stream);
    if (out.sym) return nbk::launch_sym_reduce(sym_sums(out), n, (P4*)c->acc.p, c->stream);
    if (out.S == 1) return hipSuccess;
    return nbk::launch_reduce(out.part, out.S, long(n), n, (P4*)c->acc.p, c->stream);
}

// acc = sum of the partials, then v += kdt a; x += c1dt v; x += c2dt v.
hipError_t finish_kick_drift(nb_ctx* c, const ForceOut& out, double kdt, double c1dt, double c2dt) {
    const int n = int(c->n);
    P4* pm = (P4*)c->pm.p;
    P4* vel = (P4*)c->vel.p;
    P4* acc = (P4*)c->acc.p;
    if (out.sym && out.grouped)
        return nbk::launch_symg_finish(out.gsums, n, pm, vel, acc, true, kdt, c1dt, c2dt, c->stream);
    if (out.sym)
        return nbk::launch_sym_kick_drift(sym_sums(out), n, pm, vel, acc, kdt, c1dt, c2dt, c->stream);
    return nbk::launch_kick_drift(out.part, out.S, long(n), n, pm, vel, acc, kdt, c1dt, c2dt, c->stream);
}

int kernel_form(int form) {
    switch (form) {
    case NB_FORM_FP64: return nbk::FORM_FP64;
    case NB_FORM_MIXED: return nbk::FORM_MIXED_NR;
    case NB_FORM_MIXED_RAW: return nbk::FORM_MIXED;
    case NB_FORM_FP32: return nbk::FORM_FP32;
    default: return -1;
    }
}

hipError_t accel_to_acc(nb_ctx* c, int kform, double eps) {
    ForceOut out;
    hipError_t e = force_eval(c, kform, eps, out);
    if (e != hipSuccess) return e;
    return finish_accel(c, out);
}

hipStream_t ctx_stream(nb_ctx* c) { return c->stream; }
size_t ctx_n(nb_ctx* c) { return c->n; }
P4* ctx_pm(nb_ctx* c) { return (P4*)c->pm.p; }
P4* ctx_vel(nb_ctx* c) { return (P4*)c->vel.p; }
P4* ctx_acc(nb_ctx* c) { return (P4*)c->acc.p; }
int ctx_num_cus(nb_ctx* c) { return c->num_cus; }
nbl::Learned*& ctx_learned(nb_ctx* c) { return c->learned; }

}  // namespace nbh

using namespace nbh;

extern "C" {

const char* nb_version(void) { return "nbody-mi355x 0.1 (gfx950)"; }

nb_status nb_create(nb_ctx** out, int device) {
    if (!out) return fail(NB_ERR_ARG, "nb_create: out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fail(NB_ERR_DEVICE, "nb_create: no HIP device available");
    if (device < 0 || device >= count) return fail(NB_ERR_ARG, "nb_create: device index out of range");
    NB_HIP(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop;
    NB_HIP(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(NB_ERR_DEVICE, std::string("nb_create: device is ") + prop.gcnArchName +
                                       ", this library is built for gfx950 only");
    nb_ctx* c = new (std::nothrow) nb_ctx;
    if (!c) return fail(NB_ERR_ALLOC, "nb_create: out of host memory");
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipStreamCreate");
    }
    for (auto& ev : c->step_ev) (void)hipEventCreate(&ev);
    // the static scratch's address on this device (each device has its own copy)
    c->scratch_status = hipGetSymbolAddress((void**)&c->scratch_ptr, HIP_SYMBOL(scratch));
    if (c->scratch_status != hipSuccess) c->scratch_ptr = nullptr;
    *out = c;
    nbh::clear_error();
    return NB_OK;
}

nb_status nb_destroy(nb_ctx* c) {
    if (!c) return NB_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    nbl::destroy(c->learned);
    c->learned = nullptr;
    for (DevBuf* b : {&c->pm, &c->vel, &c->acc, &c->part, &c->rows, &c->eblk, &c->bi, &c->bj, &c->bout, &c->bout2,
                      &c->eqstate,
                      &c->rk, &c->symg[0].ranges, &c->symg[0].rowseg, &c->symg[1].ranges, &c->symg[1].rowseg,
                      &c->sym[0].items, &c->sym[1].items})
        b->release();
    for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->step_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return NB_OK;
}

nb_status nb_init_system(int kind, size_t n, uint64_t seed, const double* params, double* mass,
                         double* const pos[3], double* const vel[3], double* const force[3]) {
    if (!mass || !pos || !vel || !pos[0] || !pos[1] || !pos[2] || !vel[0] || !vel[1] || !vel[2])
        return fail(NB_ERR_ARG, "nb_init_system: null output array");
    if (n > kMaxN) return fail(NB_ERR_ARG, "nb_init_system: N must stay below 1e8");
    switch (kind) {
    case NB_IC_RING: {
        if (n < 2) return fail(NB_ERR_ARG, "nb_init_system: the ring needs N >= 2");
        const double R = params ? params[0] : 1.0;
        const double omega = params ? params[1] : 2.0 * M_PI;
        const double ratio = params ? params[2] : 1e5;
        if (force && (!force[0] || !force[1] || !force[2]))
            return fail(NB_ERR_ARG, "nb_init_system: null force plane");
        ic_ring(n, R, omega, ratio, mass, pos, vel, force);
        break;
    }
    case NB_IC_SPHERE:
        ic_sphere(n, seed, mass, pos, vel);
        break;
    case NB_IC_PLUMMER: {
        const double a = params ? params[0] : 1.0;
        const double rmax = params ? params[1] : 20.0;
        if (!(a > 0.0) || !(rmax > 0.0)) return fail(NB_ERR_ARG, "nb_init_system: Plummer a, rmax must be > 0");
        ic_plummer(n, seed, a, rmax, mass, pos, vel);
        break;
    }
    default:
        return fail(NB_ERR_ARG, "nb_init_system: unknown initial-condition kind");
    }
    return NB_OK;
}

nb_status nb_yoshida(int order, double* c, double* d, size_t* lf_steps) {
    if (!c || !d || !lf_steps) return fail(NB_ERR_ARG, "nb_yoshida: null argument");
    *lf_steps = yoshida(order, c, d);
    return NB_OK;
}

nb_status nb_set_state(nb_ctx* c, size_t n, const double* mass, const double* const pos[3],
                       const double* const vel[3]) {
    if (!c) return fail(NB_ERR_ARG, "nb_set_state: null context");
    if (n > kMaxN) return fail(NB_ERR_ARG, "nb_set_state: N must stay below 1e8");
    if (n > 0 && (!mass || !pos || !pos[0] || !pos[1] || !pos[2]))
        return fail(NB_ERR_ARG, "nb_set_state: null mass or position array");
    if (n > 0 && vel && (!vel[0] || !vel[1] || !vel[2]))
        return fail(NB_ERR_ARG, "nb_set_state: null velocity plane");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const size_t cap = std::max<size_t>(n, 1);
    NB_HIP(c->pm.ensure(cap * sizeof(P4)), "allocate positions");
    NB_HIP(c->vel.ensure(cap * sizeof(P4)), "allocate velocities");
    NB_HIP(c->acc.ensure(cap * sizeof(P4)), "allocate accelerations");
    std::vector<P4> hp(n), hv(n);
    for (size_t i = 0; i < n; i++) {
        hp[i] = P4{pos[0][i], pos[1][i], pos[2][i], mass[i]};
        hv[i] = vel ? P4{vel[0][i], vel[1][i], vel[2][i], 0.0} : P4{0.0, 0.0, 0.0, 0.0};
    }
    if (n) {
        NB_HIP(hipMemcpyAsync(c->pm.p, hp.data(), n * sizeof(P4), hipMemcpyHostToDevice, c->stream), "upload positions");
        NB_HIP(hipMemcpyAsync(c->vel.p, hv.data(), n * sizeof(P4), hipMemcpyHostToDevice, c->stream), "upload velocities");
        NB_HIP(hipMemsetAsync(c->acc.p, 0, n * sizeof(P4), c->stream), "clear accelerations");
        NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    c->mass_h.assign(mass, mass + n);
    c->eq_mass = (n > 0 && mass[0] > 0.0 && std::isfinite(mass[0])) ? mass[0] : 0.0;
    for (size_t i = 1; i < n && c->eq_mass != 0.0; i++)
        if (mass[i] != mass[0]) c->eq_mass = 0.0;
    c->n = n;
    c->kinfo = nb_kernel_info{};
    nbl::on_new_state(c->learned);
    return NB_OK;
}

nb_status nb_get_state(nb_ctx* c, double* const pos[3], double* const vel[3]) {
    if (!c) return fail(NB_ERR_ARG, "nb_get_state: null context");
    if ((pos && (!pos[0] || !pos[1] || !pos[2])) || (vel && (!vel[0] || !vel[1] || !vel[2])))
        return fail(NB_ERR_ARG, "nb_get_state: null plane");
    const size_t n = c->n;
    if (n == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    std::vector<P4> h(n);
    if (pos) {
        NB_HIP(hipMemcpyAsync(h.data(), c->pm.p, n * sizeof(P4), hipMemcpyDeviceToHost, c->stream), "download positions");
        NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (size_t i = 0; i < n; i++) {
            pos[0][i] = h[i].x; pos[1][i] = h[i].y; pos[2][i] = h[i].z;
        }
    }
    if (vel) {
        NB_HIP(hipMemcpyAsync(h.data(), c->vel.p, n * sizeof(P4), hipMemcpyDeviceToHost, c->stream), "download velocities");
        NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        for (size_t i = 0; i < n; i++) {
            vel[0][i] = h[i].x; vel[1][i] = h[i].y; vel[2][i] = h[i].z;
        }
    }
    return NB_OK;
}

static nb_status check_eps(double eps, const char* fn) {
    if (!(eps >= 0.0) || !std::isfinite(eps)) return fail(NB_ERR_ARG, std::string(fn) + ": eps must be finite and >= 0");
    return NB_OK;
}

static nb_status download_acc(nb_ctx* c, double* const out[3], bool times_mass) {
    const size_t n = c->n;
    std::vector<P4> h(n);
    NB_HIP(hipMemcpyAsync(h.data(), c->acc.p, n * sizeof(P4), hipMemcpyDeviceToHost, c->stream), "download accelerations");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < n; i++) {
        const double s = times_mass ? c->mass_h[i] : 1.0;
        out[0][i] = s * h[i].x;
        out[1][i] = s * h[i].y;
        out[2][i] = s * h[i].z;
    }
    return NB_OK;
}

static nb_status accel_common(nb_ctx* c, int form, double eps, double* const out[3], bool times_mass,
                              const char* fn) {
    if (!c) return fail(NB_ERR_ARG, std::string(fn) + ": null context");
    if (!out || !out[0] || !out[1] || !out[2]) return fail(NB_ERR_ARG, std::string(fn) + ": null output plane");
    if (nb_status s = check_eps(eps, fn)) return s;
    const int kf = kernel_form(form);
    if (kf < 0) return fail(NB_ERR_ARG, std::string(fn) + ": form must be NB_FORM_FP64, NB_FORM_MIXED, NB_FORM_MIXED_RAW or NB_FORM_FP32");
    if (c->n == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(accel_to_acc(c, kf, eps), "force kernel");
    return download_acc(c, out, times_mass);
}

nb_status nb_accel(nb_ctx* c, int form, double eps, double* const acc[3]) {
    return accel_common(c, form, eps, acc, false, "nb_accel");
}

nb_status nb_compute_forces(nb_ctx* c, int form, double eps, double* const force[3]) {
    return accel_common(c, form, eps, force, true, "nb_compute_forces");
}

nb_status nb_accel_device(nb_ctx* c, int form, double eps) {
    if (!c) return fail(NB_ERR_ARG, "nb_accel_device: null context");
    if (nb_status s = check_eps(eps, "nb_accel_device")) return s;
    const int kf = kernel_form(form);
    if (kf < 0) return fail(NB_ERR_ARG, "nb_accel_device: unknown form");
    if (c->n == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(accel_to_acc(c, kf, eps), "force kernel");
    return NB_OK;
}

// Shared body of the ring-share entry points: d_pi (ni records) against d_pj
// (nj records) in j-blocks of jb, partials then one ordered reduction.
// Equal j masses (the pair terms then leave the mass out): known on the host
// (eq_mass > 0) or found on the device into *eq_dev (when set) before the
// force kernels, which read it there.
static nb_status block_eval(nb_ctx* c, int kf, double eps, long ni, const P4* d_pi, long nj,
                            const P4* d_pj, long jb, P4* d_out, P4* fixed = nullptr, size_t fixed_bytes = 0,
                            double eq_mass = 0.0, nbk::EqState* eq_dev = nullptr) {
    const int ipt = choose_ipt(c, ni);
    const long nb = (nj + jb - 1) / jb;
    int S = choose_splits(c, ni, jb, ipt);
    long jchunk = (jb + S - 1) / S;
    jchunk = (jchunk + 63) & ~63L;
    S = int((jb + jchunk - 1) / jchunk);
    const long parts = nb * S;
    P4* part = d_out;
    if (parts > 1) {
        const size_t need = size_t(parts) * size_t(ni) * sizeof(P4);
        if (fixed) {   // partial sums in caller-provided memory: nothing allocated, nothing timed
            if (need > fixed_bytes)
                return fail(NB_ERR_ARG, "nb_accel_block_device_noalloc: " + std::to_string(parts) + " partial sums of " +
                                            std::to_string(ni) + " records need " + std::to_string(need) +
                                            " bytes; the static scratch holds " + std::to_string(fixed_bytes));
            part = fixed;
        } else {
            NB_HIP(c->part.ensure(need), "allocate partial sums");
            part = (P4*)c->part.p;
        }
    }
    if (eq_dev) NB_HIP(nbk::launch_mass_check(d_pj, int(nj), eq_dev, c->stream), "mass check kernels");
    for (long b = 0; b < nb; b++) {
        nbk::AccelLaunch L;
        L.form = kf;
        L.ipt = ipt;
        L.pi = d_pi;
        L.ni = int(ni);
        L.pj = d_pj;
        L.jbeg = int(b * jb);
        L.jend = int(std::min(nj, (b + 1) * jb));
        L.jchunk = int(jchunk);
        L.splits = S;
        L.eps2 = eps * eps;
        L.G = 1.0;
        L.out = part;
        L.out_stride = ni;
        L.split_base = int(b * S);
        L.eq_mass = eq_mass;
        L.eq_dev = eq_dev;
        const long ev = fixed ? -1 : timing_begin(c);
        hipError_t e = nbk::launch_accel(L, c->stream);
        timing_end(c, ev);
        NB_HIP(e, "force kernel");
    }
    if (parts > 1) NB_HIP(nbk::launch_reduce(part, int(parts), ni, int(ni), d_out, c->stream), "reduce kernel");
    return NB_OK;
}

static nb_status block_args(nb_ctx* c, int form, double eps, size_t ni, const double* pi, size_t nj,
                            const double* pj, const void* out, const char* fn, int& kf) {
    if (!c) return fail(NB_ERR_ARG, std::string(fn) + ": null context");
    if ((ni && (!pi || !out)) || (nj && !pj)) return fail(NB_ERR_ARG, std::string(fn) + ": null buffer");
    if (ni > kMaxN || nj > kMaxN) return fail(NB_ERR_ARG, std::string(fn) + ": sizes must stay below 1e8");
    if (nb_status s = check_eps(eps, fn)) return s;
    kf = kernel_form(form);
    if (kf < 0) return fail(NB_ERR_ARG, std::string(fn) + ": unknown form");
    return NB_OK;
}

nb_status nb_accel_block(nb_ctx* c, int form, double eps, size_t ni, const double* pi, size_t nj,
                         const double* pj, size_t j_block, double* acc) {
    int kf;
    if (nb_status s = block_args(c, form, eps, ni, pi, nj, pj, acc, "nb_accel_block", kf)) return s;
    if (ni == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    const long jb = (j_block == 0 || j_block > nj) ? long(std::max<size_t>(nj, 1)) : long(j_block);
    NB_HIP(c->bi.ensure(ni * sizeof(P4)), "allocate i block");
    NB_HIP(c->bj.ensure(std::max<size_t>(nj, 1) * sizeof(P4)), "allocate j set");
    NB_HIP(c->bout.ensure(ni * sizeof(P4)), "allocate output");
    NB_HIP(hipMemcpyAsync(c->bi.p, pi, ni * sizeof(P4), hipMemcpyHostToDevice, c->stream), "upload i block");
    if (nj) NB_HIP(hipMemcpyAsync(c->bj.p, pj, nj * sizeof(P4), hipMemcpyHostToDevice, c->stream), "upload j set");
    // the j records are at hand: when their masses are all equal the pair
    // terms leave the mass out (the device entry points keep it in)
    double eq = (nj > 0 && pj[3] > 0.0 && std::isfinite(pj[3])) ? pj[3] : 0.0;
    for (size_t j = 1; j < nj && eq != 0.0; j++)
        if (pj[4 * j + 3] != pj[3]) eq = 0.0;
    if (nj == 0) {
        NB_HIP(hipMemsetAsync(c->bout.p, 0, ni * sizeof(P4), c->stream), "clear output");
    } else if (nb_status s = block_eval(c, kf, eps, long(ni), (const P4*)c->bi.p, long(nj), (const P4*)c->bj.p, jb,
                                        (P4*)c->bout.p, nullptr, 0, eq)) {
        return s;
    }
    std::vector<P4> h(ni);
    NB_HIP(hipMemcpyAsync(h.data(), c->bout.p, ni * sizeof(P4), hipMemcpyDeviceToHost, c->stream), "download output");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < ni; i++) {
        acc[3 * i] = h[i].x;
        acc[3 * i + 1] = h[i].y;
        acc[3 * i + 2] = h[i].z;
    }
    return NB_OK;
}

nb_status nb_accel_block_device(nb_ctx* c, int form, double eps, size_t ni, const double* d_pi, size_t nj,
                                const double* d_pj, size_t j_block, double* d_acc4) {
    int kf;
    if (nb_status s = block_args(c, form, eps, ni, d_pi, nj, d_pj, d_acc4, "nb_accel_block_device", kf)) return s;
    if (ni == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    if (nj == 0) {
        NB_HIP(hipMemsetAsync(d_acc4, 0, ni * sizeof(P4), c->stream), "clear output");
        return NB_OK;
    }
    const long jb = (j_block == 0 || j_block > nj) ? long(nj) : long(j_block);
    NB_HIP(c->eqstate.ensure(sizeof(nbk::EqState)), "allocate mass check");
    return block_eval(c, kf, eps, long(ni), (const P4*)d_pi, long(nj), (const P4*)d_pj, jb, (P4*)d_acc4, nullptr, 0,
                      0.0, (nbk::EqState*)c->eqstate.p);
}

// One ring round on the device: the local set d_l against the received set
// d_r (n records each, masses in the records) with the symmetric kernel's
// two-set form; m a accumulated into the force records d_fl, d_fr.
static nb_status ring_round_eval(nb_ctx* c, int kf, double eps, long n, const P4* d_l, const P4* d_r, bool incl,
                                 P4* d_fl, P4* d_fr) {
    const int K = (n >= 131072) ? (kf == nbk::FORM_FP64 ? 8 : 6) : 4;
    if (grouped_size(n)) {
        const nb_ctx::SymgPlan* plan = nullptr;
        NB_HIP(symg_plan(c, size_t(n), K, plan), "symmetric plan");
        const long S = 64L * K;
        const SymRegions R = symg_regions(plan->nB, S, plan->nseg, kf);
        NB_HIP(c->part.ensure(R.act_bytes + R.react_bytes), "allocate partial sums");
        char* const react = (char*)c->part.p + R.act_bytes;
        nbk::SymgLaunch L;
        L.form = kf;
        L.K = K;
        L.pm = d_l;
        L.pj = d_r;
        L.incl = incl;
        L.n = int(n);
        L.ranges = (const int4*)plan->ranges.p;
        L.nwg = plan->nwg;
        L.nB = plan->nB;
        L.eps2 = eps * eps;
        L.act = (double*)c->part.p;
        if (kf == nbk::FORM_FP64) L.react64 = (double*)react;
        else L.react32 = (float*)react;
        const long ev = timing_begin(c);
        hipError_t e = nbk::launch_accel_symg(L, c->stream);
        timing_end(c, ev);
        NB_HIP(e, "force kernel");
        nbk::SymgSums P;
        P.act = L.act;
        P.rowseg = (const int2*)plan->rowseg.p;
        P.react64 = L.react64;
        P.react32 = L.react32;
        P.nB = plan->nB;
        P.S = int(S);
        NB_HIP(nbk::launch_symg_ring_finish(P, int(n), d_l, d_r, d_fl, d_fr, c->stream), "ring finish kernel");
        return NB_OK;
    }
    const nb_ctx::SymPlan* plan = nullptr;
    NB_HIP(sym_plan(c, size_t(n), K, plan), "symmetric plan");
    const long S = 64L * K;
    SymRegions R = sym_regions(plan->nB, S, plan->L, kf);
    R.react_bytes += size_t(plan->nB) * size_t(S) * (kf == nbk::FORM_FP64 ? sizeof(P4) : sizeof(float4));
    NB_HIP(c->part.ensure(R.act_bytes + R.react_bytes), "allocate partial sums");
    char* const react = (char*)c->part.p + R.act_bytes;
    nbk::SymLaunch L;
    L.form = kf;
    L.K = K;
    L.pm = d_l;
    L.pj = d_r;
    L.incl = incl;
    L.n = int(n);
    L.items = (const int4*)plan->items.p;
    L.nitems = plan->nitems;
    L.nB = plan->nB;
    L.eps2 = eps * eps;
    L.part = (P4*)c->part.p;
    if (kf == nbk::FORM_FP64) L.react64 = (P4*)react;
    else L.react32 = (float4*)react;
    const long ev = timing_begin(c);
    hipError_t e = nbk::launch_accel_sym(L, c->stream);
    timing_end(c, ev);
    NB_HIP(e, "force kernel");
    nbk::SymSums P;
    P.part = L.part;
    P.react64 = L.react64;
    P.react32 = L.react32;
    P.nB = plan->nB;
    P.S = int(S);
    P.L = plan->L;
    NB_HIP(nbk::launch_ring_finish(P, int(n), d_l, d_r, d_fl, d_fr, c->stream), "ring finish kernel");
    return NB_OK;
}

static nb_status ring_args(nb_ctx* c, int form, double eps, size_t n, size_t rank, size_t comm_sz, const char* fn,
                           int& kf) {
    if (!c) return fail(NB_ERR_ARG, std::string(fn) + ": null context");
    if (comm_sz == 0 || rank >= comm_sz) return fail(NB_ERR_ARG, std::string(fn) + ": need 0 <= rank < comm_sz");
    if (n > kMaxN || comm_sz > kMaxN || n * comm_sz > kMaxN)
        return fail(NB_ERR_ARG, std::string(fn) + ": comm_sz * local_n must stay below 1e8");
    if (nb_status s = check_eps(eps, fn)) return s;
    kf = kernel_form(form);
    if (kf < 0 || kf == nbk::FORM_FP32)
        return fail(NB_ERR_ARG, std::string(fn) + ": form must be NB_FORM_FP64, NB_FORM_MIXED or NB_FORM_MIXED_RAW");
    return NB_OK;
}

nb_status nb_accel_block_device_noalloc(nb_ctx* c, int form, double eps, size_t ni, const double* d_pi, size_t nj,
                                        const double* d_pj, size_t j_block, double* d_acc4) {
    int kf;
    if (nb_status s = block_args(c, form, eps, ni, d_pi, nj, d_pj, d_acc4, "nb_accel_block_device_noalloc", kf))
        return s;
    if (ni == 0) return NB_OK;
    if (!c->scratch_ptr)
        return fail(NB_ERR_DEVICE, std::string("nb_accel_block_device_noalloc: no address of the static scratch on "
                                               "this device: ") + hipGetErrorString(c->scratch_status));
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    if (nj == 0) {   // a fill kernel, not a memset node, so the call can be captured
        hipLaunchKernelGGL(k_zero_records, dim3(unsigned((ni + 255) / 256)), dim3(256), 0, c->stream, (P4*)d_acc4,
                           long(ni));
        NB_HIP(hipGetLastError(), "zero kernel");
        return NB_OK;
    }
    const long jb = (j_block == 0 || j_block > nj) ? long(nj) : long(j_block);
    // the scratch's first 256 bytes hold the mass check, the partial sums follow
    return block_eval(c, kf, eps, long(ni), (const P4*)d_pi, long(nj), (const P4*)d_pj, jb, (P4*)d_acc4,
                      c->scratch_ptr + kScratchHead / sizeof(P4), kStaticScratchBytes, 0.0,
                      (nbk::EqState*)c->scratch_ptr);
}

nb_status nb_forces_ring_round(nb_ctx* c, int form, double eps, size_t local_n, const double* mass,
                               const double* const local_pos[3], double* const local_force[3],
                               const double* const recv_pos[3], double* const recv_force[3], size_t rank,
                               size_t comm_sz, size_t recv_round) {
    int kf;
    if (nb_status s = ring_args(c, form, eps, local_n, rank, comm_sz, "nb_forces_ring_round", kf)) return s;
    if (local_n == 0) return NB_OK;
    auto planes = [](const void* const* v) { return v && v[0] && v[1] && v[2]; };
    if (!mass || !planes((const void* const*)local_pos) || !planes((const void* const*)local_force) ||
        !planes((const void* const*)recv_pos) || !planes((const void* const*)recv_force))
        return fail(NB_ERR_ARG, "nb_forces_ring_round: null array");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    const size_t n = local_n;
    const size_t partner = (rank + recv_round) % comm_sz;
    const double* ml = mass + rank * n;
    const double* mr = mass + partner * n;
    std::vector<P4> h(4 * n);
    for (size_t i = 0; i < n; i++) {
        h[i] = P4{local_pos[0][i], local_pos[1][i], local_pos[2][i], ml[i]};
        h[n + i] = P4{recv_pos[0][i], recv_pos[1][i], recv_pos[2][i], mr[i]};
        h[2 * n + i] = P4{local_force[0][i], local_force[1][i], local_force[2][i], 0.0};
        h[3 * n + i] = P4{recv_force[0][i], recv_force[1][i], recv_force[2][i], 0.0};
    }
    NB_HIP(c->bi.ensure(2 * n * sizeof(P4)), "allocate blocks");
    NB_HIP(c->bout2.ensure(2 * n * sizeof(P4)), "allocate forces");
    P4* d_sets = (P4*)c->bi.p;
    P4* d_forces = (P4*)c->bout2.p;
    NB_HIP(hipMemcpyAsync(d_sets, h.data(), 2 * n * sizeof(P4), hipMemcpyHostToDevice, c->stream), "upload blocks");
    NB_HIP(hipMemcpyAsync(d_forces, h.data() + 2 * n, 2 * n * sizeof(P4), hipMemcpyHostToDevice, c->stream),
           "upload forces");
    if (nb_status s = ring_round_eval(c, kf, eps, long(n), d_sets, d_sets + n, rank < partner, d_forces, d_forces + n))
        return s;
    NB_HIP(hipMemcpyAsync(h.data(), d_forces, 2 * n * sizeof(P4), hipMemcpyDeviceToHost, c->stream),
           "download forces");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            local_force[d][i] = (&h[i].x)[d];
            recv_force[d][i] = (&h[n + i].x)[d];
        }
    return NB_OK;
}

nb_status nb_forces_ring_round_device(nb_ctx* c, int form, double eps, size_t local_n, const double* d_local,
                                      double* d_local_force4, const double* d_recv, double* d_recv_force4,
                                      size_t rank, size_t comm_sz, size_t recv_round) {
    int kf;
    if (nb_status s = ring_args(c, form, eps, local_n, rank, comm_sz, "nb_forces_ring_round_device", kf)) return s;
    if (local_n == 0) return NB_OK;
    if (!d_local || !d_local_force4 || !d_recv || !d_recv_force4)
        return fail(NB_ERR_ARG, "nb_forces_ring_round_device: null buffer");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    const size_t partner = (rank + recv_round) % comm_sz;
    return ring_round_eval(c, kf, eps, long(local_n), (const P4*)d_local, (const P4*)d_recv, rank < partner,
                           (P4*)d_local_force4, (P4*)d_recv_force4);
}

nb_status nb_step(nb_ctx* c, size_t time_steps, double dt, const double* cc, const double* dd, size_t lf_steps,
                  int form, double eps) {
    if (!c) return fail(NB_ERR_ARG, "nb_step: null context");
    if (!cc || !dd) return fail(NB_ERR_ARG, "nb_step: null coefficient array");
    if (lf_steps < 2 || lf_steps > 16) return fail(NB_ERR_ARG, "nb_step: lf_steps must be in [2, 16]");
    if (!std::isfinite(dt)) return fail(NB_ERR_ARG, "nb_step: dt must be finite");
    if (nb_status s = check_eps(eps, "nb_step")) return s;
    if (form == NB_FORM_LEARNED) return nbl::step_learned(c, time_steps, dt, cc, dd, lf_steps, eps);
    const int kf = kernel_form(form);
    if (kf < 0) return fail(NB_ERR_ARG, "nb_step: unknown form");
    if (c->n == 0 || time_steps == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    nbl::on_new_state(c->learned);   // the learned form's history no longer describes the state
    const int n = int(c->n);
    P4* pm = (P4*)c->pm.p;
    P4* vel = (P4*)c->vel.p;
    if (c->timing) (void)hipEventRecord(c->step_ev[0], c->stream);
    NB_HIP(nbk::launch_drift(n, pm, vel, cc[0] * dt, c->stream), "drift kernel");
    for (size_t t = 0; t < time_steps; t++) {
        for (size_t s = 0; s + 1 < lf_steps; s++) {
            ForceOut out;
            NB_HIP(force_eval(c, kf, eps, out), "force kernel");
            double c1, c2 = 0.0;
            if (s + 2 < lf_steps) {
                c1 = cc[s + 1] * dt;
            } else {
                c1 = cc[lf_steps - 1] * dt;
                if (t + 1 < time_steps) c2 = cc[0] * dt;
            }
            NB_HIP(finish_kick_drift(c, out, dd[s] * dt, c1, c2), "kick/drift kernel");
        }
    }
    if (c->timing) {
        (void)hipEventRecord(c->step_ev[1], c->stream);
        NB_HIP(hipEventSynchronize(c->step_ev[1]), "hipEventSynchronize");
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->step_ev[0], c->step_ev[1]) == hipSuccess) c->step_ms += ms;
    }
    return NB_OK;
}

// Whole run in the FPGA host's record form (lrbd_solver_no_sync,
// FPGA/host/solver.cpp:1112): the records become the resident state with the
// pair weights as masses, the kick's factor K coeff_i / mass_i (one value)
// scales the kick coefficients, nb_step runs the steps, and the records come
// back with force = K coeff_i a_i of the last evaluation.
nb_status nb_run_records(nb_ctx* c, int form, double eps, double K, const double* mass, const double* coeff,
                         double* pos4, double* vel4, double* force4, size_t n, size_t time_steps, double dt,
                         const double* cc, const double* dd, size_t lf_steps, double* seconds) {
    const char* fn = "nb_run_records";
    if (seconds) *seconds = 0.0;
    if (!c) return fail(NB_ERR_ARG, "nb_run_records: null context");
    if (n > kMaxN) return fail(NB_ERR_ARG, "nb_run_records: N must stay below 1e8");
    if (n > 0 && (!mass || !pos4 || !vel4)) return fail(NB_ERR_ARG, "nb_run_records: null mass, position or velocity records");
    if (!cc || !dd) return fail(NB_ERR_ARG, "nb_run_records: null coefficient array");
    if (lf_steps < 2 || lf_steps > 16) return fail(NB_ERR_ARG, "nb_run_records: lf_steps must be in [2, 16]");
    if (!std::isfinite(dt)) return fail(NB_ERR_ARG, "nb_run_records: dt must be finite");
    if (!std::isfinite(K)) return fail(NB_ERR_ARG, "nb_run_records: K must be finite");
    if (nb_status s = check_eps(eps, fn)) return s;
    if (form != NB_FORM_LEARNED && kernel_form(form) < 0) return fail(NB_ERR_ARG, "nb_run_records: unknown form");
    const double* w = coeff ? coeff : mass;
    double f = 1.0;
    for (size_t i = 0; i < n; i++) {
        if (!(mass[i] > 0.0) || !std::isfinite(mass[i]) || !std::isfinite(w[i]))
            return fail(NB_ERR_ARG, "nb_run_records: masses must be positive and finite, coefficients finite");
        const double fi = K * w[i] / mass[i];
        if (i == 0) f = fi;
        else if (fi != f)
            return fail(NB_ERR_ARG, "nb_run_records: K coeff[i] / mass[i] must be the same for every particle "
                                    "(the reference's host passes coeff = mass)");
    }
    if (!std::isfinite(f)) return fail(NB_ERR_ARG, "nb_run_records: K coeff[i] / mass[i] is not finite");
    if (n == 0 || time_steps == 0) return NB_OK;
    std::vector<double> P(3 * n), V(3 * n);
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            P[k * n + i] = pos4[4 * i + k];
            V[k * n + i] = vel4[4 * i + k];
        }
    double* const pp[3] = {P.data(), P.data() + n, P.data() + 2 * n};
    double* const vp[3] = {V.data(), V.data() + n, V.data() + 2 * n};
    if (nb_status s = nb_set_state(c, n, w, pp, vp)) return s;
    std::vector<double> dk(lf_steps - 1);
    for (size_t s = 0; s + 1 < lf_steps; s++) dk[s] = f * dd[s];
    hipEvent_t ev[2] = {nullptr, nullptr};
    NB_HIP(hipEventCreate(&ev[0]), "hipEventCreate");
    if (hipError_t e = hipEventCreate(&ev[1]); e != hipSuccess) {
        (void)hipEventDestroy(ev[0]);
        return fail(NB_ERR_DEVICE, std::string("nb_run_records: hipEventCreate: ") + hipGetErrorString(e));
    }
    auto drop = [&]() {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    };
    hipError_t e = hipEventRecord(ev[0], c->stream);
    nb_status s = e == hipSuccess ? nb_step(c, time_steps, dt, cc, dk.data(), lf_steps, form, eps) : NB_ERR_DEVICE;
    if (s == NB_OK && (e = hipEventRecord(ev[1], c->stream)) == hipSuccess) e = hipEventSynchronize(ev[1]);
    float ms = 0.0f;
    if (s == NB_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    drop();
    if (s != NB_OK) return s;
    if (e != hipSuccess) return fail(NB_ERR_DEVICE, std::string("nb_run_records: timing events: ") + hipGetErrorString(e));
    if (seconds) *seconds = 1e-3 * double(ms);
    if (nb_status g = nb_get_state(c, pp, vp)) return g;
    for (size_t i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            pos4[4 * i + k] = P[k * n + i];
            vel4[4 * i + k] = V[k * n + i];
        }
    if (force4) {
        if (nb_status g = download_acc(c, pp, false)) return g;
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) force4[4 * i + k] = K * w[i] * P[k * n + i];
    }
    return NB_OK;
}

// Shared argument checks of the Euler and Runge-Kutta drivers.
static nb_status driver_args(nb_ctx* c, double dt, int form, double eps, const char* fn, int& kf) {
    if (!c) return fail(NB_ERR_ARG, std::string(fn) + ": null context");
    if (!std::isfinite(dt)) return fail(NB_ERR_ARG, std::string(fn) + ": dt must be finite");
    if (nb_status s = check_eps(eps, fn)) return s;
    kf = kernel_form(form);
    if (kf < 0) return fail(NB_ERR_ARG, std::string(fn) + ": form must be NB_FORM_FP64, NB_FORM_MIXED, NB_FORM_MIXED_RAW or NB_FORM_FP32");
    return NB_OK;
}

static void step_timer_begin(nb_ctx* c) {
    if (c->timing) (void)hipEventRecord(c->step_ev[0], c->stream);
}

static nb_status step_timer_end(nb_ctx* c) {
    if (!c->timing) return NB_OK;
    (void)hipEventRecord(c->step_ev[1], c->stream);
    NB_HIP(hipEventSynchronize(c->step_ev[1]), "hipEventSynchronize");
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->step_ev[0], c->step_ev[1]) == hipSuccess) c->step_ms += ms;
    return NB_OK;
}

nb_status nb_step_euler(nb_ctx* c, size_t time_steps, double dt, int form, double eps) {
    int kf;
    if (nb_status s = driver_args(c, dt, form, eps, "nb_step_euler", kf)) return s;
    if (c->n == 0 || time_steps == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    nbl::on_new_state(c->learned);
    const int n = int(c->n);
    step_timer_begin(c);
    for (size_t t = 0; t < time_steps; t++) {
        NB_HIP(accel_to_acc(c, kf, eps), "force kernel");
        NB_HIP(nbk::launch_euler(n, (P4*)c->pm.p, (P4*)c->vel.p, (const P4*)c->acc.p, dt, c->stream), "Euler kernel");
    }
    return step_timer_end(c);
}

nb_status nb_runge_kutta(int order, double* a, double* b, size_t* rk_steps) {
    if (!a || !b || !rk_steps) return fail(NB_ERR_ARG, "nb_runge_kutta: null argument");
    *rk_steps = runge_kutta(order, a, b);
    return NB_OK;
}

nb_status nb_step_rk(nb_ctx* c, size_t time_steps, double dt, const double* a, const double* b, size_t rk_steps,
                     int form, double eps) {
    int kf;
    if (nb_status s = driver_args(c, dt, form, eps, "nb_step_rk", kf)) return s;
    if (rk_steps < 1 || rk_steps > size_t(nbk::kMaxRK)) return fail(NB_ERR_ARG, "nb_step_rk: rk_steps must be in [1, 8]");
    if (!b || (rk_steps > 1 && !a)) return fail(NB_ERR_ARG, "nb_step_rk: null coefficient array");
    if (c->n == 0 || time_steps == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    nbl::on_new_state(c->learned);
    const long n = long(c->n);
    const size_t s = rk_steps;
    NB_HIP(c->rk.ensure((2 * s + 1) * size_t(n) * sizeof(P4)), "allocate Runge-Kutta stages");
    P4* x0 = (P4*)c->rk.p;
    P4* kv = x0 + n;
    P4* ka = kv + s * n;
    P4* pm = (P4*)c->pm.p;
    P4* vel = (P4*)c->vel.p;
    const size_t bytes = size_t(n) * sizeof(P4);
    const size_t ad = s - 1;   // row stride of a (CPU/omp_solver.cpp:514)
    step_timer_begin(c);
    for (size_t t = 0; t < time_steps; t++) {
        NB_HIP(hipMemcpyAsync(x0, pm, bytes, hipMemcpyDeviceToDevice, c->stream), "save positions");
        NB_HIP(hipMemcpyAsync(kv, vel, bytes, hipMemcpyDeviceToDevice, c->stream), "save velocities");
        for (size_t k = 0; k < s; k++) {
            // stage k: acceleration at the stage positions now in pm
            NB_HIP(accel_to_acc(c, kf, eps), "force kernel");
            NB_HIP(hipMemcpyAsync(ka + k * n, c->acc.p, bytes, hipMemcpyDeviceToDevice, c->stream), "keep stage");
            if (k + 1 == s) break;
            nbk::RKTerms T;
            T.count = int(k + 1);
            for (size_t j = 0; j <= k; j++) T.w[j] = a[k * ad + j];
            NB_HIP(nbk::launch_rk_combine(int(n), x0, kv, ka, n, T, dt, pm, kv + (k + 1) * n, c->stream),
                   "Runge-Kutta stage kernel");
        }
        nbk::RKTerms T;
        T.count = int(s);
        for (size_t j = 0; j < s; j++) T.w[j] = b[j];
        NB_HIP(nbk::launch_rk_combine(int(n), x0, kv, ka, n, T, dt, pm, vel, c->stream), "Runge-Kutta step kernel");
    }
    return step_timer_end(c);
}

nb_status nb_energy(nb_ctx* c, double eps, double* T, double* U) {
    if (!c) return fail(NB_ERR_ARG, "nb_energy: null context");
    if (!T || !U) return fail(NB_ERR_ARG, "nb_energy: null output");
    if (nb_status s = check_eps(eps, "nb_energy")) return s;
    *T = 0.0;
    *U = 0.0;
    const long n = long(c->n);
    if (n == 0) return NB_OK;
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    // Workgroups of paired tile rows, split so the grid holds ~8 per CU.
    const long nblk = (n + nbk::BLOCK - 1) / nbk::BLOCK;
    const long pairs = (nblk + 1) / 2;
    const long S = std::min(nblk + 1, std::max(1L, (long(c->num_cus) * 8 + pairs - 1) / pairs));
    NB_HIP(c->rows.ensure(size_t(S) * n * 16), "allocate energy rows");
    NB_HIP(c->eblk.ensure(size_t(nblk) * 4 * sizeof(double)), "allocate energy blocks");
    NB_HIP(nbk::launch_energy((const P4*)c->pm.p, (const P4*)c->vel.p, int(n), int(S), eps * eps, c->rows.p,
                              (double*)c->eblk.p, c->stream),
           "energy kernels");
    std::vector<double> h(size_t(nblk) * 4);
    NB_HIP(hipMemcpyAsync(h.data(), c->eblk.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream),
           "download energy blocks");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    // Neumaier sum of the block partials in block order.
    auto nsum = [&](int off) {
        double s = 0.0, comp = 0.0;
        for (long b = 0; b < nblk; b++)
            for (int k = 0; k < 2; k++) {
                const double v = h[4 * b + off + k];
                const double t = s + v;
                comp += (std::fabs(s) >= std::fabs(v)) ? (s - t) + v : (v - t) + s;
                s = t;
            }
        return s + comp;
    };
    *U = -nsum(0);   // sum_i m_i row_i counts every pair once; G = 1
    *T = 0.5 * nsum(2);
    return NB_OK;
}

nb_status nb_timing_enable(nb_ctx* c, int enable) {
    if (!c) return fail(NB_ERR_ARG, "nb_timing_enable: null context");
    c->timing = enable != 0;
    return NB_OK;
}

nb_status nb_timing_read(nb_ctx* c, double* force_ms, uint64_t* evals, double* step_ms) {
    if (!c) return fail(NB_ERR_ARG, "nb_timing_read: null context");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    for (size_t k = 0; k + 1 < c->ev_used; k += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->events[k], c->events[k + 1]) == hipSuccess) c->force_ms += ms;
    }
    c->ev_used = 0;
    if (force_ms) *force_ms = c->force_ms;
    if (evals) *evals = c->force_evals;
    if (step_ms) *step_ms = c->step_ms;
    return NB_OK;
}

nb_status nb_timing_reset(nb_ctx* c) {
    if (!c) return fail(NB_ERR_ARG, "nb_timing_reset: null context");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->ev_used = 0;
    c->force_ms = 0.0;
    c->force_evals = 0;
    c->step_ms = 0.0;
    return NB_OK;
}

nb_status nb_set_kernel(nb_ctx* c, int kernel) {
    if (!c) return fail(NB_ERR_ARG, "nb_set_kernel: null context");
    if (kernel != NB_KERNEL_AUTO && kernel != NB_KERNEL_TILE && kernel != NB_KERNEL_SYMMETRIC)
        return fail(NB_ERR_ARG, "nb_set_kernel: unknown kernel");
    c->kernel = kernel;
    return NB_OK;
}

nb_status nb_kernel_info_read(nb_ctx* c, nb_kernel_info* out) {
    if (!c || !out) return fail(NB_ERR_ARG, "nb_kernel_info_read: null argument");
    *out = c->kinfo;
    return NB_OK;
}

nb_status nb_synchronize(nb_ctx* c) {
    if (!c) return fail(NB_ERR_ARG, "nb_synchronize: null context");
    NB_HIP(hipSetDevice(c->device), "hipSetDevice");
    NB_HIP(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return NB_OK;
}

}  // extern "C"

