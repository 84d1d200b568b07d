// valu_ceiling.hip -- issue-rate probe of the vector instructions the force
// kernels are made of, on the device they run on.
//
// Each measurement runs one instruction kind in 8 independent dependency
// chains per lane, 8 waves per SIMD (so latency is hidden), for a fixed
// number of iterations, timed with HIP events; the result is lane-operations
// per CU per cycle at the given clock (default 2.4 GHz, the MI355X peak engine
// clock).  64 is the full rate of a wave64 VALU instruction on the 4 SIMDs of
// a CU (16 lanes each per cycle).  With the instruction mix of a kernel's
// inner loop these rates give the time that mix needs at best: the ceiling
// the force kernels are compared with.
//
// Usage: valu_ceiling [clock_GHz]   -- one JSON line per instruction kind.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kChains = 8;
constexpr int kThreads = 256;
constexpr int kDppWaveRol1 = 0x134;

enum Op { FMA_F64, MUL_F64, ADD_F64, RSQ_F64, FMA_F32, MUL_F32, RSQ_F32, CVT_F64_F32_PAIR, DPP_B32, NUM_OPS };
const char* kNames[NUM_OPS] = {"v_fma_f64", "v_mul_f64", "v_add_f64", "v_rsq_f64", "v_fma_f32",
                               "v_mul_f32", "v_rsq_f32", "v_cvt_f32_f64+v_cvt_f64_f32", "v_mov_b32_dpp"};
// instructions per chain step
const int kPerStep[NUM_OPS] = {1, 1, 1, 1, 1, 1, 1, 2, 1};

template <int OP>
__global__ __launch_bounds__(kThreads) void k_probe(double* __restrict__ out, int iters, double seed) {
    double a[kChains];
    float b[kChains];
    int c[kChains];
#pragma unroll
    for (int q = 0; q < kChains; q++) {
        a[q] = seed + 1e-3 * threadIdx.x + q;
        b[q] = (float)a[q];
        c[q] = (int)threadIdx.x + q;
    }
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int q = 0; q < kChains; q++) {
            if constexpr (OP == FMA_F64) a[q] = __builtin_fma(a[q], 0.9999999, 1e-7);
            if constexpr (OP == MUL_F64) a[q] = a[q] * 0.9999999;
            if constexpr (OP == ADD_F64) a[q] = a[q] + 1e-7;
            if constexpr (OP == RSQ_F64) a[q] = __builtin_amdgcn_rsq(a[q]);
            if constexpr (OP == FMA_F32) b[q] = __builtin_fmaf(b[q], 0.9999999f, 1e-7f);
            if constexpr (OP == MUL_F32) b[q] = b[q] * 0.9999999f;
            if constexpr (OP == RSQ_F32) b[q] = __builtin_amdgcn_rsqf(b[q]);
            if constexpr (OP == CVT_F64_F32_PAIR) a[q] = (double)(float)a[q];
            if constexpr (OP == DPP_B32) c[q] = __builtin_amdgcn_mov_dpp(c[q], kDppWaveRol1, 0xF, 0xF, false);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < kChains; q++) s += a[q] + (double)b[q] + (double)c[q];
    out[blockIdx.x * kThreads + threadIdx.x] = s;
}

template <int OP>
hipError_t launch(int blocks, double* out, int iters) {
    hipLaunchKernelGGL(k_probe<OP>, dim3(blocks), dim3(kThreads), 0, 0, out, iters, 1.5);
    return hipGetLastError();
}

hipError_t run(int op, int blocks, double* out, int iters) {
    switch (op) {
    case FMA_F64: return launch<FMA_F64>(blocks, out, iters);
    case MUL_F64: return launch<MUL_F64>(blocks, out, iters);
    case ADD_F64: return launch<ADD_F64>(blocks, out, iters);
    case RSQ_F64: return launch<RSQ_F64>(blocks, out, iters);
    case FMA_F32: return launch<FMA_F32>(blocks, out, iters);
    case MUL_F32: return launch<MUL_F32>(blocks, out, iters);
    case RSQ_F32: return launch<RSQ_F32>(blocks, out, iters);
    case CVT_F64_F32_PAIR: return launch<CVT_F64_F32_PAIR>(blocks, out, iters);
    case DPP_B32: return launch<DPP_B32>(blocks, out, iters);
    default: return hipErrorInvalidValue;
    }
}

#define CK(call)                                                                 \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));    \
            return 1;                                                            \
        }                                                                        \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    const double ghz = argc > 1 ? atof(argv[1]) : 2.4;
    int dev = 0;
    CK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    CK(hipGetDeviceProperties(&prop, dev));
    const int cus = prop.multiProcessorCount;
    const int blocks = cus * 8;   // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const int iters = 4096;
    double* out = nullptr;
    CK(hipMalloc(&out, sizeof(double) * blocks * kThreads));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    for (int op = 0; op < NUM_OPS; op++) {
        float ms = 0.f;
        for (int rep = 0; rep < 3; rep++) {   // the first launch warms up; the best of the other two counts
            CK(hipEventRecord(e0));
            CK(run(op, blocks, out, iters));
            CK(hipEventRecord(e1));
            CK(hipEventSynchronize(e1));
            float t = 0.f;
            CK(hipEventElapsedTime(&t, e0, e1));
            if (rep == 1 || (rep > 1 && t < ms)) ms = t;
        }
        const double lane_ops = double(blocks) * kThreads * iters * kChains * kPerStep[op];
        const double per_cu_cycle = lane_ops / (ms * 1e-3) / (double(cus) * ghz * 1e9);
        printf("{\"op\": \"%s\", \"ms\": %.4f, \"lane_ops_per_cu_cycle\": %.2f, \"fraction_of_full_rate\": %.3f, "
               "\"clock_ghz\": %.2f, \"cus\": %d}\n",
               kNames[op], ms, per_cu_cycle, per_cu_cycle / 64.0, ghz, cus);
    }
    CK(hipFree(out));
    return 0;
}
