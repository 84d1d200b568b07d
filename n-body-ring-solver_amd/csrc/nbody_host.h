// nbody_host.h -- host-side helpers shared inside the library (not public).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "nbody.h"
#include "nbody_kernels.h"

namespace nbl {
struct Learned;
}

namespace nbh {

// Initial conditions and integrator coefficients (nbody_ic.cpp).
void ic_sphere(size_t n, uint64_t seed, double* m, double* const pos[3], double* const vel[3]);
void ic_plummer(size_t n, uint64_t seed, double a, double rmax, double* m, double* const pos[3],
                double* const vel[3]);
void ic_ring(size_t n_total, double R, double omega, double ratio, double* m, double* const pos[3],
             double* const vel[3], double* const force[3]);
size_t yoshida(int order, double* c, double* d);
size_t runge_kutta(int order, double* a, double* b);

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b);
    void release();
};

// Where one force evaluation left its partial sums.
struct ForceOut {
    bool sym = false;
    nbk::P4* part = nullptr;              // tile: split partials; symmetric: action slots
    const nbk::P4* react64 = nullptr;     // symmetric, FP64 form: reaction slots
    const float4* react32 = nullptr;      // symmetric, mixed forms: FP32 reaction slots
    int S = 1;          // tile kernel: number of j-splits; symmetric: block size
    int nB = 0, L = 1;  // symmetric: blocks, tiles per work item
    double scale = 1.0;             // symmetric: factor of the sums (the mass an equal-mass evaluation left out)
    bool grouped = false;           // grouped symmetric kernel: its sums are `gsums` (the fields above unused)
    nbk::SymgSums gsums;
};

// Context internals used by the learned-form module.
nb_status set_error(nb_status s, const std::string& msg);   // nbody_status.cpp: text for nb_last_error()
void clear_error();
int choose_splits(const nb_ctx* c, long ni, long nj, int ipt);
int kernel_form(int form);
hipError_t force_eval(nb_ctx* c, int kform, double eps, ForceOut& out);
hipError_t finish_accel(nb_ctx* c, const ForceOut& out);
hipError_t finish_kick_drift(nb_ctx* c, const ForceOut& out, double kdt, double c1dt, double c2dt);
hipStream_t ctx_stream(nb_ctx* c);
size_t ctx_n(nb_ctx* c);
nbk::P4* ctx_pm(nb_ctx* c);
nbk::P4* ctx_vel(nb_ctx* c);
nbk::P4* ctx_acc(nb_ctx* c);
int ctx_num_cus(nb_ctx* c);
nbl::Learned*& ctx_learned(nb_ctx* c);

}  // namespace nbh
