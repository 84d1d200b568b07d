// nbody_near.h -- short-range accelerations on a sorted uniform grid
// (nbody_near.hip), used by the learned form to split the force.
#pragma once

#include <hip/hip_runtime.h>

#include "nbody_host.h"
#include "nbody_kernels.h"

namespace nbn {

struct Near {
    nbh::DevBuf keys_in, keys_out, idx_in, idx_out, sorted, temp, scan_temp, cols, chunks, cur;
    int resident[4] = {0, 0, 0, 0};   // persistent grid of each k_near variant (0: not yet queried)
    // the grid of the last build: its particle count and radius (0: none)
    int built_n = 0;
    double built_rc = 0.0;

    // out[i] = sum over j with 0 < r_ij < rc of m_j S(r_ij) d_ij / (r_ij^2 + eps^2)^(3/2);
    // enqueued on st without synchronising it
    // fp32: FP32 pair arithmetic on FP64 differences (the mixed form's precision)
    // rebuild: sort the particles into a new grid; otherwise the grid of the
    // last build (same n and rc) is kept and searched wider by the largest
    // displacement since that build, measured on the device (exact either way)
    hipError_t near_accel(const nbk::P4* pm, int n, double rc, double eps, nbk::P4* out, hipStream_t st,
                          bool fp32 = false, bool rebuild = true);
    void release();
};

}  // namespace nbn
