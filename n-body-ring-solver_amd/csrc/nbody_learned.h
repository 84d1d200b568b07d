// nbody_learned.h -- the learned acceleration extrapolator (internal).
#pragma once

#include <cstddef>

#include "nbody.h"

namespace nbl {

struct Learned;

void destroy(Learned* l);
void on_new_state(Learned* l);
nb_status step_learned(nb_ctx* c, size_t time_steps, double dt, const double* cc, const double* dd,
                       size_t lf_steps, double eps);

}  // namespace nbl
