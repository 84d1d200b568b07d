// nbody_learned.hip -- learned acceleration extrapolator (placeholder until
// the MLP kernels land): every entry point reports that it is unavailable.
#include "nbody_host.h"
#include "nbody_learned.h"

namespace nbl {

struct Learned {};

void destroy(Learned* l) { delete l; }
void on_new_state(Learned*) {}

nb_status step_learned(nb_ctx*, size_t, double, const double*, const double*, size_t, double) {
    return nbh::set_error(NB_ERR_STATE, "nb_step: the learned form needs a trained extrapolator");
}

}  // namespace nbl

static nb_status unavailable(const char* fn) {
    return nbh::set_error(NB_ERR_STATE, std::string(fn) + ": learned extrapolator not available");
}

extern "C" {
nb_status nb_learned_configure(nb_ctx*, int, double, int) { return unavailable("nb_learned_configure"); }
nb_status nb_learned_set_mode(nb_ctx*, int) { return unavailable("nb_learned_set_mode"); }
nb_status nb_learned_collect(nb_ctx*, size_t, double, const double*, const double*, size_t, int, double) {
    return unavailable("nb_learned_collect");
}
nb_status nb_learned_train(nb_ctx*, int, int, double, uint64_t, double*) { return unavailable("nb_learned_train"); }
nb_status nb_learned_stats_read(nb_ctx*, nb_learned_stats*) { return unavailable("nb_learned_stats_read"); }
nb_status nb_mlp_param_count(size_t*) { return unavailable("nb_mlp_param_count"); }
nb_status nb_mlp_get_params(nb_ctx*, float*) { return unavailable("nb_mlp_get_params"); }
nb_status nb_mlp_set_params(nb_ctx*, const float*) { return unavailable("nb_mlp_set_params"); }
nb_status nb_mlp_forward(nb_ctx*, size_t, const float*, float*) { return unavailable("nb_mlp_forward"); }
nb_status nb_mlp_grad(nb_ctx*, size_t, const float*, const float*, float*, double*) {
    return unavailable("nb_mlp_grad");
}
}
