// nbody_status.cpp -- the library's error state: the text of the calling
// thread's last failure, kept for nb_last_error().  Host-only, shared by every
// unit through nbh::set_error.
#include <string>

#include "nbody.h"
#include "nbody_host.h"

namespace {
thread_local std::string g_err;
}  // namespace

namespace nbh {

nb_status set_error(nb_status s, const std::string& msg) {
    g_err = msg;
    return s;
}

void clear_error() { g_err.clear(); }

}  // namespace nbh

extern "C" const char* nb_last_error(void) { return g_err.c_str(); }
