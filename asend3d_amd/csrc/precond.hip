// Multigrid preconditioner (placeholder until the hierarchy is built).
#include "precond.hpp"

#include <string>

#include "asend_gpu.h"

namespace asend {

struct Multigrid {
    std::string err;
};

int mg_create(Multigrid** out, int, int, int, const int32_t*, const double*, const int32_t*, const int32_t*,
              const int32_t*, const uint8_t*, const int32_t*, const int32_t*, int64_t, int64_t) {
    *out = new Multigrid();
    (*out)->err = "multigrid preconditioner not available in this build";
    return ASEND_ERR_UNSUPPORTED;
}
void mg_destroy(Multigrid* mg) { delete mg; }
const char* mg_last_error(const Multigrid* mg) { return mg ? mg->err.c_str() : ""; }
int mg_setup(Multigrid*, hipStream_t, const double*, const double*, const float*, const float*, double) {
    return ASEND_ERR_UNSUPPORTED;
}
int mg_fcg_solve(Multigrid*, hipStream_t, const float*, float*, double, int64_t, int, int64_t*, int*) {
    return ASEND_ERR_UNSUPPORTED;
}

}  // namespace asend
