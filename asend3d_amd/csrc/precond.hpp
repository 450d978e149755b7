// Multigrid / learned preconditioner interface used by the mixed-precision
// solver (definitions in precond.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace asend {

struct Multigrid;

int mg_create(Multigrid** out, int nx, int ny, int nz, const int32_t* rank_of_grid, const double* d_xyz_r,
              const int32_t* d_conn_r, const int32_t* d_rp, const int32_t* d_bc, const uint8_t* d_slot,
              const int32_t* h_rp, const int32_t* h_bc, int64_t n_nodes, int64_t n_el);
void mg_destroy(Multigrid* mg);
const char* mg_last_error(const Multigrid* mg);
// (Re)build the hierarchy for the current fine operator.
int mg_setup(Multigrid* mg, hipStream_t st, const double* d_E, const double* d_nu, const float* d_K32,
             const float* d_dinv32, double sf);
// Flexible CG on the FP32 fine operator preconditioned by one V-cycle.
int mg_fcg_solve(Multigrid* mg, hipStream_t st, const float* b, float* x, double rel_tol, int64_t max_it,
                 int check_every, int64_t* it_out, int* fallback_used);

}  // namespace asend
