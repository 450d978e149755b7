// Multigrid / learned preconditioner interface used by the mixed-precision
// solver (definitions in precond.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace asend {

struct Multigrid;

// Structured hierarchy over an nx x ny x nz brick8 grid.  rank_of_grid maps
// the lexicographic grid node (i + (nx+1)(j + (ny+1)k)) to the fine operator's
// block row.  Device pointers are borrowed from the context and must outlive
// the multigrid object.
int mg_create(Multigrid** out, int nx, int ny, int nz, const int32_t* rank_of_grid, const double* d_xyz_r,
              const int32_t* d_conn_r, const Sell3& fine, int64_t n_nodes, int64_t n_el);
void mg_destroy(Multigrid* mg);
const char* mg_last_error(const Multigrid* mg);

// (Re)build coarse operators, smoother bounds and the coarsest inverse for the
// current fine operator (FP32 copy with the penalty folded in).  pen / mult
// list the penalised fine DOFs (unique) and their multiplicities.
int mg_setup(Multigrid* mg, hipStream_t st, const double* d_E, const double* d_nu, const float* d_K32,
             const float* d_dinv32, double sf, int n_pen, const int64_t* h_pen, const int32_t* h_mult);

// Smoother of the finest level: 0 = Chebyshev-Jacobi, 1 = Chebyshev with the
// learned node-block preconditioner (set by mg_set_learned_blocks).
int mg_set_fine_smoother(Multigrid* mg, int kind);
// Learned SPD node blocks for the fine smoother: d_blocks holds n_nodes
// row-major 3x3 blocks M_r (FP32, device), applied as z_r = M_r r_r.
int mg_set_learned_blocks(Multigrid* mg, const float* d_blocks);

// Flexible CG (one truncation vector) on the FP32 fine operator
// preconditioned by one V-cycle: solves A x = b to ||r|| <= rel_tol ||b||.
int mg_fcg_solve(Multigrid* mg, hipStream_t st, const float* b, float* x, double rel_tol, int64_t max_it,
                 int check_every, int64_t* it_out, int* fallback_used);

// Diagnostics for tests: y = A_l x on level l (natural node order for l >= 1,
// fine order for l = 0), and the transfer operators between l and l+1.
int mg_levels(const Multigrid* mg);
int64_t mg_level_ndof(const Multigrid* mg, int l);
int mg_apply_level(Multigrid* mg, hipStream_t st, int l, const float* d_x, float* d_y);
int mg_restrict(Multigrid* mg, hipStream_t st, int l, const float* d_fine, float* d_coarse);
int mg_prolong(Multigrid* mg, hipStream_t st, int l, const float* d_coarse, float* d_fine);
int mg_vcycle(Multigrid* mg, hipStream_t st, const float* d_b, float* d_x);

}  // namespace asend
