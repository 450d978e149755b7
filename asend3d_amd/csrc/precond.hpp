// Multigrid / learned preconditioner interface used by the mixed-precision
// solver (definitions in precond.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <cstdint>

#include "kernels.hpp"
#include "learned.hpp"

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
// fine_elsewhere: the finest level's block inverses and bound are enqueued by
// mg_fine_setup on another stream (joined by the caller before the cycle is
// used); otherwise they are formed here and the bound on first use.  The
// caller then enqueues mg_setup_finish (after the join, if any), which copies
// the device-computed bounds back for mg_check_bounds.
int mg_setup(Multigrid* mg, hipStream_t st, const double* d_E, const double* d_nu, const float* d_K32,
             const float* d_dinv32, const int32_t* d_kdiag0, double sf, int n_pen, const int64_t* h_pen,
             const int32_t* h_mult, bool fine_elsewhere);
int mg_fine_setup(Multigrid* mg, hipStream_t st, const float* d_K32, const float* d_dinv32, const int32_t* d_kdiag0);
int mg_setup_finish(Multigrid* mg, hipStream_t st);

// Smoother of the finest level: 0 = Chebyshev with point Jacobi, 1 = Chebyshev
// with the learned factored preconditioner P = L L^T (mg_set_learned_factor).
int mg_set_fine_smoother(Multigrid* mg, int kind);
// Chebyshev smoother of every smoothed level: point (kind 0) or 3x3 block
// (kind 1) Jacobi, degree 1..8, interval [lo_frac, hi_frac] * lambda_max
// (default block, 2, [0.1, 1.15]).  Changing the kind needs a new mg_setup.
int mg_set_smoother(Multigrid* mg, int kind, int degree, double lo_frac, double hi_frac);
// Re-derive the Chebyshev step coefficients after mg_set_smoother changed the
// degree or interval (same Jacobi kind: the stored bounds stay valid).
// Waits for the coarse-level smoother bounds the last setup estimated on the
// device, keeps them on the host and reports a broken-down estimate.
int mg_check_bounds(Multigrid* mg);
int mg_refresh_coefficients(Multigrid* mg, hipStream_t st);
// New finest-level operator values on an unchanged hierarchy (same moduli,
// penalty and pattern below it): block-Jacobi inverse recomputed, Chebyshev
// bound re-estimated on first use, coarse levels kept.
int mg_refresh_fine(Multigrid* mg, hipStream_t st, const float* d_K32, const float* d_dinv32,
                    const int32_t* d_kdiag0);
int mg_smoother_degree(const Multigrid* mg);
// Workgroup geometry of the finest-level SpMV (waves per workgroup).
void mg_set_fine_geometry(Multigrid* mg, int waves_per_block);
// FP16 row-scaled copy of the finest operator (launch_to_f16_rowscaled) and
// the FP32 Jacobi diagonal that restores its rows; with cycle_half on, the
// cycle's finest-level residuals and the smoother bounds read it (the FCG's
// own operator products stay FP32).
void mg_set_fine_half(Multigrid* mg, const __half* val16, const float* dinv32);
void mg_set_cycle_half(Multigrid* mg, bool on);
// Register the learned factor (device arrays borrowed, must stay valid while
// it is in use); with smoother_bounds also estimate lambda_max(P A), which the
// Chebyshev smoother needs.
int mg_set_learned_factor(Multigrid* mg, hipStream_t st, const Factor& F, bool smoother_bounds);
// The same with the bound's power iteration overlapping the hierarchy setup:
// begin registers F and enqueues the power iteration on st (finest level,
// own scratch; d_K32 / d_dinv32 are the finest operator the setup will use);
// mg_setup may then run on another stream; end (after st has been ordered
// behind that setup) waits for st and uploads the Chebyshev coefficients.
int mg_learned_bound_begin(Multigrid* mg, hipStream_t st, const Factor& F, const float* d_K32,
                           const float* d_dinv32);
int mg_learned_bound_end(Multigrid* mg, hipStream_t st);
// Training through one V-cycle with the learned smoother: loss = energy of the
// error left by one cycle on b = A e, relative to e; fills the factor uses
// (inputs v_k of P and dLoss/dz_k) for learned_step_from_uses.  The returned
// pointers stay valid until the next call.
int mg_vcycle_uses(Multigrid* mg, hipStream_t st, const float* e, FactorUses* uses, double* loss);
bool mg_factor_has_bounds(const Multigrid* mg);
// Diagnostics: keep the previous factor's lambda_max when a new factor is set
// (finite-difference checks of the training gradient, which holds it fixed).
void mg_keep_factor_bounds(Multigrid* mg, bool keep);
// Replay each block of check_every FCG iterations as a captured hipGraph
// (default on); off launches the same kernels one by one.
void mg_use_graphs(Multigrid* mg, bool on);
bool mg_ready(const Multigrid* mg);
void mg_clear_learned_factor(Multigrid* mg);

// Reliable updates of a mixed-precision FCG solve.  run() adds the FP32
// correction x to the caller's FP64 solution and zeroes x, forms the FP64
// residual, writes r / |r| into the right-hand side b, and returns
// |r_old| / |r_new| (ratio), the recursive-residual level |r|^2 (relative to
// the new normalisation) that triggers the next refresh (next), and whether
// the FP64 residual meets the tolerance (converged).
struct FcgRefresh {
    double next = 1e-4;  // first trigger, relative to |b|^2 = 1
    std::function<int(double* ratio, double* next, bool* converged)> run;
    int refreshes = 0;
    bool converged = false;
};

// Flexible CG (one truncation vector) on the FP32 fine operator
// preconditioned by one V-cycle (precond 0) or by the learned factor alone
// (precond 1): solves A x = b to ||r|| <= rel_tol ||b||.  With ru, b must be
// normalised (|b| = 1); the solve instead runs until ru->run reports
// convergence (or max_it), refreshing the residual from FP64 whenever the
// recursive |r|^2 falls below ru->next: r = b, the search direction rescaled
// by the ratio, its previous directions kept (FCG takes alpha and beta from
// fresh inner products each iteration, so nothing else carries over).
int mg_fcg_solve(Multigrid* mg, hipStream_t st, const float* b, float* x, double rel_tol, int64_t max_it,
                 int check_every, int64_t* it_out, int* fallback_used, int precond, FcgRefresh* ru = nullptr);

// Diagnostics for tests: y = A_l x on level l (natural node order for l >= 1,
// fine order for l = 0), and the transfer operators between l and l+1.
int mg_levels(const Multigrid* mg);
int64_t mg_level_ndof(const Multigrid* mg, int l);
int mg_apply_level(Multigrid* mg, hipStream_t st, int l, const float* d_x, float* d_y);
int mg_restrict(Multigrid* mg, hipStream_t st, int l, const float* d_fine, float* d_coarse);
int mg_prolong(Multigrid* mg, hipStream_t st, int l, const float* d_coarse, float* d_fine);
int mg_vcycle(Multigrid* mg, hipStream_t st, const float* d_b, float* d_x);
// x = A_c^-1 b on the coarsest level (its explicit FP64 inverse)
int mg_coarse_solve(Multigrid* mg, hipStream_t st, const float* d_b, float* d_x);

}  // namespace asend
