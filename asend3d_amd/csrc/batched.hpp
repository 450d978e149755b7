// Batched multi-right-hand-side solver over many design variants (definitions
// in batched.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

#include "kernels.hpp"

namespace asend {

constexpr int kNrhs = 16;  // right-hand sides per variant advanced together

// Device view of one assembled variant (owned by its context).
struct OperatorView {
    int n_rows = 0;                 // block rows (nodes)
    int n_slices = 0;
    int64_t n_slots = 0;
    const int32_t* h_sptr = nullptr;  // host copy of the slice pointer
    const int32_t* scol = nullptr;
    const float* K32 = nullptr;
    const double* K64 = nullptr;
    const float* dinv32 = nullptr;
};

struct BatchStats {
    int64_t outer_iterations;
    int64_t lockstep_iterations;    // inner iterations of the slowest system, summed over refinements
    int64_t system_iterations;      // inner iterations summed over all systems
    int64_t unconverged;
    double max_rel_residual;
    double device_ms;
};

struct Batch;
int batch_create(Batch** out, const std::vector<OperatorView>& ops);
void batch_destroy(Batch* b);
const char* batch_error(const Batch* b);
int64_t batch_rows_total(const Batch* b);
int64_t batch_row_offset(const Batch* b, int v);
// h_F / h_U: [rows_total][3][kNrhs] (FP64, host); systems with rhs index >= nrhs
// or zero right-hand side are skipped.
// verbose != 0 prints one progress line per refinement to stderr.
int batch_solve(Batch* b, const double* h_F, double* h_U, int nrhs, double tol, double inner_tol, int64_t max_it,
                int check_every, int verbose, BatchStats* stats);

}  // namespace asend
