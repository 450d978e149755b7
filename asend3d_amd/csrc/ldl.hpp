// Device LDL^T factor of the assembled operator restricted to diagonal blocks
// of the DOF numbering (the reference's LowerTriMat, S/lower_tri_mat_meth.rs):
// one block covering every DOF is an exact (direct) factor, smaller blocks
// give the block preconditioner of its conjugate gradient solver.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

namespace asend {

constexpr int kLdlTile = 64;  // tile edge of the banded storage (DOFs)
constexpr int kLdlTileElems = kLdlTile * kLdlTile;
// Most tiles a tile row may hold below the diagonal: the solve keeps the last
// kLdlMaxWt + 1 solution tiles of its block in LDS (64 doubles each).
constexpr int kLdlMaxWt = 239;
// Below this many blocks the solve runs one launch per tile row with the
// tiles of a row spread over workgroups; from it on, one workgroup streams
// each block (enough blocks to fill the chip).
constexpr int kLdlStepBlocks = 64;

// Banded tile storage of the factor.  Block b covers DOFs [b bs, min((b+1) bs,
// n)); inside a block, tile (I, J) (rows 64 I.., columns 64 J.., I - wt <= J
// <= I) is 64 x 64 row-major at
//   L + b stride + (row_off(I) + J - max(0, I - wt)) * 4096,
// row_off(I) = sum over I' < I of (min(I', wt) + 1).  Diagonal tiles hold
// D on their diagonal and the unit lower factor below it (the reference's
// storage of d_i at the end of each row); rows past the block end are padded
// with identity rows.
struct LdlDesc {
    int64_t n = 0;       // DOFs
    int64_t bs = 0;      // DOFs per block (the last block may be shorter)
    int n_blocks = 0;
    int nt_full = 0;     // tiles of a full block
    int wt = 0;          // tiles below the diagonal kept in a tile row
    int64_t stride = 0;  // doubles per block
};

__host__ __device__ inline int64_t ldl_row_off(int I, int wt) {
    return I <= wt ? (int64_t)I * (I + 1) / 2 : (int64_t)wt * (wt + 1) / 2 + (int64_t)(I - wt) * (wt + 1);
}

__host__ __device__ inline int ldl_block_nt(const LdlDesc& d, int b) {
    int64_t s = (int64_t)b * d.bs;
    int64_t m = d.n - s < d.bs ? d.n - s : d.bs;
    return (int)((m + kLdlTile - 1) / kLdlTile);
}

__host__ __device__ inline int64_t ldl_tile(const LdlDesc& d, int b, int I, int J) {
    int j0 = I - d.wt > 0 ? I - d.wt : 0;
    return (int64_t)b * d.stride + (ldl_row_off(I, d.wt) + (J - j0)) * (int64_t)kLdlTileElems;
}

struct LdlFactor {
    LdlDesc d;
    int64_t max_bandwidth = 0;  // the reference's maxBandwidth (row span incl. the diagonal)
    double* L = nullptr;        // banded tiles
    double* P = nullptr;        // per block: D^-1 L_kk^-1 of the current step (panel scratch)
    int32_t* neg = nullptr;     // count of negative pivots (pos_def false when > 0)
    double* work = nullptr;     // n doubles: the solve's work vector
    int64_t asm_id = -1;        // assembly the factor was computed from
    int64_t bytes = 0;
    double factor_ms = 0.0;
    int neg_pivots = 0;
};

// Plan the storage from the block-row pattern (h_rp/h_bc: CSR of 3x3 blocks,
// block row r = DOFs 3r..3r+2): the reference's envelope (first column of
// each row inside its block, allocate_from_sparse_mat), its max bandwidth and
// the tile band.  Returns 0 or an error message.
const char* ldl_plan(LdlDesc* d, int64_t* max_bw, int64_t n_dof, int64_t block_dofs, const std::vector<int32_t>& h_rp,
                     const std::vector<int32_t>& h_bc);

// Fill the banded tiles from the sliced-ELL FP64 operator (values of the
// envelope of each block), then factor them.  Every launch on st.
void ldl_fill(hipStream_t st, const LdlDesc& d, int n_rows, int n_slices, const int32_t* sptr, const int32_t* scol,
              const double* K, double* L);
void ldl_factor_launch(hipStream_t st, const LdlDesc& d, double* L, double* P, int32_t* neg);

// x = (L D L^T)^-1 b over every block (ldl_solve).  x may alias b; work is
// n doubles of scratch (distinct from both).  When done is given and *done
// is set the launches do nothing (solver loops).
void ldl_solve_launch(hipStream_t st, const LdlDesc& d, const double* L, const double* b, double* x, double* work,
                      const int* done);

}  // namespace asend
