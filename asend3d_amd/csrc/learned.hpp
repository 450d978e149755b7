// Learned factored preconditioner: a 3-layer message-passing network on the
// mesh node graph (hidden 64, BF16 weights and activations on MFMA, FP32
// accumulation) predicts a sparse factor L with a 3x3 block per node and a
// 3x3 coupling per mesh-edge neighbour; the preconditioner is P = L L^T (SPD
// whenever L is nonsingular).  L is scaled by the Cholesky factors C_r of the
// operator's 3x3 diagonal blocks, L_rs = C_r^-T Lambda_rs, so Lambda = I is
// block Jacobi and the network only sees the dimensionless operator
// C^-1 A C^-T.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "kernels.hpp"

namespace asend {

constexpr int kFeat = 8;     // node input features
constexpr int kHid = 64;     // hidden width
constexpr int kLayers = 3;   // message-passing layers
constexpr int kHead = 24;    // head outputs: 6 node-block + 9 "from" + 9 "to" edge terms
constexpr int kMaxEdges = 16;  // edge-neighbour slots per node (hex meshes use at most 6)
// flat parameter layout (FP32 master copy)
constexpr int kOffW0 = 0;                                   // W0 [64][8]
constexpr int kOffB0 = kOffW0 + kHid * kFeat;               // b0 [64]
constexpr int kLayerSize = kHid * 2 * kHid + kHid;          // Wcat [64][128] + b [64]
constexpr int kOffL1 = kOffB0 + kHid;                       // layers 1..3
constexpr int kOffWh = kOffL1 + kLayers * kLayerSize;       // head W [24][64]
constexpr int kOffBh = kOffWh + kHead * kHid;               // head b [24]
constexpr int kOffWa = kOffBh + kHead;                      // edge map of C^-1 A_rs C^-T: [9][9]
constexpr int kNumParams = kOffWa + 81;

struct LearnedGraph {
    int n_nodes = 0;
    const int32_t* rp = nullptr;       // node-block row pointer (incl. self)
    const int32_t* bc = nullptr;       // neighbour ranks
    const int32_t* n2e_ptr = nullptr;  // node -> elements
    const int32_t* n2e = nullptr;
    const double* E = nullptr;         // per-element modulus
    const uint8_t* clamped = nullptr;  // per node: any penalised DOF
    int n_edge = 0;                    // edge slots per node (padded width)
    const int32_t* enb = nullptr;      // [n_edge][n] edge neighbour rank, -1 = none
    const int32_t* erev = nullptr;     // [n_edge][n] slot of this node in the neighbour's list
    const int32_t* eslot = nullptr;    // [n_edge][n] in-row slot of the neighbour in the operator
    Sell3 op;                          // fine operator layout
    const float* K32 = nullptr;        // fine operator values (FP32, penalty folded)
    const int32_t* kdiag = nullptr;    // in-row slot of the diagonal block
    double mean_log_e = 0.0;           // mean of log E over elements
};

// Device view of a factor for the apply kernels: Ls[(s*9 + c)*n + r] holds
// block s of row r (s = 0 the node block, s = 1 + k the k-th edge), row-major c.
struct Factor {
    int n = 0;
    int n_edge = 0;
    const float* Ls = nullptr;
    const float* LsT = nullptr;  // L^T in the same row layout (block s of row j = (L_ij)^T)
    const int32_t* enb = nullptr;
    const int32_t* erev = nullptr;
    // FP16 copies of Ls and LsT for the smoother, each node's row scaled by
    // the largest magnitude in it (sc16 / scT16 restore the scale)
    const __half* Ls16 = nullptr;
    const __half* LsT16 = nullptr;
    const float* sc16 = nullptr;
    const float* scT16 = nullptr;
};

// z = L L^T r (t is an n-vector of scratch): two coalesced row gathers, one
// over L^T and one over L; no atomics.
void launch_factor_apply(hipStream_t st, const Factor& F, const float* r, float* t, float* z);
// The same product with the FP16 copies (FP32 arithmetic).
void launch_factor_apply16(hipStream_t st, const Factor& F, const float* r, float* t, float* z);
// One Chebyshev step with P = L L^T from the FP16 copies: t = L^T r, then,
// fused into the second product, z = L t and (coef = {c1, c2} in device
// memory) first: d = c2 z, x = d; otherwise d = c1 d + c2 z, x += d.
void launch_factor_cheb16(hipStream_t st, const Factor& F, const float* r, float* t, const float* coef, float* d,
                          float* x, bool first);

// The factor's uses inside a larger linear computation: P was applied to
// v[k] (z_k = P v_k) and zbar[k] is dLoss/dz_k, so dLoss/dP = sum_k zbar_k v_k^T.
constexpr int kMaxUses = 8;
struct FactorUses {
    int m = 0;
    const float* v[kMaxUses];
    const float* zbar[kMaxUses];
};

struct Learned;

int learned_create(Learned** out, uint64_t seed);
void learned_destroy(Learned* lp);
const char* learned_error(const Learned* lp);
int learned_get_params(Learned* lp, float* host, int n);
int learned_set_params(Learned* lp, const float* host, int n);
// Forward pass on graph g; *F describes the factor (owned by the Learned
// object, valid until its next call).
int learned_infer(Learned* lp, hipStream_t st, const LearnedGraph& g, Factor* F);
// One training step on the pair (r = A e, e): forward, loss
// min_w ||e - w L L^T r||_A^2 / ||e||_A^2 (scale-free in P), backward, Adam.
// d_e = null draws e from a seeded Gaussian (zero on penalised nodes).
// loss_out may be null.
int learned_train_step(Learned* lp, hipStream_t st, const LearnedGraph& g, const float* d_e, uint64_t sample_seed,
                       float lr, double* loss_out);
// Backward pass and Adam step for a loss whose dependence on P is given by
// its uses (after learned_infer on the same graph and parameters).
int learned_step_from_uses(Learned* lp, hipStream_t st, const LearnedGraph& g, const FactorUses& uses, float lr);
// The same gradient without the update, copied to h_grad (kNumParams) when non-null.
int learned_grad_from_uses(Learned* lp, hipStream_t st, const LearnedGraph& g, const FactorUses& uses,
                           float* h_grad);
// Seeded Gaussian sample (zero on penalised nodes) into d_e (3n floats).
void learned_sample(hipStream_t st, const LearnedGraph& g, uint64_t seed, float* d_e);
// Diagnostics: loss and parameter gradient for a given sample e (device
// vector), without updating the parameters; features (n x 8) and head outputs
// (n x 24) copied to host when the pointers are non-null.
int learned_eval(Learned* lp, hipStream_t st, const LearnedGraph& g, const float* d_e, double* loss, float* h_grad,
                 float* h_feat, float* h_out);

}  // namespace asend
