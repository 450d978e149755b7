// Launch wrappers of the solve-path kernels (definitions in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace asend {

constexpr int kSpmvBlocks = 4096;  // blocks of the block-row SpMV (each a contiguous row range)
constexpr int kVecBlocks = 2048;   // blocks of the streaming vector kernels
constexpr int kMaxPartials = 8192; // partial-sum slots per reduction value

// Scalars of one Krylov loop (and assembly), resident in device memory.
struct Scalars {
    double rz, pq, rr, ff, alpha, beta, tol, tol2, maxabs, sf, dot, zq;
    int64_t it, max_it;
    int done, stop_abs, pad0, pad1;
};

void launch_hex8_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_r,
                          const double* xyz_r, const double* E, const double* nu, const double* rho, double c_mass,
                          const int32_t* rp, const uint8_t* slot, double* K);
void launch_maxabs(hipStream_t st, const double* v, int64_t n, double* partials, Scalars* sc);
void launch_penalty_diag(hipStream_t st, const int64_t* pen, const int32_t* mult, int n_pen, const int32_t* rp,
                         const int32_t* kdiag, int bs, double* K, const Scalars* sc);
void launch_to_f32(hipStream_t st, const double* a, float* b, int64_t n);
void launch_diag_inv(hipStream_t st, int n_rows, int bs, const int32_t* rp, const int32_t* kdiag, const double* K,
                     double* d64, float* d32);

// mode 0: y = K x; 1: y = K x, sc->pq = x.y, sc->alpha = sc->rz/pq; 2: y = f - K x, sc->rr = |y|^2
// 3: y = f - K x (no norm).  Modes 1 and 2 launch their single-block finalize kernel after the SpMV.
template <typename T>
void launch_spmv3(hipStream_t st, int mode, int n_rows, const int32_t* rp, const int32_t* bc, const T* val,
                  const T* x, T* y, const T* f, double* partials, Scalars* sc);

template <typename T>
void launch_pcg_init(hipStream_t st, int64_t n, const T* b, const T* dinv, T* x, T* r, T* p, double* partials,
                     Scalars* sc);
template <typename T>
void launch_pcg_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, const T* p, const T* q,
                       double* partials, Scalars* sc);
template <typename T>
void launch_pcg_dir(hipStream_t st, int64_t n, const T* dinv, const T* r, T* p, const Scalars* sc);

void launch_scale_to_f32(hipStream_t st, int64_t n, const double* r, float* r32, const Scalars* sc);
void launch_axpy_from_f32(hipStream_t st, int64_t n, double* u, const float* d, const Scalars* sc_outer);
void launch_dot64(hipStream_t st, int64_t n, const double* a, const double* b, double* partials, Scalars* sc);
void launch_axpy64(hipStream_t st, int64_t n, double alpha, const double* x, double* y);

void launch_hex8_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                      const double* nu, const double* u, const double* lam, double* out);

}  // namespace asend
