// Launch wrappers of the solve-path kernels (definitions in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <cstdint>

namespace asend {

constexpr int kSpmvBlocks = 4096;  // blocks of the block-row SpMV (each a contiguous row range)
constexpr int kVecBlocks = 2048;   // blocks of the streaming vector kernels
constexpr int kMaxPartials = 8192; // partial-sum slots per reduction value

// Scalars of one Krylov loop (and assembly), resident in device memory.
struct Scalars {
    double rz, pq, rr, ff, alpha, beta, tol, tol2, maxabs, sf, dot, zq;
    int64_t it, max_it;
    int done, stop_abs;
    int gate;     // mixed step solve: a refinement transition is due
    int n_resid;  // mixed step solve: FP64 residual evaluations so far
};

// Sliced-ELL fine operator: 64-row slices of 3x3 blocks (see kernels.hip).
struct Sell3 {
    int n_rows = 0, n_slices = 0;
    const int32_t* sptr = nullptr;  // slot offset of each slice (n_slices + 1)
    const int32_t* scol = nullptr;  // column (node row) of slot k, lane l at 64k + l
    int waves_per_block = 4;        // slices (wavefronts) per workgroup of the SpMV: 1, 2, 4 or 8
    int64_t n_slots = 0;            // slots over all slices (sptr[n_slices])
};
// Operators whose values exceed this are streamed with non-temporal loads
// (they cannot stay in the 256 MB Infinity Cache, and the loads then keep the
// gathered x in L2); smaller ones are read with plain loads so that they stay
// cache-resident between SpMVs.
constexpr int64_t kSellStreamBytes = 192ll << 20;
constexpr int kSellMaxWaves = 8;

// Index of component c (= 3i + j) of block k of row r in the value array.
__host__ __device__ inline int64_t sell_index(const int32_t* sptr, int r, int k, int c) {
    return ((int64_t)9 * (sptr[r >> 6] + k) + c) * 64 + (r & 63);
}

// brick8 general meshes, element-once: the blocks k_scale K_e + c_mass
// rho_e M_e (+ c_damp D_e, the B^T C_d B integral of the element's damping
// matrix Cd[36 cd_idx[e]], cd_idx null: matrix 0) of every element as row
// records into Kr (8 kRecStride doubles per element), then their row gather
// into K (every slot, padding as zeros);
// n2rec lists the records 8 e + a of each row (CSR by n2e_ptr), slot maps
// block (e, a, b) to its in-row slot.  With Re set the gather also forms the
// internal force rint of the element residuals Re (24 per element) and, with
// rhs set, rhs = fact f - rint.
void launch_hex8_elem_rec(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                          const double* nu, const double* rho, double k_scale, double c_mass, const double* Cd,
                          const int32_t* cd_idx, double c_damp, const double* kcond, const double* Cs,
                          const int32_t* cs_idx, double* Kr);
void launch_rec_gather(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const int32_t* n2e_ptr,
                       const int32_t* n2rec, const uint8_t* slot, const double* Kr, const double* Re, double fact,
                       const double* f, double* K, double* rint, double* rhs);
// Khat (64 x 9) and Mhat (64) of the element conn_r[0..7] with E = rho = 1
void launch_hex8_reference(hipStream_t st, const int32_t* conn_r, const double* xyz_r, double nu, double* Khat,
                           double* Mhat);
// brick8 / brickIM by rows on a mesh of translated copies of one element (at
// most 8 elements per node): K = sum k_scale E_e Khat + c_mass rho_e Mhat,
// every slot written
void launch_hex8_assemble_uniform(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const int32_t* scol,
                                  const int32_t* rp, const int32_t* n2e_ptr, const int32_t* n2e,
                                  const int32_t* conn_r, const double* Khat, const double* Mhat, const double* E,
                                  const double* rho, double k_scale, double c_mass, double* K);
// brickIM (condensed bubbles): adds k_scale K_e + c_mass M_e per element colour;
// reference blocks of the element conn_r[0..7] (E = rho = 1, its nu);
// dJ/dE_e = -lam_e^T (K_e / E_e) u_e for every element
void launch_brickim_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_r,
                             const double* xyz_r, const double* E, const double* nu, const double* rho,
                             double k_scale, double c_mass, const int32_t* sptr, const uint8_t* slot, double* K);
void launch_brickim_reference(hipStream_t st, const int32_t* conn_r, const double* xyz_r, const double* nu,
                              double* Khat, double* Mhat);
// geometrically nonlinear brickIM: condensed tangent records and residuals at
// (u, alpha) plus the internal-update maps (225 doubles per element), and the
// update alpha -= w + W du
void launch_brickim_nl(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                       const double* nu, const double* u, const double* alpha, double* Ke, double* Re, double* Wmap);
void launch_brickim_nl_internal(hipStream_t st, int n_el, const int32_t* conn_r, const double* du,
                                const double* Wmap, double* alpha);
void launch_brickim_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                         const double* nu, const double* u, const double* lam, double* sens);
// shell4: adds k_scale K_e + m_scale M_e; conn_p holds the 8 pseudo-node rows
// (2 per node: translations, rotations) of each element, xyz_r is per node
void launch_shell4_assemble(hipStream_t st, const int32_t* elems, int n_elems, const int32_t* conn_p,
                            const double* xyz_r, const double* E, const double* nu, const double* thk,
                            const double* rho, double k_scale, double m_scale, const int32_t* sptr,
                            const uint8_t* slot, double* K);
// Geometrically nonlinear brick8 (nonlinear.hip): element tangents Ke
// (row records, 8 kRecStride per element) and residuals Re (24) at
// displacement u (gathered by launch_rec_gather); the penalty load
// rhs -= mult sf u.
void launch_hex8_nl_element(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                            const double* nu, const double* u, double* Ke, double* Re);
void launch_nl_penalty_load(hipStream_t st, const int64_t* pen, const int32_t* mult, int n_pen, const Scalars* sc,
                            const double* u, double* rhs);
void launch_maxabs(hipStream_t st, const double* v, int64_t n, double* partials, Scalars* sc);
void launch_penalty_diag(hipStream_t st, const int64_t* pen, const int32_t* mult, int n_pen, const int32_t* sptr,
                         const int32_t* kdiag, double* K, const Scalars* sc);
void launch_to_f32(hipStream_t st, const double* a, float* b, int64_t n);
void launch_diag_inv(hipStream_t st, int n_rows, const int32_t* sptr, const int32_t* kdiag, const double* K,
                     double* d64, float* d32);

// Sliced-ELL SpMV, same modes as launch_spmv3, plus mode 4: y = K x with
// x = u, f = r and the single-reduction PCG scalars (launch_cg1_*).
template <typename T>
void launch_sell3(hipStream_t st, int mode, const Sell3& A, const T* val, const T* x, T* y, const T* f,
                  double* partials, Scalars* sc);

// FP16 copy of the fine operator for the multigrid cycle: every value of DOF
// row d is stored as K_dj / K_dd (so entries are O(1) and within FP16 range
// whatever the moduli and the penalty scale), one wavefront per slice.
void launch_to_f16_rowscaled(hipStream_t st, int n_rows, int n_slices, const int32_t* sptr, const double* K,
                             const double* dinv64, __half* K16);
// Event timing of the mode-3 FP16 SpMV launches (outside stream capture)
// between start and stop: launch count and summed kernel time.
void kernel_timer_start();
int kernel_timer_stop(int64_t* launches, double* total_ms);
// SpMV with that copy, FP32 accumulation, the row scale restored from the
// FP32 Jacobi diagonal: mode 0: y = K x; mode 3: y = f - K x.
void launch_sell3_h(hipStream_t st, int mode, const Sell3& A, const __half* val, const float* dinv32, const float* x,
                    float* y, const float* f);

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
// Flexible Jacobi PCG step (after launch_sell3 mode 7): x += alpha p,
// r -= alpha q, beta = -(Dinv r).q / p.q; then launch_pcg_dir.
template <typename T>
void launch_fpcg_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, const T* p, const T* q,
                        double* partials, Scalars* sc);
// reliable update of a flexible PCG: r = b; p *= ratio (ratio = 0: p = Dinv r)
template <typename T>
void launch_ru_reload(hipStream_t st, int64_t n, const T* b, const T* dinv, T* r, T* p, double ratio);
// Single-reduction Jacobi PCG: start (cold: x = 0, r = b; warm: r = b - K x
// formed by the caller), then per iteration launch_sell3 mode 4 (w = K u)
// followed by launch_cg1_update.
template <typename T>
void launch_cg1_init(hipStream_t st, int64_t n, const T* b, bool cold, const T* dinv, T* x, T* r, T* u, T* p, T* s,
                     double* partials, Scalars* sc);
template <typename T>
void launch_cg1_update(hipStream_t st, int64_t n, const T* dinv, T* x, T* r, T* u, const T* w, T* p, T* s,
                       const Scalars* sc);

// Mixed-precision solve of K x = rhs under device-side control, for launch
// sequences captured once and replayed: single-reduction Jacobi PCG in FP32
// (K32, Dinv32; r, u = Dinv r, w = K u, p, s = K p) with the solution x
// accumulated in FP64 and reliable updates: when the recursive residual has
// fallen by ru_factor since the last refresh, or below the tolerance, a
// refresh slot recomputes r = rhs - K64 x in FP64 and reloads r and u in FP32,
// keeping the search directions.  Convergence is decided on the FP64
// residual only: sc->rr = |rhs - K64 x|^2 of the last refresh, sc->ff =
// |rhs|^2, done when rr <= tol2 ff, or when it >= max_it, or after
// kMixedMaxRefresh refreshes.  sc->gate is set while a refresh is pending
// (the iteration kernels pause).
struct MixedSolve {
    const Sell3* A = nullptr;
    const double* K64 = nullptr;
    const float *K32 = nullptr, *dinv32 = nullptr;
    const double* rhs = nullptr;
    double *x64 = nullptr, *r64 = nullptr;
    float *r32 = nullptr, *u32 = nullptr, *w32 = nullptr, *p32 = nullptr, *s32 = nullptr;
    double* partials = nullptr;
    Scalars* sc = nullptr;
    int64_t n = 0;
    double ru_factor = 0.01;
};
constexpr int kMixedMaxRefresh = 200;
// start from the x64 passed in: |rhs|^2, first FP64 residual, FP32 start
void launch_mixed_start(hipStream_t st, const MixedSolve& m);
// one FP32 iteration (a no-op once done or while a refresh is pending)
void launch_mixed_iteration(hipStream_t st, const MixedSolve& m);
// refresh slot (a no-op unless a refresh is pending)
void launch_mixed_refresh(hipStream_t st, const MixedSolve& m);

void launch_scale_to_f32(hipStream_t st, int64_t n, const double* r, float* r32, const Scalars* sc);
void launch_axpy_from_f32(hipStream_t st, int64_t n, double* u, const float* d, const Scalars* sc_outer);
// PCG whose preconditioner runs as separate launches (z = M^-1 r in between):
// start x = 0, r = b; dots: r.z and r.r into the loop scalars (init: also
// p = z, start values; else beta, iteration count and the stop test);
// update x += alpha p, r -= alpha q; dir p = z + beta p.
void launch_pcgz_start(hipStream_t st, int64_t n, const double* b, double* x, double* r);
void launch_pcgz_dots(hipStream_t st, int64_t n, const double* r, const double* z, double* p, bool init,
                      double* partials, Scalars* sc);
void launch_pcgz_update(hipStream_t st, int64_t n, double* x, double* r, const double* p, const double* q,
                        const Scalars* sc);
void launch_pcgz_dir(hipStream_t st, int64_t n, const double* z, double* p, const Scalars* sc);

void launch_dot64(hipStream_t st, int64_t n, const double* a, const double* b, double* partials, Scalars* sc);
void launch_axpy64(hipStream_t st, int64_t n, double alpha, const double* x, double* y);

// per-element partials of -lambda^T K_e u w.r.t. (E1, E2, E3, nu12, nu13, nu23,
// G12, G13, G23) of the compliance, out[9 e + p]
// brick8 shape sensitivities: out[24 e + 3 a + c] = lam_e^T (dK_e/dX_{a,c}) u_e
// for the 8 nodes a and coordinates c of every element (dual numbers).
void launch_hex8_shape_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                            const double* nu, const double* u, const double* lam, double* out);
// brick8 centre-point strains, stresses (isotropic C of each element) and
// element volumes from u: out[13 e ..] = 6 strains, 6 stresses, volume.
void launch_hex8_centre_fields(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                               const double* nu, const double* u, double* out);
// y = sum over elements of B_c^T h_e (strain-space weights h, 6 per element)
// gathered per node row (n2e_ptr / n2rec: the records 8 e + a of each row);
// re is 24 n_el doubles of scratch.
void launch_hex8_centre_adjoint(hipStream_t st, int n_el, int n_rows, const int32_t* conn_r, const double* xyz_r,
                                const int32_t* n2e_ptr, const int32_t* n2rec, const double* h, double* re,
                                double* y);
void launch_hex8_sens_mat(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                          const double* nu, const double* u, const double* lam, double* out);
void launch_hex8_sens(hipStream_t st, int n_el, const int32_t* conn_r, const double* xyz_r, const double* E,
                      const double* nu, const double* u, const double* lam, double* out);

}  // namespace asend

namespace asend {
void launch_copy_f32(hipStream_t st, int64_t n4, const float* a, float* b);

// Newmark step kernels (dynamics.hip)
void launch_nm_predict(hipStream_t st, int64_t n, const double* u, const double* v, const double* a, double dt,
                       double c2, double* w);
void launch_nm_rhs(hipStream_t st, int64_t n, const int64_t* step, const double* fac, const double* f_shape,
                   double c_m, const double* q, double* rhs);
void launch_nm_update(hipStream_t st, int64_t n, const double* du, const double* w, double c_m, double dt,
                      double gamma, double* u, double* v, double* a);
void launch_nm_record(hipStream_t st, int n_watch, const int64_t* watch, const double* u, double* hist,
                      int64_t* step, int64_t* its, const Scalars* sc);
void launch_nm_save(hipStream_t st, int64_t n, const int64_t* step, const double* u, double* uhist);
// yhist[step] = u + ck v (the state dR/dE acts on under stiffness-proportional damping)
void launch_nm_save_y(hipStream_t st, int64_t n, const int64_t* step, const double* u, const double* v, double ck,
                      double* yhist);
// z = v + dt (1 - gamma) a - cd w: the velocity part of the damped residual
void launch_nm_damp_z(hipStream_t st, int64_t n, const double* v, const double* a, const double* w, double dt,
                      double gamma, double cd, double* z);
// modal analysis (shift-invert Lanczos) vector kernels
int modal_max_vectors();
void launch_add_diag(hipStream_t st, int n_rows, const int32_t* sptr, const int32_t* kdiag, double* K, const double* d,
                     double alpha);
void launch_vmul(hipStream_t st, int64_t n, const double* a, const double* x, double* y);
void launch_vtw(hipStream_t st, int64_t n, int m, const double* V, const double* w, double* partials, int nb,
                double* c);
void launch_add_vc(hipStream_t st, int64_t n, int m, const double* V, const double* c, double alpha, double* w);
// dlam = D lam_{n+1} (null without damping); dvu, dva, dvv: derivatives of
// the velocity of step n+1 with respect to (u, a, v) of step n
void launch_adj_prep(hipStream_t st, int64_t n, const double* q, double wn, double c1, double c2, double dt,
                     double gamma, int last, const double* mlam, const double* dlam, double dvu, double dva,
                     double dvv, double* aadj, double* vadj, double* rhs);

}  // namespace asend
