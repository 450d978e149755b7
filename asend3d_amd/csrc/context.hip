// Device context and the C ABI of include/asend_gpu.h.
//
// Host orchestration of the solve path: mesh upload (node-block pattern,
// element colouring, element-to-block map), assembly, Krylov drivers with
// device-resident scalars, sensitivities.  The reference runs all of this in
// one serial Rust process (Model::solve / Model::get_obj_gradient,
// S/model_analysis.rs:560-673, :1544-1706); here every per-element and
// per-iteration step is a HIP kernel on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "asend_gpu.h"
#include "kernels.hpp"
#include "precond.hpp"
#include "learned.hpp"
#include "batched.hpp"
#include "host_prep.hpp"
#include "ldl.hpp"
#include "device_common.hpp"

using namespace asend;

struct asend_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;

    int etype = 0, bs = 3;
    int64_t n_nodes = 0, n_dof = 0, n_el = 0, nnzb = 0;
    int64_t grid[3] = {0, 0, 0};
    std::vector<int32_t> h_rp, h_bc, h_color_ptr;

    int32_t *d_conn_r = nullptr, *d_rp = nullptr, *d_bc = nullptr, *d_kdiag = nullptr, *d_elems = nullptr;
    // fine operator storage: sliced ELL of 3x3 blocks, 64-row slices
    int64_t n_slices = 0, n_slots = 0;
    std::vector<int32_t> h_sptr;
    int32_t *d_sptr = nullptr, *d_scol = nullptr;
    Sell3 fine() const {
        Sell3 A;
        A.n_rows = (int)n_nodes;
        A.n_slices = (int)n_slices;
        A.sptr = d_sptr;
        A.scol = d_scol;
        A.waves_per_block = spmv_wpb;
        A.n_slots = n_slots;
        return A;
    }
    int64_t nval() const { return 9 * 64 * n_slots; }
    uint8_t* d_slot = nullptr;
    double *d_xyz_r = nullptr, *d_E = nullptr, *d_nu = nullptr, *d_rho = nullptr;
    // shell4 meshes: per-element thickness (asend_set_thickness); each node is
    // two block rows (translations 2r, rotations 2r+1) of the operator
    double* d_thk = nullptr;
    bool thk_set = false;
    // brick8 meshes of translated copies of one element with one Poisson
    // ratio: reference blocks Khat (E = 1) and Mhat (rho = 1) of that element
    bool uniform_hex = false;
    double *d_khat = nullptr, *d_mhat = nullptr;
    double* d_K64 = nullptr;
    float* d_K32 = nullptr;
    // FP16 row-scaled copy read by the multigrid cycle (allocated with the
    // hierarchy); cycle_half selects it (asend_mg_set_cycle_precision)
    __half* d_K16 = nullptr;
    bool cycle_half = true;
    double* d_dinv64 = nullptr;
    float* d_dinv32 = nullptr;

    int64_t* d_pen = nullptr;
    int32_t* d_pen_mult = nullptr;
    int n_pen = 0;
    std::vector<int64_t> h_pen;
    std::vector<int32_t> h_pen_mult;
    double sf_user = 0.0, sf = 0.0;
    bool sf_fixed = false, assembled = false;

    // work vectors
    double *d_f = nullptr, *d_u = nullptr, *d_r = nullptr, *d_p = nullptr, *d_q = nullptr, *d_lam = nullptr,
           *d_sens = nullptr, *d_x64 = nullptr, *d_r2 = nullptr;
    float *d_r32 = nullptr, *d_x32 = nullptr, *d_p32 = nullptr, *d_q32 = nullptr, *d_w32 = nullptr;

    int64_t partials_cap = kMaxPartials;  // blocks whose partial sums d_partials can hold
    Scalars* d_sc = nullptr;  // [0] outer/FP64 loop, [1] inner FP32 loop, [2] misc
    Scalars* h_sc = nullptr;  // pinned mirror
    double* d_partials = nullptr;
    unsigned* d_counter = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side stream (default priority) for the multigrid setup that overlaps
    // the learned smoother's inference and bound, with its fork/join events
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    Multigrid* mg = nullptr;

    // learned preconditioner: network, graph data, and the assembly its factor belongs to
    Learned* learned = nullptr;
    int32_t *d_n2e_ptr = nullptr, *d_n2e = nullptr;
    int32_t* d_n2rec = nullptr;  // brick8: record 8 e + a of each node's elements (CSR by d_n2e_ptr)
    int max_valence = 0;  // most elements sharing one node
    // set inside a nonlinear solve after its first Newton iteration: a new
    // assembly changes only the finest multigrid level (the coarse levels are
    // integrated from the moduli, which a Newton iteration does not change)
    bool mg_fine_only = false;
    // geometrically nonlinear statics: element tangents (576 per element),
    // element residuals (24), displacement state, applied load, internal force
    double *d_kb = nullptr, *d_nl_re = nullptr, *d_nl_u = nullptr, *d_nl_f = nullptr, *d_nl_r = nullptr;
    // brickIM under geometric nonlinearity: the 9 bubble amplitudes of each
    // element (kept between calls, zero after an upload) and the maps of
    // their update (K_ii^-1 K_ie and K_ii^-1 r_i, 225 per element)
    double *d_nl_alpha = nullptr, *d_nl_wmap = nullptr;
    uint8_t* d_clamped = nullptr;
    int32_t *d_enb = nullptr, *d_erev = nullptr, *d_eslot = nullptr;
    int n_edge = 0;
    double mean_log_e = 0.0;
    std::vector<double> h_nu;  // per-element Poisson ratio (moduli updates are checked against it)
    // LDL^T factor over DOF blocks (asend_ldl_factor): the block size asked
    // for (< 0: none), the factor and the preconditioned vector of its PCG
    LdlFactor ldl;
    int64_t ldl_block_dofs = -1;
    double* d_ldl_z = nullptr;
    int64_t asm_count = 0, factor_asm = -1;
    int64_t mg_asm = -1;  // assembly the multigrid hierarchy was last set up for
    double asm_cmass = 0.0;   // mass coefficient of the current assembly
    double asm_kscale = 1.0;  // stiffness factor of the current assembly
    // Rayleigh damping D = ray_cm M + ray_ck K of the next Newmark setup
    double ray_cm = 0.0, ray_ck = 0.0;
    // material damping matrices (n_dmat x 36) and each element's matrix
    // (null: all matrix 0), added to D of the next Newmark setup
    double* d_dmat = nullptr;
    int32_t* d_dmat_idx = nullptr;
    int64_t n_dmat = 0;
    double asm_cdamp = 0.0;  // weight of the material damping in the current assembly
    // heat conduction (asend_set_conductivity): 6 conductivities per element;
    // while set, assemblies integrate the conduction operator (temperature at
    // the first DOF of each node, the same scalar operator on the other two)
    double* d_kcond = nullptr;
    // general material stiffness (asend_set_material_stiffness): n_smat 6x6
    // matrices and each element's matrix; while set, brick8 assembly uses
    // E_e times them in place of the isotropic material
    double* d_smat = nullptr;
    int32_t* d_smat_idx = nullptr;
    int64_t n_smat = 0;
    // Newmark step solver (asend_newmark_set_solver)
    int nm_mode = ASEND_NM_AUTO;
    double nm_inner_tol = 1e-3;  // reliable-update factor of the mixed step solver
    bool use_graphs = true;
    // autotuned choices: SpMV workgroup geometry, inner tolerance and check
    // interval used when a solve passes 0 for them
    int spmv_wpb = 4;
    double tuned_inner_tol = 0.0;
    int tuned_check = 0;
    // set by the learned autotuning when the pipeline without the network
    // was the faster one: ASEND_PC_LEARNED solves then run the plain
    // multigrid smoother (stats.fallback_used = 2); a changed network clears it
    bool learned_aside = false;

    // Newmark time stepping: mass operator, state and per-run buffers
    struct Newmark {
        bool ready = false;
        double dt = 0, beta = 0, gamma = 0, c_m = 0, c2 = 0;
        double* M64 = nullptr;
        double *v = nullptr, *a = nullptr, *w = nullptr, *q = nullptr, *rhs0 = nullptr, *rhs = nullptr,
               *fshape = nullptr;
        double *fac = nullptr, *hist = nullptr;
        int64_t *step = nullptr, *its = nullptr, *watch = nullptr;
        int64_t fac_cap = 0, hist_cap = 0, watch_cap = 0;
        hipGraphExec_t exec_head = nullptr, exec_more = nullptr, exec_tail = nullptr;
        int graph_iters = -1, graph_watch = -1;
        // displacement history for the adjoint sweep (save_u: where the step
        // tail stores u, null when not recording) and the adjoint vectors
        double* save_u = nullptr;
        const double* graph_save = nullptr;
        double* uhist = nullptr;
        int64_t uhist_cap = 0;
        double *lam = nullptr, *aadj = nullptr, *vadj = nullptr, *mlam = nullptr, *qv = nullptr,
               *sens_acc = nullptr;
        // single-reduction PCG of the step solves: u = Dinv r and s = K p
        double *cg_u = nullptr, *cg_s = nullptr;
        // Rayleigh damping: coefficients, cd = c_m dt gamma (the weight of D
        // in K_eff), D itself, the damped load, and the history of
        // u + ck v the adjoint's dR/dE acts on (save_y: where the tail
        // stores it, null when not recording)
        double cm = 0, ck = 0, cd = 0;
        double keff_k = 1.0, keff_m = 0.0;  // K_eff = keff_k K + keff_m M (+ keff_d D_mat + penalty)
        double keff_d = 0.0;
        bool mat = false;  // material damping matrices are part of D
        double *D64 = nullptr, *rhsd = nullptr, *yhist = nullptr, *save_y = nullptr;
        int64_t yhist_cap = 0;
        const double* graph_save_y = nullptr;
        bool damped() const { return cm != 0.0 || ck != 0.0 || mat; }
        // step solver (ASEND_NM_*): FP32 single-reduction PCG with an FP64
        // increment and reliable updates, FP64 single-reduction PCG, or by
        // operator size; the FP32 work vectors of the mixed solver (5 n
        // floats)
        int mixed = ASEND_NM_AUTO;
        double inner_tol = 1e-3;  // reliable-update factor
        float* f32 = nullptr;
        int graph_mixed = -1;
        // the step graphs also hold the reliable-update factor and the SpMV
        // workgroup geometry as launch arguments
        double graph_ru = -1.0;
        int graph_wpb = -1;
    } nm;
};

namespace {

int fail(asend_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return code;
}

#define HIPCHK_CTX(expr)                      \
    do {                                      \
        if ((expr) != hipSuccess) {           \
            delete ab;                        \
            return ASEND_ERR_HIP;             \
        }                                     \
    } while (0)

#define HIPCHK(expr)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(ctx, ASEND_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(asend_ctx* ctx, T** p, size_t n) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void**)p, n * sizeof(T)));
    return ASEND_OK;
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

int ensure_work(asend_ctx* ctx) {
    int64_t n = ctx->n_dof;
    if (ctx->d_f) return ASEND_OK;
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_f, n)) || (rc = dalloc(ctx, &ctx->d_u, n)) || (rc = dalloc(ctx, &ctx->d_r, n)) ||
        (rc = dalloc(ctx, &ctx->d_p, n)) || (rc = dalloc(ctx, &ctx->d_q, n)) || (rc = dalloc(ctx, &ctx->d_lam, n)) ||
        (rc = dalloc(ctx, &ctx->d_x64, n)) || (rc = dalloc(ctx, &ctx->d_sens, ctx->n_el)) ||
        (rc = dalloc(ctx, &ctx->d_r2, n)) ||
        (rc = dalloc(ctx, &ctx->d_r32, n)) || (rc = dalloc(ctx, &ctx->d_x32, n)) ||
        (rc = dalloc(ctx, &ctx->d_p32, n)) || (rc = dalloc(ctx, &ctx->d_q32, n)) ||
        (rc = dalloc(ctx, &ctx->d_w32, n)))
        return rc;
    return ASEND_OK;
}

// Copy device scalars `dev` into (or from) the pinned host mirror slot idx.
int read_scalars_from(asend_ctx* ctx, const Scalars* dev, int idx) {
    HIPCHK(hipMemcpyAsync(&ctx->h_sc[idx], dev, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ASEND_OK;
}

int write_scalars_to(asend_ctx* ctx, Scalars* dev, int idx) {
    HIPCHK(hipMemcpyAsync(dev, &ctx->h_sc[idx], sizeof(Scalars), hipMemcpyHostToDevice, ctx->st));
    return ASEND_OK;
}

int read_scalars(asend_ctx* ctx, int idx) { return read_scalars_from(ctx, ctx->d_sc + idx, idx); }

int write_scalars(asend_ctx* ctx, int idx) { return write_scalars_to(ctx, ctx->d_sc + idx, idx); }

void newmark_free(asend_ctx* ctx) {
    auto& nm = ctx->nm;
    for (double** q : {&nm.M64, &nm.v, &nm.a, &nm.w, &nm.q, &nm.rhs0, &nm.rhs, &nm.fshape, &nm.fac, &nm.hist,
                       &nm.uhist, &nm.lam, &nm.aadj, &nm.vadj, &nm.mlam, &nm.qv, &nm.sens_acc, &nm.cg_u, &nm.cg_s,
                       &nm.D64, &nm.rhsd, &nm.yhist})
        dfree(*q);
    dfree(nm.f32);
    nm.graph_mixed = -1;
    nm.yhist_cap = 0;
    nm.save_y = nullptr;
    nm.graph_save_y = nullptr;
    for (int64_t** q : {&nm.step, &nm.its, &nm.watch}) dfree(*q);
    nm.fac_cap = nm.hist_cap = nm.watch_cap = nm.uhist_cap = 0;
    nm.save_u = nullptr;
    nm.graph_save = nullptr;
    for (hipGraphExec_t* e : {&nm.exec_head, &nm.exec_more, &nm.exec_tail}) {
        if (*e) (void)hipGraphExecDestroy(*e);
        *e = nullptr;
    }
    nm.graph_iters = nm.graph_watch = -1;
    nm.graph_ru = -1.0;
    nm.graph_wpb = -1;
    nm.ready = false;
}

void ldl_release(asend_ctx* ctx) {
    dfree(ctx->ldl.L);
    dfree(ctx->ldl.P);
    dfree(ctx->ldl.neg);
    dfree(ctx->ldl.work);
    dfree(ctx->d_ldl_z);
    ctx->ldl = LdlFactor();
}

void free_mesh(asend_ctx* ctx) {
    ldl_release(ctx);
    ctx->ldl_block_dofs = -1;
    dfree(ctx->d_dmat);
    dfree(ctx->d_dmat_idx);
    ctx->n_dmat = 0;
    dfree(ctx->d_kcond);
    dfree(ctx->d_smat);
    dfree(ctx->d_smat_idx);
    ctx->n_smat = 0;
    dfree(ctx->d_conn_r);
    dfree(ctx->d_rp);
    dfree(ctx->d_bc);
    dfree(ctx->d_sptr);
    dfree(ctx->d_scol);
    dfree(ctx->d_kdiag);
    dfree(ctx->d_elems);
    dfree(ctx->d_slot);
    dfree(ctx->d_xyz_r);
    dfree(ctx->d_E);
    dfree(ctx->d_nu);
    dfree(ctx->d_rho);
    // penalty DOFs are numbered on the mesh they were set for
    dfree(ctx->d_pen);
    dfree(ctx->d_pen_mult);
    ctx->n_pen = 0;
    ctx->h_pen.clear();
    ctx->h_pen_mult.clear();
    ctx->sf_user = ctx->sf = 0.0;
    ctx->sf_fixed = false;
    dfree(ctx->d_thk);
    ctx->thk_set = false;
    dfree(ctx->d_khat);
    dfree(ctx->d_mhat);
    ctx->uniform_hex = false;
    dfree(ctx->d_K64);
    dfree(ctx->d_K32);
    dfree(ctx->d_K16);
    dfree(ctx->d_dinv64);
    dfree(ctx->d_dinv32);
    dfree(ctx->d_f);
    dfree(ctx->d_u);
    dfree(ctx->d_r);
    dfree(ctx->d_p);
    dfree(ctx->d_q);
    dfree(ctx->d_lam);
    dfree(ctx->d_x64);
    dfree(ctx->d_r2);
    dfree(ctx->d_sens);
    dfree(ctx->d_r32);
    dfree(ctx->d_x32);
    dfree(ctx->d_p32);
    dfree(ctx->d_q32);
    dfree(ctx->d_w32);
    newmark_free(ctx);
    dfree(ctx->d_n2e_ptr);
    dfree(ctx->d_n2e);
    dfree(ctx->d_n2rec);
    dfree(ctx->d_kb);
    dfree(ctx->d_nl_re);
    dfree(ctx->d_nl_u);
    dfree(ctx->d_nl_f);
    dfree(ctx->d_nl_r);
    dfree(ctx->d_nl_alpha);
    dfree(ctx->d_nl_wmap);
    dfree(ctx->d_clamped);
    dfree(ctx->d_enb);
    dfree(ctx->d_erev);
    dfree(ctx->d_eslot);
    ctx->n_edge = 0;
    ctx->factor_asm = -1;
    if (ctx->mg) {
        mg_destroy(ctx->mg);
        ctx->mg = nullptr;
    }
    ctx->assembled = false;
}

__global__ void k_scale_inplace(int64_t n, double s, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= s;
}

__global__ void k_sub_f32(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                          float* __restrict__ c) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c[i] = a[i] - b[i];
}

// ---------------------------------------------------------------------------
// Krylov drivers
// ---------------------------------------------------------------------------

// Device buffers of the FP64 Jacobi PCG solve (ASEND_SOLVE_PCG_FP64): the
// right-hand side f, the solution u, the outer residual r = f - K u, the
// correction x of one PCG run, its residual r2, direction p and product q
// (n_dof doubles each), the reduction partials (2 kMaxPartials doubles) and
// two loop scalars (sc[0] the outer FP64 residual, sc[1] the PCG run).
struct Fp64Bufs {
    double *f, *u, *r, *x, *r2, *p, *q, *partials;
    Scalars* sc;
};

Fp64Bufs ctx_fp64_bufs(asend_ctx* ctx) {
    return {ctx->d_f, ctx->d_u, ctx->d_r, ctx->d_x64, ctx->d_r2, ctx->d_p, ctx->d_q, ctx->d_partials, ctx->d_sc};
}

__global__ void k_fill64(int64_t n, double v, double* __restrict__ x) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// One Jacobi PCG run on the FP64 operator: B.x solves K x = B.r from x = 0.
// Host mirror slot 1 receives the run's scalars (iterations, g^T M^-1 g).
int run_jacobi_pcg64(asend_ctx* ctx, const Fp64Bufs& B, double tol, bool stop_abs, int64_t max_it, int check_every) {
    const int64_t n = ctx->n_dof;
    const double* K = ctx->d_K64;
    const double* dinv = ctx->d_dinv64;
    Scalars& h = ctx->h_sc[1];
    std::memset(&h, 0, sizeof(Scalars));
    h.tol = tol;
    h.tol2 = tol * tol;
    h.stop_abs = stop_abs ? 1 : 0;
    h.max_it = max_it;
    Scalars* sc = B.sc + 1;
    int rc;
    if ((rc = write_scalars_to(ctx, sc, 1))) return rc;
    launch_pcg_init<double>(ctx->st, n, B.r, dinv, B.x, B.r2, B.p, B.partials, sc);
    for (;;) {
        for (int k = 0; k < check_every; ++k) {
            launch_sell3<double>(ctx->st, 1, ctx->fine(), K, B.p, B.q, nullptr, B.partials, sc);
            launch_pcg_update<double>(ctx->st, n, dinv, B.x, B.r2, B.p, B.q, B.partials, sc);
            launch_pcg_dir<double>(ctx->st, n, dinv, B.r2, B.p, sc);
        }
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars_from(ctx, sc, 1))) return rc;
        if (h.done) break;
    }
    return ASEND_OK;
}

// The LDL^T factor of the current assembly over the requested DOF blocks
// (LowerTriMat allocate/populate/ldl_factor, S/lower_tri_mat_meth.rs:15-190,
// done before each static solve at S/model_analysis.rs:584-589): planned and
// allocated once per mesh and block size, refilled and refactored whenever
// the operator was assembled again.
int ldl_ensure(asend_ctx* ctx) {
    if (ctx->ldl_block_dofs < 0)
        return fail(ctx, ASEND_ERR_STATE, "no LDL^T factor requested (asend_ldl_factor)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    LdlFactor& F = ctx->ldl;
    if (F.L && F.asm_id == ctx->asm_count) return ASEND_OK;
    int rc;
    if (!F.L) {
        LdlDesc d;
        int64_t W = 0;
        const char* msg = ldl_plan(&d, &W, ctx->n_dof, ctx->ldl_block_dofs, ctx->h_rp, ctx->h_bc);
        if (msg) return fail(ctx, ASEND_ERR_UNSUPPORTED, "LDL^T factor: %s", msg);
        const size_t bytes = (size_t)d.n_blocks * (size_t)d.stride * sizeof(double);
        size_t free_b = 0, total_b = 0;
        HIPCHK(hipMemGetInfo(&free_b, &total_b));
        if (bytes + ((size_t)1 << 30) > free_b)
            return fail(ctx, ASEND_ERR_UNSUPPORTED,
                        "LDL^T factor of %lld blocks, band %lld: %.2f GB of device memory needed, %.2f GB free",
                        (long long)d.n_blocks, (long long)W, bytes / 1e9, free_b / 1e9);
        if ((rc = dalloc(ctx, &F.L, (size_t)d.n_blocks * d.stride)) ||
            (rc = dalloc(ctx, &F.P, (size_t)d.n_blocks * kLdlTileElems)) || (rc = dalloc(ctx, &F.neg, 1)) ||
            (rc = dalloc(ctx, &F.work, (size_t)ctx->n_dof)) ||
            (rc = dalloc(ctx, &ctx->d_ldl_z, (size_t)ctx->n_dof))) {
            ldl_release(ctx);
            return rc;
        }
        F.d = d;
        F.max_bandwidth = W;
        F.bytes = (int64_t)bytes;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    const auto t0 = std::chrono::steady_clock::now();
    ldl_fill(ctx->st, F.d, ctx->fine().n_rows, (int)ctx->n_slices, ctx->d_sptr, ctx->d_scol, ctx->d_K64, F.L);
    ldl_factor_launch(ctx->st, F.d, F.L, F.P, F.neg);
    HIPCHK(hipGetLastError());
    int32_t neg = 0;
    HIPCHK(hipMemcpyAsync(&neg, F.neg, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    F.factor_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    F.neg_pivots = neg;
    F.asm_id = ctx->asm_count;
    return ASEND_OK;
}

// One PCG run on the FP64 operator preconditioned by the LDL^T factor
// (conj_grad_sparse, S/matrix_functions.rs:285-362): B.x solves K x = B.r
// from x = 0; host mirror slot 1 receives the iteration count and the final
// g^T M^-1 g, the quantity the absolute stop rule tests.
int run_ldl_pcg64(asend_ctx* ctx, const Fp64Bufs& B, double tol, bool stop_abs, int64_t max_it, int check_every) {
    const int64_t n = ctx->n_dof;
    Scalars& h = ctx->h_sc[1];
    std::memset(&h, 0, sizeof(Scalars));
    h.tol = tol;
    h.tol2 = tol * tol;
    h.stop_abs = stop_abs ? 1 : 0;
    h.max_it = max_it;
    Scalars* sc = B.sc + 1;
    int rc;
    if ((rc = write_scalars_to(ctx, sc, 1))) return rc;
    const LdlFactor& F = ctx->ldl;
    double* z = ctx->d_ldl_z;
    launch_pcgz_start(ctx->st, n, B.r, B.x, B.r2);
    ldl_solve_launch(ctx->st, F.d, F.L, B.r2, z, F.work, nullptr);
    launch_pcgz_dots(ctx->st, n, B.r2, z, B.p, true, B.partials, sc);
    for (;;) {
        for (int k = 0; k < check_every; ++k) {
            launch_sell3<double>(ctx->st, 1, ctx->fine(), ctx->d_K64, B.p, B.q, nullptr, B.partials, sc);
            launch_pcgz_update(ctx->st, n, B.x, B.r2, B.p, B.q, sc);
            ldl_solve_launch(ctx->st, F.d, F.L, B.r2, z, F.work, &sc->done);
            launch_pcgz_dots(ctx->st, n, B.r2, z, B.p, false, B.partials, sc);
            launch_pcgz_dir(ctx->st, n, z, B.p, sc);
        }
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars_from(ctx, sc, 1))) return rc;
        if (h.done) break;
    }
    return ASEND_OK;
}

// FP64 Jacobi PCG solve of K u = B.f into B.u (conj_grad_sparse with x0 = 0,
// S/matrix_functions.rs:285-362, with the Jacobi preconditioner).  Under the
// relative rule a run whose recursive residual drifted from the true one is
// restarted from the true residual.  Allocates nothing: every buffer it
// writes is in B.
int solve_fp64(asend_ctx* ctx, const Fp64Bufs& B, const asend_solve_opts* o, asend_solve_stats* st) {
    if (o->precond != ASEND_PC_JACOBI && o->precond != ASEND_PC_LDL)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "FP64 PCG supports the Jacobi and LDL^T preconditioners only");
    const bool ldl = o->precond == ASEND_PC_LDL;
    int rc;
    if (ldl && (rc = ldl_ensure(ctx))) return rc;
    const int64_t n = ctx->n_dof;
    const int64_t max_it = o->max_it > 0 ? o->max_it : n;
    const int check = o->check_every > 0 ? o->check_every : (ctx->tuned_check > 0 ? ctx->tuned_check : 16);
    const int64_t max_outer = o->max_outer > 0 ? o->max_outer : 20;
    asend_solve_stats s;
    std::memset(&s, 0, sizeof(s));
    // ||f||^2 with u = 0: r = f - K*0
    hipLaunchKernelGGL(k_fill64, dim3(kVecBlocks), dim3(256), 0, ctx->st, n, 0.0, B.u);
    std::memset(&ctx->h_sc[0], 0, sizeof(Scalars));
    if ((rc = write_scalars_to(ctx, B.sc, 0))) return rc;
    launch_sell3<double>(ctx->st, 2, ctx->fine(), ctx->d_K64, B.u, B.r, B.f, B.partials, B.sc);
    HIPCHK(hipGetLastError());
    if ((rc = read_scalars_from(ctx, B.sc, 0))) return rc;
    const double ff = ctx->h_sc[0].rr;
    if (ff == 0.0) {
        s.converged = 1;
        *st = s;
        return ASEND_OK;
    }
    double rr = ff;
    int64_t total_it = 0, outer = 0;
    for (; outer < max_outer; ++outer) {
        if (o->stop_rule == ASEND_STOP_REL_RESIDUAL && std::sqrt(rr / ff) <= o->tol) break;
        const double tol_in = (o->stop_rule == ASEND_STOP_REL_RESIDUAL) ? o->tol * std::sqrt(ff / rr) : o->tol;
        if ((rc = (ldl ? run_ldl_pcg64 : run_jacobi_pcg64)(ctx, B, tol_in, o->stop_rule == ASEND_STOP_REF_ABS,
                                                            max_it - total_it, check)))
            return rc;
        total_it += ctx->h_sc[1].it;
        s.ref_residual = ctx->h_sc[1].rz;
        // u += x ; r = f - K u
        launch_axpy64(ctx->st, n, 1.0, B.x, B.u);
        launch_sell3<double>(ctx->st, 2, ctx->fine(), ctx->d_K64, B.u, B.r, B.f, B.partials, B.sc);
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars_from(ctx, B.sc, 0))) return rc;
        rr = ctx->h_sc[0].rr;
        if (o->stop_rule == ASEND_STOP_REF_ABS || total_it >= max_it) {
            ++outer;
            break;
        }
    }
    s.outer_iterations = outer;
    s.iterations = total_it;
    s.rel_residual = std::sqrt(rr / ff);
    s.converged = (o->stop_rule == ASEND_STOP_REL_RESIDUAL) ? (s.rel_residual <= o->tol) : (s.ref_residual <= o->tol);
    *st = s;
    return ASEND_OK;
}

// FP64 single-reduction Jacobi PCG on the fine operator, relative tolerance
// (|b - K x| <= tol |b|); warm: start from the x passed in, else from zero.
// u and s are work vectors of n_dof doubles.
int run_cg1_pcg(asend_ctx* ctx, int si, const double* b, double* x, bool warm, double tol, int64_t max_it,
                int check_every, double* u, double* s) {
    const int64_t n = ctx->n_dof;
    Scalars& h = ctx->h_sc[si];
    std::memset(&h, 0, sizeof(Scalars));
    h.tol = tol;
    h.tol2 = tol * tol;
    h.max_it = max_it;
    int rc;
    if ((rc = write_scalars(ctx, si))) return rc;
    Scalars* sc = ctx->d_sc + si;
    if (warm) launch_sell3<double>(ctx->st, 3, ctx->fine(), ctx->d_K64, x, ctx->d_r2, b, nullptr, nullptr);
    launch_cg1_init<double>(ctx->st, n, b, !warm, ctx->d_dinv64, x, ctx->d_r2, u, ctx->d_p, s, ctx->d_partials, sc);
    for (;;) {
        for (int k = 0; k < check_every; ++k) {
            launch_sell3<double>(ctx->st, 4, ctx->fine(), ctx->d_K64, u, ctx->d_q, ctx->d_r2, ctx->d_partials, sc);
            launch_cg1_update<double>(ctx->st, n, ctx->d_dinv64, x, ctx->d_r2, u, ctx->d_q, ctx->d_p, s, sc);
        }
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars(ctx, si))) return rc;
        if (h.done) break;
    }
    return ASEND_OK;
}

// FP32 flexible Jacobi PCG on K32 with the FP64 reliable updates of ru (see
// solve_device): starts on the normalised residual in d_r32, refreshes when
// the recursive residual has fallen to ru.next (or claims convergence), and
// restarts the search direction where the FP32 recursion drifted (FP64
// residual more than twice the recursive one).  rr: the caller's FP64 |r|^2,
// updated by the refreshes; ff = |f|^2.
int run_jacobi_fpcg_ru(asend_ctx* ctx, FcgRefresh& ru, double& rr, double ff, double tol, int64_t max_it,
                       int check_every, int64_t* it_out) {
    const int64_t n = ctx->n_dof;
    float *x = ctx->d_x32, *r = ctx->d_w32, *p = ctx->d_p32, *q = ctx->d_q32;
    const float* dinv = ctx->d_dinv32;
    Scalars& h = ctx->h_sc[1];
    std::memset(&h, 0, sizeof(Scalars));
    h.tol2 = tol * tol * ff / rr;  // recursive convergence claim, relative to |b|^2 = 1
    h.max_it = max_it;
    int rc;
    if ((rc = write_scalars(ctx, 1))) return rc;
    Scalars* sc = ctx->d_sc + 1;
    launch_pcg_init<float>(ctx->st, n, ctx->d_r32, dinv, x, r, p, ctx->d_partials, sc);
    *it_out = 0;
    for (;;) {
        for (int k = 0; k < check_every; ++k) {
            launch_sell3<float>(ctx->st, 7, ctx->fine(), ctx->d_K32, p, q, r, ctx->d_partials, sc);
            launch_fpcg_update<float>(ctx->st, n, dinv, x, r, p, q, ctx->d_partials, sc);
            launch_pcg_dir<float>(ctx->st, n, dinv, r, p, sc);
        }
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars(ctx, 1))) return rc;
        *it_out = h.it;
        if (h.it >= max_it) break;
        if (!(h.done || h.rr <= ru.next * h.ff)) continue;
        const double rec = h.rr / h.ff * rr;  // recursive residual, FP64 units
        double ratio = 1.0, next = ru.next;
        bool conv = false;
        if ((rc = ru.run(&ratio, &next, &conv))) return rc;
        ru.refreshes += 1;
        if (conv) {
            ru.converged = true;
            break;
        }
        if (rr > 4.0 * rec) ratio = 0.0;  // drifted: restart the direction
        launch_ru_reload<float>(ctx->st, n, ctx->d_r32, dinv, r, p, ratio);
        HIPCHK(hipGetLastError());
        ru.next = next;
        h.ff = 1.0;
        h.rr = 1.0;
        h.done = 0;
        h.tol2 = tol * tol * ff / rr;
        if ((rc = write_scalars(ctx, 1))) return rc;
    }
    return ASEND_OK;
}

// FP64 residual r = f - K u, sc[0].rr = ||r||^2.
int residual64(asend_ctx* ctx, const double* f, const double* u, double* r) {
    launch_sell3<double>(ctx->st, 2, ctx->fine(), ctx->d_K64, u, r, f, ctx->d_partials, ctx->d_sc + 0);
    HIPCHK(hipGetLastError());
    return ASEND_OK;
}

// Inner-iteration budget of one learned-preconditioned FCG solve; an inner
// solve that needs more has failed to cut iterations and the solve falls back
// to Jacobi PCG.
constexpr int64_t kLearnedInnerBudget = 300;
constexpr int kNewmarkBlock = 24;  // PCG iterations per captured block of a Newmark step
constexpr int kMixedRefreshEvery = 4;  // mixed step solver: iterations between reliable-update slots
// ASEND_NM_AUTO takes the mixed step solver above this FP64 operator size
constexpr int64_t kMixedAutoBytes = 128ll << 20;
constexpr int64_t kLearnedPcgBudget = 20000;

LearnedGraph learned_graph(const asend_ctx* ctx) {
    LearnedGraph g;
    g.n_nodes = (int)ctx->n_nodes;
    g.rp = ctx->d_rp;
    g.bc = ctx->d_bc;
    g.n2e_ptr = ctx->d_n2e_ptr;
    g.n2e = ctx->d_n2e;
    g.E = ctx->d_E;
    g.clamped = ctx->d_clamped;
    g.n_edge = ctx->n_edge;
    g.enb = ctx->d_enb;
    g.erev = ctx->d_erev;
    g.eslot = ctx->d_eslot;
    g.op = ctx->fine();
    g.K32 = ctx->d_K32;
    g.kdiag = ctx->d_kdiag;
    g.mean_log_e = ctx->mean_log_e;
    return g;
}

// Any call that reuses the network's buffers (training, diagnostics)
// invalidates the factor the multigrid smoother borrows.
void invalidate_factor(asend_ctx* ctx) {
    ctx->factor_asm = -1;
    if (ctx->mg) mg_clear_learned_factor(ctx->mg);
}

// Multigrid setup (coarse operators, smoother bounds, coarsest inverse) for
// the current assembly, on first use: a solve that never asks for the
// hierarchy does not pay for it.
int ensure_mg(asend_ctx* ctx) {
    if (!ctx->mg) return fail(ctx, ASEND_ERR_STATE, "multigrid preconditioner not built (asend_set_grid)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    if (ctx->mg_asm == ctx->asm_count && mg_ready(ctx->mg)) return ASEND_OK;
    int rc;
    if (ctx->mg_fine_only && mg_ready(ctx->mg)) {
        if ((rc = mg_refresh_fine(ctx->mg, ctx->st, ctx->d_K32, ctx->d_dinv32, ctx->d_kdiag)))
            return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
        ctx->mg_asm = ctx->asm_count;
        return ASEND_OK;
    }
    // the finest level's block inverses and Chebyshev bound (finest-level
    // SpMVs) on the context stream, the coarse levels' setup (small,
    // latency-bound kernels) on the side stream, joined before the cycle;
    // every bound is formed on the device, so the host does not wait here
    HIPCHK(hipEventRecord(ctx->ev_fork, ctx->st));
    HIPCHK(hipStreamWaitEvent(ctx->st2, ctx->ev_fork, 0));
    rc = mg_fine_setup(ctx->mg, ctx->st, ctx->d_K32, ctx->d_dinv32, ctx->d_kdiag);
    const int rc2 = mg_setup(ctx->mg, ctx->st2, ctx->d_E, ctx->d_nu, ctx->d_K32, ctx->d_dinv32, ctx->d_kdiag,
                             ctx->sf, ctx->n_pen, ctx->h_pen.data(), ctx->h_pen_mult.data(), true);
    HIPCHK(hipEventRecord(ctx->ev_join, ctx->st2));
    HIPCHK(hipStreamWaitEvent(ctx->st, ctx->ev_join, 0));
    if (rc || (rc = rc2) || (rc = mg_setup_finish(ctx->mg, ctx->st))) {
        (void)hipStreamSynchronize(ctx->st);
        return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    }
    ctx->mg_asm = ctx->asm_count;
    return ASEND_OK;
}

// Network inference for the current operator (once per assembly) and
// registration of the factor with the multigrid object; the Chebyshev bounds
// lambda_max(P A) are estimated when the factor is used as the smoother.
int ensure_factor(asend_ctx* ctx, bool smoother) {
    if (ctx->factor_asm == ctx->asm_count && (!smoother || mg_factor_has_bounds(ctx->mg))) return ASEND_OK;
    invalidate_factor(ctx);
    Factor F;
    int rc = learned_infer(ctx->learned, ctx->st, learned_graph(ctx), &F);
    if (rc) return fail(ctx, rc, "%s", learned_error(ctx->learned));
    if ((rc = mg_set_learned_factor(ctx->mg, ctx->st, F, smoother)))
        return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    ctx->factor_asm = ctx->asm_count;
    return ASEND_OK;
}

// Multigrid setup plus the learned smoother (inference and its Chebyshev
// bound) for the current assembly, overlapped: the hierarchy's small,
// latency-bound setup kernels run on the side stream while the network and
// the bound's power iteration (finest-level SpMVs and factor products) run on
// the context stream.  Sequential when only one of them is due.  A setup
// failure sets *mg_failed; a failed bound returns ASEND_ERR_NUMERIC with the
// hierarchy ready (the caller falls back to Jacobi PCG).
int ensure_mg_and_factor(asend_ctx* ctx, bool* mg_failed) {
    *mg_failed = false;
    int rc;
    if (!ctx->mg) {
        *mg_failed = true;
        return fail(ctx, ASEND_ERR_STATE, "multigrid preconditioner not built (asend_set_grid)");
    }
    const bool need_mg = !(ctx->mg_asm == ctx->asm_count && mg_ready(ctx->mg));
    const bool need_factor = !(ctx->factor_asm == ctx->asm_count && mg_factor_has_bounds(ctx->mg));
    if (!need_mg || !need_factor || ctx->mg_fine_only) {
        if ((rc = ensure_mg(ctx))) {
            *mg_failed = true;
            return rc;
        }
        return ensure_factor(ctx, true);
    }
    if (!ctx->assembled) {
        *mg_failed = true;
        return fail(ctx, ASEND_ERR_STATE, "assemble first");
    }
    invalidate_factor(ctx);
    // the side stream starts after everything queued so far (the assembly)
    HIPCHK(hipEventRecord(ctx->ev_fork, ctx->st));
    HIPCHK(hipStreamWaitEvent(ctx->st2, ctx->ev_fork, 0));
    Factor F;
    if ((rc = learned_infer(ctx->learned, ctx->st, learned_graph(ctx), &F))) {
        *mg_failed = true;
        return fail(ctx, rc, "%s", learned_error(ctx->learned));
    }
    if ((rc = mg_learned_bound_begin(ctx->mg, ctx->st, F, ctx->d_K32, ctx->d_dinv32))) {
        *mg_failed = true;
        return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    }
    rc = mg_setup(ctx->mg, ctx->st2, ctx->d_E, ctx->d_nu, ctx->d_K32, ctx->d_dinv32, ctx->d_kdiag, ctx->sf,
                  ctx->n_pen, ctx->h_pen.data(), ctx->h_pen_mult.data(), false);
    // join: nothing later on the context stream runs before the setup ends
    HIPCHK(hipEventRecord(ctx->ev_join, ctx->st2));
    HIPCHK(hipStreamWaitEvent(ctx->st, ctx->ev_join, 0));
    if (rc == ASEND_OK) rc = mg_setup_finish(ctx->mg, ctx->st);
    if (rc) {
        *mg_failed = true;
        (void)hipStreamSynchronize(ctx->st);
        return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    }
    const int rc_bound = mg_learned_bound_end(ctx->mg, ctx->st);
    if ((rc = mg_check_bounds(ctx->mg))) {
        *mg_failed = true;
        return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    }
    ctx->mg_asm = ctx->asm_count;
    if (rc_bound) return fail(ctx, rc_bound, "%s", mg_last_error(ctx->mg));
    ctx->factor_asm = ctx->asm_count;
    return ASEND_OK;
}

// Direct solve with the one-block factor: u = (L D L^T)^-1 f (the
// reference's "direct" branch, ldl_solve at S/model_analysis.rs:495-496);
// the stats carry the FP64 relative residual of the result.
int solve_direct(asend_ctx* ctx, asend_solve_stats* st) {
    int rc;
    if ((rc = ldl_ensure(ctx))) return rc;
    if (ctx->ldl.d.n_blocks != 1)
        return fail(ctx, ASEND_ERR_STATE, "direct solve needs a factor over one block (asend_ldl_factor with "
                                          "block_dofs 0 or >= n_dof)");
    const int64_t n = ctx->n_dof;
    asend_solve_stats s;
    std::memset(&s, 0, sizeof(s));
    ldl_solve_launch(ctx->st, ctx->ldl.d, ctx->ldl.L, ctx->d_f, ctx->d_u, ctx->ldl.work, nullptr);
    std::memset(&ctx->h_sc[0], 0, sizeof(Scalars));
    if ((rc = write_scalars(ctx, 0))) return rc;
    launch_dot64(ctx->st, n, ctx->d_f, ctx->d_f, ctx->d_partials, ctx->d_sc);
    HIPCHK(hipGetLastError());
    if ((rc = read_scalars(ctx, 0))) return rc;
    const double ff = ctx->h_sc[0].dot;
    if ((rc = residual64(ctx, ctx->d_f, ctx->d_u, ctx->d_r))) return rc;
    if ((rc = read_scalars(ctx, 0))) return rc;
    s.rel_residual = ff > 0.0 ? std::sqrt(ctx->h_sc[0].rr / ff) : 0.0;
    s.outer_iterations = 1;
    s.converged = std::isfinite(s.rel_residual) ? 1 : 0;
    *st = s;
    return ASEND_OK;
}

int solve_device(asend_ctx* ctx, const asend_solve_opts* o, asend_solve_stats* st) {
    if ((ctx->d_kcond || ctx->n_smat > 0) && o->precond != ASEND_PC_JACOBI &&
        o->precond != ASEND_PC_BLOCK_JACOBI && o->precond != ASEND_PC_LDL)
        return fail(ctx, ASEND_ERR_UNSUPPORTED,
                    "conduction solves and general material matrices take the Jacobi or LDL^T preconditioners "
                    "(the multigrid hierarchy integrates isotropic elasticity)");
    if (o->method == ASEND_SOLVE_PCG_FP64) return solve_fp64(ctx, ctx_fp64_bufs(ctx), o, st);
    if (o->method == ASEND_SOLVE_DIRECT) return solve_direct(ctx, st);
    const int64_t n = ctx->n_dof;
    int64_t max_it = o->max_it > 0 ? o->max_it : n;
    int check = o->check_every > 0 ? o->check_every : (ctx->tuned_check > 0 ? ctx->tuned_check : 16);
    int64_t max_outer = o->max_outer > 0 ? o->max_outer : 20;
    int rc;
    asend_solve_stats s;
    std::memset(&s, 0, sizeof(s));
    // ||f||^2 with u = 0: r = f - K*0
    HIPCHK(hipMemsetAsync(ctx->d_u, 0, n * sizeof(double), ctx->st));
    {
        Scalars& h = ctx->h_sc[0];
        std::memset(&h, 0, sizeof(Scalars));
        if ((rc = write_scalars(ctx, 0))) return rc;
    }
    if ((rc = residual64(ctx, ctx->d_f, ctx->d_u, ctx->d_r))) return rc;
    if ((rc = read_scalars(ctx, 0))) return rc;
    const double ff = ctx->h_sc[0].rr;
    if (ff == 0.0) {
        s.converged = 1;
        s.outer_iterations = 0;
        *st = s;
        return ASEND_OK;
    }
    double rr = ff;
    int64_t total_it = 0;

    if (o->method != ASEND_SOLVE_IR_MIXED) return fail(ctx, ASEND_ERR_ARG, "unknown solve method %d", o->method);
    if (o->stop_rule != ASEND_STOP_REL_RESIDUAL)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "mixed solver supports the relative-residual rule only");
    double inner_tol = o->inner_tol > 0 ? o->inner_tol : (ctx->tuned_inner_tol > 0 ? ctx->tuned_inner_tol : 1e-4);
    // Reliable updates for the multigrid FCG (plain and learned smoother):
    // one FCG run keeps its search directions through the FP64 refreshes,
    // which come whenever the recursive FP32 residual has fallen by inner_tol
    // (and when it claims convergence); the FP64 residual decides.
    FcgRefresh ru;
    ru.next = inner_tol * inner_tol;
    ru.run = [&](double* ratio, double* next, bool* conv) -> int {
        const double s_old = std::sqrt(rr);
        launch_axpy_from_f32(ctx->st, n, ctx->d_u, ctx->d_x32, ctx->d_sc + 0);  // u += |r_old| x
        HIPCHK(hipMemsetAsync(ctx->d_x32, 0, n * sizeof(float), ctx->st));
        int rc2;
        if ((rc2 = residual64(ctx, ctx->d_f, ctx->d_u, ctx->d_r))) return rc2;
        if ((rc2 = read_scalars(ctx, 0))) return rc2;
        rr = ctx->h_sc[0].rr;
        *conv = !(rr > 0.0) || std::sqrt(rr / ff) <= o->tol;
        if (*conv) return ASEND_OK;
        launch_scale_to_f32(ctx->st, n, ctx->d_r, ctx->d_r32, ctx->d_sc + 0);
        *ratio = s_old / std::sqrt(rr);
        *next = std::max(inner_tol * inner_tol, o->tol * o->tol * ff / rr);
        return ASEND_OK;
    };
    int64_t outer = 0;
    for (; outer < max_outer; ++outer) {
        if (std::sqrt(rr / ff) <= o->tol) break;
        if (total_it >= max_it) break;
        launch_scale_to_f32(ctx->st, n, ctx->d_r, ctx->d_r32, ctx->d_sc + 0);
        int64_t it_in = 0;
        const bool aside = (o->precond == ASEND_PC_LEARNED && ctx->learned_aside);
        const bool learned_pc = !aside && (o->precond == ASEND_PC_LEARNED || o->precond == ASEND_PC_LEARNED_PCG);
        if (aside) s.fallback_used = 2;
        bool use_learned = learned_pc && !s.fallback_used;
        int rc_pre = ASEND_OK;
        if (learned_pc) {
            if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
            if (use_learned && o->precond == ASEND_PC_LEARNED) {
                bool mg_failed = false;
                rc_pre = ensure_mg_and_factor(ctx, &mg_failed);
                if (mg_failed) return rc_pre;
            } else if ((rc = ensure_mg(ctx))) {
                return rc;
            }
        }
        if (use_learned) {
            // learned factor as the multigrid smoother (PC_LEARNED) or alone
            // (PC_LEARNED_PCG); Jacobi PCG takes over for the rest of the solve
            // if it breaks down or needs more than the inner budget
            const bool smoother = (o->precond == ASEND_PC_LEARNED);
            const int64_t budget = smoother ? kLearnedInnerBudget : kLearnedPcgBudget;
            const int rc_factor = rc_pre ? rc_pre : ensure_factor(ctx, smoother);
            rc = rc_factor;
            if (rc == ASEND_OK) {
                if (smoother) mg_set_fine_smoother(ctx->mg, 1);
                int fb = 0;
                rc = mg_fcg_solve(ctx->mg, ctx->st, ctx->d_r32, ctx->d_x32, inner_tol,
                                  std::min<int64_t>(budget, max_it - total_it), check, &it_in, &fb, smoother ? 0 : 1,
                                  smoother ? &ru : nullptr);
                mg_set_fine_smoother(ctx->mg, 0);
            }
            if (rc == ASEND_ERR_NUMERIC || (rc == ASEND_OK && it_in >= budget)) {
                total_it += it_in;
                s.fallback_used = 1;
                use_learned = false;
                it_in = 0;
                launch_scale_to_f32(ctx->st, n, ctx->d_r, ctx->d_r32, ctx->d_sc + 0);
            } else if (rc) {
                return rc_factor ? rc_factor : fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            }
        }
        if (use_learned) {
            // done above
        } else if (o->precond == ASEND_PC_JACOBI || learned_pc) {
            if ((rc = run_jacobi_fpcg_ru(ctx, ru, rr, ff, o->tol, max_it - total_it, check, &it_in))) return rc;
        } else {
            if (o->precond != ASEND_PC_MG && !aside)
                return fail(ctx, ASEND_ERR_ARG, "unknown preconditioner %d", o->precond);
            if ((rc = ensure_mg(ctx))) return rc;
            mg_set_fine_smoother(ctx->mg, 0);
            int fb = 0;
            if ((rc = mg_fcg_solve(ctx->mg, ctx->st, ctx->d_r32, ctx->d_x32, inner_tol, max_it - total_it, check,
                                   &it_in, &fb, 0, &ru)))
                return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
        }
        total_it += it_in;
        if (ru.converged) {  // the last refresh took the solution and its residual
            outer += ru.refreshes;
            break;
        }
        launch_axpy_from_f32(ctx->st, n, ctx->d_u, ctx->d_x32, ctx->d_sc + 0);
        if ((rc = residual64(ctx, ctx->d_f, ctx->d_u, ctx->d_r))) return rc;
        if ((rc = read_scalars(ctx, 0))) return rc;
        rr = ctx->h_sc[0].rr;
    }
    s.outer_iterations = outer;
    s.iterations = total_it;
    s.rel_residual = std::sqrt(rr / ff);
    s.converged = s.rel_residual <= o->tol;
    *st = s;
    return ASEND_OK;
}

// Element integration into K: k_scale K_e(E) + c_mass M_e of every element.
// Translated-element meshes (brick8, brickIM): one row-wise launch scaling the
// reference blocks.  Other brick8 meshes: each element integrated once into
// row records, then one row-wise gather.  brickIM and shell4: K is cleared, then one scatter
// launch per element colour.
int integrate_elements(asend_ctx* ctx, const double* E, double k_scale, double c_mass, double* K,
                       double c_damp = 0.0) {
    if (c_damp != 0.0 && (ctx->etype != ASEND_BRICK8 || ctx->n_dmat == 0))
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "material damping is implemented for brick8 meshes");
    if (ctx->d_kcond && c_damp != 0.0)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the conduction operator has no damping term");
    if (ctx->uniform_hex && c_damp == 0.0 && !ctx->d_kcond && ctx->n_smat == 0) {
        launch_hex8_assemble_uniform(ctx->st, (int)ctx->n_nodes, (int)ctx->n_slices, ctx->d_sptr, ctx->d_scol,
                                     ctx->d_rp, ctx->d_n2e_ptr, ctx->d_n2e, ctx->d_conn_r, ctx->d_khat, ctx->d_mhat, E,
                                     ctx->d_rho, k_scale, c_mass, K);
    } else if (ctx->etype == ASEND_BRICKIM) {
        HIPCHK(hipMemsetAsync(K, 0, ctx->nval() * sizeof(double), ctx->st));
        const int ncol = (int)ctx->h_color_ptr.size() - 1;
        for (int c = 0; c < ncol; ++c) {
            const int b = ctx->h_color_ptr[c], e = ctx->h_color_ptr[c + 1];
            launch_brickim_assemble(ctx->st, ctx->d_elems + b, e - b, ctx->d_conn_r, ctx->d_xyz_r, E, ctx->d_nu,
                                    ctx->d_rho, k_scale, c_mass, ctx->d_sptr, ctx->d_slot, K);
        }
    } else if (ctx->etype == ASEND_SHELL4) {
        HIPCHK(hipMemsetAsync(K, 0, ctx->nval() * sizeof(double), ctx->st));
        const int ncol = (int)ctx->h_color_ptr.size() - 1;
        for (int c = 0; c < ncol; ++c) {
            const int b = ctx->h_color_ptr[c], e = ctx->h_color_ptr[c + 1];
            launch_shell4_assemble(ctx->st, ctx->d_elems + b, e - b, ctx->d_conn_r, ctx->d_xyz_r, E, ctx->d_nu,
                                   ctx->d_thk, ctx->d_rho, k_scale, c_mass, ctx->d_sptr, ctx->d_slot, K);
        }
    } else {
        if (!ctx->d_kb) {
            const int rc = dalloc(ctx, &ctx->d_kb, (int64_t)8 * kRecStride * ctx->n_el);
            if (rc) return rc;
        }
        launch_hex8_elem_rec(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, E, ctx->d_nu, ctx->d_rho, k_scale,
                             c_mass, ctx->d_dmat, ctx->d_dmat_idx, c_damp, ctx->d_kcond,
                             ctx->n_smat ? ctx->d_smat : nullptr, ctx->d_smat_idx, ctx->d_kb);
        launch_rec_gather(ctx->st, (int)ctx->n_nodes, (int)ctx->n_slices, ctx->d_sptr, ctx->d_n2e_ptr, ctx->d_n2rec,
                          ctx->d_slot, ctx->d_kb, nullptr, 0.0, nullptr, K, nullptr, nullptr);
    }
    HIPCHK(hipGetLastError());
    return ASEND_OK;
}

// dJ/dE_e = -lam_e^T (dK_e/dE_e) u_e for every element into d_sens
void launch_sens(asend_ctx* ctx, const double* u, const double* lam) {
    if (ctx->etype == ASEND_BRICKIM)
        launch_brickim_sens(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, u, lam,
                            ctx->d_sens);
    else
        launch_hex8_sens(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, u, lam,
                         ctx->d_sens);
}

// FP16 row-scaled copy of the current operator for the multigrid cycle,
// when the hierarchy exists (after the Jacobi diagonal).
void refresh_f16(asend_ctx* ctx) {
    if (ctx->d_K16)
        launch_to_f16_rowscaled(ctx->st, (int)ctx->n_nodes, (int)ctx->n_slices, ctx->d_sptr, ctx->d_K64,
                                ctx->d_dinv64, ctx->d_K16);
}

// Penalty scale (fixed on the first assembly), penalty diagonal, Jacobi
// diagonal and FP32 copy of the element operator held in d_K64
int finish_assembly(asend_ctx* ctx, double k_scale, double c_mass) {
    const int64_t nval = ctx->nval();
    int rc;
    if (ctx->n_pen > 0) {
        if (ctx->h_pen.back() >= ctx->n_dof) return fail(ctx, ASEND_ERR_STATE, "penalty DOF beyond the mesh");
        if (!ctx->sf_fixed) {
            if (ctx->sf_user > 0) {
                ctx->sf = ctx->sf_user;
            } else {
                // reference rule: sf = 1e5 * max|K_ij| of the first assembly (S/model_analysis.rs:255-259)
                launch_maxabs(ctx->st, ctx->d_K64, nval, ctx->d_partials, ctx->d_sc + 2);
                HIPCHK(hipGetLastError());
                if ((rc = read_scalars(ctx, 2))) return rc;
                ctx->sf = 1.0e5 * ctx->h_sc[2].maxabs;
            }
            ctx->sf_fixed = true;
        }
        ctx->h_sc[2].sf = ctx->sf;
        if ((rc = write_scalars(ctx, 2))) return rc;
        launch_penalty_diag(ctx->st, ctx->d_pen, ctx->d_pen_mult, ctx->n_pen, ctx->d_sptr, ctx->d_kdiag,
                            ctx->d_K64, ctx->d_sc + 2);
    }
    launch_diag_inv(ctx->st, (int)ctx->n_nodes, ctx->d_sptr, ctx->d_kdiag, ctx->d_K64, ctx->d_dinv64,
                    ctx->d_dinv32);
    launch_to_f32(ctx->st, ctx->d_K64, ctx->d_K32, nval);
    refresh_f16(ctx);
    HIPCHK(hipGetLastError());
    ctx->assembled = true;
    ctx->asm_count += 1;
    ctx->asm_cmass = c_mass;
    ctx->asm_kscale = k_scale;
    ctx->asm_cdamp = 0.0;  // assemble_device sets it after this for its own assemblies
    return ASEND_OK;
}

// K_eff = k_scale K + c_mass M (+ penalty), its Jacobi diagonal and FP32 copy
int assemble_device(asend_ctx* ctx, double k_scale, double c_mass, double c_damp = 0.0) {
    if (ctx->etype == ASEND_SHELL4 && !ctx->thk_set)
        return fail(ctx, ASEND_ERR_STATE, "shell4 mesh without thickness (asend_set_thickness)");
    int rc = integrate_elements(ctx, ctx->d_E, k_scale, c_mass, ctx->d_K64, c_damp);
    if (rc || (rc = finish_assembly(ctx, k_scale, c_mass))) return rc;
    ctx->asm_cdamp = c_damp;
    return ASEND_OK;
}

// Geometrically nonlinear brick8: tangent of the internal force at the
// displacement state d_nl_u into d_K64 (then penalty, diagonal, FP32 copy),
// the internal force into d_nl_r and, when rhs is set, the Newton load
// rhs = fact f - r(u) - sf P u (build_elastic_soln_load(true) with
// nonlinear_geom, S/model_analysis.rs:179-201, and build_elastic_const_load,
// :275-287).
// Mesh checks and buffers of the nonlinear path (allocated on first use).
int nl_prepare(asend_ctx* ctx) {
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8 && ctx->etype != ASEND_BRICKIM)
        return fail(ctx, ASEND_ERR_UNSUPPORTED,
                    "geometrically nonlinear analysis is implemented for brick8 and brickIM meshes");
    if (ctx->max_valence > 8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "nonlinear assembly supports at most 8 elements per node (mesh has %d)",
                    ctx->max_valence);
    int rc;
    const int64_t n = ctx->n_dof;
    if (!ctx->d_kb && (rc = dalloc(ctx, &ctx->d_kb, (int64_t)8 * kRecStride * ctx->n_el))) return rc;
    if (!ctx->d_nl_re &&
        ((rc = dalloc(ctx, &ctx->d_nl_re, 24 * ctx->n_el)) ||
         (rc = dalloc(ctx, &ctx->d_nl_u, n)) || (rc = dalloc(ctx, &ctx->d_nl_f, n)) ||
         (rc = dalloc(ctx, &ctx->d_nl_r, n))))
        return rc;
    if (ctx->etype == ASEND_BRICKIM && !ctx->d_nl_alpha) {
        if ((rc = dalloc(ctx, &ctx->d_nl_alpha, 9 * ctx->n_el)) || (rc = dalloc(ctx, &ctx->d_nl_wmap, 225 * ctx->n_el)))
            return rc;
        HIPCHK(hipMemset(ctx->d_nl_alpha, 0, 9 * ctx->n_el * sizeof(double)));
    }
    return ASEND_OK;
}

int nl_assemble(asend_ctx* ctx, double fact, double* rhs) {
    int rc;
    if (ctx->etype == ASEND_BRICKIM)
        launch_brickim_nl(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, ctx->d_nl_u,
                          ctx->d_nl_alpha, ctx->d_kb, ctx->d_nl_re, ctx->d_nl_wmap);
    else
        launch_hex8_nl_element(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu,
                               ctx->d_nl_u, ctx->d_kb, ctx->d_nl_re);
    launch_rec_gather(ctx->st, (int)ctx->n_nodes, (int)ctx->n_slices, ctx->d_sptr, ctx->d_n2e_ptr, ctx->d_n2rec,
                      ctx->d_slot, ctx->d_kb, ctx->d_nl_re, fact, ctx->d_nl_f, ctx->d_K64, ctx->d_nl_r, rhs);
    HIPCHK(hipGetLastError());
    if ((rc = finish_assembly(ctx, 1.0, 0.0))) return rc;
    if (rhs) {
        launch_nl_penalty_load(ctx->st, ctx->d_pen, ctx->d_pen_mult, ctx->n_pen, ctx->d_sc + 2, ctx->d_nl_u, rhs);
        HIPCHK(hipGetLastError());
    }
    return ASEND_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Scratch of the allocation-free FP64 solve (asend_solve_noalloc): the seven
// work vectors of the FP64 Jacobi PCG for up to kNoAllocMaxDof DOFs, the
// reduction partials and the two loop scalars, in one device array that the
// code object holds on every GPU the library runs on until the process ends
// (14.8 MB).  Host code never passes `scratch` itself to a kernel: its name
// is a host-side stand-in, not a device address.  Each device's address is
// looked up once (scratch_setup, called by asend_ctx_create with that device
// current) and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
namespace {
constexpr int64_t kNoAllocMaxDof = 1 << 18;
constexpr int64_t kNoAllocPartials = 2 * (int64_t)kMaxPartials;
constexpr int64_t kNoAllocScalarDoubles = 64;  // room for the two Scalars
static_assert(2 * sizeof(Scalars) <= kNoAllocScalarDoubles * sizeof(double), "scalar slot too small");
constexpr int64_t kNoAllocDoubles = 7 * kNoAllocMaxDof + kNoAllocPartials + kNoAllocScalarDoubles;
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};

// Device address of `scratch` on `device` (which must be current), looked up
// on the first call for that device.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return hipSuccess;  // asend_solve_noalloc then refuses
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = (double*)ptr;
    return hipSuccess;
}

double* scratch_ptr(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// The FP64 solve's buffers carved out of the device's scratch array.
Fp64Bufs scratch_fp64_bufs(double* base) {
    const int64_t v = kNoAllocMaxDof;
    Fp64Bufs B;
    B.f = base;
    B.u = base + v;
    B.r = base + 2 * v;
    B.x = base + 3 * v;
    B.r2 = base + 4 * v;
    B.p = base + 5 * v;
    B.q = base + 6 * v;
    B.partials = base + 7 * v;
    B.sc = (Scalars*)(base + 7 * v + kNoAllocPartials);
    return B;
}

// Entry points whose kernels integrate the isotropic material refuse a
// context whose elements carry general material matrices.
int iso_only(asend_ctx* ctx, const char* what) {
    if (ctx->n_smat == 0) return ASEND_OK;
    return fail(ctx, ASEND_ERR_UNSUPPORTED,
                "%s integrates isotropic materials; this context holds general material matrices "
                "(asend_set_material_stiffness)",
                what);
}

// Entry points built on the elastic operator refuse a context that
// assembles heat conduction.
int elastic_only(asend_ctx* ctx, const char* what) {
    if (!ctx->d_kcond) return ASEND_OK;
    return fail(ctx, ASEND_ERR_UNSUPPORTED,
                "%s needs the elastic operator; this context assembles heat conduction (asend_set_conductivity)", what);
}
}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int asend_ctx_create(int device, asend_ctx** out) {
    if (!out) return ASEND_ERR_ARG;
    *out = nullptr;
    asend_ctx* ctx = new asend_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        delete ctx;
        return ASEND_ERR_HIP;
    }
    if (scratch_setup(device) != hipSuccess) {
        delete ctx;
        return ASEND_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&ctx->d_sc, 4 * sizeof(Scalars)) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_sc, 4 * sizeof(Scalars), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&ctx->d_partials, kMaxPartials * 2 * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&ctx->d_counter, 16 * sizeof(unsigned)) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->st2, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming) != hipSuccess) {
        asend_ctx_destroy(ctx);
        return ASEND_ERR_HIP;
    }
    if (hipMemset(ctx->d_counter, 0, 16 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(ctx->d_sc, 0, 4 * sizeof(Scalars)) != hipSuccess) {
        asend_ctx_destroy(ctx);
        return ASEND_ERR_HIP;
    }
    std::memset(ctx->h_sc, 0, 4 * sizeof(Scalars));
    *out = ctx;
    return ASEND_OK;
}

void asend_ctx_destroy(asend_ctx* ctx) {
    if (!ctx) return;
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    free_mesh(ctx);
    if (ctx->learned) learned_destroy(ctx->learned);
    ctx->learned = nullptr;
    dfree(ctx->d_pen);
    dfree(ctx->d_pen_mult);
    dfree(ctx->d_sc);
    dfree(ctx->d_partials);
    dfree(ctx->d_counter);
    if (ctx->h_sc) (void)hipHostFree(ctx->h_sc);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->st2) (void)hipStreamSynchronize(ctx->st2);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->st2) (void)hipStreamDestroy(ctx->st2);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

const char* asend_last_error(const asend_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int asend_sync(asend_ctx* ctx) {
    if (!ctx) return ASEND_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ASEND_OK;
}

int asend_upload_mesh(asend_ctx* ctx, int elem_type, int64_t n_mesh_nodes, const double* xyz,
                      const int64_t* dof_index, int64_t n_dof, int64_t n_el, const int32_t* conn, const double* E,
                      const double* nu, const double* rho) {
    if (!ctx) return ASEND_ERR_ARG;
    if (elem_type != ASEND_BRICK8 && elem_type != ASEND_BRICKIM && elem_type != ASEND_SHELL4)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element type %d has no device integration kernel", elem_type);
    if (n_mesh_nodes <= 0 || n_el <= 0 || !xyz || !dof_index || !conn || !E || !nu)
        return fail(ctx, ASEND_ERR_ARG, "bad mesh arguments");
    // a brick8 node is one 3x3 block row; a shell4 node two (translations,
    // rotations), so every element touches 8 block rows either way
    const bool shell = (elem_type == ASEND_SHELL4);
    const int dpn = shell ? 6 : 3, npe = shell ? 4 : 8;
    const int64_t n_nodes = shell ? 2 * n_mesh_nodes : n_mesh_nodes;  // block rows
    if (n_dof != dpn * n_mesh_nodes)
        return fail(ctx, ASEND_ERR_ARG, "n_dof must equal %d*n_nodes for this element type", dpn);
    if (n_nodes >= (1ll << 31) / 9) return fail(ctx, ASEND_ERR_UNSUPPORTED, "mesh too large for int32 indices");
    if (shell) {
        // flat elements only: the diagonals' normal must be parallel to z
        for (int64_t e = 0; e < n_el; ++e) {
            double d1[3], d2[3];
            for (int c = 0; c < 4; ++c)
                if (conn[4 * e + c] < 0 || conn[4 * e + c] >= n_mesh_nodes)
                    return fail(ctx, ASEND_ERR_ARG, "connectivity out of range");
            for (int k = 0; k < 3; ++k) {
                d1[k] = xyz[3 * (int64_t)conn[4 * e + 2] + k] - xyz[3 * (int64_t)conn[4 * e + 0] + k];
                d2[k] = xyz[3 * (int64_t)conn[4 * e + 3] + k] - xyz[3 * (int64_t)conn[4 * e + 1] + k];
            }
            const double nx = d1[1] * d2[2] - d1[2] * d2[1], ny = d1[2] * d2[0] - d1[0] * d2[2];
            if (std::sqrt(nx * nx + ny * ny) >= 1e-12)
                return fail(ctx, ASEND_ERR_UNSUPPORTED,
                            "shell4 element %lld is not normal to z (only the identity element frame is supported)",
                            (long long)e);
        }
    }
    {
        int64_t first = -1;
        const int64_t nbad = compliance_clamp_count(elem_type, n_el, E, nu, &first);
        if (nbad > 0)
            return fail(ctx, ASEND_ERR_UNSUPPORTED,
                        "%lld element(s) (first %lld: E = %.6g, nu = %.6g) are past the pivot clamp of the reference's "
                        "compliance inversion (pivots below 1e-12 replaced by 1e-12), where its stiffness is not the "
                        "isotropic C integrated here; pose the model in units where 1/E stays above 1e-12 (e.g. MPa)",
                        (long long)nbad, (long long)first, E[first], nu[first]);
    }
    HIPCHK(hipSetDevice(ctx->device));
    free_mesh(ctx);
    ctx->etype = elem_type;
    ctx->bs = 3;
    ctx->n_nodes = n_nodes;
    ctx->n_dof = n_dof;
    ctx->n_el = n_el;
    ctx->sf_fixed = false;
    // node rank from the DOF numbering: dof_index[6n+c] = dpn*rank + c
    std::vector<int32_t> rank(n_mesh_nodes);
    std::vector<uint8_t> seen(n_mesh_nodes, 0);
    for (int64_t nd = 0; nd < n_mesh_nodes; ++nd) {
        int64_t d0 = dof_index[6 * nd];
        bool ok = d0 >= 0 && d0 % dpn == 0 && d0 < n_dof;
        for (int c = 1; c < dpn && ok; ++c) ok = dof_index[6 * nd + c] == d0 + c;
        if (!ok) return fail(ctx, ASEND_ERR_ARG, "dof_index of node %lld is not a %d-DOF node block", (long long)nd, dpn);
        int64_t r = d0 / dpn;
        if (seen[r]) return fail(ctx, ASEND_ERR_ARG, "dof_index is not a permutation");
        seen[r] = 1;
        rank[nd] = (int32_t)r;
    }
    std::vector<int32_t> conn_r(8 * n_el);
    for (int64_t i = 0; i < npe * n_el; ++i) {
        int32_t v = conn[i];
        if (v < 0 || v >= n_mesh_nodes) return fail(ctx, ASEND_ERR_ARG, "connectivity out of range");
        if (shell) {
            conn_r[2 * i] = 2 * rank[v];
            conn_r[2 * i + 1] = 2 * rank[v] + 1;
        } else {
            conn_r[i] = rank[v];
        }
    }
    std::vector<double> xyz_r(3 * n_mesh_nodes);
    for (int64_t nd = 0; nd < n_mesh_nodes; ++nd)
        for (int c = 0; c < 3; ++c) xyz_r[3 * (int64_t)rank[nd] + c] = xyz[3 * nd + c];
    // node -> element adjacency
    std::vector<int32_t> n2e_ptr(n_nodes + 1, 0), n2e, n2rec;
    for (int64_t i = 0; i < 8 * n_el; ++i) n2e_ptr[conn_r[i] + 1]++;
    for (int64_t r = 0; r < n_nodes; ++r) n2e_ptr[r + 1] += n2e_ptr[r];
    n2e.resize(n2e_ptr[n_nodes]);
    n2rec.resize(n2e_ptr[n_nodes]);
    {
        std::vector<int32_t> fill(n2e_ptr.begin(), n2e_ptr.end() - 1);
        for (int64_t e = 0; e < n_el; ++e)
            for (int a = 0; a < 8; ++a) {
                const int32_t pos = fill[conn_r[8 * e + a]]++;
                n2e[pos] = (int32_t)e;
                n2rec[pos] = (int32_t)(8 * e + a);
            }
    }
    ctx->max_valence = 0;
    for (int64_t r = 0; r < n_nodes; ++r) ctx->max_valence = std::max(ctx->max_valence, n2e_ptr[r + 1] - n2e_ptr[r]);
    // brick8 meshes whose elements are all translates of the first one (to
    // 1e-12 of the mesh extent) with one Poisson ratio and at most 8 elements
    // per node assemble from scaled reference blocks
    bool uniform = !shell;
    if (uniform) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t nd = 0; nd < n_mesh_nodes; ++nd)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], xyz[3 * nd + c]);
                hi[c] = std::max(hi[c], xyz[3 * nd + c]);
            }
        const double tol = 1e-12 * std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
        double ref[8][3];
        for (int t = 0; t < 8; ++t)
            for (int c = 0; c < 3; ++c) ref[t][c] = xyz[3 * (int64_t)conn[t] + c] - xyz[3 * (int64_t)conn[0] + c];
        for (int64_t e = 0; e < n_el && uniform; ++e) {
            if (nu[e] != nu[0]) uniform = false;
            for (int t = 0; t < 8 && uniform; ++t)
                for (int c = 0; c < 3; ++c)
                    if (std::fabs(xyz[3 * (int64_t)conn[8 * e + t] + c] - xyz[3 * (int64_t)conn[8 * e] + c] - ref[t][c]) >
                        tol)
                        uniform = false;
        }
        for (int64_t r = 0; r < n_nodes && uniform; ++r)
            if (n2e_ptr[r + 1] - n2e_ptr[r] > 8) uniform = false;
    }
    // block rows: sorted unique neighbour ranks
    ctx->h_rp.assign(n_nodes + 1, 0);
    std::vector<std::vector<int32_t>> rows(n_nodes);
    {
        std::vector<int32_t> tmp;
        for (int64_t r = 0; r < n_nodes; ++r) {
            tmp.clear();
            for (int32_t k = n2e_ptr[r]; k < n2e_ptr[r + 1]; ++k) {
                int32_t e = n2e[k];
                for (int a = 0; a < 8; ++a) tmp.push_back(conn_r[8 * (int64_t)e + a]);
            }
            if (tmp.empty()) tmp.push_back((int32_t)r);  // isolated node keeps its diagonal
            std::sort(tmp.begin(), tmp.end());
            tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
            if (tmp.size() > 255) return fail(ctx, ASEND_ERR_UNSUPPORTED, "node with more than 255 neighbours");
            rows[r] = tmp;
            ctx->h_rp[r + 1] = ctx->h_rp[r] + (int32_t)tmp.size();
        }
    }
    ctx->nnzb = ctx->h_rp[n_nodes];
    ctx->h_bc.resize(ctx->nnzb);
    std::vector<int32_t> kdiag(n_nodes);
    for (int64_t r = 0; r < n_nodes; ++r) {
        std::copy(rows[r].begin(), rows[r].end(), ctx->h_bc.begin() + ctx->h_rp[r]);
        kdiag[r] = (int32_t)(std::lower_bound(rows[r].begin(), rows[r].end(), (int32_t)r) - rows[r].begin());
    }
    // mesh-edge neighbours (the 12 edges of each hex) for the learned factor:
    // [slot][rank] tables, -1 padded, with the reverse slot and the operator slot
    std::vector<int32_t> enb, erev, eslot;
    int n_edge = 0;
    {
        static const int hex_edges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                             {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
        std::vector<std::vector<int32_t>> en(n_nodes);
        for (int64_t e = 0; e < (shell ? 0 : n_el); ++e)
            for (const auto& ed : hex_edges) {
                int32_t a = conn_r[8 * e + ed[0]], b = conn_r[8 * e + ed[1]];
                if (a == b) continue;
                en[a].push_back(b);
                en[b].push_back(a);
            }
        for (auto& v : en) {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            n_edge = std::max<int>(n_edge, (int)v.size());
        }
        if (n_edge > kMaxEdges) return fail(ctx, ASEND_ERR_UNSUPPORTED, "node with more than %d edge neighbours", kMaxEdges);
        enb.assign((size_t)std::max(n_edge, 1) * n_nodes, -1);
        erev.assign(enb.size(), -1);
        eslot.assign(enb.size(), -1);
        for (int64_t r = 0; r < n_nodes; ++r)
            for (size_t s = 0; s < en[r].size(); ++s) {
                int32_t j = en[r][s];
                enb[s * n_nodes + r] = j;
                erev[s * n_nodes + r] = (int32_t)(std::lower_bound(en[j].begin(), en[j].end(), (int32_t)r) - en[j].begin());
                eslot[s * n_nodes + r] = (int32_t)(std::lower_bound(rows[r].begin(), rows[r].end(), j) - rows[r].begin());
            }
    }
    // element -> in-row slot of each (a, b) block
    std::vector<uint8_t> slot(64 * n_el);
    for (int64_t e = 0; e < n_el; ++e)
        for (int a = 0; a < 8; ++a) {
            const auto& row = rows[conn_r[8 * e + a]];
            for (int b = 0; b < 8; ++b)
                slot[64 * e + 8 * a + b] =
                    (uint8_t)(std::lower_bound(row.begin(), row.end(), conn_r[8 * e + b]) - row.begin());
        }
    // sliced-ELL layout of the fine operator: 64-row slices padded to the
    // widest row of the slice; padding slots point at the row itself with
    // zero values
    ctx->n_slices = (n_nodes + 63) / 64;
    ctx->h_sptr.assign(ctx->n_slices + 1, 0);
    for (int64_t sl = 0; sl < ctx->n_slices; ++sl) {
        int w = 0;
        for (int64_t r = 64 * sl; r < std::min<int64_t>(64 * sl + 64, n_nodes); ++r) w = std::max<int>(w, (int)rows[r].size());
        ctx->h_sptr[sl + 1] = ctx->h_sptr[sl] + w;
    }
    ctx->n_slots = ctx->h_sptr[ctx->n_slices];
    // the SpMV with one slice per workgroup writes one partial per slice
    if (ctx->n_slices > ctx->partials_cap) {
        int rc0 = dalloc(ctx, &ctx->d_partials, 2 * (size_t)ctx->n_slices);
        if (rc0) return rc0;
        ctx->partials_cap = ctx->n_slices;
    }
    std::vector<int32_t> scol(64 * ctx->n_slots, 0);
    for (int64_t sl = 0; sl < ctx->n_slices; ++sl)
        for (int lane = 0; lane < 64; ++lane) {
            int64_t r = 64 * sl + lane;
            int w = ctx->h_sptr[sl + 1] - ctx->h_sptr[sl];
            for (int k = 0; k < w; ++k) {
                int32_t c = 0;
                if (r < n_nodes) c = (k < (int)rows[r].size()) ? rows[r][k] : (int32_t)r;
                scol[64 * (int64_t)(ctx->h_sptr[sl] + k) + lane] = c;
            }
        }
    rows.clear();
    rows.shrink_to_fit();
    // greedy element colouring: no two elements of a colour share a node
    std::vector<uint64_t> mask(n_nodes, 0);
    std::vector<int8_t> col(n_el);
    int ncol = 0;
    for (int64_t e = 0; e < n_el; ++e) {
        uint64_t used = 0;
        for (int a = 0; a < 8; ++a) used |= mask[conn_r[8 * e + a]];
        if (~used == 0) return fail(ctx, ASEND_ERR_UNSUPPORTED, "more than 64 element colours");
        int c = __builtin_ctzll(~used);
        col[e] = (int8_t)c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < 8; ++a) mask[conn_r[8 * e + a]] |= (1ull << c);
    }
    ctx->h_color_ptr.assign(ncol + 1, 0);
    for (int64_t e = 0; e < n_el; ++e) ctx->h_color_ptr[col[e] + 1]++;
    for (int c = 0; c < ncol; ++c) ctx->h_color_ptr[c + 1] += ctx->h_color_ptr[c];
    std::vector<int32_t> elems(n_el);
    {
        std::vector<int32_t> fill(ctx->h_color_ptr.begin(), ctx->h_color_ptr.end() - 1);
        for (int64_t e = 0; e < n_el; ++e) elems[fill[col[e]]++] = (int32_t)e;
    }
    int rc;
    const int64_t nval = ctx->nval();
    if ((rc = dalloc(ctx, &ctx->d_sptr, ctx->n_slices + 1)) || (rc = dalloc(ctx, &ctx->d_scol, scol.size())) ||
        (rc = dalloc(ctx, &ctx->d_conn_r, 8 * n_el)) || (rc = dalloc(ctx, &ctx->d_rp, n_nodes + 1)) ||
        (rc = dalloc(ctx, &ctx->d_bc, ctx->nnzb)) || (rc = dalloc(ctx, &ctx->d_kdiag, n_nodes)) ||
        (rc = dalloc(ctx, &ctx->d_elems, n_el)) || (rc = dalloc(ctx, &ctx->d_slot, 64 * n_el)) ||
        (rc = dalloc(ctx, &ctx->d_xyz_r, 3 * n_nodes)) || (rc = dalloc(ctx, &ctx->d_E, n_el)) ||
        (rc = dalloc(ctx, &ctx->d_nu, n_el)) || (rc = dalloc(ctx, &ctx->d_rho, n_el)) ||
        (rc = dalloc(ctx, &ctx->d_K64, nval)) || (rc = dalloc(ctx, &ctx->d_K32, nval)) ||
        (rc = dalloc(ctx, &ctx->d_dinv64, n_dof)) || (rc = dalloc(ctx, &ctx->d_dinv32, n_dof)) ||
        (rc = dalloc(ctx, &ctx->d_n2e_ptr, n_nodes + 1)) || (rc = dalloc(ctx, &ctx->d_n2e, n2e.size())) ||
        (rc = dalloc(ctx, &ctx->d_n2rec, n2rec.size())) ||
        (rc = dalloc(ctx, &ctx->d_clamped, n_nodes)) || (rc = dalloc(ctx, &ctx->d_enb, enb.size())) ||
        (rc = dalloc(ctx, &ctx->d_erev, erev.size())) || (rc = dalloc(ctx, &ctx->d_eslot, eslot.size())))
        return rc;
    ctx->n_edge = n_edge;
    HIPCHK(hipMemcpy(ctx->d_n2e_ptr, n2e_ptr.data(), n2e_ptr.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_n2e, n2e.data(), n2e.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_n2rec, n2rec.data(), n2rec.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(ctx->d_clamped, 0, n_nodes));
    HIPCHK(hipMemcpy(ctx->d_enb, enb.data(), enb.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_erev, erev.data(), erev.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_eslot, eslot.data(), eslot.size() * 4, hipMemcpyHostToDevice));
    {
        double s = 0.0;
        for (int64_t e = 0; e < n_el; ++e) s += std::log(E[e]);
        ctx->mean_log_e = s / (double)n_el;
        ctx->h_nu.assign(nu, nu + n_el);
    }
    if (shell && (rc = dalloc(ctx, &ctx->d_thk, n_el))) return rc;
    HIPCHK(hipMemcpy(ctx->d_sptr, ctx->h_sptr.data(), ctx->h_sptr.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_scol, scol.data(), scol.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_conn_r, conn_r.data(), conn_r.size() * 4, hipMemcpyHostToDevice));
    // padding values must read as zero in every later assembly
    HIPCHK(hipMemset(ctx->d_K64, 0, nval * sizeof(double)));
    HIPCHK(hipMemcpy(ctx->d_rp, ctx->h_rp.data(), ctx->h_rp.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_bc, ctx->h_bc.data(), ctx->h_bc.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_kdiag, kdiag.data(), kdiag.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_elems, elems.data(), elems.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_slot, slot.data(), slot.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_xyz_r, xyz_r.data(), xyz_r.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_E, E, n_el * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_nu, nu, n_el * 8, hipMemcpyHostToDevice));
    if (rho) HIPCHK(hipMemcpy(ctx->d_rho, rho, n_el * 8, hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(ctx->d_rho, 0, n_el * 8));
    if (uniform) {
        if ((rc = dalloc(ctx, &ctx->d_khat, 64 * 9)) || (rc = dalloc(ctx, &ctx->d_mhat, 64))) return rc;
        if (elem_type == ASEND_BRICKIM)
            launch_brickim_reference(ctx->st, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_nu, ctx->d_khat, ctx->d_mhat);
        else
            launch_hex8_reference(ctx->st, ctx->d_conn_r, ctx->d_xyz_r, nu[0], ctx->d_khat, ctx->d_mhat);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->st));
        ctx->uniform_hex = true;
    }
    if ((rc = ensure_work(ctx))) return rc;
    return ASEND_OK;
}

int asend_set_thickness(asend_ctx* ctx, const double* t) {
    if (!ctx || !t) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_SHELL4) return fail(ctx, ASEND_ERR_UNSUPPORTED, "thickness applies to shell4 meshes");
    for (int64_t e = 0; e < ctx->n_el; ++e)
        if (!(t[e] > 0.0)) return fail(ctx, ASEND_ERR_ARG, "thickness of element %lld is not positive", (long long)e);
    HIPCHK(hipMemcpy(ctx->d_thk, t, ctx->n_el * 8, hipMemcpyHostToDevice));
    ctx->thk_set = true;
    ctx->assembled = false;
    return ASEND_OK;
}

int asend_set_grid(asend_ctx* ctx, int64_t nx, int64_t ny, int64_t nz) {
    if (!ctx) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the multigrid hierarchy is built for brick8 grids");
    if (nx * ny * nz != ctx->n_el || (nx + 1) * (ny + 1) * (nz + 1) != ctx->n_nodes)
        return fail(ctx, ASEND_ERR_ARG, "grid %lldx%lldx%lld does not match the mesh", (long long)nx, (long long)ny,
                    (long long)nz);
    ctx->grid[0] = nx;
    ctx->grid[1] = ny;
    ctx->grid[2] = nz;
    if (ctx->mg) mg_destroy(ctx->mg);
    ctx->mg = nullptr;
    // rank of structured node (i,j,k) -> needed by the transfer operators
    std::vector<int32_t> rank_of_grid(ctx->n_nodes, -1);
    std::vector<int32_t> conn_r(8 * ctx->n_el);
    HIPCHK(hipMemcpy(conn_r.data(), ctx->d_conn_r, conn_r.size() * 4, hipMemcpyDeviceToHost));
    // element e = i + nx*(j + ny*k) has node 0 at grid (i,j,k), node 6 at (i+1,j+1,k+1)
    const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j)
            for (int64_t i = 0; i < nx; ++i) {
                int64_t e = i + nx * (j + ny * k);
                for (int a = 0; a < 8; ++a) {
                    int64_t gi = i + corner[a][0], gj = j + corner[a][1], gk = k + corner[a][2];
                    int32_t& slotr = rank_of_grid[gi + (nx + 1) * (gj + (ny + 1) * gk)];
                    if (slotr >= 0 && slotr != conn_r[8 * e + a])
                        return fail(ctx, ASEND_ERR_ARG, "mesh is not the structured grid %lldx%lldx%lld in label order",
                                    (long long)nx, (long long)ny, (long long)nz);
                    slotr = conn_r[8 * e + a];
                }
            }
    int rc = mg_create(&ctx->mg, (int)nx, (int)ny, (int)nz, rank_of_grid.data(), ctx->d_xyz_r, ctx->d_conn_r,
                       ctx->fine(), ctx->n_nodes, ctx->n_el);
    if (ctx->mg) mg_use_graphs(ctx->mg, ctx->use_graphs);
    if (rc) {
        std::string m = ctx->mg ? mg_last_error(ctx->mg) : "multigrid create failed";
        if (ctx->mg) mg_destroy(ctx->mg);
        ctx->mg = nullptr;
        return fail(ctx, rc, "%s", m.c_str());
    }
    if (!ctx->d_K16 && (rc = dalloc(ctx, &ctx->d_K16, ctx->nval()))) return rc;
    mg_set_fine_half(ctx->mg, ctx->d_K16, ctx->d_dinv32);
    mg_set_cycle_half(ctx->mg, ctx->cycle_half);
    if (ctx->assembled) {
        refresh_f16(ctx);
        HIPCHK(hipGetLastError());
    }
    return ASEND_OK;
}

int asend_mg_set_cycle_precision(asend_ctx* ctx, int bits) {
    if (!ctx || (bits != 16 && bits != 32)) return ASEND_ERR_ARG;
    ctx->cycle_half = (bits == 16);
    if (ctx->mg) mg_set_cycle_half(ctx->mg, ctx->cycle_half);
    return ASEND_OK;
}

int asend_update_moduli(asend_ctx* ctx, const double* E) {
    if (!ctx || !E) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_update_moduli")) return rc0;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    {
        int64_t first = -1;
        const int64_t nbad = compliance_clamp_count(ctx->etype, ctx->n_el, E, ctx->h_nu.data(), &first);
        if (nbad > 0)
            return fail(ctx, ASEND_ERR_UNSUPPORTED,
                        "%lld element(s) (first %lld: E = %.6g) are past the pivot clamp of the reference's compliance "
                        "inversion; moduli not updated",
                        (long long)nbad, (long long)first, E[first]);
    }
    HIPCHK(hipMemcpyAsync(ctx->d_E, E, ctx->n_el * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    double s = 0.0;
    for (int64_t e = 0; e < ctx->n_el; ++e) s += std::log(E[e]);
    ctx->mean_log_e = s / (double)ctx->n_el;
    ctx->assembled = false;
    return ASEND_OK;
}

int asend_set_penalty(asend_ctx* ctx, int64_t n, const int64_t* dofs, double sf) {
    if (!ctx || n < 0 || (n > 0 && !dofs)) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    std::vector<int64_t> v(dofs, dofs + n);
    for (int64_t d : v)
        if (d < 0 || d >= ctx->n_dof) return fail(ctx, ASEND_ERR_ARG, "penalty DOF out of range");
    std::sort(v.begin(), v.end());
    std::vector<int64_t> u;
    std::vector<int32_t> m;
    for (size_t i = 0; i < v.size(); ++i) {
        if (!u.empty() && u.back() == v[i]) m.back()++;
        else {
            u.push_back(v[i]);
            m.push_back(1);
        }
    }
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_pen, u.size())) || (rc = dalloc(ctx, &ctx->d_pen_mult, m.size()))) return rc;
    if (!u.empty()) {
        HIPCHK(hipMemcpy(ctx->d_pen, u.data(), u.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_pen_mult, m.data(), m.size() * 4, hipMemcpyHostToDevice));
    }
    {
        std::vector<uint8_t> cl(ctx->n_nodes, 0);
        for (int64_t d : u) cl[d / 3] = 1;
        HIPCHK(hipMemcpy(ctx->d_clamped, cl.data(), cl.size(), hipMemcpyHostToDevice));
    }
    ctx->n_pen = (int)u.size();
    ctx->h_pen = u;
    ctx->h_pen_mult = m;
    ctx->sf_user = sf;
    ctx->sf_fixed = false;
    ctx->assembled = false;
    return ASEND_OK;
}

double asend_penalty_scale(const asend_ctx* ctx) { return ctx ? ctx->sf : 0.0; }

int asend_uniform_elements(const asend_ctx* ctx) { return (ctx && ctx->uniform_hex) ? 1 : 0; }

int asend_assemble(asend_ctx* ctx, double c_mass) {
    if (!ctx) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    int rc = assemble_device(ctx, 1.0, c_mass);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ASEND_OK;
}

int asend_get_matrix(asend_ctx* ctx, int64_t* rowptr, int32_t* cols, double* vals, int64_t* nnz, int with_penalty) {
    if (!ctx || !nnz) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    const int64_t n_nodes = ctx->n_nodes;
    *nnz = ctx->nnzb * 9;
    if (!rowptr) return ASEND_OK;
    for (int64_t r = 0; r < n_nodes; ++r) {
        int64_t nb = ctx->h_rp[r + 1] - ctx->h_rp[r];
        for (int i = 0; i < 3; ++i) rowptr[3 * r + i] = 9 * (int64_t)ctx->h_rp[r] + 3 * nb * i;
    }
    rowptr[3 * n_nodes] = 9 * ctx->nnzb;
    if (!cols || !vals) return ASEND_OK;
    std::vector<double> K(ctx->nval());
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipMemcpy(K.data(), ctx->d_K64, K.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < n_nodes; ++r) {
        int64_t s = ctx->h_rp[r], nb = ctx->h_rp[r + 1] - s;
        for (int i = 0; i < 3; ++i) {
            int64_t o = rowptr[3 * r + i];
            for (int64_t k = 0; k < nb; ++k)
                for (int j = 0; j < 3; ++j) {
                    cols[o + 3 * k + j] = 3 * ctx->h_bc[s + k] + j;
                    vals[o + 3 * k + j] = K[sell_index(ctx->h_sptr.data(), (int)r, (int)k, 3 * i + j)];
                }
        }
    }
    if (!with_penalty && ctx->n_pen > 0) {
        std::vector<int64_t> pen(ctx->n_pen);
        std::vector<int32_t> mult(ctx->n_pen);
        HIPCHK(hipMemcpy(pen.data(), ctx->d_pen, pen.size() * 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(mult.data(), ctx->d_pen_mult, mult.size() * 4, hipMemcpyDeviceToHost));
        for (int i = 0; i < ctx->n_pen; ++i) {
            int64_t d = pen[i];
            for (int64_t k = rowptr[d]; k < rowptr[d + 1]; ++k)
                if (cols[k] == d) vals[k] -= mult[i] * ctx->sf;
        }
    }
    return ASEND_OK;
}

int asend_spmv(asend_ctx* ctx, int precision, const double* x, double* y) {
    if (!ctx || !x || !y) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    const int64_t n = ctx->n_dof;
    if (precision == 64) {
        HIPCHK(hipMemcpy(ctx->d_p, x, n * 8, hipMemcpyHostToDevice));
        launch_sell3<double>(ctx->st, 0, ctx->fine(), ctx->d_K64, ctx->d_p, ctx->d_q, nullptr, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->st));
        HIPCHK(hipMemcpy(y, ctx->d_q, n * 8, hipMemcpyDeviceToHost));
    } else if (precision == 32) {
        std::vector<float> xf(x, x + n), yf(n);
        HIPCHK(hipMemcpy(ctx->d_p32, xf.data(), n * 4, hipMemcpyHostToDevice));
        launch_sell3<float>(ctx->st, 0, ctx->fine(), ctx->d_K32, ctx->d_p32, ctx->d_q32, nullptr, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->st));
        HIPCHK(hipMemcpy(yf.data(), ctx->d_q32, n * 4, hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < n; ++i) y[i] = yf[i];
    } else {
        return fail(ctx, ASEND_ERR_ARG, "precision must be 32 or 64");
    }
    return ASEND_OK;
}

int asend_solve(asend_ctx* ctx, const double* rhs, double* x, const asend_solve_opts* opts, asend_solve_stats* stats) {
    if (!ctx || !rhs || !x || !opts) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    const int64_t n = ctx->n_dof;
    HIPCHK(hipMemcpy(ctx->d_f, rhs, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    asend_solve_stats s;
    int rc = solve_device(ctx, opts, &s);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    s.device_ms = ms;
    HIPCHK(hipMemcpy(x, ctx->d_u, n * 8, hipMemcpyDeviceToHost));
    if (stats) *stats = s;
    return ASEND_OK;
}

int asend_cycle_spmv_timing(asend_ctx* ctx, int enable, int64_t* launches, double* total_ms) {
    if (!ctx) return ASEND_ERR_ARG;
    if (enable) {
        HIPCHK(hipStreamSynchronize(ctx->st));
        kernel_timer_start();
        return ASEND_OK;
    }
    if (!launches || !total_ms) return ASEND_ERR_ARG;
    if (kernel_timer_stop(launches, total_ms)) return fail(ctx, ASEND_ERR_HIP, "cycle SpMV timing events failed");
    return ASEND_OK;
}

int asend_ldl_factor(asend_ctx* ctx, int64_t block_dofs, asend_ldl_info* info) {
    if (!ctx) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    const int64_t bs = (block_dofs <= 0 || block_dofs >= ctx->n_dof) ? ctx->n_dof : block_dofs;
    if (bs != ctx->ldl_block_dofs) {
        ldl_release(ctx);
        ctx->ldl_block_dofs = bs;
    }
    int rc;
    if ((rc = ldl_ensure(ctx))) {
        ldl_release(ctx);
        ctx->ldl_block_dofs = -1;
        return rc;
    }
    if (info) {
        const LdlFactor& F = ctx->ldl;
        info->block_dofs = F.d.bs;
        info->n_blocks = F.d.n_blocks;
        info->max_bandwidth = F.max_bandwidth;
        info->bytes = F.bytes;
        info->neg_pivots = F.neg_pivots;
        info->pos_def = F.neg_pivots == 0 ? 1 : 0;
        info->factor_ms = F.factor_ms;
    }
    return ASEND_OK;
}

int asend_ldl_solve(asend_ctx* ctx, const double* rhs, double* x) {
    if (!ctx || !rhs || !x) return ASEND_ERR_ARG;
    int rc;
    if ((rc = ldl_ensure(ctx))) return rc;
    const int64_t n = ctx->n_dof;
    HIPCHK(hipMemcpyAsync(ctx->d_r, rhs, n * 8, hipMemcpyHostToDevice, ctx->st));
    ldl_solve_launch(ctx->st, ctx->ldl.d, ctx->ldl.L, ctx->d_r, ctx->d_ldl_z, ctx->ldl.work, nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(x, ctx->d_ldl_z, n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    return ASEND_OK;
}

int asend_ldl_release(asend_ctx* ctx) {
    if (!ctx) return ASEND_ERR_ARG;
    HIPCHK(hipStreamSynchronize(ctx->st));
    ldl_release(ctx);
    ctx->ldl_block_dofs = -1;
    return ASEND_OK;
}

int asend_solve_noalloc(asend_ctx* ctx, const double* rhs, double* x, const asend_solve_opts* opts,
                        asend_solve_stats* stats) {
    if (!ctx || !rhs || !x || !opts) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    if (opts->method != ASEND_SOLVE_PCG_FP64 || opts->precond != ASEND_PC_JACOBI)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "asend_solve_noalloc runs FP64 Jacobi PCG only");
    const int64_t n = ctx->n_dof;
    if (n > kNoAllocMaxDof)
        return fail(ctx, ASEND_ERR_ARG, "asend_solve_noalloc: %lld DOFs exceed its scratch limit of %lld DOFs",
                    (long long)n, (long long)kNoAllocMaxDof);
    double* base = scratch_ptr(ctx->device);
    if (!base) return fail(ctx, ASEND_ERR_STATE, "no scratch address for device %d", ctx->device);
    const Fp64Bufs B = scratch_fp64_bufs(base);
    HIPCHK(hipMemcpyAsync(B.f, rhs, n * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    asend_solve_stats s;
    int rc = solve_fp64(ctx, B, opts, &s);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    s.device_ms = ms;
    HIPCHK(hipMemcpyAsync(x, B.u, n * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (stats) *stats = s;
    return ASEND_OK;
}

int asend_element_sens(asend_ctx* ctx, const double* u, const double* lam, double* dJdE) {
    if (!ctx || !u || !lam || !dJdE) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_element_sens")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_element_sens")) return rc1;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype == ASEND_SHELL4)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element sensitivities are implemented for brick8 and brickIM meshes");
    const int64_t n = ctx->n_dof;
    HIPCHK(hipMemcpy(ctx->d_x64, u, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_lam, lam, n * 8, hipMemcpyHostToDevice));
    launch_sens(ctx, ctx->d_x64, ctx->d_lam);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->st));
    HIPCHK(hipMemcpy(dJdE, ctx->d_sens, ctx->n_el * 8, hipMemcpyDeviceToHost));
    return ASEND_OK;
}

int asend_element_sens_material(asend_ctx* ctx, const double* u, const double* lam, double* dJdp) {
    if (!ctx || !u || !lam || !dJdp) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_element_sens_material")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_element_sens_material")) return rc1;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "material-parameter sensitivities are implemented for brick8 meshes");
    const int64_t n = ctx->n_dof;
    double* d_out = nullptr;
    int rc = dalloc(ctx, &d_out, 9 * ctx->n_el);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->d_x64, u, n * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->d_lam, lam, n * 8, hipMemcpyHostToDevice, ctx->st));
    launch_hex8_sens_mat(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, ctx->d_x64,
                         ctx->d_lam, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(dJdp, d_out, 9 * ctx->n_el * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    dfree(d_out);
    return ASEND_OK;
}

int asend_element_sens_coords(asend_ctx* ctx, const double* u, const double* lam, double* out) {
    if (!ctx || !u || !lam || !out) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_element_sens_coords")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_element_sens_coords")) return rc1;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "shape sensitivities are implemented for brick8 meshes");
    const int64_t n = ctx->n_dof;
    double* d_out = nullptr;
    int rc = dalloc(ctx, &d_out, 24 * ctx->n_el);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->d_x64, u, n * 8, hipMemcpyHostToDevice, ctx->st));
    HIPCHK(hipMemcpyAsync(ctx->d_lam, lam, n * 8, hipMemcpyHostToDevice, ctx->st));
    launch_hex8_shape_sens(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, ctx->d_x64,
                           ctx->d_lam, d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out, 24 * ctx->n_el * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    dfree(d_out);
    return ASEND_OK;
}

int asend_element_centre_fields(asend_ctx* ctx, const double* u, double* out) {
    if (!ctx || !u || !out) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_element_centre_fields")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_element_centre_fields")) return rc1;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element centre fields are implemented for brick8 meshes");
    double* d_out = nullptr;
    int rc = dalloc(ctx, &d_out, 13 * ctx->n_el);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->d_x64, u, ctx->n_dof * 8, hipMemcpyHostToDevice, ctx->st));
    launch_hex8_centre_fields(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, ctx->d_x64,
                              d_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, d_out, 13 * ctx->n_el * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    dfree(d_out);
    return ASEND_OK;
}

int asend_element_centre_adjoint(asend_ctx* ctx, const double* h, double* y) {
    if (!ctx || !h || !y) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_element_centre_adjoint")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_element_centre_adjoint")) return rc1;
    if (!ctx->d_E) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element centre fields are implemented for brick8 meshes");
    double *d_h = nullptr, *d_re = nullptr;
    int rc;
    if ((rc = dalloc(ctx, &d_h, 6 * ctx->n_el)) || (rc = dalloc(ctx, &d_re, 24 * ctx->n_el))) {
        dfree(d_h);
        return rc;
    }
    HIPCHK(hipMemcpyAsync(d_h, h, 6 * ctx->n_el * 8, hipMemcpyHostToDevice, ctx->st));
    launch_hex8_centre_adjoint(ctx->st, (int)ctx->n_el, (int)ctx->n_nodes, ctx->d_conn_r, ctx->d_xyz_r,
                               ctx->d_n2e_ptr, ctx->d_n2rec, d_h, d_re, ctx->d_x64);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, ctx->d_x64, ctx->n_dof * 8, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    dfree(d_h);
    dfree(d_re);
    return ASEND_OK;
}

int asend_static_compliance(asend_ctx* ctx, const double* f, double* u, double* dJdE, double* compliance,
                            const asend_solve_opts* opts, asend_solve_stats* stats) {
    if (!ctx || !f || !opts) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_static_compliance")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_static_compliance")) return rc1;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype == ASEND_SHELL4)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element sensitivities are implemented for brick8 and brickIM meshes");
    const int64_t n = ctx->n_dof;
    HIPCHK(hipMemcpy(ctx->d_f, f, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    int rc = assemble_device(ctx, 1.0, 0.0);
    if (rc) return rc;
    asend_solve_stats s;
    if ((rc = solve_device(ctx, opts, &s))) return rc;
    // compliance J = f.u ; adjoint rhs dJ/du = f, so lambda = u (same operator)
    launch_dot64(ctx->st, n, ctx->d_f, ctx->d_u, ctx->d_partials, ctx->d_sc + 2);
    launch_sens(ctx, ctx->d_u, ctx->d_u);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    s.device_ms = ms;
    if ((rc = read_scalars(ctx, 2))) return rc;
    if (compliance) *compliance = ctx->h_sc[2].dot;
    if (u) HIPCHK(hipMemcpy(u, ctx->d_u, n * 8, hipMemcpyDeviceToHost));
    if (dJdE) HIPCHK(hipMemcpy(dJdE, ctx->d_sens, ctx->n_el * 8, hipMemcpyDeviceToHost));
    if (stats) *stats = s;
    return ASEND_OK;
}

int asend_nl_assemble(asend_ctx* ctx, const double* u, double* r_int) {
    if (!ctx || !u) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_nl_assemble")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_nl_assemble")) return rc1;
    int rc;
    if ((rc = nl_prepare(ctx))) return rc;
    HIPCHK(hipMemcpy(ctx->d_nl_u, u, ctx->n_dof * 8, hipMemcpyHostToDevice));
    if ((rc = nl_assemble(ctx, 0.0, nullptr))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    if (r_int) HIPCHK(hipMemcpy(r_int, ctx->d_nl_r, ctx->n_dof * 8, hipMemcpyDeviceToHost));
    return ASEND_OK;
}

int asend_brickim_internal(asend_ctx* ctx, const double* u, double* alpha) {
    if (!ctx || !alpha) return ASEND_ERR_ARG;
    if (ctx->etype != ASEND_BRICKIM) return fail(ctx, ASEND_ERR_STATE, "the mesh is not a brickIM mesh");
    if (const int rc0 = elastic_only(ctx, "asend_brickim_internal")) return rc0;
    int rc;
    if ((rc = nl_prepare(ctx))) return rc;
    const int64_t na = 9 * ctx->n_el;
    if (!u) {
        HIPCHK(hipMemcpy(alpha, ctx->d_nl_alpha, na * sizeof(double), hipMemcpyDeviceToHost));
        return ASEND_OK;
    }
    // the maps at rest (u = 0, amplitudes 0) give W = K_ii^-1 K_ie of the
    // linear element; then alpha = 0 - W u_e by the update kernel
    double* tmp = nullptr;
    if ((rc = dalloc(ctx, &tmp, na))) return rc;
    HIPCHK(hipMemsetAsync(tmp, 0, na * sizeof(double), ctx->st));
    HIPCHK(hipMemsetAsync(ctx->d_nl_u, 0, ctx->n_dof * sizeof(double), ctx->st));
    launch_brickim_nl(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_xyz_r, ctx->d_E, ctx->d_nu, ctx->d_nl_u, tmp,
                      ctx->d_kb, ctx->d_nl_re, ctx->d_nl_wmap);
    HIPCHK(hipMemcpyAsync(ctx->d_nl_r, u, ctx->n_dof * sizeof(double), hipMemcpyHostToDevice, ctx->st));
    launch_brickim_nl_internal(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_nl_r, ctx->d_nl_wmap, tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(alpha, tmp, na * sizeof(double), hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    dfree(tmp);
    return ASEND_OK;
}

int asend_brickim_set_internal(asend_ctx* ctx, const double* alpha) {
    if (!ctx || !alpha) return ASEND_ERR_ARG;
    if (ctx->etype != ASEND_BRICKIM) return fail(ctx, ASEND_ERR_STATE, "the mesh is not a brickIM mesh");
    if (const int rc0 = elastic_only(ctx, "asend_brickim_set_internal")) return rc0;
    int rc;
    if ((rc = nl_prepare(ctx))) return rc;
    HIPCHK(hipMemcpy(ctx->d_nl_alpha, alpha, 9 * ctx->n_el * sizeof(double), hipMemcpyHostToDevice));
    return ASEND_OK;
}

int asend_static_nonlinear(asend_ctx* ctx, const double* f, double* u, const asend_nl_opts* nl,
                           const asend_solve_opts* opts, asend_nl_stats* stats) {
    if (!ctx || !f || !u || !opts) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_static_nonlinear")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_static_nonlinear")) return rc1;
    int rc;
    if ((rc = nl_prepare(ctx))) return rc;
    asend_nl_opts o;
    o.load_steps = 1;
    o.max_newton = 50;
    o.du_tol = 1e-12;
    o.forcing = 0.0;
    if (nl) {
        if (nl->load_steps > 0) o.load_steps = nl->load_steps;
        if (nl->max_newton > 0) o.max_newton = nl->max_newton;
        if (nl->du_tol > 0) o.du_tol = nl->du_tol;
        // a forcing term is a relative tolerance: it applies to the relative stop rule only
        if (nl->forcing > 0 && opts->stop_rule == ASEND_STOP_REL_RESIDUAL) o.forcing = std::min(nl->forcing, 0.5);
    }
    asend_solve_opts lo = *opts;  // per-iteration tolerance under a forcing term
    struct FineOnly {
        asend_ctx* c;
        ~FineOnly() { c->mg_fine_only = false; }
    } fine_only{ctx};
    const int64_t n = ctx->n_dof;
    HIPCHK(hipMemcpy(ctx->d_nl_u, u, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_nl_f, f, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    asend_nl_stats st;
    std::memset(&st, 0, sizeof(st));
    st.load_steps = o.load_steps;
    const int L = o.load_steps;
    for (int s = 0; s < L; ++s) {
        // load ramp of a static nonlinear solve (S/model_analysis.rs:636-643)
        const double q = (double)(L - 1 - s);
        const double fact = 1.0 - q * q / ((double)L * (double)L);
        int it = 0;
        bool done = false;
        double eta = o.forcing, load_prev = 0.0;
        while (it < o.max_newton && !done) {
            if ((rc = nl_assemble(ctx, fact, ctx->d_f))) return rc;
            if (o.forcing > 0) {
                // Eisenstat-Walker choice 2 (gamma 0.9, exponent 2) with its
                // safeguard; never tighter than opts->tol
                launch_dot64(ctx->st, n, ctx->d_f, ctx->d_f, ctx->d_partials, ctx->d_sc + 2);
                HIPCHK(hipGetLastError());
                if ((rc = read_scalars(ctx, 2))) return rc;
                const double load = std::sqrt(ctx->h_sc[2].dot);
                if (it > 0 && load_prev > 0) {
                    const double ratio = load / load_prev;
                    const double e_new = 0.9 * ratio * ratio, e_safe = 0.9 * eta * eta;
                    eta = std::min(o.forcing, e_safe > 0.1 ? std::max(e_new, e_safe) : e_new);
                }
                load_prev = load;
                lo.tol = std::max(eta, opts->tol);
            }
            asend_solve_stats ls;
            if ((rc = solve_device(ctx, &lo, &ls))) return rc;
            ctx->mg_fine_only = true;
            st.linear_iterations += ls.iterations;
            if (!ls.converged) st.unconverged_solves += 1;
            if (ctx->etype == ASEND_BRICKIM)  // bubble amplitudes follow the increment
                launch_brickim_nl_internal(ctx->st, (int)ctx->n_el, ctx->d_conn_r, ctx->d_u, ctx->d_nl_wmap,
                                           ctx->d_nl_alpha);
            launch_axpy64(ctx->st, n, 1.0, ctx->d_u, ctx->d_nl_u);
            launch_maxabs(ctx->st, ctx->d_u, n, ctx->d_partials, ctx->d_sc + 2);
            HIPCHK(hipGetLastError());
            if ((rc = read_scalars(ctx, 2))) return rc;
            st.last_du = ctx->h_sc[2].maxabs;
            ++it;
            done = st.last_du <= o.du_tol;
        }
        st.newton_iterations += it;
        st.max_step_newton = std::max(st.max_step_newton, it);
        if (!done) st.unconverged_steps += 1;
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    st.device_ms = ms;
    HIPCHK(hipMemcpy(u, ctx->d_nl_u, n * 8, hipMemcpyDeviceToHost));
    if (stats) *stats = st;
    return ASEND_OK;
}

int asend_bench_spmv(asend_ctx* ctx, int precision, int n_iter, double* avg_ms) {
    if (!ctx || n_iter <= 0 || !avg_ms) return ASEND_ERR_ARG;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    // warm-up launch, then n_iter timed launches between two events
    for (int pass = 0; pass < 2; ++pass) {
        int cnt = pass == 0 ? 1 : n_iter;
        if (pass == 1) HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
        for (int i = 0; i < cnt; ++i) {
            if (precision == 64)
                launch_sell3<double>(ctx->st, 0, ctx->fine(), ctx->d_K64, ctx->d_u, ctx->d_q, nullptr, nullptr,
                                     nullptr);
            else
                launch_sell3<float>(ctx->st, 0, ctx->fine(), ctx->d_K32, ctx->d_x32, ctx->d_q32, nullptr, nullptr,
                                    nullptr);
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *avg_ms = ms / n_iter;
    return ASEND_OK;
}

// ---------------------------------------------------------------------------
// Batched multi-RHS solves over variants
// ---------------------------------------------------------------------------
}  // extern "C"

struct asend_batch {
    Batch* b = nullptr;
    std::vector<int64_t> n_dof;
    std::string err;
};

extern "C" {

int asend_batch_create(asend_ctx* const* variants, int n_variants, asend_batch** out) {
    if (!variants || n_variants < 1 || !out) return ASEND_ERR_ARG;
    *out = nullptr;
    for (int v = 0; v < n_variants; ++v)
        if (variants[v] && variants[v]->d_kcond)
            return fail(variants[v], ASEND_ERR_UNSUPPORTED, "batched solves need elastic contexts");
    std::vector<OperatorView> ops(n_variants);
    asend_batch* ab = new asend_batch();
    for (int v = 0; v < n_variants; ++v) {
        asend_ctx* c = variants[v];
        if (!c || !c->assembled || (c->etype != ASEND_BRICK8 && c->etype != ASEND_BRICKIM)) {
            delete ab;
            return ASEND_ERR_STATE;
        }
        OperatorView& o = ops[v];
        o.n_rows = (int)c->n_nodes;
        o.n_slices = (int)c->n_slices;
        o.n_slots = c->n_slots;
        o.h_sptr = c->h_sptr.data();
        o.scol = c->d_scol;
        o.K32 = c->d_K32;
        o.K64 = c->d_K64;
        o.dinv32 = c->d_dinv32;
        ab->n_dof.push_back(c->n_dof);
        HIPCHK_CTX(hipStreamSynchronize(c->st));
    }
    int rc = batch_create(&ab->b, ops);
    if (rc) {
        ab->err = batch_error(ab->b);
        batch_destroy(ab->b);
        ab->b = nullptr;
    }
    *out = ab;
    return rc;
}

void asend_batch_destroy(asend_batch* ab) {
    if (!ab) return;
    batch_destroy(ab->b);
    delete ab;
}

const char* asend_batch_error(const asend_batch* ab) { return ab ? ab->err.c_str() : "null batch"; }

int asend_batch_solve(asend_batch* ab, int nrhs, const double* const* rhs, double* const* x, double tol,
                      double inner_tol, int64_t max_it, int check_every, int verbose, asend_batch_stats* stats) {
    if (!ab || !ab->b || !rhs || !x || nrhs < 1 || nrhs > kNrhs) return ASEND_ERR_ARG;
    Batch* b = ab->b;
    const int nv = (int)ab->n_dof.size();
    const int64_t ne = batch_rows_total(b) * 3 * kNrhs;
    std::vector<double> F(ne, 0.0), U(ne, 0.0);
    for (int v = 0; v < nv; ++v) {
        const int64_t nd = ab->n_dof[v], off = batch_row_offset(b, v) * 3;
        for (int j = 0; j < nrhs; ++j)
            for (int64_t d = 0; d < nd; ++d) F[(off + d) * kNrhs + j] = rhs[v][j * nd + d];
    }
    BatchStats bs;
    int rc = batch_solve(b, F.data(), U.data(), nrhs, tol, inner_tol, max_it, check_every, verbose, &bs);
    if (rc) {
        ab->err = batch_error(b);
        return rc;
    }
    for (int v = 0; v < nv; ++v) {
        const int64_t nd = ab->n_dof[v], off = batch_row_offset(b, v) * 3;
        for (int j = 0; j < nrhs; ++j)
            for (int64_t d = 0; d < nd; ++d) x[v][j * nd + d] = U[(off + d) * kNrhs + j];
    }
    if (stats) {
        stats->outer_iterations = bs.outer_iterations;
        stats->lockstep_iterations = bs.lockstep_iterations;
        stats->system_iterations = bs.system_iterations;
        stats->unconverged = bs.unconverged;
        stats->max_rel_residual = bs.max_rel_residual;
        stats->device_ms = bs.device_ms;
    }
    return ASEND_OK;
}

// ---------------------------------------------------------------------------
// Newmark dynamics
// ---------------------------------------------------------------------------
int asend_set_material_damping(asend_ctx* ctx, int64_t n_mat, const double* dmat, const int32_t* el_mat) {
    if (!ctx || n_mat < 0 || (n_mat > 0 && !dmat)) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (n_mat == 0) {
        dfree(ctx->d_dmat);
        dfree(ctx->d_dmat_idx);
        ctx->n_dmat = 0;
        return ASEND_OK;
    }
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "material damping is implemented for brick8 meshes");
    for (int64_t m = 0; m < n_mat; ++m)
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < i; ++j)
                if (dmat[36 * m + 6 * i + j] != dmat[36 * m + 6 * j + i])
                    return fail(ctx, ASEND_ERR_ARG, "damping matrix %lld is not symmetric", (long long)m);
    if (el_mat)
        for (int64_t e = 0; e < ctx->n_el; ++e)
            if (el_mat[e] < 0 || el_mat[e] >= n_mat)
                return fail(ctx, ASEND_ERR_ARG, "element %lld: damping matrix index %d out of range", (long long)e,
                            el_mat[e]);
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_dmat, 36 * n_mat))) return rc;
    HIPCHK(hipMemcpy(ctx->d_dmat, dmat, 36 * n_mat * sizeof(double), hipMemcpyHostToDevice));
    if (el_mat) {
        if ((rc = dalloc(ctx, &ctx->d_dmat_idx, ctx->n_el))) return rc;
        HIPCHK(hipMemcpy(ctx->d_dmat_idx, el_mat, ctx->n_el * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
        dfree(ctx->d_dmat_idx);
    }
    ctx->n_dmat = n_mat;
    return ASEND_OK;
}

int asend_set_material_stiffness(asend_ctx* ctx, int64_t n_mat, const double* cmat, const int32_t* el_mat) {
    if (!ctx || n_mat < 0 || (n_mat > 0 && !cmat)) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    ctx->assembled = false;
    if (n_mat == 0) {
        dfree(ctx->d_smat);
        dfree(ctx->d_smat_idx);
        ctx->n_smat = 0;
        return ASEND_OK;
    }
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "general material matrices are implemented for brick8 meshes");
    for (int64_t m = 0; m < n_mat; ++m) {
        // symmetric to rounding (an inverted compliance is symmetric to a few ulps)
        double amax = 0.0;
        for (int i = 0; i < 36; ++i) amax = std::max(amax, std::fabs(cmat[36 * m + i]));
        for (int i = 0; i < 6; ++i) {
            if (!std::isfinite(cmat[36 * m + 7 * i]) || !(cmat[36 * m + 7 * i] > 0.0))
                return fail(ctx, ASEND_ERR_ARG, "material matrix %lld: diagonal entry %d is not positive",
                            (long long)m, i);
            for (int j = 0; j < i; ++j)
                if (!(std::fabs(cmat[36 * m + 6 * i + j] - cmat[36 * m + 6 * j + i]) <= 1e-12 * amax))
                    return fail(ctx, ASEND_ERR_ARG, "material matrix %lld is not symmetric", (long long)m);
        }
    }
    if (el_mat)
        for (int64_t e = 0; e < ctx->n_el; ++e)
            if (el_mat[e] < 0 || el_mat[e] >= n_mat)
                return fail(ctx, ASEND_ERR_ARG, "element %lld: material matrix index %d out of range", (long long)e,
                            el_mat[e]);
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_smat, 36 * n_mat))) return rc;
    HIPCHK(hipMemcpy(ctx->d_smat, cmat, 36 * n_mat * sizeof(double), hipMemcpyHostToDevice));
    if (el_mat) {
        if ((rc = dalloc(ctx, &ctx->d_smat_idx, ctx->n_el))) return rc;
        HIPCHK(hipMemcpy(ctx->d_smat_idx, el_mat, ctx->n_el * sizeof(int32_t), hipMemcpyHostToDevice));
    } else {
        dfree(ctx->d_smat_idx);
    }
    ctx->n_smat = n_mat;
    return ASEND_OK;
}

int asend_set_conductivity(asend_ctx* ctx, const double* kcond) {
    if (!ctx) return ASEND_ERR_ARG;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    // the operator changes kind: the penalty scale is taken again from the
    // next assembly (each field has its own constraint list and scale)
    ctx->sf_fixed = false;
    ctx->assembled = false;
    if (!kcond) {
        dfree(ctx->d_kcond);
        return ASEND_OK;
    }
    if (ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "heat conduction is implemented for brick8 meshes");
    for (int64_t i = 0; i < 6 * ctx->n_el; ++i)
        if (!std::isfinite(kcond[i]))
            return fail(ctx, ASEND_ERR_ARG, "element %lld: conductivity is not finite", (long long)(i / 6));
    int rc;
    if ((rc = dalloc(ctx, &ctx->d_kcond, 6 * ctx->n_el))) return rc;
    HIPCHK(hipMemcpy(ctx->d_kcond, kcond, 6 * ctx->n_el * sizeof(double), hipMemcpyHostToDevice));
    return ASEND_OK;
}

int asend_set_rayleigh_damping(asend_ctx* ctx, double cm, double ck) {
    if (!ctx || !(cm >= 0.0) || !(ck >= 0.0) || !std::isfinite(cm) || !std::isfinite(ck)) return ASEND_ERR_ARG;
    ctx->ray_cm = cm;
    ctx->ray_ck = ck;
    return ASEND_OK;
}

int asend_newmark_set_solver(asend_ctx* ctx, int mode, double ru_factor) {
    if (!ctx || (mode != ASEND_NM_MIXED && mode != ASEND_NM_FP64 && mode != ASEND_NM_AUTO) ||
        (mode != ASEND_NM_FP64 && !(ru_factor >= 1e-6 && ru_factor < 1.0)))
        return ASEND_ERR_ARG;
    ctx->nm_mode = mode;
    if (mode != ASEND_NM_FP64) ctx->nm_inner_tol = ru_factor;
    ctx->nm.mixed = mode;
    ctx->nm.inner_tol = ctx->nm_inner_tol;
    return ASEND_OK;
}

int asend_newmark_setup(asend_ctx* ctx, double dt, double beta, double gamma) {
    if (!ctx || !(dt > 0.0)) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_newmark_setup")) return rc0;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (!(gamma > beta)) return fail(ctx, ASEND_ERR_ARG, "Newmark needs gamma > beta for an SPD effective operator");
    HIPCHK(hipSetDevice(ctx->device));
    newmark_free(ctx);
    auto& nm = ctx->nm;
    nm.mixed = ctx->nm_mode;
    nm.inner_tol = ctx->nm_inner_tol;
    nm.dt = dt;
    nm.beta = beta;
    nm.gamma = gamma;
    nm.c_m = 1.0 / (dt * dt * (gamma - beta));  // = -c1 of update_vel_acc
    nm.c2 = dt * dt * (0.5 + beta - gamma);
    nm.cm = ctx->ray_cm;
    nm.ck = ctx->ray_ck;
    nm.cd = nm.c_m * dt * gamma;
    nm.mat = ctx->n_dmat > 0;
    nm.keff_d = nm.mat ? nm.cd : 0.0;
    // The reference adds stiffness-proportional damping to the external block
    // of elements with condensed internal DOFs before condensing
    // (S/element_equations.rs:991-1003); that form is not built here.
    if (nm.ck != 0.0 && ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED,
                    "stiffness-proportional damping is implemented for brick8 meshes (mass-proportional for all)");
    const int64_t n = ctx->n_dof, nval = ctx->nval();
    int rc;
    if ((rc = dalloc(ctx, &nm.M64, nval)) || (rc = dalloc(ctx, &nm.v, n)) || (rc = dalloc(ctx, &nm.a, n)) ||
        (rc = dalloc(ctx, &nm.w, n)) || (rc = dalloc(ctx, &nm.q, n)) || (rc = dalloc(ctx, &nm.rhs0, n)) ||
        (rc = dalloc(ctx, &nm.rhs, n)) || (rc = dalloc(ctx, &nm.fshape, n)) || (rc = dalloc(ctx, &nm.step, 1)) ||
        (rc = dalloc(ctx, &nm.cg_u, n)) || (rc = dalloc(ctx, &nm.cg_s, n)) || (rc = dalloc(ctx, &nm.f32, 5 * n)))
        return rc;
    // consistent mass alone: the element kernels with c_mass = 1 and no
    // stiffness (zero moduli for the brick8 kernels, k_scale = 0 for the
    // kernels that condense internal DOFs, whose stiffness must stay regular)
    if (ctx->etype == ASEND_SHELL4 || ctx->etype == ASEND_BRICKIM) {
        if (ctx->etype == ASEND_SHELL4 && !ctx->thk_set)
            return fail(ctx, ASEND_ERR_STATE, "shell4 mesh without thickness (asend_set_thickness)");
        if ((rc = integrate_elements(ctx, ctx->d_E, 0.0, 1.0, nm.M64))) return rc;
        HIPCHK(hipStreamSynchronize(ctx->st));
    } else {
        double* zeroE = nullptr;
        if ((rc = dalloc(ctx, &zeroE, ctx->n_el))) return rc;
        HIPCHK(hipMemsetAsync(zeroE, 0, ctx->n_el * 8, ctx->st));
        rc = integrate_elements(ctx, zeroE, 1.0, 1.0, nm.M64);
        HIPCHK(hipStreamSynchronize(ctx->st));
        dfree(zeroE);
        if (rc) return rc;
    }
    // D = cm M + ck K (+ the material damping matrices), the damped load's
    // work vector
    if (nm.damped()) {
        if ((rc = dalloc(ctx, &nm.D64, nval)) || (rc = dalloc(ctx, &nm.rhsd, n))) return rc;
        if ((rc = integrate_elements(ctx, ctx->d_E, nm.ck, nm.cm, nm.D64, nm.mat ? 1.0 : 0.0))) return rc;
    }
    // K_eff = K + c_m M + cd D (+ penalty) = (1 + cd ck) K + (c_m + cd cm) M.
    // The reference scales its penalty once, on the first matrix a solve
    // step assembles, which for a dynamic job already holds the mass and
    // damping terms (S/model_analysis.rs:255-259, :486-488)
    nm.keff_k = 1.0 + nm.cd * nm.ck;
    nm.keff_m = nm.c_m + nm.cd * nm.cm;
    ctx->sf_fixed = false;
    if ((rc = assemble_device(ctx, nm.keff_k, nm.keff_m, nm.keff_d))) return rc;
    HIPCHK(hipStreamSynchronize(ctx->st));
    nm.ready = true;
    return ASEND_OK;
}

int asend_newmark_run(asend_ctx* ctx, int64_t n_steps, const double* f_shape, const double* load_fac,
                      const double* u0, const double* v0, const double* a0, int64_t n_watch, const int64_t* watch,
                      double* hist, double* u_out, double* v_out, double* a_out, double tol, int max_it_step,
                      asend_newmark_stats* stats) {
    if (!ctx || n_steps < 0 || !f_shape || !load_fac || n_watch < 0 || (n_watch > 0 && (!watch || !hist)) ||
        !(tol > 0.0) || max_it_step < 1)
        return ASEND_ERR_ARG;
    auto& nm = ctx->nm;
    if (!nm.ready) return fail(ctx, ASEND_ERR_STATE, "call asend_newmark_setup first");
    const int64_t n = ctx->n_dof;
    for (int64_t k = 0; k < n_watch; ++k)
        if (watch[k] < 0 || watch[k] >= n) return fail(ctx, ASEND_ERR_ARG, "watched DOF out of range");
    int rc;
    // the operator must be K_eff of this setup (a static assembly may have replaced it)
    if (!ctx->assembled || ctx->asm_cmass != nm.keff_m || ctx->asm_kscale != nm.keff_k ||
        ctx->asm_cdamp != nm.keff_d) {
        if ((rc = assemble_device(ctx, nm.keff_k, nm.keff_m, nm.keff_d))) return rc;
    }
    bool realloc = false;
    if (n_steps > nm.fac_cap) {
        if ((rc = dalloc(ctx, &nm.fac, n_steps)) || (rc = dalloc(ctx, &nm.its, n_steps))) return rc;
        nm.fac_cap = n_steps;
        realloc = true;
    }
    if (n_steps * std::max<int64_t>(n_watch, 1) > nm.hist_cap) {
        if ((rc = dalloc(ctx, &nm.hist, n_steps * std::max<int64_t>(n_watch, 1)))) return rc;
        nm.hist_cap = n_steps * std::max<int64_t>(n_watch, 1);
        realloc = true;
    }
    if (std::max<int64_t>(n_watch, 1) > nm.watch_cap) {
        if ((rc = dalloc(ctx, &nm.watch, std::max<int64_t>(n_watch, 1)))) return rc;
        nm.watch_cap = std::max<int64_t>(n_watch, 1);
        realloc = true;
    }
    if (realloc) nm.graph_iters = -1;  // recapture: the step graphs hold the old buffers
    // initial state, ordered on the context stream after the assembly queued
    // there (outside any captured region)
    hipStream_t s0 = ctx->st;
    HIPCHK(hipMemcpyAsync(nm.fshape, f_shape, n * 8, hipMemcpyHostToDevice, s0));
    if (n_steps) HIPCHK(hipMemcpyAsync(nm.fac, load_fac, n_steps * 8, hipMemcpyHostToDevice, s0));
    if (n_watch) HIPCHK(hipMemcpyAsync(nm.watch, watch, n_watch * 8, hipMemcpyHostToDevice, s0));
    if (u0) HIPCHK(hipMemcpyAsync(ctx->d_u, u0, n * 8, hipMemcpyHostToDevice, s0));
    else HIPCHK(hipMemsetAsync(ctx->d_u, 0, n * 8, s0));
    if (v0) HIPCHK(hipMemcpyAsync(nm.v, v0, n * 8, hipMemcpyHostToDevice, s0));
    else HIPCHK(hipMemsetAsync(nm.v, 0, n * 8, s0));
    if (a0) HIPCHK(hipMemcpyAsync(nm.a, a0, n * 8, hipMemcpyHostToDevice, s0));
    else HIPCHK(hipMemsetAsync(nm.a, 0, n * 8, s0));
    HIPCHK(hipMemsetAsync(nm.step, 0, 8, s0));
    HIPCHK(hipMemsetAsync(ctx->d_x64, 0, n * 8, s0));  // increment of the step before the first
    // the host arrays may be pageable and are the caller's once this returns
    HIPCHK(hipStreamSynchronize(s0));
    // per-step solve: FP64 Jacobi PCG on K_eff, relative tolerance, at most
    // max_it_step iterations (the kernels skip their work once converged)
    {
        Scalars& h = ctx->h_sc[1];
        std::memset(&h, 0, sizeof(Scalars));
        h.tol = tol;
        h.tol2 = tol * tol;
        h.max_it = max_it_step;
        if ((rc = write_scalars(ctx, 1))) return rc;
    }
    Scalars* sc = ctx->d_sc + 1;
    const Sell3 A = ctx->fine();
    MixedSolve mx;
    mx.A = &A;
    mx.K64 = ctx->d_K64;
    mx.K32 = ctx->d_K32;
    mx.dinv32 = ctx->d_dinv32;
    mx.rhs = nm.rhs;
    mx.x64 = ctx->d_x64;
    mx.r64 = ctx->d_r2;
    float* f = nm.f32;
    mx.r32 = f;
    mx.u32 = f + n;
    mx.w32 = f + 2 * n;
    mx.p32 = f + 3 * n;
    mx.s32 = f + 4 * n;
    mx.partials = ctx->d_partials;
    mx.sc = sc;
    mx.n = n;
    mx.ru_factor = nm.inner_tol;
    const int nw = (int)n_watch;
    // A step is three launch sequences: head (effective load, PCG start and a
    // first block of iterations), more (another block, repeated until the
    // step's solve has converged: one small device-to-host read per block),
    // tail (state update and history).  Each is captured once as a hipGraph.
    const int blk = std::min(max_it_step, kNewmarkBlock);
    // single-reduction Jacobi PCG: x = du, r, u = Dinv r, w = K u, p, s = K p;
    // mixed: the same recurrence in FP32 with x in FP64, and a reliable-update
    // slot after every kMixedRefreshEvery iterations (a no-op unless a refresh
    // is pending; the iterations pause while one is)
    const bool mixed = nm.mixed == ASEND_NM_MIXED || (nm.mixed == ASEND_NM_AUTO && 8 * ctx->nval() > kMixedAutoBytes);
    auto iterations = [&]() {
        for (int k = 0; k < blk; ++k) {
            if (mixed) {
                launch_mixed_iteration(ctx->st, mx);
                if ((k + 1) % kMixedRefreshEvery == 0 || k + 1 == blk) launch_mixed_refresh(ctx->st, mx);
            } else {
                launch_sell3<double>(ctx->st, 4, A, ctx->d_K64, nm.cg_u, ctx->d_q, ctx->d_r2, ctx->d_partials, sc);
                launch_cg1_update<double>(ctx->st, n, ctx->d_dinv64, ctx->d_x64, ctx->d_r2, nm.cg_u, ctx->d_q,
                                          ctx->d_p, nm.cg_s, sc);
            }
        }
    };
    auto head = [&]() {
        launch_nm_predict(ctx->st, n, ctx->d_u, nm.v, nm.a, nm.dt, nm.c2, nm.w);
        launch_sell3<double>(ctx->st, 0, ctx->fine(), nm.M64, nm.w, nm.q, nullptr, nullptr, nullptr);
        launch_nm_rhs(ctx->st, n, nm.step, nm.fac, nm.fshape, nm.c_m, nm.q, nm.rhs0);
        const double* load = nm.rhs0;
        if (nm.damped()) {  // rhsd = rhs0 - D z (q is free again)
            launch_nm_damp_z(ctx->st, n, nm.v, nm.a, nm.w, nm.dt, nm.gamma, nm.cd, nm.q);
            launch_sell3<double>(ctx->st, 3, ctx->fine(), nm.D64, nm.q, nm.rhsd, nm.rhs0, nullptr, nullptr);
            load = nm.rhsd;
        }
        launch_sell3<double>(ctx->st, 3, ctx->fine(), ctx->d_K64, ctx->d_u, nm.rhs, load, nullptr, nullptr);
        // warm start from the previous step's increment (zero before the first
        // step); the tolerance stays relative to |rhs|
        if (mixed) {
            launch_mixed_start(ctx->st, mx);
        } else {
            launch_sell3<double>(ctx->st, 3, A, ctx->d_K64, ctx->d_x64, ctx->d_r2, nm.rhs, nullptr, nullptr);
            launch_cg1_init<double>(ctx->st, n, nm.rhs, false, ctx->d_dinv64, ctx->d_x64, ctx->d_r2, nm.cg_u,
                                    ctx->d_p, nm.cg_s, ctx->d_partials, sc);
        }
        iterations();
    };
    auto tail = [&]() {
        launch_nm_update(ctx->st, n, ctx->d_x64, nm.w, nm.c_m, nm.dt, nm.gamma, ctx->d_u, nm.v, nm.a);
        if (nm.save_u) launch_nm_save(ctx->st, n, nm.step, ctx->d_u, nm.save_u);
        if (nm.save_y) launch_nm_save_y(ctx->st, n, nm.step, ctx->d_u, nm.v, nm.ck, nm.save_y);
        launch_nm_record(ctx->st, nw, nm.watch, ctx->d_u, nm.hist, nm.step, nm.its, sc);
    };
    auto capture = [&](hipGraphExec_t* exec, const std::function<void()>& seq) -> int {
        if (*exec) (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
        seq();
        hipError_t e1 = hipGetLastError();
        hipError_t e2 = hipStreamEndCapture(ctx->st, &g);
        if (e1 != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(ctx, ASEND_ERR_HIP, "capturing a Newmark step graph failed");
        }
        hipError_t e3 = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e3 != hipSuccess) {
            *exec = nullptr;
            return fail(ctx, ASEND_ERR_HIP, "instantiating a Newmark step graph failed");
        }
        return ASEND_OK;
    };
    const bool graph = ctx->use_graphs;
    if (graph && (!nm.exec_head || nm.graph_iters != blk || nm.graph_watch != nw || nm.graph_save != nm.save_u ||
                  nm.graph_save_y != nm.save_y || nm.graph_mixed != (int)mixed || nm.graph_ru != mx.ru_factor ||
                  nm.graph_wpb != A.waves_per_block)) {
        if ((rc = capture(&nm.exec_head, head)) || (rc = capture(&nm.exec_more, iterations)) ||
            (rc = capture(&nm.exec_tail, tail)))
            return rc;
        nm.graph_iters = blk;
        nm.graph_watch = nw;
        nm.graph_save = nm.save_u;
        nm.graph_save_y = nm.save_y;
        nm.graph_mixed = (int)mixed;
        nm.graph_ru = mx.ru_factor;
        nm.graph_wpb = A.waves_per_block;
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    for (int64_t s = 0; s < n_steps; ++s) {
        if (graph) HIPCHK(hipGraphLaunch(nm.exec_head, ctx->st));
        else head();
        for (;;) {
            if ((rc = read_scalars(ctx, 1))) return rc;
            if (ctx->h_sc[1].done) break;
            if (graph) HIPCHK(hipGraphLaunch(nm.exec_more, ctx->st));
            else iterations();
        }
        if (graph) HIPCHK(hipGraphLaunch(nm.exec_tail, ctx->st));
        else tail();
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::vector<int64_t> its(std::max<int64_t>(n_steps, 1));
    if (n_steps) HIPCHK(hipMemcpy(its.data(), nm.its, n_steps * 8, hipMemcpyDeviceToHost));
    if (n_watch && n_steps) HIPCHK(hipMemcpy(hist, nm.hist, n_steps * n_watch * 8, hipMemcpyDeviceToHost));
    if (u_out) HIPCHK(hipMemcpy(u_out, ctx->d_u, n * 8, hipMemcpyDeviceToHost));
    if (v_out) HIPCHK(hipMemcpy(v_out, nm.v, n * 8, hipMemcpyDeviceToHost));
    if (a_out) HIPCHK(hipMemcpy(a_out, nm.a, n * 8, hipMemcpyDeviceToHost));
    if (stats) {
        asend_newmark_stats st;
        std::memset(&st, 0, sizeof(st));
        st.steps = n_steps;
        for (int64_t s = 0; s < n_steps; ++s) {
            const int64_t it = its[s] >= 0 ? its[s] : -its[s] - 1;
            st.total_iterations += it;
            st.max_step_iterations = std::max(st.max_step_iterations, it);
            if (its[s] < 0) st.unconverged_steps += 1;
        }
        st.graph_used = graph ? 1 : 0;
        st.device_ms = ms;
        st.step_solver = mixed ? ASEND_NM_MIXED : ASEND_NM_FP64;
        *stats = st;
    }
    return ASEND_OK;
}

int asend_newmark_adjoint(asend_ctx* ctx, int64_t n_steps, const double* f_shape, const double* load_fac,
                          const double* q, const double* w, double tol, int max_it_step, double* J, double* dJdE,
                          asend_newmark_stats* fwd_stats, asend_newmark_stats* adj_stats) {
    if (!ctx || n_steps < 1 || !f_shape || !load_fac || !q || !w || !(tol > 0.0) || max_it_step < 1)
        return ASEND_ERR_ARG;
    if (const int rc1 = iso_only(ctx, "asend_newmark_adjoint")) return rc1;
    auto& nm = ctx->nm;
    if (!nm.ready) return fail(ctx, ASEND_ERR_STATE, "call asend_newmark_setup first");
    if (ctx->etype == ASEND_SHELL4)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "element sensitivities are implemented for brick8 and brickIM meshes");
    const int64_t n = ctx->n_dof;
    int rc;
    if (n_steps * n > nm.uhist_cap) {
        if ((rc = dalloc(ctx, &nm.uhist, n_steps * n))) return rc;
        nm.uhist_cap = n_steps * n;
    }
    // under stiffness-proportional damping dR/dE acts on u + ck v
    if (nm.ck != 0.0 && n_steps * n > nm.yhist_cap) {
        if ((rc = dalloc(ctx, &nm.yhist, n_steps * n))) return rc;
        nm.yhist_cap = n_steps * n;
    }
    if (!nm.lam) {
        if ((rc = dalloc(ctx, &nm.lam, n)) || (rc = dalloc(ctx, &nm.aadj, n)) || (rc = dalloc(ctx, &nm.vadj, n)) ||
            (rc = dalloc(ctx, &nm.mlam, n)) || (rc = dalloc(ctx, &nm.qv, n)) || (rc = dalloc(ctx, &nm.sens_acc, ctx->n_el)))
            return rc;
    }
    // forward sweep from rest, recording u after every step
    nm.save_u = nm.uhist;
    nm.save_y = nm.ck != 0.0 ? nm.yhist : nullptr;
    rc = asend_newmark_run(ctx, n_steps, f_shape, load_fac, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr,
                           nullptr, nullptr, tol, max_it_step, fwd_stats);
    nm.save_u = nullptr;
    nm.save_y = nullptr;
    if (rc) return rc;
    // backward sweep (Model::get_obj_gradient, dynamic branch,
    // S/model_analysis.rs:1563-1631): one adjoint solve with K_eff per step
    const double c1 = -nm.c_m;  // 1 / (dt^2 (beta - gamma))
    // velocity of step n+1 with respect to (u, a, v) of step n (d_vd_up,
    // d_vd_ap, d_vd_vp of augmentd_ld_u, S/model_analysis.rs:953-955)
    const double dvu = nm.dt * nm.gamma * c1;
    const double dva = nm.dt * (1.0 - nm.gamma) + nm.dt * nm.gamma * c1 * nm.c2;
    const double dvv = 1.0 + nm.dt * nm.gamma * c1 * nm.dt;
    double* dlam = nm.damped() ? nm.rhsd : nullptr;  // D lam_{n+1}
    HIPCHK(hipMemcpy(nm.qv, q, n * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(nm.aadj, 0, n * 8, ctx->st));
    HIPCHK(hipMemsetAsync(nm.vadj, 0, n * 8, ctx->st));
    HIPCHK(hipMemsetAsync(nm.sens_acc, 0, ctx->n_el * 8, ctx->st));
    const int check = ctx->tuned_check > 0 ? ctx->tuned_check : 16;
    double Jsum = 0.0;
    asend_newmark_stats st;
    std::memset(&st, 0, sizeof(st));
    st.steps = n_steps;
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    for (int64_t k = n_steps - 1; k >= 0; --k) {
        const double* uk = nm.uhist + k * n;
        const double* yk = nm.ck != 0.0 ? nm.yhist + k * n : uk;
        const int last = (k == n_steps - 1) ? 1 : 0;
        if (!last) {
            launch_sell3<double>(ctx->st, 0, ctx->fine(), nm.M64, nm.lam, nm.mlam, nullptr, nullptr, nullptr);
            if (dlam) launch_sell3<double>(ctx->st, 0, ctx->fine(), nm.D64, nm.lam, dlam, nullptr, nullptr, nullptr);
        }
        launch_adj_prep(ctx->st, n, nm.qv, w[k], c1, nm.c2, nm.dt, nm.gamma, last, nm.mlam, dlam, dvu, dva, dvv,
                        nm.aadj, nm.vadj, nm.rhs);
        HIPCHK(hipGetLastError());
        if ((rc = run_cg1_pcg(ctx, 1, nm.rhs, nm.lam, !last, tol, max_it_step, check, nm.cg_u, nm.cg_s))) return rc;
        const Scalars& h = ctx->h_sc[1];
        st.total_iterations += h.it;
        st.max_step_iterations = std::max<int64_t>(st.max_step_iterations, h.it);
        if (!(h.rr <= h.tol2 * h.ff)) st.unconverged_steps += 1;
        // dJ/dE += -lam_k^T (dK/dE) (u_k + ck v_k) ; J += w_k q.u_k
        launch_sens(ctx, yk, nm.lam);
        launch_axpy64(ctx->st, ctx->n_el, 1.0, ctx->d_sens, nm.sens_acc);
        launch_dot64(ctx->st, n, nm.qv, uk, ctx->d_partials, ctx->d_sc + 2);
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars(ctx, 2))) return rc;
        Jsum += w[k] * ctx->h_sc[2].dot;
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    st.device_ms = ms;
    if (J) *J = Jsum;
    if (dJdE) HIPCHK(hipMemcpy(dJdE, nm.sens_acc, ctx->n_el * 8, hipMemcpyDeviceToHost));
    if (adj_stats) *adj_stats = st;
    return ASEND_OK;
}

// Eigenpairs of a small dense symmetric matrix A (m x m, row-major, destroyed)
// by cyclic Jacobi rotations: values into w, eigenvectors into the columns of
// Z.  Used on the Lanczos tridiagonal (m <= a few hundred).
static void jacobi_eigen(int m, std::vector<double>& A, std::vector<double>& w, std::vector<double>& Z) {
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                const double a = A[(size_t)i * m + j] * A[(size_t)i * m + j];
                tot += a;
                if (i != j) off += a;
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[(size_t)p * m + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
                // rotation angle that zeroes A_pq: tan(2 phi) = 2 A_pq / (A_qq - A_pp)
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < m; ++k) {
                    const double akp = A[(size_t)k * m + p], akq = A[(size_t)k * m + q];
                    A[(size_t)k * m + p] = c * akp - s * akq;
                    A[(size_t)k * m + q] = s * akp + c * akq;
                }
                for (int k = 0; k < m; ++k) {
                    const double apk = A[(size_t)p * m + k], aqk = A[(size_t)q * m + k];
                    A[(size_t)p * m + k] = c * apk - s * aqk;
                    A[(size_t)q * m + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < m; ++k) {
                    const double zkp = Z[(size_t)k * m + p], zkq = Z[(size_t)k * m + q];
                    Z[(size_t)k * m + p] = c * zkp - s * zkq;
                    Z[(size_t)k * m + q] = s * zkp + c * zkq;
                }
            }
    }
    w.resize(m);
    for (int i = 0; i < m; ++i) w[i] = A[(size_t)i * m + i];
}

// Model::eigen_solve (S/model_analysis.rs:713-888).  kind 0 (frequency):
// lumped mass; kind 1 (buckling): identity "mass" and, per mode, the load
// factor phi^T K phi / (phi^T K phi - lambda) with K the stiffness at zero
// displacement and |phi| = 1.  With u the operator is the tangent at u
// (build_elastic_soln_load with nonlinear_geom, as eigen_solve builds it),
// else the linear stiffness.  out_fact: frequencies (Hz, -1 for a negative
// eigenvalue) or load factors.
static int modal_impl(asend_ctx* ctx, int kind, const double* u, int n_modes, double shift,
                      const asend_solve_opts* opts, double* eig_vals, double* out_fact, double* eig_vecs,
                      asend_modal_stats* stats) {
    if (!ctx || n_modes < 1 || !opts || !eig_vals || (kind != 0 && kind != 1)) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_modal")) return rc0;
    if (!ctx->d_K64) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    if (ctx->etype == ASEND_SHELL4 && !ctx->thk_set)
        return fail(ctx, ASEND_ERR_STATE, "shell4 mesh without thickness (asend_set_thickness)");
    const int64_t n = ctx->n_dof, nval = ctx->nval();
    const int m = (int)std::min<int64_t>({(int64_t)std::max(5 * n_modes, n_modes + 20), (int64_t)modal_max_vectors() - 1, n});
    if (n_modes > m) return fail(ctx, ASEND_ERR_ARG, "too many modes requested");
    int rc;
    // lumped mass: row sums of the consistent mass, small entries raised to
    // 1e-6 of the mean as the reference does (Model::eigen_solve,
    // S/model_analysis.rs:744-773)
    double *Mtmp = nullptr, *ones = nullptr, *dm = nullptr, *ds = nullptr, *dsi = nullptr, *V = nullptr,
           *w = nullptr, *c = nullptr, *part = nullptr;
    auto cleanup = [&]() {
        for (double** p : {&Mtmp, &ones, &dm, &ds, &dsi, &V, &w, &c, &part}) dfree(*p);
    };
    const int nb = (int)std::min<int64_t>((n + 255) / 256, 1024);
    if ((rc = dalloc(ctx, &Mtmp, nval)) || (rc = dalloc(ctx, &ones, n)) || (rc = dalloc(ctx, &dm, n)) ||
        (rc = dalloc(ctx, &ds, n)) || (rc = dalloc(ctx, &dsi, n)) || (rc = dalloc(ctx, &V, (size_t)(m + 1) * n)) ||
        (rc = dalloc(ctx, &w, n)) || (rc = dalloc(ctx, &c, m + 1)) || (rc = dalloc(ctx, &part, (size_t)nb * (m + 1)))) {
        cleanup();
        return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    // consistent mass alone (zero moduli for the brick8 kernels, no stiffness
    // term for the kernels that condense internal DOFs); buckling: none
    if (kind == 1) {
        rc = ASEND_OK;
    } else if (ctx->etype == ASEND_SHELL4 || ctx->etype == ASEND_BRICKIM) {
        rc = integrate_elements(ctx, ctx->d_E, 0.0, 1.0, Mtmp);
    } else {
        double* zeroE = nullptr;
        if (!(rc = dalloc(ctx, &zeroE, ctx->n_el))) {
            HIPCHK(hipMemsetAsync(zeroE, 0, ctx->n_el * 8, ctx->st));
            rc = integrate_elements(ctx, zeroE, 1.0, 1.0, Mtmp);
            HIPCHK(hipStreamSynchronize(ctx->st));
            dfree(zeroE);
        }
    }
    if (rc) {
        cleanup();
        return rc;
    }
    {
        std::vector<double> h1(n, 1.0);
        HIPCHK(hipMemcpy(ones, h1.data(), n * 8, hipMemcpyHostToDevice));
    }
    std::vector<double> hm(n, 1.0);
    if (kind == 0) {
        launch_sell3<double>(ctx->st, 0, ctx->fine(), Mtmp, ones, dm, nullptr, nullptr, nullptr);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(ctx->st));  // the context stream does not order null-stream copies
        HIPCHK(hipMemcpy(hm.data(), dm, n * 8, hipMemcpyDeviceToHost));
    }
    double m_avg = 0.0;
    for (double v : hm) m_avg += std::fabs(v);
    m_avg /= (double)n;
    if (!(m_avg > 0.0)) {
        cleanup();
        return fail(ctx, ASEND_ERR_STATE, "modal analysis needs element densities (upload rho)");
    }
    int n_raised = 0;
    for (double& v : hm)
        if (v < 1e-6 * m_avg) {
            v = 1e-6 * m_avg;
            ++n_raised;
        }
    {
        std::vector<double> hs(n), hsi(n);
        for (int64_t i = 0; i < n; ++i) {
            hs[i] = std::sqrt(hm[i]);
            hsi[i] = 1.0 / hs[i];
        }
        HIPCHK(hipMemcpy(dm, hm.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ds, hs.data(), n * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dsi, hsi.data(), n * 8, hipMemcpyHostToDevice));
    }
    dfree(Mtmp);
    // shifted operator K_eff - shift diag(m) (penalty included)
    if (u) {
        HIPCHK(hipMemcpy(ctx->d_nl_u, u, n * 8, hipMemcpyHostToDevice));
        rc = nl_assemble(ctx, 0.0, nullptr);
    } else {
        rc = assemble_device(ctx, 1.0, 0.0);
    }
    if (rc) {
        cleanup();
        return rc;
    }
    if (shift != 0.0) {
        launch_add_diag(ctx->st, (int)ctx->n_nodes, ctx->d_sptr, ctx->d_kdiag, ctx->d_K64, dm, -shift);
        launch_diag_inv(ctx->st, (int)ctx->n_nodes, ctx->d_sptr, ctx->d_kdiag, ctx->d_K64, ctx->d_dinv64,
                        ctx->d_dinv32);
        launch_to_f32(ctx->st, ctx->d_K64, ctx->d_K32, nval);
        refresh_f16(ctx);
        HIPCHK(hipGetLastError());
    }
    // Lanczos on B = S (K - shift M)^-1 S, S = diag(sqrt(m)), with full
    // reorthogonalisation (two classical Gram-Schmidt passes per step)
    {
        std::vector<double> v0(n);
        double nrm = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            v0[i] = std::sin((double)(i + 1));
            nrm += v0[i] * v0[i];
        }
        nrm = 1.0 / std::sqrt(nrm);
        for (double& x : v0) x *= nrm;
        HIPCHK(hipMemcpy(V, v0.data(), n * 8, hipMemcpyHostToDevice));
    }
    std::vector<double> alpha, beta;
    int64_t solve_its = 0, unconverged = 0;
    int steps = 0;
    std::vector<double> hc(m + 1), hc2(m + 1);
    for (int j = 0; j < m; ++j) {
        double* vj = V + (int64_t)j * n;
        launch_vmul(ctx->st, n, ds, vj, ctx->d_f);
        asend_solve_stats st;
        if ((rc = solve_device(ctx, opts, &st))) {
            cleanup();
            return rc;
        }
        if (!st.converged) ++unconverged;  // reported, as the reference warns and goes on
        solve_its += st.iterations;
        launch_vmul(ctx->st, n, ds, ctx->d_u, w);
        launch_vtw(ctx->st, n, j + 1, V, w, part, nb, c);
        HIPCHK(hipMemcpyAsync(hc.data(), c, (j + 1) * 8, hipMemcpyDeviceToHost, ctx->st));
        launch_add_vc(ctx->st, n, j + 1, V, c, -1.0, w);
        launch_vtw(ctx->st, n, j + 1, V, w, part, nb, c);
        HIPCHK(hipMemcpyAsync(hc2.data(), c, (j + 1) * 8, hipMemcpyDeviceToHost, ctx->st));
        launch_add_vc(ctx->st, n, j + 1, V, c, -1.0, w);
        launch_dot64(ctx->st, n, w, w, ctx->d_partials, ctx->d_sc + 2);
        HIPCHK(hipGetLastError());
        if ((rc = read_scalars(ctx, 2))) {
            cleanup();
            return rc;
        }
        const double a = hc[j] + hc2[j];
        const double b = std::sqrt(ctx->h_sc[2].dot);
        alpha.push_back(a);
        steps = j + 1;
        if (j == m - 1 || !(b > 1e-13 * std::fabs(a))) break;  // Krylov space exhausted
        beta.push_back(b);
        HIPCHK(hipMemcpyAsync(V + (int64_t)(j + 1) * n, w, n * 8, hipMemcpyDeviceToDevice, ctx->st));
        hipLaunchKernelGGL(k_scale_inplace, dim3(256), dim3(256), 0, ctx->st, n, 1.0 / b, V + (int64_t)(j + 1) * n);
    }
    const double b_last = std::sqrt(ctx->h_sc[2].dot);
    // Ritz pairs of the tridiagonal: largest theta = eigenvalues nearest the shift
    std::vector<double> T((size_t)steps * steps, 0.0), th, Z;
    for (int i = 0; i < steps; ++i) {
        T[(size_t)i * steps + i] = alpha[i];
        if (i + 1 < steps) T[(size_t)i * steps + i + 1] = T[(size_t)(i + 1) * steps + i] = beta[i];
    }
    jacobi_eigen(steps, T, th, Z);
    std::vector<int> order(steps);
    for (int i = 0; i < steps; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return th[x] > th[y]; });
    const int nm_out = std::min(n_modes, steps);
    double max_res = 0.0;
    std::vector<double> vec_tmp;
    if (kind == 1 && !eig_vecs) {  // the load factors need the modes
        vec_tmp.resize((size_t)n_modes * n);
        eig_vecs = vec_tmp.data();
    }
    for (int k = 0; k < nm_out; ++k) {
        const int i = order[k];
        const double lam = shift + 1.0 / th[i];
        eig_vals[k] = lam;
        if (out_fact && kind == 0) out_fact[k] = lam >= 0.0 ? std::sqrt(lam) / (2.0 * M_PI) : -1.0;
        // residual estimate of the Ritz pair in B: |b_last * z_last| / theta
        max_res = std::max(max_res, std::fabs(b_last * Z[(size_t)(steps - 1) * steps + i]) / std::fabs(th[i]));
        if (eig_vecs) {
            std::vector<double> zc(steps);
            for (int j = 0; j < steps; ++j) zc[j] = Z[(size_t)j * steps + i];
            HIPCHK(hipMemcpy(c, zc.data(), steps * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemsetAsync(w, 0, n * 8, ctx->st));
            launch_add_vc(ctx->st, n, steps, V, c, 1.0, w);
            launch_vmul(ctx->st, n, dsi, w, ctx->d_f);
            launch_dot64(ctx->st, n, ctx->d_f, ctx->d_f, ctx->d_partials, ctx->d_sc + 2);
            if ((rc = read_scalars(ctx, 2))) {
                cleanup();
                return rc;
            }
            const double sc = 1.0 / std::sqrt(ctx->h_sc[2].dot);
            hipLaunchKernelGGL(k_scale_inplace, dim3(256), dim3(256), 0, ctx->st, n, sc, ctx->d_f);
            HIPCHK(hipStreamSynchronize(ctx->st));
            HIPCHK(hipMemcpy(eig_vecs + (int64_t)k * n, ctx->d_f, n * 8, hipMemcpyDeviceToHost));
        }
    }
    for (int k = nm_out; k < n_modes; ++k) {
        eig_vals[k] = 0.0;
        if (out_fact) out_fact[k] = kind == 0 ? -1.0 : 1.0e100;
    }
    if (kind == 1) {
        // v_kv = phi^T K phi with the stiffness at zero displacement, without
        // the constraint penalty (elastic_mat alone)
        if ((rc = assemble_device(ctx, 1.0, 0.0))) {
            cleanup();
            return rc;
        }
        std::vector<double> y(n);
        for (int k = 0; k < nm_out; ++k) {
            const double* ph = eig_vecs + (int64_t)k * n;
            HIPCHK(hipMemcpy(w, ph, n * 8, hipMemcpyHostToDevice));
            launch_sell3<double>(ctx->st, 0, ctx->fine(), ctx->d_K64, w, ctx->d_f, nullptr, nullptr, nullptr);
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(ctx->st));
            HIPCHK(hipMemcpy(y.data(), ctx->d_f, n * 8, hipMemcpyDeviceToHost));
            double vkv = 0.0;
            for (int64_t i = 0; i < n; ++i) vkv += ph[i] * y[i];
            for (size_t q = 0; q < ctx->h_pen.size(); ++q)
                vkv -= ctx->h_pen_mult[q] * ctx->sf * ph[ctx->h_pen[q]] * ph[ctx->h_pen[q]];
            if (out_fact) out_fact[k] = (vkv - eig_vals[k]) != 0.0 ? vkv / (vkv - eig_vals[k]) : 1.0e100;
        }
    }
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    cleanup();
    // the operator now holds the shift (or the tangent): a static solve must assemble again
    if (shift != 0.0 || u) ctx->assembled = false;
    if (kind == 1) ctx->assembled = true;  // buckling ends with the linear stiffness assembled
    if (stats) {
        asend_modal_stats s;
        std::memset(&s, 0, sizeof(s));
        s.lanczos_steps = steps;
        s.modes_found = nm_out;
        s.solve_iterations = solve_its;
        s.max_ritz_residual = max_res;
        s.masses_raised = n_raised;
        s.unconverged_solves = unconverged;
        s.device_ms = ms;
        *stats = s;
    }
    return ASEND_OK;
}

int asend_modal(asend_ctx* ctx, int n_modes, double shift, const asend_solve_opts* opts, double* eig_vals,
                double* freqs_hz, double* eig_vecs, asend_modal_stats* stats) {
    return modal_impl(ctx, 0, nullptr, n_modes, shift, opts, eig_vals, freqs_hz, eig_vecs, stats);
}

int asend_eigen(asend_ctx* ctx, int kind, const double* u, int n_modes, double shift, const asend_solve_opts* opts,
                double* eig_vals, double* factors, double* eig_vecs, asend_modal_stats* stats) {
    if (ctx && u) {
        int rc;
        if ((rc = nl_prepare(ctx))) return rc;
    }
    return modal_impl(ctx, kind, u, n_modes, shift, opts, eig_vals, factors, eig_vecs, stats);
}

// ---------------------------------------------------------------------------
// Learned preconditioner
// ---------------------------------------------------------------------------
int asend_learned_num_params(void) { return kNumParams; }

int asend_learned_init(asend_ctx* ctx, uint64_t seed) {
    if (!ctx) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_learned_init")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_learned_init")) return rc1;
    ctx->learned_aside = false;
    HIPCHK(hipSetDevice(ctx->device));
    invalidate_factor(ctx);
    if (ctx->learned) learned_destroy(ctx->learned);
    ctx->learned = nullptr;
    Learned* lp = nullptr;
    int rc = learned_create(&lp, seed);
    if (rc) {
        std::string m = learned_error(lp);
        learned_destroy(lp);
        return fail(ctx, rc, "%s", m.c_str());
    }
    ctx->learned = lp;
    return ASEND_OK;
}

int asend_learned_get_params(asend_ctx* ctx, float* out, int n) {
    if (!ctx || !out) return ASEND_ERR_ARG;
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (n != kNumParams) return fail(ctx, ASEND_ERR_ARG, "expected %d parameters", kNumParams);
    int rc = learned_get_params(ctx->learned, out, n);
    return rc ? fail(ctx, rc, "%s", learned_error(ctx->learned)) : ASEND_OK;
}

int asend_learned_set_params(asend_ctx* ctx, const float* in, int n) {
    if (!ctx || !in) return ASEND_ERR_ARG;
    ctx->learned_aside = false;
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (n != kNumParams) return fail(ctx, ASEND_ERR_ARG, "expected %d parameters", kNumParams);
    invalidate_factor(ctx);
    int rc = learned_set_params(ctx->learned, in, n);
    return rc ? fail(ctx, rc, "%s", learned_error(ctx->learned)) : ASEND_OK;
}

int asend_learned_train(asend_ctx* ctx, int n_steps, float lr, uint64_t seed, int mode, double* losses) {
    if (!ctx || n_steps < 0 || !(lr > 0.f) || mode < 0 || mode > 3) return ASEND_ERR_ARG;
    ctx->learned_aside = false;
    if (ctx->d_K64 && ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the learned preconditioner is defined on brick8 meshes");
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    if (mode > 0 && !ctx->mg) return fail(ctx, ASEND_ERR_STATE, "this training mode needs the multigrid hierarchy");
    int rc0;
    if (mode > 0 && (rc0 = ensure_mg(ctx))) return rc0;
    invalidate_factor(ctx);
    const LearnedGraph g = learned_graph(ctx);
    const int64_t n = ctx->n_dof;
    // mode 1: solver samples only, alternating the error a V-cycle leaves and
    // a V-cycle solution of a random load; mode 3 puts a Gaussian sample
    // before every such pair
    const bool solver_samples = (mode == 1 || mode == 3);
    double ema = 0.0;  // running level of the accepted mode-2 losses
    for (int k = 0; k < n_steps; ++k) {
        const uint64_t s = seed * 1000003ull + (uint64_t)k;
        if (mode == 2) {
            // end to end through one V-cycle with the current factor as the
            // finest smoother (its Chebyshev bounds re-estimated every step)
            learned_sample(ctx->st, g, s, ctx->d_p32);
            Factor F;
            int rc = learned_infer(ctx->learned, ctx->st, g, &F);
            if (rc) return fail(ctx, rc, "%s", learned_error(ctx->learned));
            if ((rc = mg_set_learned_factor(ctx->mg, ctx->st, F, true)))
                return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            FactorUses uses;
            double loss = 0.0;
            rc = mg_vcycle_uses(ctx->mg, ctx->st, ctx->d_p32, &uses, &loss);
            if (rc && rc != ASEND_ERR_NUMERIC) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            if (losses) losses[k] = rc ? NAN : loss;
            // a sample whose cycle diverged (loss far above the running level,
            // typically an underestimated lambda_max) is skipped, not learned from
            if (rc || !(loss < 4.0 * ema || k < 8)) continue;
            ema = (k == 0) ? loss : 0.95 * ema + 0.05 * loss;
            if ((rc = learned_step_from_uses(ctx->learned, ctx->st, g, uses, lr)))
                return fail(ctx, rc, "%s", learned_error(ctx->learned));
            continue;
        }
        const float* e = nullptr;
        const int kind = !solver_samples ? 0 : (mode == 1 ? 1 + k % 2 : k % 3);
        if (kind == 1) {
            // error left by one V-cycle B of the solver: e = (I - B A) e0
            learned_sample(ctx->st, g, s, ctx->d_p32);
            launch_sell3<float>(ctx->st, 0, ctx->fine(), ctx->d_K32, ctx->d_p32, ctx->d_r32, nullptr, nullptr,
                                nullptr);
            mg_set_fine_smoother(ctx->mg, 0);
            int rc = mg_vcycle(ctx->mg, ctx->st, ctx->d_r32, ctx->d_x32);
            if (rc) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            hipLaunchKernelGGL(k_sub_f32, dim3(kVecBlocks), dim3(256), 0, ctx->st, n, ctx->d_p32, ctx->d_x32,
                               ctx->d_q32);
            e = ctx->d_q32;
        } else if (kind == 2) {
            // approximate solution of A e = b for a random load b: e = B b
            learned_sample(ctx->st, g, s, ctx->d_r32);
            mg_set_fine_smoother(ctx->mg, 0);
            int rc = mg_vcycle(ctx->mg, ctx->st, ctx->d_r32, ctx->d_q32);
            if (rc) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            e = ctx->d_q32;
        }
        double loss = 0.0;
        int rc = learned_train_step(ctx->learned, ctx->st, g, e, s, lr, losses ? &loss : nullptr);
        if (rc) return fail(ctx, rc, "%s", learned_error(ctx->learned));
        if (losses) losses[k] = loss;
    }
    HIPCHK(hipStreamSynchronize(ctx->st));
    invalidate_factor(ctx);
    return ASEND_OK;
}

int asend_learned_eval(asend_ctx* ctx, const double* e, double* loss, float* grad, float* feat, float* out) {
    if (!ctx || !e) return ASEND_ERR_ARG;
    if (ctx->d_K64 && ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the learned preconditioner is defined on brick8 meshes");
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    invalidate_factor(ctx);
    const int64_t n = ctx->n_dof;
    std::vector<float> ef(e, e + n);
    HIPCHK(hipMemcpy(ctx->d_q32, ef.data(), n * 4, hipMemcpyHostToDevice));
    int rc = learned_eval(ctx->learned, ctx->st, learned_graph(ctx), ctx->d_q32, loss, grad, feat, out);
    return rc ? fail(ctx, rc, "%s", learned_error(ctx->learned)) : ASEND_OK;
}

int asend_learned_vcycle_eval(asend_ctx* ctx, const double* e, int keep_bounds, double* loss, float* grad) {
    if (!ctx || !e || !loss) return ASEND_ERR_ARG;
    if (ctx->d_K64 && ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the learned preconditioner is defined on brick8 meshes");
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (!ctx->assembled || !ctx->mg) return fail(ctx, ASEND_ERR_STATE, "assemble with a multigrid hierarchy first");
    {
        int rc0 = ensure_mg(ctx);
        if (rc0) return rc0;
    }
    const LearnedGraph g = learned_graph(ctx);
    const int64_t n = ctx->n_dof;
    std::vector<float> ef(e, e + n);
    HIPCHK(hipMemcpy(ctx->d_p32, ef.data(), n * 4, hipMemcpyHostToDevice));
    ctx->factor_asm = -1;
    Factor F;
    int rc = learned_infer(ctx->learned, ctx->st, g, &F);
    if (rc) return fail(ctx, rc, "%s", learned_error(ctx->learned));
    mg_keep_factor_bounds(ctx->mg, keep_bounds != 0);
    rc = mg_set_learned_factor(ctx->mg, ctx->st, F, true);
    mg_keep_factor_bounds(ctx->mg, false);
    if (rc) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    FactorUses uses;
    if ((rc = mg_vcycle_uses(ctx->mg, ctx->st, ctx->d_p32, &uses, loss))) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    if ((rc = learned_grad_from_uses(ctx->learned, ctx->st, g, uses, grad)))
        return fail(ctx, rc, "%s", learned_error(ctx->learned));
    return ASEND_OK;
}

int asend_learned_apply(asend_ctx* ctx, const double* r, double* z) {
    if (!ctx || !r || !z) return ASEND_ERR_ARG;
    if (ctx->d_K64 && ctx->etype != ASEND_BRICK8)
        return fail(ctx, ASEND_ERR_UNSUPPORTED, "the learned preconditioner is defined on brick8 meshes");
    if (!ctx->learned) return fail(ctx, ASEND_ERR_STATE, "no learned preconditioner (asend_learned_init)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    invalidate_factor(ctx);
    const int64_t n = ctx->n_dof;
    Factor F;
    int rc = learned_infer(ctx->learned, ctx->st, learned_graph(ctx), &F);
    if (rc) return fail(ctx, rc, "%s", learned_error(ctx->learned));
    std::vector<float> buf(r, r + n);
    HIPCHK(hipMemcpyAsync(ctx->d_r32, buf.data(), n * 4, hipMemcpyHostToDevice, ctx->st));
    launch_factor_apply(ctx->st, F, ctx->d_r32, ctx->d_w32, ctx->d_x32);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(buf.data(), ctx->d_x32, n * 4, hipMemcpyDeviceToHost, ctx->st));
    HIPCHK(hipStreamSynchronize(ctx->st));
    for (int64_t i = 0; i < n; ++i) z[i] = buf[i];
    return ASEND_OK;
}

int asend_learned_edges(asend_ctx* ctx, int32_t* enb, int32_t* erev, int* n_edge) {
    if (!ctx || !n_edge) return ASEND_ERR_ARG;
    if (!ctx->d_enb) return fail(ctx, ASEND_ERR_STATE, "upload a mesh first");
    *n_edge = ctx->n_edge;
    const size_t cnt = (size_t)std::max(ctx->n_edge, 1) * ctx->n_nodes;
    if (enb) HIPCHK(hipMemcpy(enb, ctx->d_enb, cnt * 4, hipMemcpyDeviceToHost));
    if (erev) HIPCHK(hipMemcpy(erev, ctx->d_erev, cnt * 4, hipMemcpyDeviceToHost));
    return ASEND_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Multigrid diagnostics (host arrays, FP32): used by the parity tests to check
// the Galerkin identity A_{l+1} = P^T A_l P and the V-cycle on the device.
// ---------------------------------------------------------------------------
extern "C" {

int asend_mg_info(asend_ctx* ctx, int* n_levels, int64_t* level_ndof) {
    if (!ctx || !n_levels) return ASEND_ERR_ARG;
    if (!ctx->mg) return fail(ctx, ASEND_ERR_STATE, "no multigrid hierarchy (asend_set_grid)");
    *n_levels = mg_levels(ctx->mg);
    if (level_ndof)
        for (int l = 0; l < *n_levels && l < 16; ++l) level_ndof[l] = mg_level_ndof(ctx->mg, l);
    return ASEND_OK;
}

static int mg_io(asend_ctx* ctx, int op, int level, const float* in, float* out) {
    if (!ctx || !in || !out) return ASEND_ERR_ARG;
    if (!ctx->mg) return fail(ctx, ASEND_ERR_STATE, "no multigrid hierarchy (asend_set_grid)");
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    {
        int rc0 = ensure_mg(ctx);
        if (rc0) return rc0;
    }
    int nl = mg_levels(ctx->mg);
    if (op == 4) level = nl - 1;
    if (level < 0 || level >= nl || (op >= 1 && op <= 3 && level + 1 >= nl))
        return fail(ctx, ASEND_ERR_ARG, "bad level %d", level);
    int64_t nin = 0, nout = 0;
    if (op == 0) nin = nout = mg_level_ndof(ctx->mg, level);
    if (op == 1) { nin = mg_level_ndof(ctx->mg, level); nout = mg_level_ndof(ctx->mg, level + 1); }
    if (op == 2) { nin = mg_level_ndof(ctx->mg, level + 1); nout = mg_level_ndof(ctx->mg, level); }
    if (op == 3) nin = nout = mg_level_ndof(ctx->mg, 0);
    if (op == 4) nin = nout = mg_level_ndof(ctx->mg, level);
    float *din = nullptr, *dout = nullptr;
    HIPCHK(hipMalloc((void**)&din, nin * sizeof(float)));
    HIPCHK(hipMalloc((void**)&dout, nout * sizeof(float)));
    HIPCHK(hipMemcpy(din, in, nin * sizeof(float), hipMemcpyHostToDevice));
    int rc = 0;
    if (op == 0) rc = mg_apply_level(ctx->mg, ctx->st, level, din, dout);
    if (op == 1) rc = mg_restrict(ctx->mg, ctx->st, level, din, dout);
    if (op == 2) rc = mg_prolong(ctx->mg, ctx->st, level, din, dout);
    if (op == 3) rc = mg_vcycle(ctx->mg, ctx->st, din, dout);
    if (op == 4) rc = mg_coarse_solve(ctx->mg, ctx->st, din, dout);
    hipError_t e1 = hipStreamSynchronize(ctx->st);
    hipError_t e2 = hipMemcpy(out, dout, nout * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(din);
    (void)hipFree(dout);
    if (rc) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    if (e1 != hipSuccess || e2 != hipSuccess) return fail(ctx, ASEND_ERR_HIP, "multigrid diagnostic copy failed");
    return ASEND_OK;
}

int asend_mg_apply(asend_ctx* ctx, int level, const float* x, float* y) { return mg_io(ctx, 0, level, x, y); }
int asend_mg_restrict(asend_ctx* ctx, int level, const float* fine, float* coarse) {
    return mg_io(ctx, 1, level, fine, coarse);
}
int asend_mg_prolong(asend_ctx* ctx, int level, const float* coarse, float* fine) {
    return mg_io(ctx, 2, level, coarse, fine);
}
int asend_mg_vcycle(asend_ctx* ctx, const float* b, float* x) { return mg_io(ctx, 3, 0, b, x); }
int asend_mg_coarse_solve(asend_ctx* ctx, const float* b, float* x) { return mg_io(ctx, 4, 0, b, x); }

// ---------------------------------------------------------------------------
// Autotuner
// ---------------------------------------------------------------------------
int asend_autotune(asend_ctx* ctx, const double* rhs, int what, asend_tuning* out) {
    if (!ctx || !out) return ASEND_ERR_ARG;
    if (const int rc0 = elastic_only(ctx, "asend_autotune")) return rc0;
    if (const int rc1 = iso_only(ctx, "asend_autotune")) return rc1;
    if (!ctx->assembled) return fail(ctx, ASEND_ERR_STATE, "assemble first");
    std::memset(out, 0, sizeof(*out));
    int rc;
    if (what & ASEND_TUNE_SPMV) {
        // SpMV workgroup geometry: 1, 2, 4 or 8 slices per workgroup, timed
        // with HIP events on the context stream (FP32 decides, FP64 reported)
        double best = 1e300;
        int best_w = ctx->spmv_wpb;
        for (int k = 0; k < 4; ++k) {
            const int w = 1 << k;
            if ((ctx->n_slices + w - 1) / w > ctx->partials_cap) continue;
            ctx->spmv_wpb = w;
            double ms32 = 0, ms64 = 0;
            if ((rc = asend_bench_spmv(ctx, 32, 50, &ms32)) || (rc = asend_bench_spmv(ctx, 64, 20, &ms64))) return rc;
            out->spmv_ms_fp32[k] = ms32;
            out->spmv_ms_fp64[k] = ms64;
            out->n_trials += 1;
            if (ms32 < best) {
                best = ms32;
                best_w = w;
            }
        }
        ctx->spmv_wpb = best_w;
        if (ctx->mg) mg_set_fine_geometry(ctx->mg, best_w);
    }
    out->spmv_waves_per_block = ctx->spmv_wpb;
    if (what & ASEND_TUNE_SOLVER) {
        // multigrid-preconditioned mixed solve: Chebyshev degree, reliable-
        // update factor and convergence-check interval, timed on full solves
        // of rhs
        if (!rhs) return fail(ctx, ASEND_ERR_ARG, "solver tuning needs a right-hand side");
        if (!ctx->mg) return fail(ctx, ASEND_ERR_STATE, "solver tuning needs the multigrid hierarchy");
        const bool learned = (what & ASEND_TUNE_LEARNED) != 0;
        if (learned && !ctx->learned)
            return fail(ctx, ASEND_ERR_STATE, "learned tuning needs the learned preconditioner (asend_learned_init)");
        HIPCHK(hipMemcpy(ctx->d_f, rhs, ctx->n_dof * 8, hipMemcpyHostToDevice));
        const int degrees[3] = {1, 2, 3};
        const double tols[4] = {1e-1, 1e-2, 1e-3, 1e-4};
        const int checks[3] = {1, 2, 4};
        double best = 1e300;
        int bd = 2, bc = 4;
        double bt = 1e-3;
        // learned tuning also tunes the same solve without the network and
        // keeps the faster pipeline: the network steps aside when it does not
        // pay for itself (ASEND_PC_LEARNED solves then report fallback_used 2)
        ctx->learned_aside = false;
        double best_pc[2] = {1e300, 1e300};  // [0] multigrid smoother, [1] learned
        bool best_learned = learned;
        for (int pass = learned ? 1 : 0; pass >= 0; --pass)
        for (int d : degrees) {
            if ((rc = mg_set_smoother(ctx->mg, 1, d, 0.1, 1.15)) || (rc = mg_refresh_coefficients(ctx->mg, ctx->st)))
                return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
            if ((rc = ensure_mg(ctx))) return rc;
            for (double t : tols)
                for (int c : checks) {
                    asend_solve_opts o;
                    std::memset(&o, 0, sizeof(o));
                    o.method = ASEND_SOLVE_IR_MIXED;
                    o.precond = pass == 1 ? ASEND_PC_LEARNED : ASEND_PC_MG;
                    o.stop_rule = ASEND_STOP_REL_RESIDUAL;
                    o.tol = 1e-8;
                    o.inner_tol = t;
                    o.check_every = c;
                    o.max_it = 100000;
                    o.max_outer = 60;
                    // the faster of two solves: the first also captures the
                    // iteration blocks of this choice as graphs
                    float ms = 1e30f;
                    bool ok = true;
                    for (int rep = 0; rep < 2; ++rep) {
                        asend_solve_stats st;
                        HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
                        if ((rc = solve_device(ctx, &o, &st))) return rc;
                        HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
                        HIPCHK(hipEventSynchronize(ctx->ev1));
                        float m = 0;
                        HIPCHK(hipEventElapsedTime(&m, ctx->ev0, ctx->ev1));
                        ms = std::min(ms, m);
                        ok = ok && st.converged && !st.fallback_used;
                    }
                    out->n_trials += 1;
                    if (ok) best_pc[pass] = std::min(best_pc[pass], (double)ms);
                    if (ok && ms < best) {
                        best = ms;
                        bd = d;
                        bt = t;
                        bc = c;
                        best_learned = (pass == 1);
                    }
                }
        }
        if ((rc = mg_set_smoother(ctx->mg, 1, bd, 0.1, 1.15)) || (rc = mg_refresh_coefficients(ctx->mg, ctx->st)))
            return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
        ctx->tuned_inner_tol = bt;
        ctx->tuned_check = bc;
        ctx->learned_aside = learned && !best_learned;
        out->solve_ms = best;
        out->learned_active = learned && best_learned ? 1 : 0;
        out->solve_ms_learned = learned ? best_pc[1] : 0.0;
        out->solve_ms_no_network = best_pc[0];
    }
    if (ctx->mg) out->cheb_degree = mg_smoother_degree(ctx->mg);
    out->inner_tol = ctx->tuned_inner_tol;
    out->check_every = ctx->tuned_check;
    return ASEND_OK;
}

int asend_set_graphs(asend_ctx* ctx, int on) {
    if (!ctx) return ASEND_ERR_ARG;
    ctx->use_graphs = (on != 0);
    if (ctx->mg) mg_use_graphs(ctx->mg, ctx->use_graphs);
    return ASEND_OK;
}

int asend_mg_set_smoother(asend_ctx* ctx, int kind, int degree, double lo_frac, double hi_frac) {
    if (!ctx) return ASEND_ERR_ARG;
    if (!ctx->mg) return fail(ctx, ASEND_ERR_STATE, "no multigrid hierarchy (asend_set_grid)");
    int rc = mg_set_smoother(ctx->mg, kind, degree, lo_frac, hi_frac);
    // same Jacobi kind: the step coefficients of the new degree / interval
    // from the stored bounds (a new kind re-runs the setup at the next solve)
    if (rc == ASEND_OK) rc = mg_refresh_coefficients(ctx->mg, ctx->st);
    if (rc) return fail(ctx, rc, "%s", mg_last_error(ctx->mg));
    return ASEND_OK;
}

}  // extern "C"

extern "C" int asend_bench_copy(asend_ctx* ctx, int64_t bytes, int n_iter, double* avg_ms) {
    if (!ctx || bytes <= 0 || n_iter <= 0 || !avg_ms) return ASEND_ERR_ARG;
    const int64_t n4 = bytes / 16;
    float *a = nullptr, *b = nullptr;
    HIPCHK(hipMalloc((void**)&a, n4 * 16));
    HIPCHK(hipMalloc((void**)&b, n4 * 16));
    HIPCHK(hipMemsetAsync(a, 0, n4 * 16, ctx->st));
    launch_copy_f32(ctx->st, n4, a, b);
    HIPCHK(hipEventRecord(ctx->ev0, ctx->st));
    for (int i = 0; i < n_iter; ++i) launch_copy_f32(ctx->st, n4, a, b);
    HIPCHK(hipEventRecord(ctx->ev1, ctx->st));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *avg_ms = ms / n_iter;
    (void)hipFree(a);
    (void)hipFree(b);
    return ASEND_OK;
}
