/*
 * asend_gpu.h -- C ABI of the MI355X (gfx950) solve path for ASenD3D models.
 *
 * The reference has no plugin/FFI layer: its hot path is the in-process call
 * chain Model::solve -> solve_step -> build_elastic_soln_load / get_ru_dfd0 ->
 * conj_grad_sparse, and Model::get_obj_gradient -> solve_for_adjoint ->
 * d_relasticd_d (reference paths below are relative to the reference checkout,
 * S/ = solverSrc/asend_run_job/src/).  Each entry point here replaces one link
 * of that chain and is what a thin Rust FFI crate (bindgen over this header)
 * would bind in place of the Rust routine named in its comment.
 *
 * Conventions
 *   - Plain C types only; every array is a host pointer owned by the caller,
 *     read-only unless documented as output.  No pointer is retained past the
 *     call.  The context owns every device buffer.
 *   - Node coordinates xyz[3*n + c]; element connectivity conn[nn*e + a] in
 *     the reference's local node order; dof_index[6*n + d] = global DOF of
 *     node n, component d (-1 when the node has fewer DOFs), exactly as
 *     produced by asend_reorder_nodes.
 *   - Global vectors are in the reference's DOF numbering.
 *   - Every function returns ASEND_OK (0) or a negative ASEND_ERR_* code and
 *     never throws across the ABI; asend_last_error() gives the message.
 *     Non-convergence is not an error: the call returns ASEND_OK and sets
 *     stats->converged = 0 (the reference only warns, S/matrix_functions.rs:354-356).
 *   - A context is not thread-safe; its device work runs on one HIP stream,
 *     except the multigrid setup of a learned-smoother solve, which runs on
 *     a second stream of the context joined back before the solve; both are
 *     created at default priority.
 */
#ifndef ASEND_GPU_H
#define ASEND_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ASEND_OK 0
#define ASEND_ERR_ARG (-1)
#define ASEND_ERR_HIP (-2)
#define ASEND_ERR_STATE (-3)
#define ASEND_ERR_UNSUPPORTED (-4)
#define ASEND_ERR_NUMERIC (-5)

/* element type codes of the reference (S/model_input.rs:405-423) */
#define ASEND_BRICK8 8
#define ASEND_BRICKIM 81
#define ASEND_SHELL4 41

/* solver methods */
#define ASEND_SOLVE_PCG_FP64 0    /* FP64 PCG on the assembled FP64 operator          */
#define ASEND_SOLVE_IR_MIXED 1    /* FP32 flexible PCG with FP64 reliable updates */
#define ASEND_SOLVE_DIRECT 2      /* u = (L D L^T)^-1 f with the one-block factor of
                                     asend_ldl_factor (the reference's "direct"
                                     solverMethod, S/model_analysis.rs:495-496) */

/* preconditioners */
#define ASEND_PC_JACOBI 0         /* point Jacobi                                       */
#define ASEND_PC_BLOCK_JACOBI 1   /* inverse of the per-node diagonal blocks            */
#define ASEND_PC_MG 2             /* geometric multigrid, Galerkin coarse operators     */
#define ASEND_PC_LEARNED 3        /* multigrid whose fine smoother is the learned L L^T */
#define ASEND_PC_LEARNED_PCG 4    /* the learned L L^T alone                            */
#define ASEND_PC_LDL 5            /* the block LDL^T factor of asend_ldl_factor (the
                                     reference's LowerTriMat preconditioner of
                                     conj_grad_sparse); FP64 PCG only             */

/* stopping rules */
#define ASEND_STOP_REL_RESIDUAL 0 /* ||f - K u||_2 / ||f||_2 <= tol, evaluated in FP64 */
#define ASEND_STOP_REF_ABS 1      /* g^T M^-1 g <= tol (conj_grad_sparse rule,
                                     S/matrix_functions.rs:328), FP64 PCG only       */

typedef struct asend_ctx asend_ctx;

typedef struct asend_solve_opts {
    int method;           /* ASEND_SOLVE_*                                   */
    int precond;          /* ASEND_PC_*                                      */
    int stop_rule;        /* ASEND_STOP_*                                    */
    int check_every;      /* iterations between convergence checks (>=1)    */
    double tol;           /* outer tolerance for the stop rule               */
    double inner_tol;     /* mixed: recursive FP32 residual drop that triggers
                             an FP64 reliable update (refresh)               */
    int64_t max_it;       /* max total (inner) iterations; <=0: n_dof        */
    int64_t max_outer;    /* mixed: max restarts of the FP32 solve (each
                             runs through its reliable updates)              */
} asend_solve_opts;

typedef struct asend_solve_stats {
    int64_t iterations;       /* total Krylov iterations (inner for mixed)   */
    int64_t outer_iterations; /* FP64 residual refreshes (mixed), 1 for FP64 */
    double rel_residual;      /* final ||f - K u|| / ||f|| in FP64            */
    double ref_residual;      /* final g^T M^-1 g of the last Krylov loop     */
    int converged;            /* 1 when the stop rule was met                */
    int fallback_used;        /* 1 when the learned preconditioner was
                                 rejected and Jacobi PCG took over; 2 when
                                 the learned autotuning measured the
                                 multigrid smoother without the network
                                 faster and the solve ran it              */
    double device_ms;         /* HIP-event time of the solve on the stream   */
} asend_solve_stats;

/* ---- context ------------------------------------------------------------ */
int asend_ctx_create(int device, asend_ctx** out);
void asend_ctx_destroy(asend_ctx* ctx);
const char* asend_last_error(const asend_ctx* ctx);
const char* asend_version(void);
/* Block until all work queued on the context's stream has finished. */
int asend_sync(asend_ctx* ctx);

/* ---- host-side model preparation (no device work) ----------------------- */

/* Replaces Model::reorder_nodes + Model::add_constraint_conn
 * (S/model_prep.rs:56-303, :20-54): level-order node insertion from the
 * least-connected node with a restart at the nearest uninserted node every
 * block_dim nodes; DOFs numbered node by node in insertion order.
 * Constraint c has terms [cterm_ptr[c], cterm_ptr[c+1]); term t's node set is
 * set_nodes[term_set_ptr[t] .. term_set_ptr[t+1]).
 * Out: ordered[n_nodes], dof_index[6*n_nodes], *n_dof_out. */
int asend_reorder_nodes(int64_t n_nodes, const double* xyz, int64_t n_el, int nn, int dof_per_nd,
                        const int32_t* conn, int64_t n_const, const int64_t* cterm_ptr,
                        const int64_t* term_set_ptr, const int32_t* set_nodes, int64_t block_dim,
                        int32_t* ordered, int64_t* dof_index, int64_t* n_dof_out);

/* Replaces the iterative block-size rule of Model::analysis_prep
 * (S/model_prep.rs:888-902). */
int64_t asend_block_dim_auto(int64_t n_nodes);

/* Replaces the surfaceTraction branch of Element::get_app_load_dfd0
 * (S/element_equations.rs:1833-2056) with find_surface_faces
 * (S/model_prep.rs:510-554): accumulates (+=) the nodal loads of a uniform
 * traction on the surface faces of el_set whose outward normal lies within
 * norm_tol_deg of norm_dir (normalised here, as the reference's loads are).
 * Solid hex types only. */
int asend_surface_traction(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el,
                           const int32_t* conn, const int64_t* dof_index, int64_t n_set,
                           const int32_t* el_set, const double* load, const double* norm_dir,
                           double norm_tol_deg, int64_t n_dof, double* f);

/* The surfacePressure branch of the same routine (S/element_equations.rs:
 * 1970-2043): accumulates (+=) the loads of the traction -pressure n on the
 * surface faces of el_set whose outward unit normal n lies within
 * norm_tol_deg of norm_dir (normalised here, as Load::set_norm_dir does,
 * S/load_meth.rs:22-32).  Solid hex types only. */
int asend_surface_pressure(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el,
                           const int32_t* conn, const int64_t* dof_index, int64_t n_set,
                           const int32_t* el_set, double pressure, const double* norm_dir,
                           double norm_tol_deg, int64_t n_dof, double* f);

/* The bodyForce (kind 0), gravitational (kind 1) and centrifugal (kind 2)
 * branches of Element::get_app_load_dfd0 (S/element_equations.rs:1876-1969)
 * for solid hex elements: accumulates (+=) the consistent nodal loads of the
 * elements of el_set.  bodyForce: load = force per unit volume (each
 * element's vector rescaled to |b| V as the reference does); gravitational:
 * load = acceleration, times the element density rho[e]; centrifugal:
 * rho[e] angular_velocity^2 times the distance vector from the axis (through
 * center along axis, normalised here as Load::set_axis does) to the element
 * centroid.  rho may be NULL for kind 0. */
/* surfaceFlux (Element::get_app_therm_load_dfd0, S/element_equations.rs:
 * 2113-2150, and Face::get_area_normal_dfd0): accumulates (+=) into
 * f[dof_index[6 n]] the nodal heat input of a uniform flux q on the surface
 * faces of el_set that pass the reference's normal test: the element's
 * running sum of norm_dir . n over its surface faces (the reference never
 * resets it between faces) exceeds cos(norm_tol_deg).  Each such face's
 * vector is the unit-capacity heat-rate integral of a field equal to q on the
 * face's nodes (and on the nodes of earlier passing faces), kept on the
 * face's nodes and rescaled to q times the face area; the element adds the
 * vector of its last passing face.  bodyHeatGen is asend_body_load kind 0
 * with load (Q, 0, 0) (the same consistent integral, rescaled to Q V), and
 * nodalHeatGen adds its value at each node's temperature DOF.  Solid hex
 * types only. */
int asend_surface_flux(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                       const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double q,
                       const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f);

/* Thermal expansion load (get_ruk_dfd0, S/element_equations.rs:272-298: the
 * strain at each Gauss point loses the material's expansion coefficients
 * times the interpolated nodal temperature, with no reference temperature):
 * accumulates (+=) f_a = sum_gp w detJ S g_a, S the stress of
 * C (alpha T_gp).  C: isotropic from (E_e, nu_e), or E_e times the
 * element's matrix of cmat (n x 36, cmat_idx per element; NULL: matrix 0)
 * when cmat is given.  alpha: 6 per element (strain order 11, 22, 33, 12,
 * 13, 23); temp: per node.  A linear static step then solves
 * K u = f + this load.  brick8 only. */
int asend_thermal_strain_load(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                              const int64_t* dof_index, const double* E, const double* nu, const double* cmat,
                              const int32_t* cmat_idx, const double* alpha, const double* temp, int64_t n_dof,
                              double* f);

int asend_body_load(int elem_type, int kind, int64_t n_nodes, const double* xyz, int64_t n_el,
                    const int32_t* conn, const int64_t* dof_index, int64_t n_set, const int32_t* el_set,
                    const double* rho, const double* load, const double* center, const double* axis,
                    double angular_velocity, int64_t n_dof, double* f);

/* Seeded per-element moduli E_e = median * exp(sigma_ln * z_e), z_e standard
 * normal from splitmix64(seed) + Box-Muller (cosine branch, one pair per
 * element).  Input synthesis for heterogeneous designs; no reference
 * counterpart (the reference takes per-element values through design
 * variables, S/model_input.rs:1377-1401). */
int asend_lognormal_moduli(uint64_t seed, int64_t n, double median, double sigma_ln, double* out);

/* Host-only check of the isotropic materials of n_el elements against the
 * pivot clamp of the reference's compliance inversion: get_solid_stiff_dfd0
 * (S/element_properties.rs:728-759; 6x6, brick8/brickIM) and
 * get_layer_q_dfd0 (:196-210; 3x3 plane stress, shell4) invert
 * S(E, nu, E/(2(1+nu))) by the Givens QR of q_rfactor
 * (S/matrix_functions.rs:167-215), which replaces pivots below 1e-12 by
 * 1e-12.  *n_clamped counts the elements where that fires (for nu = 0.3,
 * E above ~1e12 in the model's units), where the reference's stiffness is no
 * longer the isotropic C that the device integrates; *first (may be NULL)
 * gets the first such label or -1.  asend_upload_mesh and
 * asend_update_moduli refuse such materials with ASEND_ERR_UNSUPPORTED. */
int asend_compliance_clamp_check(int elem_type, int64_t n_el, const double* E, const double* nu,
                                 int64_t* n_clamped, int64_t* first);

/* ---- device model ------------------------------------------------------- */

/* Upload a homogeneous mesh: the replacement for the element/node arrays that
 * Model::analysis_prep hands to build_elastic_soln_load
 * (S/model_prep.rs:881-952, S/model_analysis.rs:179-201).  E, nu, rho are per
 * element (isotropic; G = E/(2(1+nu)) as P/model/Material.py:24-26 writes
 * it).  rho may be NULL for static work.  Builds the node-block sparsity
 * pattern, the element colouring and the element-to-block map.
 * elem_type ASEND_BRICK8 or ASEND_BRICKIM (brick8 + 3 incompatible bubble modes
 * condensed per element): conn holds 8 nodes per element, n_dof = 3 n_nodes.
 * elem_type ASEND_SHELL4: conn holds 4 nodes per element, n_dof = 6 n_nodes
 * (dof_index[6n + c] = 6 rank + c), one isotropic layer per element whose
 * thickness asend_set_thickness supplies; every element must lie in a plane
 * normal to z (the identity element frame).  A new mesh clears the penalty
 * set (asend_set_penalty numbers DOFs on the mesh it is called for).
 * Materials past the compliance-inversion pivot clamp of the reference
 * (asend_compliance_clamp_check) are refused with ASEND_ERR_UNSUPPORTED: pose
 * such models in units where 1/E stays above 1e-12 (mm-N-MPa, say). */
int asend_upload_mesh(asend_ctx* ctx, int elem_type, int64_t n_nodes, const double* xyz,
                      const int64_t* dof_index, int64_t n_dof, int64_t n_el, const int32_t* conn,
                      const double* E, const double* nu, const double* rho);

/* Per-element thickness of a shell4 mesh: the single layer of the shell
 * section (Layer.thickness, S/section.rs:93, read by get_abd_dfd0,
 * S/element_properties.rs:766-876), centred on the mid-plane.  Required
 * before the first assembly of a shell4 mesh. */
int asend_set_thickness(asend_ctx* ctx, const double* t);

/* Optional structured-grid description (element (i,j,k) is label
 * i + nx*(j + ny*k), nodes likewise with nx+1, ny+1), used by the multigrid
 * preconditioners.  No reference counterpart. */
int asend_set_grid(asend_ctx* ctx, int64_t nx, int64_t ny, int64_t nz);

/* Replace the per-element moduli on the device (design update).  Refuses,
 * like asend_upload_mesh, moduli past the compliance pivot clamp, leaving the
 * previous moduli in place. */
int asend_update_moduli(asend_ctx* ctx, const double* E);

/* Penalty constraints of single-DOF clamp type: K_eff = K + sf * sum C^T C
 * with one unit entry per constrained DOF (Constraint::build_mat,
 * S/const_meth.rs:30-90; get_total_vec_mult :161-168).  sf <= 0 selects the
 * reference rule sf = 1e5 * max|K_ij| evaluated on the next assembly
 * (Model::scale_const, S/model_analysis.rs:255-259). */
int asend_set_penalty(asend_ctx* ctx, int64_t n, const int64_t* dofs, double sf);
double asend_penalty_scale(const asend_ctx* ctx);

/* 1 when the uploaded brick8 mesh consists of translated copies of one
 * element with one Poisson ratio (at most 8 elements per node): assembly then
 * scales that element's E = 1 / rho = 1 blocks by E_e / rho_e instead of
 * integrating every element.  No reference counterpart (diagnostic). */
int asend_uniform_elements(const asend_ctx* ctx);

/* Element integration + colored atomic-free assembly of the elastic operator:
 * replaces build_elastic_soln_load(true) with get_ruk_dfd0 / get_ru_dfd0
 * (S/model_analysis.rs:179-201, S/element_equations.rs:206-418, :852-1062).
 * c_mass != 0 adds c_mass * M (consistent mass) for the Newmark operator
 * K + M/(dt^2 (gamma - beta)) (S/element_equations.rs:977-1005).
 * Produces the FP64 operator (penalty folded into its diagonal), its FP32
 * copy and the Jacobi diagonal. */
int asend_assemble(asend_ctx* ctx, double c_mass);

/* Orthotropic and other anisotropic solid materials (get_solid_stiff_dfd0,
 * S/element_properties.rs:648-764: the stiffness is the inverse of the
 * compliance built from E1..3, nu12, nu13, nu23, G12, G13, G23, or the
 * material's stiffness entries): n_mat 6 x 6 matrices, symmetric to 1e-12 (strain order
 * 11, 22, 33, 12, 13, 23, engineering shear, in the global frame) and each
 * element's matrix (el_mat; NULL: matrix 0).  While set, brick8 assembly
 * integrates E_e times the element's matrix in place of the isotropic
 * material (upload E = 1 to take the matrices as they are; the mass and the
 * Rayleigh and material damping terms are unchanged).  Entry points whose
 * kernels integrate the isotropic material (sensitivities, element centre
 * fields, compliance pipeline, nonlinear, transient adjoint, multigrid,
 * learned, autotuning) refuse such a context.  n_mat = 0 removes them. */
int asend_set_material_stiffness(asend_ctx* ctx, int64_t n_mat, const double* cmat, const int32_t* el_mat);

/* Heat conduction statics (the thermal branch of Model::solve_step,
 * S/model_analysis.rs:331-370; Element::get_rtk_dfd0,
 * S/element_equations.rs:1064-1123; conductivity of
 * get_conductivity_dfd0, S/element_properties.rs:1715-1760): kcond holds 6
 * conductivities per element (k11 k22 k33 k12 k13 k23, the material key
 * thermal: conductivity), after which asend_assemble(c_mass) integrates
 * sum over the Gauss points of w detJ (G^T k G + c_mass rho_e N N^T) per
 * element: conduction plus c_mass times the heat capacity, rho_e being the
 * uploaded density times the specific heat (get_rtm_dfd0,
 * S/element_equations.rs:1125-1204; c_mass = 1/(gamma dt) gives the
 * operator of a transient step, get_rt_dfd0 :1237-1250).  The temperature
 * of a node is its first DOF (dof_index[6 n]); the same scalar operator sits
 * on its other two DOFs, which carry zero loads and, penalised at the
 * constrained nodes like the first, solve to zero: a PCG over the three
 * copies takes the iterates of the one-DOF system, and a block of 3 b DOFs
 * is the reference's thermal LDL^T block of b nodes.  The penalty scale is
 * taken again from the next assembly (1e5 max|K|, scale_const on the
 * thermal matrix).  Solves take the (block) Jacobi or LDL^T preconditioners;
 * entry points built on the elastic operator (sensitivities, Newmark, modal,
 * nonlinear, multigrid, learned) refuse such a context.  kcond = NULL
 * returns the context to elasticity.  brick8 meshes. */
int asend_set_conductivity(asend_ctx* ctx, const double* kcond);

/* Download the assembled FP64 operator as scalar CSR (columns ascending):
 * rowptr[n_dof+1]; pass cols = vals = NULL to query *nnz only. */
int asend_get_matrix(asend_ctx* ctx, int64_t* rowptr, int32_t* cols, double* vals, int64_t* nnz,
                     int with_penalty);

/* y = K_eff x with the FP64 (precision 64) or FP32 (precision 32) operator:
 * SparseMat::vector_multiply + ConstraintList::get_total_vec_mult
 * (S/list_ent_meth.rs:64-79, S/const_meth.rs:92-103). */
int asend_spmv(asend_ctx* ctx, int precision, const double* x, double* y);

/* Solve K_eff x = rhs: replaces conj_grad_sparse (S/matrix_functions.rs:285-362)
 * and the direct ldl_solve branch of solve_step (S/model_analysis.rs:495-501). */
int asend_solve(asend_ctx* ctx, const double* rhs, double* x, const asend_solve_opts* opts,
                asend_solve_stats* stats);

/* Measurement (no reference counterpart): with enable = 1, every launch of
 * the multigrid cycle's FP16 SpMV with residual epilogue that is not being
 * captured into a graph is bracketed by HIP events on its stream; enable = 0
 * stops and returns the number of launches and their summed kernel time.
 * For timing that kernel inside a whole pipeline (run it with graphs off). */
int asend_cycle_spmv_timing(asend_ctx* ctx, int enable, int64_t* launches, double* total_ms);

/* LDL^T factor of the assembled K_eff restricted to diagonal blocks of
 * block_dofs consecutive DOFs (the last block may be shorter; block_dofs <= 0
 * or >= n_dof: one block, an exact factor).  Replaces LowerTriMat
 * allocate_from_sparse_mat / populate_from_sparse_mat / ldl_factor
 * (S/lower_tri_mat_meth.rs:15-190) with the reference's block size
 * 6 * solverBlockDim (S/model_analysis.rs:475-478, S/model_prep.rs:888-903):
 * each row keeps its columns from its first nonzero inside its block, as the
 * reference's envelope, stored as a band of 64 x 64 tiles on the device.  The
 * factor follows the operator: a solve after a new assembly refactors it
 * with the same blocks.  It is used by ASEND_PC_LDL and ASEND_SOLVE_DIRECT.
 * Fails with ASEND_ERR_UNSUPPORTED when its band does not fit the device
 * memory or the solve's window (a band of more than 240 tiles). */
typedef struct asend_ldl_info {
    int64_t block_dofs;     /* DOFs per block                                   */
    int64_t n_blocks;
    int64_t max_bandwidth;  /* the reference's maxBandwidth: widest row of the
                               envelope, diagonal included                     */
    int64_t bytes;          /* device storage of the factor                     */
    int32_t neg_pivots;     /* pivots d_i < 0                                   */
    int32_t pos_def;        /* LowerTriMat::pos_def (S/lower_tri_mat_meth.rs:229) */
    double factor_ms;       /* fill + factorization, wall clock                 */
} asend_ldl_info;
int asend_ldl_factor(asend_ctx* ctx, int64_t block_dofs, asend_ldl_info* info);

/* x = (L D L^T)^-1 rhs with the current factor, block by block (ldl_solve,
 * S/lower_tri_mat_meth.rs:194-227); host arrays of n_dof doubles. */
int asend_ldl_solve(asend_ctx* ctx, const double* rhs, double* x);

/* Free the factor's device memory (a later ASEND_PC_LDL or direct solve needs
 * asend_ldl_factor again). */
int asend_ldl_release(asend_ctx* ctx);

/* Allocation-free variant of asend_solve for ASEND_SOLVE_PCG_FP64 with
 * ASEND_PC_JACOBI (other methods: ASEND_ERR_UNSUPPORTED).  No reference
 * counterpart: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc/hipFree: the solve's work vectors,
 * reduction partials and loop scalars live in one static device array per
 * GPU, sized for at most 262,144 DOFs (14.8 MB, held by the library on every
 * GPU it runs on until the process ends); a larger model returns
 * ASEND_ERR_ARG.  Same kernels, iterates and stats as asend_solve.  Calls on
 * one GPU share that array, from every context: calls that overlap in time on
 * one GPU race. */
int asend_solve_noalloc(asend_ctx* ctx, const double* rhs, double* x, const asend_solve_opts* opts,
                        asend_solve_stats* stats);

/* Compliance adjoint sensitivities for per-element modulus design variables:
 * dJdE[e] = -lambda_e^T (dK_e/dE_e) u_e, FP64 reduction.  Replaces the DV
 * loop of Model::get_obj_gradient with d_relasticd_d
 * (S/model_analysis.rs:1679-1701, :1379-1457).  brick8 and brickIM (with the
 * condensed bubbles: lambda_e^T (K_e / E_e) u_e, as update_internal's
 * condensed internal DOFs give); not shell4. */
int asend_element_sens(asend_ctx* ctx, const double* u, const double* lam, double* dJdE);

/* Per-element material-parameter sensitivities of brick8 meshes:
 * dJdp[9 e + p] = -lambda_e^T (dK_e/dp) u_e for p = E1, E2, E3, nu12, nu13,
 * nu23, G12, G13, G23 of the orthotropic compliance the reference inverts
 * (get_solid_stiff_dfd0, S/element_properties.rs:728-759), evaluated at the
 * element's isotropic material.  The per-element part of the reference's
 * gradient for design variables of category modulus, poissonRatio and
 * shearModulus (d_relasticd_d, S/model_analysis.rs:1379-1457, in
 * get_obj_gradient :1544-1706); the DV sum over elements and coefficients
 * is the caller's. */
int asend_element_sens_material(asend_ctx* ctx, const double* u, const double* lam, double* dJdp);

/* Shape sensitivities for nodeCoord design variables (the node-coordinate
 * duals of d_relasticd_d, S/model_analysis.rs:1379-1457, with
 * Node::get_crd_dfd0, S/node_meth.rs:255-282): out[24 e + 3 a + c] =
 * lam_e^T (dK_e/dX_{a,c}) u_e for node a (local index) and coordinate c of
 * every element e, u and lam in the DOF numbering (n_dof each).  A nodeCoord
 * DV moving node n along c with coefficient k has dL/dD = -k sum over the
 * (e, a) at node n of out[24 e + 3 a + c], plus lam^T df/dD for loads that
 * depend on the geometry.  brick8. */
int asend_element_sens_coords(asend_ctx* ctx, const double* u, const double* lam, double* out);

/* Element quantities at the brick8 stress point (the element centre, s_cent,
 * S/element_meth.rs:81-83) for stress, strain and strainEnergy objective
 * terms (get_obj_val, S/objective_meth.rs:651-700, with
 * get_stress_strain_dfd0, S/element_soln_fields.rs:1463-1530, and
 * get_volume_dfd0): out[13 e + k] = strain k (11, 22, 33, 12, 13, 23,
 * engineering shear), out[13 e + 6 + k] = stress k (C of the element's E, nu),
 * out[13 e + 12] = element volume (sum over the 8 Gauss points of detJ).
 * u: n_dof displacements.  brick8. */
int asend_element_centre_fields(asend_ctx* ctx, const double* u, double* out);

/* dL/du of centre-point objective terms (d_qd_u of getd_ld_u for the element
 * categories, S/objective_meth.rs:1000-1160): y = sum over elements of
 * B_c^T h_e with h[6 e ..] the strain-space weights of element e (for a
 * stress term, C^T times the stress weights; for strainEnergy, the weight
 * times the stress).  y: n_dof.  brick8. */
int asend_element_centre_adjoint(asend_ctx* ctx, const double* h, double* y);

/* The static compliance pipeline on resident data: assemble, solve K u = f,
 * compliance J = f^T u, adjoint K lambda = dJ/du (= f, so lambda = u without a
 * second solve when the objective is the compliance), dJ/dE.  Replaces
 * Model::solve (static) + Model::get_obj_gradient for a compliance objective
 * (S/model_analysis.rs:560-673, :1544-1706).  device_ms in stats covers the
 * whole pipeline. */
int asend_static_compliance(asend_ctx* ctx, const double* f, double* u, double* dJdE, double* compliance,
                            const asend_solve_opts* opts, asend_solve_stats* stats);

/* ---- geometrically nonlinear statics (brick8) ---------------------------
 * Job flags nonlinearGeom and loadRampSteps (S/job.rs:12,17).  Green-Lagrange
 * strain with the quadratic displacement-gradient term
 * (S/element_soln_fields.rs:1345-1461), residual and consistent tangent with
 * its geometric part (S/element_equations.rs:206-418), isotropic material,
 * dead loads.  Meshes with at most 8 elements per node. */
typedef struct {
    int load_steps;  /* L >= 1: load factors 1 - ((L-1-s)/L)^2, s = 0..L-1
                        (S/model_analysis.rs:636-643); 0 = 1 */
    int max_newton;  /* Newton iterations per load step; 0 = 50, the reference's */
    double du_tol;   /* a load step ends when max|du| <= du_tol (absolute);
                        0 = 1e-12, the reference's */
    double forcing;  /* 0: every Newton solve to opts->tol (the reference's
                        full solves).  > 0: inexact Newton, the first solve of
                        a load step to this relative tolerance (at most 0.5),
                        later ones by Eisenstat-Walker choice 2 (0.9 times the
                        squared ratio of successive load norms, safeguarded),
                        never tighter than opts->tol.  Ignored unless
                        opts->stop_rule is ASEND_STOP_REL_RESIDUAL.  No
                        reference counterpart; the stopping rule is
                        unchanged. */
} asend_nl_opts;

typedef struct {
    int load_steps;
    int newton_iterations;     /* over all load steps */
    int max_step_newton;       /* most Newton iterations of one load step */
    int unconverged_steps;     /* load steps that ended at max_newton */
    int64_t linear_iterations; /* PCG iterations of all Newton solves */
    int unconverged_solves;    /* Newton solves whose PCG missed its tolerance */
    double last_du;            /* max|du| of the last Newton iteration */
    double device_ms;          /* whole solve, HIP events */
} asend_nl_stats;

/* Tangent stiffness of the internal force at displacement u (n_dof, DOF
 * order) assembled into the context operator with the penalty (as
 * asend_assemble does for the linear one, so asend_get_matrix / asend_solve /
 * asend_spmv see it), and the internal force r(u) into r_int (may be NULL):
 * build_elastic_soln_load(true) with nonlinearGeom (S/model_analysis.rs:179-201). */
int asend_nl_assemble(asend_ctx* ctx, const double* u, double* r_int);

/* Static solve with geometric nonlinearity: for each load step, Newton
 * iterations K_T(u) du = fact f - r(u) - sf P u, u += du, until
 * max|du| <= du_tol (the elastic branch of Model::solve_step with
 * nonlinear_geom, S/model_analysis.rs:450-546, under the load ramp of
 * Model::solve, :633-645).  u: starting displacement on entry (zeros for an
 * unloaded model), converged displacement on return.  Each Newton solve uses
 * opts (any method and preconditioner of asend_solve; the tangent is
 * re-assembled every iteration, so multigrid and learned setups are redone).
 * The penalty scale is fixed on the context's first assembly.
 * brick8 and brickIM meshes.  brickIM: the 9 bubble amplitudes of each
 * element enter the displacement gradient like nodal values; every Newton
 * iteration condenses them (K_ee - K_ei K_ii^-1 K_ie, r_e - K_ei K_ii^-1 r_i:
 * condense_mat and update_external, S/element_equations.rs:19-121) and moves
 * them by -K_ii^-1 (r_i + K_ie du) after the solve (update_internal,
 * :123-194).  They stay on the device between calls, as the reference's
 * elements keep them, and are zero after asend_upload_mesh; asend_nl_assemble
 * reports the condensed tangent and internal force at the amplitudes held. */
int asend_static_nonlinear(asend_ctx* ctx, const double* f, double* u, const asend_nl_opts* nl,
                           const asend_solve_opts* opts, asend_nl_stats* stats);

/* brickIM bubble amplitudes, 9 per element in element order, 3 k + component
 * (the reference's internal DOFs Element::internal_disp in dof_table order,
 * S/element_meth.rs:193-201; written after the nodes in the state files,
 * S/model_output.rs:70-78).  With u (n_dof): those that follow a linear solve
 * reaching u from rest, -K_ii^-1 K_ie u_e (update_internal,
 * S/element_equations.rs:123-194); with u = NULL: the amplitudes held by the
 * nonlinear path.  brickIM meshes. */
int asend_brickim_internal(asend_ctx* ctx, const double* u, double* alpha);
/* Replace the held amplitudes (read_time_step_soln followed by a backstep). */
int asend_brickim_set_internal(asend_ctx* ctx, const double* alpha);

/* Timing helper: average duration (ms) of n_iter back-to-back launches of the
 * SpMV kernel of the given precision, HIP events on the context stream. */
int asend_bench_spmv(asend_ctx* ctx, int precision, int n_iter, double* avg_ms);

/* Timing helper: average duration (ms) of a streaming copy of `bytes` bytes
 * (read + write), for the achievable-bandwidth figure and for calibrating
 * HBM traffic counters. */
int asend_bench_copy(asend_ctx* ctx, int64_t bytes, int n_iter, double* avg_ms);

/* ---- batched multi-RHS solves over design variants ----------------------
 * No reference counterpart (the reference solves one model and one load at a
 * time, S/model_analysis.rs:560-673).  Each variant is an assembled brick8
 * context; their operators are laid side by side as one block-diagonal
 * sliced-ELL operator and up to 16 right-hand sides per variant advance in
 * lockstep: FP32 Jacobi PCG (one SpMM launch per iteration for the whole
 * batch, per-system coefficients) inside FP64 iterative refinement. */
typedef struct asend_batch asend_batch;
typedef struct asend_batch_stats {
    int64_t outer_iterations;
    int64_t lockstep_iterations;  /* inner iterations of the slowest system, summed over refinements */
    int64_t system_iterations;    /* inner iterations summed over all systems */
    int64_t unconverged;          /* systems above tol after the last refinement */
    double max_rel_residual;      /* max over systems of ||f - K u|| / ||f|| (FP64) */
    double device_ms;
} asend_batch_stats;
/* The contexts must stay alive and assembled while the batch exists (their
 * operators are copied; their device arrays are not borrowed afterwards). */
int asend_batch_create(asend_ctx* const* variants, int n_variants, asend_batch** out);
void asend_batch_destroy(asend_batch* b);
const char* asend_batch_error(const asend_batch* b);
/* rhs[v] and x[v]: nrhs x n_dof_v, right-hand side major (rhs[v][j * n_dof_v + d]).
 * verbose != 0 prints one progress line per refinement to stderr. */
int asend_batch_solve(asend_batch* b, int nrhs, const double* const* rhs, double* const* x, double tol,
                      double inner_tol, int64_t max_it, int check_every, int verbose, asend_batch_stats* stats);

/* ---- Newmark dynamics ---------------------------------------------------
 * Replaces the dynamic branch of Model::solve (S/model_analysis.rs:560-615)
 * with solve_step's elastic solve (:312-558) and the node kinematics
 * update_vel_acc / advance_disp (S/node_meth.rs:127-199), in the reference's
 * (beta - gamma) form: c1 = 1/(dt^2 (beta - gamma)), c2 = dt^2 (0.5 + beta -
 * gamma), a(u) = c1 (u_n - u + dt v_n + c2 a_n); each step solves
 * (K + c_m M) du = f(t_n) - K u_n - M a(u_n) with c_m = -c1 (the load of the
 * step is taken at its start time, as solve_step(time) does), then
 * u += du, a = a(u), v += dt ((1 - gamma) a_n + gamma a).  State, effective
 * load and update in FP64; each step is one FP64 Jacobi PCG solve on the
 * device, and the whole step is replayed as a captured hipGraph. */
typedef struct asend_newmark_stats {
    int64_t steps;
    int64_t total_iterations;     /* PCG iterations over all steps            */
    int64_t max_step_iterations;
    int64_t unconverged_steps;    /* steps that hit max_it_step before tol    */
    int graph_used;               /* 1 when the steps ran as graph replays    */
    double device_ms;             /* HIP-event time of the step loop          */
    int step_solver;              /* ASEND_NM_FP64 or ASEND_NM_MIXED as run   */
} asend_newmark_stats;
/* Step solver of asend_newmark_run: ASEND_NM_FP64 runs a single-reduction
 * Jacobi PCG in FP64; ASEND_NM_MIXED runs the same
 * recurrence in FP32 (FP32 operator and vectors, FP64 reductions) with the
 * increment accumulated in FP64 and reliable updates: whenever the recursive
 * residual has fallen by ru_factor (1e-6 <= ru_factor < 1, default 0.001)
 * since the last refresh, the residual rhs - K_eff du is recomputed in FP64
 * and reloaded, keeping the search directions.  Either stops when the FP64
 * residual satisfies ||rhs - K_eff du|| <= tol ||rhs||.  ASEND_NM_AUTO (the
 * default) takes the mixed solver when the FP64 operator exceeds 128 MiB and
 * FP64 otherwise: an operator that stays resident in the 256 MB Infinity
 * Cache gains little from FP32 storage (measured: 256x256 shell plate,
 * 181 MB: mixed 19% faster; 100x20x20 brick8, 91 MB: FP64 14% faster).
 * Takes effect at the next asend_newmark_run. */
#define ASEND_NM_FP64 0
#define ASEND_NM_MIXED 1
#define ASEND_NM_AUTO 2
int asend_newmark_set_solver(asend_ctx* ctx, int mode, double ru_factor);
/* Rayleigh damping D = cm M + ck K for the next asend_newmark_setup (job keys
 * rayDampCM / rayDampCK, JobCommand::ray_damp_cm / ray_damp_ck,
 * S/job.rs:15-16; Element::get_rud_dfd0, S/element_equations.rs:699-753).
 * cm, ck >= 0; default 0.  ck > 0 is implemented for brick8 meshes (the
 * reference adds it to the uncondensed external block of elements with
 * internal DOFs); cm for every element type.  Material damping matrices:
 * asend_set_material_damping. */
int asend_set_rayleigh_damping(asend_ctx* ctx, double cm, double ck);
/* Material damping for the next asend_newmark_setup (material key damping,
 * Material::set_damping, S/section_meth.rs:73-77; get_solid_damp_dfd0,
 * S/element_properties.rs:1537-1575; the material-damping force of
 * get_rud_dfd0, S/element_equations.rs:758-850): n_mat symmetric 6 x 6
 * matrices (row-major, strain order 11, 22, 33, 12, 13, 23 with engineering
 * shear, as the stiffness), and the matrix of each element (el_mat, n_el
 * entries; NULL: every element uses matrix 0).  Each element adds
 * sum over its Gauss points of w detJ B^T D_m B to D (beside the Rayleigh
 * terms).  brick8 meshes; n_mat = 0 removes them. */
int asend_set_material_damping(asend_ctx* ctx, int64_t n_mat, const double* dmat, const int32_t* el_mat);
/* Assemble M (consistent, rho per element), D (with damping) and
 * K_eff = K + c_m M + c_m dt gamma D, c_m = 1/(dt^2 (gamma - beta)), with the
 * penalty of asend_set_penalty (Element::get_ru_dfd0,
 * S/element_equations.rs:948-1003); requires gamma > beta.  brick8, brickIM
 * and shell4 meshes uploaded with rho (shell4: after asend_set_thickness).
 * D = cm M + ck K plus the material damping of asend_set_material_damping. */
int asend_newmark_setup(asend_ctx* ctx, double dt, double beta, double gamma);
/* n_steps steps from (u0, v0, a0) (null = zero).  The load of step n is
 * load_fac[n] * f_shape; the step's residual is f - K u - M a(u) - D v(u) at
 * the predicted state (S/model_analysis.rs:312-558, S/node_meth.rs:127-199).  hist[n * n_watch + k] = u[watch[k]] after step n.
 * Each step's solve stops at ||r|| <= tol ||rhs|| or max_it_step iterations
 * (the mixed solver also stops after 200 FP64 residual refreshes); it starts
 * from the previous step's increment.
 * Final u, v, a into the out arrays when non-null. */
int asend_newmark_run(asend_ctx* ctx, int64_t n_steps, const double* f_shape, const double* load_fac,
                      const double* u0, const double* v0, const double* a0, int64_t n_watch, const int64_t* watch,
                      double* hist, double* u_out, double* v_out, double* a_out, double tol, int max_it_step,
                      asend_newmark_stats* stats);

/* Transient adjoint: replaces the dynamic branch of Model::get_obj_gradient
 * (S/model_analysis.rs:1563-1631) with augmentd_ld_u (:936-1063) and the
 * dynamic elastic part of solve_for_adjoint (:1089-1100), with the Rayleigh
 * damping of the setup, for the objective L = sum_n w[n] q^T u_n (displacement objective terms of exponent
 * 1, S/objective_meth.rs:618-663).  Runs the forward sweep from rest
 * (load_fac[n] f_shape as asend_newmark_run, keeping u_n of every step in
 * HBM: n_steps * n_dof doubles), then for n = N..1
 *   t = a_adj - M lam_{n+1}, d = D lam_{n+1} (not on the last step),
 *   v_adj = c1 dt t + v_adj - dvv d,
 *   a_adj = c1 c2 t + dt (1 - gamma) v_adj_old - dva d + dt gamma v_adj,
 *   K_eff lam_n = w[n] q + c1 t - dvu d - c1 a_adj  (FP64 Jacobi PCG to tol,
 *   started from lam_{n+1}),
 *   dJdE += -lam_n^T (dK/dE) (u_n + ck v_n),
 * with c1 = 1/(dt^2 (beta - gamma)), c2 = dt^2 (0.5 + beta - gamma),
 * dvu = dt gamma c1, dva = dt (1 - gamma) + dt gamma c1 c2,
 * dvv = 1 + dt^2 gamma c1.  The
 * reference evaluates every step's objective terms at the final time; for
 * these objectives the time does not enter.  brick8 and brickIM meshes. */
int asend_newmark_adjoint(asend_ctx* ctx, int64_t n_steps, const double* f_shape, const double* load_fac,
                          const double* q, const double* w, double tol, int max_it_step, double* J, double* dJdE,
                          asend_newmark_stats* fwd_stats, asend_newmark_stats* adj_stats);

/* ---- natural frequencies -------------------------------------------------
 * Replaces Model::eigen_solve for type "frequency" (S/model_analysis.rs:713-888)
 * and eigen_sparse_direct (S/matrix_functions.rs:1099-1265): the lumped mass
 * m = M 1 (row sums of the consistent mass; entries below 1e-6 of the mean
 * raised to that value), the shifted operator K_eff - shift diag(m) (penalty
 * included), and Lanczos on diag(sqrt m) (K_eff - shift diag m)^-1 diag(sqrt m)
 * with min(max(5 n_modes, n_modes + 20), 255) vectors in HBM and full
 * reorthogonalisation (the reference orthogonalises against two vectors and
 * filters repeated Ritz vectors).  Each operator application is one solve
 * with opts (e.g. the multigrid mixed solver); the solves are iterative, so
 * the shifted operator must stay positive definite (shift below the lowest
 * eigenvalue; the reference factors it directly).  eig_vals[k] = shift +
 * 1/theta_k for the n_modes largest Ritz values theta, ascending in
 * eigenvalue; freqs_hz[k] = sqrt(eig_vals[k]) / (2 pi), or -1 for a negative
 * eigenvalue; eig_vecs (optional, n_modes x n_dof) are the Ritz vectors
 * diag(1/sqrt m) y normalised to unit 2-norm.  With a nonzero shift the
 * operator stays shifted: assemble again before a static solve. */
typedef struct asend_modal_stats {
    int64_t lanczos_steps;
    int64_t modes_found;
    int64_t solve_iterations;   /* inner iterations over all operator applications */
    double max_ritz_residual;   /* max |beta_m z_m| / theta over the returned modes */
    int64_t masses_raised;      /* lumped masses raised to 1e-6 of the mean */
    int64_t unconverged_solves; /* operator applications that missed opts->tol */
    double device_ms;
} asend_modal_stats;
int asend_modal(asend_ctx* ctx, int n_modes, double shift, const asend_solve_opts* opts, double* eig_vals,
                double* freqs_hz, double* eig_vecs, asend_modal_stats* stats);

/* Model::eigen_solve with its state (S/model_analysis.rs:713-888): the
 * operator is the tangent stiffness at the displacement u (n_dof; the
 * reference builds it with nonlinear_geom at the current state, so a
 * preceding static solve stiffens or softens the modes; brickIM uses the
 * bubble amplitudes held, asend_brickim_set_internal), or the linear
 * stiffness when u is NULL.  kind 0 (type frequency): as asend_modal,
 * factors[k] = frequency in Hz.  kind 1 (type buckling): identity mass
 * (diag_mass = 1), eigenvalues of K_T(u) - shift I plus the shift, and
 * factors[k] = v / (v - eig_vals[k]) with v = phi_k^T K phi_k, K the
 * stiffness at zero displacement without the penalty and |phi_k| = 1 (the
 * reference's load_fact: the multiple of the applied load at which the mode
 * buckles; 1e100 when v = eig_vals[k]).  Buckling leaves the linear
 * stiffness assembled. */
int asend_eigen(asend_ctx* ctx, int kind, const double* u, int n_modes, double shift, const asend_solve_opts* opts,
                double* eig_vals, double* factors, double* eig_vecs, asend_modal_stats* stats);

/* ---- multigrid diagnostics (no reference counterpart) --------------------
 * Host FP32 arrays.  Level 0 uses the fine DOF numbering, coarser levels the
 * lexicographic node order of their grid.  Used to verify on the device that
 * A_{l+1} = P^T A_l P and that one V-cycle is a symmetric positive operator. */
int asend_mg_info(asend_ctx* ctx, int* n_levels, int64_t* level_ndof /* [16] */);
int asend_mg_apply(asend_ctx* ctx, int level, const float* x, float* y);
int asend_mg_restrict(asend_ctx* ctx, int level, const float* fine, float* coarse);
int asend_mg_prolong(asend_ctx* ctx, int level, const float* coarse, float* fine);
int asend_mg_vcycle(asend_ctx* ctx, const float* b, float* x);
/* x = A_c^-1 b with the explicit inverse of the coarsest operator. */
int asend_mg_coarse_solve(asend_ctx* ctx, const float* b, float* x);
/* Replay the launch sequences of the iterative loops (a block of check_every
 * FCG iterations, a Newmark time step) as captured hipGraphs (on = 1, the
 * default) or launch the same kernels one by one (on = 0).  Results are
 * identical either way. */
int asend_set_graphs(asend_ctx* ctx, int on);
/* Autotuner (no reference counterpart).  ASEND_TUNE_SPMV times the fine SpMV
 * with 1, 2, 4 and 8 slices (wavefronts) per workgroup and keeps the fastest;
 * ASEND_TUNE_SOLVER times full multigrid-preconditioned mixed solves of rhs
 * over Chebyshev degree {1, 2, 3} x reliable-update factor (inner_tol)
 * {1e-1, 1e-2, 1e-3, 1e-4} x check interval {1, 2, 4} (the faster of two
 * solves each) and keeps the fastest converged choice; a later solve uses the
 * tuned inner tolerance / check interval where its options pass 0.  With
 * ASEND_TUNE_LEARNED as well, the same grid is timed with the learned
 * smoother (ASEND_PC_LEARNED) and without it, and the fastest choice of either
 * is kept: when it is the one without the network, the network steps aside
 * (learned_active = 0; ASEND_PC_LEARNED solves report fallback_used = 2). */
#define ASEND_TUNE_SPMV 1
#define ASEND_TUNE_SOLVER 2
#define ASEND_TUNE_LEARNED 4
typedef struct asend_tuning {
    int spmv_waves_per_block;  /* chosen geometry */
    double spmv_ms_fp32[4];    /* measured, for 1, 2, 4, 8 waves (0 = not tried) */
    double spmv_ms_fp64[4];
    int cheb_degree;
    double inner_tol;
    int check_every;
    double solve_ms;           /* best tuned solve */
    int n_trials;
    /* ASEND_TUNE_LEARNED: 1 when the learned pipeline was the faster one; 0
     * when the same solve without the network was (ASEND_PC_LEARNED solves
     * then run the multigrid smoother alone and report fallback_used = 2,
     * until the network changes) */
    int learned_active;
    double solve_ms_learned;     /* best learned solve (learned tuning) */
    double solve_ms_no_network;  /* best solve without the network */
} asend_tuning;
int asend_autotune(asend_ctx* ctx, const double* rhs, int what, asend_tuning* out);

/* Chebyshev smoother of every smoothed level: M = point Jacobi (kind 0) or the
 * inverse of the 3x3 node blocks (kind 1), degree 1..8, interval
 * [lo_frac, hi_frac] * lambda_max(M A).  Defaults 1, 2, 0.1, 1.15.  Kept across
 * assemblies of the same hierarchy; a new kind rebuilds the hierarchy's
 * smoother data at the next solve. */
int asend_mg_set_smoother(asend_ctx* ctx, int kind, int degree, double lo_frac, double hi_frac);

/* Arithmetic of the finest level inside the multigrid cycle: 16 (default)
 * reads an FP16 copy of the operator whose rows are scaled by their diagonal
 * (entries O(1) whatever the moduli and penalty), FP32 accumulation, for the
 * smoothing and restriction residuals and the smoother bounds; 32 reads the
 * FP32 operator.  The flexible CG around the cycle keeps its FP32 operator
 * products and the FP64 residual refreshes decide convergence either way. */
int asend_mg_set_cycle_precision(asend_ctx* ctx, int bits);

/* ---- learned preconditioner (no reference counterpart) --------------------
 * A 3-layer message-passing network on the mesh node graph (hidden 64, BF16
 * weights and activations on MFMA, FP32 accumulation) maps node features
 * (element moduli around the node, degree, clamp flag, diagonal-block shape)
 * and edge features (the diagonally scaled operator block of each mesh edge)
 * to a sparse factor L with one 3x3 block per node and one per mesh-edge
 * neighbour; P = L L^T is SPD.  ASEND_PC_LEARNED uses P as the Chebyshev
 * preconditioner of the finest multigrid level inside flexible PCG; when the
 * learned solve breaks down or needs more than its inner budget the solver
 * switches to Jacobi PCG and sets stats.fallback_used.  The FP64 outer
 * residual decides convergence in every case.  The network lives in the
 * context and survives asend_upload_mesh, so one network is trained over
 * several meshes. */
int asend_learned_num_params(void);
/* Create the network (deterministic init from seed; untrained it gives L = C^-T,
 * block Jacobi). */
int asend_learned_init(asend_ctx* ctx, uint64_t seed);
int asend_learned_get_params(asend_ctx* ctx, float* out, int n);
int asend_learned_set_params(asend_ctx* ctx, const float* in, int n);
/* n_steps Adam steps (learning rate lr) on the current assembled mesh.
 * mode 0: one pair (r = K e, e) per step, e Gaussian (zero on penalised
 *   nodes); loss = energy left by the best damped step,
 *   min_w ||e - w P K e||_K^2 / ||e||_K^2 (independent of the scale of P).
 * mode 1: solver-generated samples only: the steps alternate the error
 *   e = (I - B K) e0 left by one multigrid V-cycle B of the solver and the
 *   approximate solution e = B b of a random load b (same loss).
 * mode 3: as mode 1 with a Gaussian e before every such pair.
 * mode 2: end to end through the solver: P is the finest Chebyshev smoother
 *   of one V-cycle B; loss = ||(I - B K) e||_K^2 / ||e||_K^2 for a Gaussian e,
 *   differentiated through every use of P in the cycle.
 * losses[k] receives each step's loss when non-null. */
int asend_learned_train(asend_ctx* ctx, int n_steps, float lr, uint64_t seed, int mode, double* losses);
/* Diagnostics: loss and parameter gradient for the sample e (n_dof), node
 * features (n_nodes x 8) and head outputs (n_nodes x 24); null outputs skipped.
 * Node order is the DOF-block order. */
int asend_learned_eval(asend_ctx* ctx, const double* e, double* loss, float* grad, float* feat, float* out);
/* Diagnostics of training mode 2: loss of one V-cycle on the sample e (n_dof)
 * and its parameter gradient (grad may be null).  keep_bounds != 0 reuses the
 * Chebyshev bounds of the previous call instead of re-estimating them, which
 * is what the gradient assumes (finite-difference checks). */
int asend_learned_vcycle_eval(asend_ctx* ctx, const double* e, int keep_bounds, double* loss, float* grad);
/* z = L L^T r with the factor of the current operator (FP32 inside). */
int asend_learned_apply(asend_ctx* ctx, const double* r, double* z);
/* Mesh-edge neighbour tables, [n_edge][n_nodes] each (-1 = none), as used by L. */
int asend_learned_edges(asend_ctx* ctx, int32_t* enb, int32_t* erev, int* n_edge);

#ifdef __cplusplus
}
#endif

#endif /* ASEND_GPU_H */
