/*
 * asend_oracle.h -- CPU FP64 restatement of the ASenD3D solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (asend3d_amd/, include/asend_gpu.h) never links, loads or calls it.
 *
 * Every function restates the reference algorithm it names (file:line relative
 * to the reference checkout; S/ = solverSrc/asend_run_job/src/).  It is written
 * from the mathematics of those routines, serially, in FP64, the way the
 * reference runs them.
 *
 * Conventions shared with the product ABI:
 *   - node coordinates xyz[3*n + c];
 *   - element connectivity conn[nn*e + a] (reference local node order);
 *   - dof_index[6*n + d] (global DOF of node n, component d; -1 if unused);
 *   - CSR matrices with int64 row pointers, int32 columns sorted ascending.
 * All entry points return 0 on success and a negative code on error.
 */
#ifndef ASEND_ORACLE_H
#define ASEND_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Model::reorder_nodes (S/model_prep.rs:56-303) + Model::add_constraint_conn
 * (S/model_prep.rs:20-54).  Constraints are given as terms: constraint c has
 * terms [cterm_ptr[c], cterm_ptr[c+1]); term t's node set is
 * set_nodes[term_set_ptr[t] .. term_set_ptr[t+1]).
 * Outputs: ordered[n_nodes] (insertion order), dof_index[6*n_nodes],
 * n_dof_out (= el_mat_dim).
 * Parity unpinned: the reference ships no node orderings and its Rust cannot
 * run here; tests pin only invariants (a permutation, a least-connected start
 * node, contiguous DOF blocks, restart blocks of block_dim), not the exact
 * order. */
int orc_reorder_nodes(int64_t n_nodes, const double* xyz,
                      int64_t n_el, int nn, int dof_per_nd, const int32_t* conn,
                      int64_t n_const, const int64_t* cterm_ptr,
                      const int64_t* term_set_ptr, const int32_t* set_nodes,
                      int64_t block_dim,
                      int32_t* ordered, int64_t* dof_index, int64_t* n_dof_out);

/* Solid material matrix: inverse of the orthotropic compliance built in
 * Element::get_solid_stiff_dfd0 (S/element_properties.rs:648-764).
 * mod[3]=E1,E2,E3  pr[3]=nu12,nu13,nu23  sm[3]=G12,G13,G23.  C is 6x6 row-major. */
int orc_solid_cmat(const double* mod, const double* pr, const double* sm, double* C);

/* Element stiffness/residual for brick8 (type 8) and brickIM (type 81),
 * Element::get_ruk_dfd0 (S/element_equations.rs:206-418) with
 * get_ip_data_dfd0 (S/element_soln_fields.rs:383-463) and, for type 81,
 * static condensation condense_mat (S/element_equations.rs:19-57).
 *   X     : element nodal coordinates SoA [x0..x7 | y0..y7 | z0..z7]
 *   ori   : 3x3 local orientation (row-major), NULL = identity
 *   C     : 6x6 material matrix
 *   ue    : 24 nodal displacements, node-major (node a, comp d at 3a+d); NULL = 0
 *   K     : out 24x24 (condensed for type 81), row-major
 *   R     : out 24 residual (internal force), may be NULL
 * For type 81 the internal DOFs are taken as zero (linear static solve from
 * the zero state, S/model_prep.rs:602-605). */
int orc_solid_element(int type, const double* X, const double* ori, const double* C,
                      const double* ue, double* K, double* R);

/* Same element, differentiated: dK/dE (24x24) for an isotropic modulus
 * parameter E entering as E1=E2=E3=E, G = E/(2(1+nu)), evaluated with
 * forward dual numbers exactly as DiffDoub1 does in d_relasticd_d
 * (S/model_analysis.rs:1379-1457, S/diff_doub_meth.rs:82-185). */
int orc_solid_element_dE(int type, const double* X, const double* ori, double E, double nu,
                         double* K, double* dK);
int orc_solid_element_dmat(int type, const double* X, const double* mod, const double* pr, const double* sm,
                           int param, double* K, double* dK);

/* Consistent solid mass (rho * int N^T N), Element::get_rum_dfd0 solid branch
 * (S/element_equations.rs:583-609).  M out 24x24. */
int orc_solid_mass(int type, const double* X, double rho, double* M);

/* Assemble the global elastic stiffness into CSR:
 * build_elastic_soln_load(true) + get_ru_dfd0 scatter (S/model_analysis.rs:179-201,
 * S/element_equations.rs:1040-1059).  Homogeneous mesh of `type` (8 or 81),
 * isotropic per-element E[e], nu[e] (G = E/(2(1+nu))).
 * rowptr[n_dof+1] and cols[] are the sorted union of element couplings; pass
 * cols=NULL/vals=NULL to query nnz_out only. */
int orc_assemble_solid(int type, int64_t n_nodes, const double* xyz, int64_t n_el,
                       const int32_t* conn, const double* E, const double* nu,
                       const int64_t* dof_index, int64_t n_dof,
                       int64_t* rowptr, int32_t* cols, double* vals, int64_t* nnz_out);

/* Consistent mass assembled on the stiffness pattern (same rowptr/cols). */
int orc_assemble_solid_mass(int type, int64_t n_nodes, const double* xyz, int64_t n_el,
                            const int32_t* conn, const double* rho,
                            const int64_t* dof_index, int64_t n_dof,
                            const int64_t* rowptr, const int32_t* cols, double* vals);

/* Surface traction -> nodal loads, Element::get_app_load_dfd0 surfaceTraction
 * branch (S/element_equations.rs:1833-2056) with find_surface_faces
 * (S/model_prep.rs:510-554) and Face::get_area_normal_dfd0 (S/face_meth.rs:174-221).
 * el_set lists the load's elements; f[n_dof] is accumulated (+=). */
int orc_surface_traction(int type, int64_t n_nodes, const double* xyz, int64_t n_el,
                         const int32_t* conn, const int64_t* dof_index,
                         int64_t n_set, const int32_t* el_set,
                         const double* load, const double* norm_dir, double norm_tol_deg,
                         double* f);
int orc_surface_pressure(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                         const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double pressure,
                         const double* norm_dir, double norm_tol_deg, double* f);
int orc_body_load(int type, int kind, int64_t n_el, const double* xyz, const int32_t* conn,
                  const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* rho,
                  const double* load, const double* center, const double* axis, double ang_vel, double* f);

/* SparseMat::get_max_abs_val (S/list_ent_meth.rs:81-94). */
double orc_max_abs(int64_t nnz, const double* vals);

/* y = (K + sf * sum_c C^T C) x for single-DOF penalty rows (clamps):
 * SparseMat::vector_multiply + ConstraintList::get_total_vec_mult
 * (S/list_ent_meth.rs:64-79, S/const_meth.rs:92-103,161-168). */
int orc_spmv_penalty(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals,
                     int64_t n_pen, const int64_t* pen_dofs, double sf,
                     const double* x, double* y);

/* LowerTriMat allocate/populate/factor (S/lower_tri_mat_meth.rs:15-192) on
 * K + sf*C^T C with blocks of block_dofs DOFs.  The factor is kept in an
 * opaque handle used by orc_ldl_solve / orc_conj_grad_sparse. */
typedef struct orc_ldl orc_ldl;
int orc_ldl_build(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals,
                  int64_t n_pen, const int64_t* pen_dofs, double sf, int64_t block_dofs,
                  orc_ldl** out);
int orc_ldl_solve(orc_ldl* L, const double* rhs, double* x);   /* S/lower_tri_mat_meth.rs:194-227 */
int orc_ldl_pos_def(orc_ldl* L);                               /* S/lower_tri_mat_meth.rs:229-241 */
int64_t orc_ldl_max_bandwidth(orc_ldl* L);
void orc_ldl_free(orc_ldl* L);

/* conj_grad_sparse (S/matrix_functions.rs:285-362): PCG from x0 = 0,
 * preconditioned by the LDL handle, stopping when g^T M^-1 g <= conv_tol
 * (absolute) or after max_it iterations.  it_out / res_out report the
 * iteration count and final g^T M^-1 g. */
int orc_conj_grad_sparse(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals,
                         int64_t n_pen, const int64_t* pen_dofs, double sf, orc_ldl* pc,
                         const double* rhs, double conv_tol, int64_t max_it,
                         double* x, int64_t* it_out, double* res_out);

/* Plain Jacobi PCG on K + sf*C^T C stopping on ||f - K x||_2 / ||f||_2 <= rel_tol
 * (the metric's stopping rule; not a reference routine, the CPU twin of the
 * device FP64 Jacobi baseline). */
/* OpenMP threads of orc_jacobi_pcg and orc_spmv_penalty (default 1: the
 * serial restatement).  For the multi-core CPU baseline only. */
int orc_set_threads(int n);

int orc_jacobi_pcg(int64_t n, const int64_t* rowptr, const int32_t* cols, const double* vals,
                   int64_t n_pen, const int64_t* pen_dofs, double sf,
                   const double* rhs, double rel_tol, int64_t max_it,
                   double* x, int64_t* it_out, double* relres_out);

/* Compliance adjoint sensitivities for per-element isotropic modulus design
 * variables: dJ/dE_e = dL/dd - lambda^T dR/dd with dR/dd from the dual-number
 * element residual (d_relasticd_d, S/model_analysis.rs:1379-1457; DV loop
 * S/model_analysis.rs:1679-1701).  u and lam are global DOF vectors. */
int orc_solid_sens(int type, int64_t n_nodes, const double* xyz, int64_t n_el,
                   const int32_t* conn, const double* E, const double* nu,
                   const int64_t* dof_index, const double* u, const double* lam, double* dJdE);


/* The reference's design-variable loop itself, for per-element modulus DVs
 * (DV d acts on element d, d < n_dv): per DV the O(N) sweeps of
 * d_relasticd_d (S/model_analysis.rs:1379-1457) -- node coordinate duals
 * twice, the dual residual vector over every DOF zeroed and negated, the
 * element flags reset, the load elements scanned -- the dual-number residual
 * of the DV's element scattered into it, then dL/dd -= lam . dR/dd over every
 * DOF and the internal-DOF adjoint term of every element (get_obj_gradient,
 * :1679-1701).  Same values as orc_solid_sens for those elements (up to the
 * order of summation); its cost is the reference's O(N_dv (N_dof + N_el +
 * N_nodes)). */
int orc_dv_loop(int type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn, const double* E,
                const double* nu, const int64_t* dof_index, int64_t n_dof, const double* u, const double* lam,
                int64_t n_load_el, const int32_t* load_el, int64_t n_dv, double* dJdE);

/* shell4 (type 41) single isotropic layer of thickness t at the mid-plane:
 * condensed 24x24 stiffness and 24x24 consistent mass of one element
 * (S/element_equations.rs:206-418, :420-626, S/element_soln_fields.rs:1160-1343,
 * S/element_properties.rs:766-876, :1354-1408).  X is 3 x 4 SoA.  Elements
 * whose normal is not parallel to z return -2 (the identity element frame is
 * the only one restated). */
int orc_shell4_element(const double* X, double E, double nu, double t, double rho, double* Ke, double* Me);
/* Shell assembly into CSR (6 DOF per node, dof_index[6n + d]); mvals (may be
 * null) receives the consistent mass on the same pattern. */
int orc_assemble_shell(int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn, const double* E,
                       const double* nu, const double* t, const double* rho, const int64_t* dof_index, int64_t n_dof,
                       int64_t* rowptr, int32_t* cols, double* vals, double* mvals, int64_t* nnz_out);

#ifdef __cplusplus
}
#endif
#endif
