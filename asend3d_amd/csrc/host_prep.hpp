// Internal declarations shared by the host-side preparation and the device
// context code.
#pragma once
#include <cstdint>

namespace asend {

int reorder_nodes(int64_t n_nodes, const double* xyz, int64_t n_el, int nn, int dof_per_nd, const int32_t* conn,
                  int64_t n_const, const int64_t* cterm_ptr, const int64_t* term_set_ptr, const int32_t* set_nodes,
                  int64_t block_dim, int32_t* ordered, int64_t* dof_index, int64_t* n_dof_out);

int surface_traction(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                     const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* load,
                     const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f);

int surface_pressure(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                     const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double pressure,
                     const double* norm_dir, double norm_tol_deg, int64_t n_dof, double* f);

int surface_flux(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                 const int64_t* dof_index, int64_t n_set, const int32_t* el_set, double q, const double* norm_dir,
                 double norm_tol_deg, int64_t n_dof, double* f);

int thermal_strain_load(int elem_type, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
                        const int64_t* dof_index, const double* E, const double* nu, const double* cmat,
                        const int32_t* cmat_idx, const double* alpha, const double* temp, int64_t n_dof, double* f);

int body_load(int elem_type, int kind, int64_t n_nodes, const double* xyz, int64_t n_el, const int32_t* conn,
              const int64_t* dof_index, int64_t n_set, const int32_t* el_set, const double* rho, const double* load,
              const double* center, const double* axis, double ang_vel, int64_t n_dof, double* f);

int lognormal_moduli(uint64_t seed, int64_t n, double median, double sigma_ln, double* out);

// Number of elements whose isotropic compliance the reference's pivot-clamped
// QR inverse would not invert exactly; *first gets the first such label
// (-1 when none).
int64_t compliance_clamp_count(int elem_type, int64_t n_el, const double* E, const double* nu, int64_t* first);

}  // namespace asend
