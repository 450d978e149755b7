"""ctypes binding of libasend_gpu.so (the C ABI declared in include/asend_gpu.h).

There is no CPU fallback: if the library is missing this module raises at
import, and device calls raise when the HIP runtime reports an error.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libasend_gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python -m asend3d_amd.build` (hipcc, gfx950)")

_lib = ctypes.CDLL(LIB_PATH)

BRICK8 = 8
BRICKIM = 81
SHELL4 = 41

SOLVE_PCG_FP64 = 0
SOLVE_IR_MIXED = 1
SOLVE_DIRECT = 2  # one-block LDL^T factor (asend_ldl_factor), the reference's direct solverMethod
NM_FP64 = 0   # Newmark step solver: FP64 single-reduction Jacobi PCG (default)
NM_MIXED = 1  # FP32 single-reduction Jacobi PCG, FP64 increment, reliable updates
NM_AUTO = 2   # mixed when the FP64 operator exceeds 128 MiB, else FP64 (default)

PC_JACOBI = 0
PC_BLOCK_JACOBI = 1
PC_MG = 2
PC_LEARNED = 3
PC_LEARNED_PCG = 4
PC_LDL = 5  # block LDL^T factor of asend_ldl_factor (the reference's LowerTriMat preconditioner)

STOP_REL_RESIDUAL = 0
STOP_REF_ABS = 1

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class SolveOpts(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("precond", ctypes.c_int), ("stop_rule", ctypes.c_int),
                ("check_every", ctypes.c_int), ("tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("max_it", ctypes.c_int64), ("max_outer", ctypes.c_int64)]


class NewmarkStats(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("total_iterations", ctypes.c_int64),
                ("max_step_iterations", ctypes.c_int64), ("unconverged_steps", ctypes.c_int64),
                ("graph_used", ctypes.c_int), ("device_ms", ctypes.c_double), ("step_solver", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ModalStats(ctypes.Structure):
    _fields_ = [("lanczos_steps", ctypes.c_int64), ("modes_found", ctypes.c_int64),
                ("solve_iterations", ctypes.c_int64), ("max_ritz_residual", ctypes.c_double),
                ("masses_raised", ctypes.c_int64), ("unconverged_solves", ctypes.c_int64),
                ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [("outer_iterations", ctypes.c_int64), ("lockstep_iterations", ctypes.c_int64),
                ("system_iterations", ctypes.c_int64), ("unconverged", ctypes.c_int64),
                ("max_rel_residual", ctypes.c_double), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


TUNE_SPMV = 1
TUNE_SOLVER = 2
TUNE_LEARNED = 4  # tune the solver with the learned smoother


class Tuning(ctypes.Structure):
    _fields_ = [("spmv_waves_per_block", ctypes.c_int), ("spmv_ms_fp32", ctypes.c_double * 4),
                ("spmv_ms_fp64", ctypes.c_double * 4), ("cheb_degree", ctypes.c_int), ("inner_tol", ctypes.c_double),
                ("check_every", ctypes.c_int), ("solve_ms", ctypes.c_double), ("n_trials", ctypes.c_int),
                ("learned_active", ctypes.c_int), ("solve_ms_learned", ctypes.c_double),
                ("solve_ms_no_network", ctypes.c_double)]

    def as_dict(self):
        d = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            d[k] = list(v) if hasattr(v, "__len__") else v
        return d


class NlOpts(ctypes.Structure):
    _fields_ = [("load_steps", ctypes.c_int), ("max_newton", ctypes.c_int), ("du_tol", ctypes.c_double),
                ("forcing", ctypes.c_double)]


class NlStats(ctypes.Structure):
    _fields_ = [("load_steps", ctypes.c_int), ("newton_iterations", ctypes.c_int),
                ("max_step_newton", ctypes.c_int), ("unconverged_steps", ctypes.c_int),
                ("linear_iterations", ctypes.c_int64), ("unconverged_solves", ctypes.c_int),
                ("last_du", ctypes.c_double), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LdlInfo(ctypes.Structure):
    _fields_ = [("block_dofs", ctypes.c_int64), ("n_blocks", ctypes.c_int64), ("max_bandwidth", ctypes.c_int64),
                ("bytes", ctypes.c_int64), ("neg_pivots", ctypes.c_int32), ("pos_def", ctypes.c_int32),
                ("factor_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SolveStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int64), ("outer_iterations", ctypes.c_int64),
                ("rel_residual", ctypes.c_double), ("ref_residual", ctypes.c_double),
                ("converged", ctypes.c_int), ("fallback_used", ctypes.c_int), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib.asend_last_error.restype = ctypes.c_char_p
_lib.asend_last_error.argtypes = [_vp]
_lib.asend_version.restype = ctypes.c_char_p
_lib.asend_block_dim_auto.restype = ctypes.c_int64
_lib.asend_block_dim_auto.argtypes = [ctypes.c_int64]
_lib.asend_learned_train.argtypes = [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_uint64, ctypes.c_int, _vp]
_lib.asend_learned_init.argtypes = [_vp, ctypes.c_uint64]
_lib.asend_batch_error.restype = ctypes.c_char_p
_lib.asend_batch_error.argtypes = [_vp]
_lib.asend_batch_destroy.argtypes = [_vp]
_lib.asend_penalty_scale.restype = ctypes.c_double
_lib.asend_penalty_scale.argtypes = [_vp]
_lib.asend_ctx_destroy.restype = None
_lib.asend_ctx_destroy.argtypes = [_vp]

EXPORTED = ["asend_ctx_create", "asend_ctx_destroy", "asend_last_error", "asend_version", "asend_sync",
            "asend_reorder_nodes", "asend_block_dim_auto", "asend_surface_traction", "asend_lognormal_moduli",
            "asend_compliance_clamp_check", "asend_surface_pressure", "asend_body_load", "asend_surface_flux",
            "asend_set_conductivity", "asend_brickim_internal", "asend_brickim_set_internal",
            "asend_eigen", "asend_set_material_stiffness", "asend_thermal_strain_load",
            "asend_upload_mesh", "asend_set_thickness", "asend_set_grid", "asend_update_moduli", "asend_set_penalty",
            "asend_penalty_scale", "asend_uniform_elements", "asend_assemble", "asend_get_matrix", "asend_spmv", "asend_solve", "asend_solve_noalloc",
            "asend_ldl_factor", "asend_ldl_solve", "asend_ldl_release", "asend_cycle_spmv_timing",
            "asend_element_sens", "asend_element_sens_material",
            "asend_element_centre_fields", "asend_element_centre_adjoint", "asend_element_sens_coords", "asend_static_compliance", "asend_bench_spmv", "asend_mg_info", "asend_mg_apply",
            "asend_mg_restrict", "asend_mg_prolong", "asend_mg_vcycle", "asend_mg_coarse_solve", "asend_bench_copy",
            "asend_learned_num_params", "asend_learned_init", "asend_learned_get_params",
            "asend_learned_set_params", "asend_learned_train", "asend_learned_eval", "asend_learned_apply",
            "asend_learned_edges", "asend_mg_set_smoother", "asend_mg_set_cycle_precision", "asend_learned_vcycle_eval", "asend_set_graphs",
            "asend_set_rayleigh_damping", "asend_newmark_set_solver", "asend_newmark_setup", "asend_newmark_run", "asend_newmark_adjoint", "asend_modal", "asend_batch_create", "asend_batch_destroy",
            "asend_batch_error", "asend_batch_solve", "asend_autotune", "asend_nl_assemble",
            "asend_static_nonlinear"]


def lib():
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64a(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class AsendError(RuntimeError):
    pass


def version():
    return _lib.asend_version().decode()


# ---------------------------------------------------------------- host prep --
def reorder_nodes(xyz, conn, dof_per_nd, constraints, block_dim):
    """Node reordering + DOF numbering (replaces Model::reorder_nodes,
    S/model_prep.rs:56-303).  Same arguments as the oracle's binding."""
    xyz = _f64(xyz)
    conn = _i32(conn)
    n_nodes = xyz.shape[0]
    n_el, nn = conn.shape
    cptr, tptr, nodes = [0], [0], []
    for c in constraints:
        for t in c:
            nodes.extend(int(v) for v in t)
            tptr.append(len(nodes))
        cptr.append(len(tptr) - 1)
    cptr = _i64a(cptr)
    tptr = _i64a(tptr)
    setn = _i32(nodes if nodes else [0])
    ordered = np.zeros(n_nodes, dtype=np.int32)
    dof_index = np.zeros((n_nodes, 6), dtype=np.int64)
    ndof = _i64(0)
    rc = _lib.asend_reorder_nodes(_i64(n_nodes), _p(xyz), _i64(n_el), ctypes.c_int(nn), ctypes.c_int(dof_per_nd),
                                  _p(conn), _i64(len(constraints)), _p(cptr), _p(tptr), _p(setn), _i64(block_dim),
                                  _p(ordered), _p(dof_index), ctypes.byref(ndof))
    if rc != 0:
        raise AsendError(f"asend_reorder_nodes failed ({rc})")
    return ordered, dof_index, int(ndof.value)


def block_dim_auto(n_nodes):
    return int(_lib.asend_block_dim_auto(n_nodes))


def surface_traction(etype, xyz, conn, dof_index, n_dof, el_set, load, norm_dir, norm_tol_deg, f=None):
    xyz = _f64(xyz)
    conn = _i32(conn)
    di = _i64a(dof_index)
    es = _i32(el_set)
    ld = _f64(load)
    nd = _f64(norm_dir)
    if f is None:
        f = np.zeros(n_dof)
    rc = _lib.asend_surface_traction(ctypes.c_int(etype), _i64(xyz.shape[0]), _p(xyz), _i64(conn.shape[0]), _p(conn),
                                     _p(di), _i64(es.shape[0]), _p(es), _p(ld), _p(nd), _dbl(norm_tol_deg),
                                     _i64(n_dof), _p(f))
    if rc != 0:
        raise AsendError(f"asend_surface_traction failed ({rc})")
    return f


def surface_pressure(etype, xyz, conn, dof_index, n_dof, el_set, pressure, norm_dir, norm_tol_deg, f=None):
    xyz = _f64(xyz)
    conn = _i32(conn)
    di = _i64a(dof_index)
    es = _i32(el_set)
    nd = _f64(norm_dir)
    f = np.zeros(n_dof) if f is None else f
    rc = _lib.asend_surface_pressure(ctypes.c_int(etype), _i64(xyz.shape[0]), _p(xyz), _i64(conn.shape[0]), _p(conn),
                                     _p(di), _i64(es.shape[0]), _p(es), _dbl(pressure), _p(nd), _dbl(norm_tol_deg),
                                     _i64(n_dof), _p(f))
    if rc != 0:
        raise AsendError(f"asend_surface_pressure failed ({rc})")
    return f


def surface_flux(etype, xyz, conn, dof_index, n_dof, el_set, q, norm_dir, norm_tol_deg, f=None):
    """asend_surface_flux: the surfaceFlux load into each node's temperature
    DOF (dof_index[:, 0])."""
    xyz = _f64(xyz)
    conn = _i32(conn)
    di = _i64a(dof_index)
    es = _i32(el_set)
    nd = _f64(norm_dir)
    f = np.zeros(n_dof) if f is None else f
    rc = _lib.asend_surface_flux(ctypes.c_int(etype), _i64(xyz.shape[0]), _p(xyz), _i64(conn.shape[0]), _p(conn),
                                 _p(di), _i64(es.shape[0]), _p(es), _dbl(q), _p(nd), _dbl(norm_tol_deg),
                                 _i64(n_dof), _p(f))
    if rc != 0:
        raise AsendError(f"asend_surface_flux failed ({rc})")
    return f


def thermal_strain_load(etype, xyz, conn, dof_index, n_dof, E, nu, alpha, temp, cmat=None, cmat_idx=None, f=None):
    """asend_thermal_strain_load: the expansion load of node temperatures
    temp with 6 expansion coefficients per element (alpha, n_el x 6)."""
    xyz = _f64(xyz)
    conn = _i32(conn)
    di = _i64a(dof_index)
    E, nu, al, tp = _f64(E), _f64(nu), _f64(np.asarray(alpha).reshape(-1, 6)), _f64(temp)
    cm = None if cmat is None else _f64(np.asarray(cmat).reshape(-1, 36))
    ci = None if cmat_idx is None else _i32(cmat_idx)
    f = np.zeros(n_dof) if f is None else f
    rc = _lib.asend_thermal_strain_load(ctypes.c_int(etype), _i64(xyz.shape[0]), _p(xyz), _i64(conn.shape[0]),
                                        _p(conn), _p(di), _p(E), _p(nu), _p(cm), _p(ci), _p(al), _p(tp),
                                        _i64(n_dof), _p(f))
    if rc != 0:
        raise AsendError(f"asend_thermal_strain_load failed ({rc})")
    return f


BODY_FORCE, GRAVITATIONAL, CENTRIFUGAL = 0, 1, 2


def body_load(etype, kind, xyz, conn, dof_index, n_dof, el_set, rho=None, load=(0.0, 0.0, 0.0),
              center=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0), ang_vel=0.0, f=None):
    """asend_body_load: bodyForce (BODY_FORCE), gravitational, centrifugal."""
    xyz = _f64(xyz)
    conn = _i32(conn)
    di = _i64a(dof_index)
    es = _i32(el_set)
    rho = None if rho is None else _f64(rho)
    ld, ce, ax = _f64(load), _f64(center), _f64(axis)
    f = np.zeros(n_dof) if f is None else f
    rc = _lib.asend_body_load(ctypes.c_int(etype), ctypes.c_int(kind), _i64(xyz.shape[0]), _p(xyz),
                              _i64(conn.shape[0]), _p(conn), _p(di), _i64(es.shape[0]), _p(es), _p(rho), _p(ld),
                              _p(ce), _p(ax), _dbl(ang_vel), _i64(n_dof), _p(f))
    if rc != 0:
        raise AsendError(f"asend_body_load failed ({rc})")
    return f


def lognormal_moduli(seed, n, median, sigma_ln):
    out = np.zeros(n)
    rc = _lib.asend_lognormal_moduli(ctypes.c_uint64(seed), _i64(n), _dbl(median), _dbl(sigma_ln), _p(out))
    if rc != 0:
        raise AsendError(f"asend_lognormal_moduli failed ({rc})")
    return out


def compliance_clamp_check(etype, E, nu):
    """(count, first) of the elements whose isotropic compliance the
    reference's pivot-clamped QR inverse would not invert exactly
    (S/matrix_functions.rs:167-215 on the matrix of
    S/element_properties.rs:728-759, or :196-210 for shell4 layers); first is
    -1 when there are none.  asend_upload_mesh refuses meshes with any."""
    E = _f64(np.atleast_1d(E))
    nu = _f64(np.broadcast_to(nu, E.shape))
    n = ctypes.c_int64(0)
    first = ctypes.c_int64(-1)
    rc = _lib.asend_compliance_clamp_check(ctypes.c_int(etype), _i64(E.size), _p(E), _p(nu), ctypes.byref(n),
                                           ctypes.byref(first))
    if rc != 0:
        raise AsendError(f"asend_compliance_clamp_check failed ({rc})")
    return int(n.value), int(first.value)


# ------------------------------------------------------------------ context --
class Context:
    """One device context (one HIP stream) holding a mesh and its operator."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = _lib.asend_ctx_create(ctypes.c_int(device), ctypes.byref(h))
        if rc != 0:
            raise AsendError(f"asend_ctx_create failed ({rc}): no usable GPU")
        self.h = h
        self.n_dof = 0
        self.n_el = 0

    def close(self):
        if self.h is not None and self.h.value:
            _lib.asend_ctx_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, name):
        if rc != 0:
            msg = _lib.asend_last_error(self.h).decode()
            raise AsendError(f"{name} failed ({rc}): {msg}")

    def upload_mesh(self, etype, xyz, dof_index, n_dof, conn, E, nu, rho=None):
        xyz = _f64(xyz)
        di = _i64a(dof_index)
        conn = _i32(conn)
        E = _f64(E)
        nu = _f64(nu)
        rho = None if rho is None else _f64(rho)
        self._chk(_lib.asend_upload_mesh(self.h, ctypes.c_int(etype), _i64(xyz.shape[0]), _p(xyz), _p(di),
                                         _i64(n_dof), _i64(conn.shape[0]), _p(conn), _p(E), _p(nu), _p(rho)),
                  "asend_upload_mesh")
        self.n_dof = n_dof
        self.n_el = conn.shape[0]

    def set_thickness(self, t):
        t = _f64(t)
        self._chk(_lib.asend_set_thickness(self.h, _p(t)), "asend_set_thickness")

    def set_grid(self, nx, ny, nz):
        self._chk(_lib.asend_set_grid(self.h, _i64(nx), _i64(ny), _i64(nz)), "asend_set_grid")

    def update_moduli(self, E):
        E = _f64(E)
        self._chk(_lib.asend_update_moduli(self.h, _p(E)), "asend_update_moduli")

    def set_penalty(self, dofs, sf=0.0):
        d = _i64a(dofs)
        self._chk(_lib.asend_set_penalty(self.h, _i64(d.shape[0]), _p(d), _dbl(sf)), "asend_set_penalty")

    def uniform_elements(self):
        return bool(_lib.asend_uniform_elements(self.h))

    def penalty_scale(self):
        return float(_lib.asend_penalty_scale(self.h))

    def assemble(self, c_mass=0.0):
        self._chk(_lib.asend_assemble(self.h, _dbl(c_mass)), "asend_assemble")

    def get_matrix(self, with_penalty=True):
        nnz = _i64(0)
        self._chk(_lib.asend_get_matrix(self.h, None, None, None, ctypes.byref(nnz), ctypes.c_int(1)),
                  "asend_get_matrix")
        rowptr = np.zeros(self.n_dof + 1, dtype=np.int64)
        cols = np.zeros(nnz.value, dtype=np.int32)
        vals = np.zeros(nnz.value)
        self._chk(_lib.asend_get_matrix(self.h, _p(rowptr), _p(cols), _p(vals), ctypes.byref(nnz),
                                        ctypes.c_int(1 if with_penalty else 0)), "asend_get_matrix")
        return rowptr, cols, vals

    def spmv(self, x, precision=64):
        x = _f64(x)
        y = np.zeros(self.n_dof)
        self._chk(_lib.asend_spmv(self.h, ctypes.c_int(precision), _p(x), _p(y)), "asend_spmv")
        return y

    @staticmethod
    def opts(method=SOLVE_PCG_FP64, precond=PC_JACOBI, stop_rule=STOP_REL_RESIDUAL, tol=1e-8, inner_tol=1e-4,
             max_it=0, max_outer=30, check_every=16):
        return SolveOpts(method, precond, stop_rule, check_every, tol, inner_tol, max_it, max_outer)

    def solve(self, rhs, opts):
        rhs = _f64(rhs)
        x = np.zeros(self.n_dof)
        st = SolveStats()
        self._chk(_lib.asend_solve(self.h, _p(rhs), _p(x), ctypes.byref(opts), ctypes.byref(st)), "asend_solve")
        return x, st

    def solve_noalloc(self, rhs, opts):
        """asend_solve_noalloc: the FP64 Jacobi PCG solve with its buffers in
        the library's static per-GPU scratch array (no device allocation)."""
        rhs = _f64(rhs)
        x = np.zeros(self.n_dof)
        st = SolveStats()
        self._chk(_lib.asend_solve_noalloc(self.h, _p(rhs), _p(x), ctypes.byref(opts), ctypes.byref(st)),
                  "asend_solve_noalloc")
        return x, st

    def element_sens(self, u, lam):
        u = _f64(u)
        lam = _f64(lam)
        out = np.zeros(self.n_el)
        self._chk(_lib.asend_element_sens(self.h, _p(u), _p(lam), _p(out)), "asend_element_sens")
        return out

    def cycle_spmv_timing(self, enable):
        """Start (True) or stop (False) the event timing of the cycle's FP16
        SpMV launches; stop returns (launches, total_ms)."""
        n = _i64(0)
        ms = _dbl(0)
        self._chk(_lib.asend_cycle_spmv_timing(self.h, ctypes.c_int(1 if enable else 0), ctypes.byref(n),
                                               ctypes.byref(ms)), "asend_cycle_spmv_timing")
        return None if enable else (int(n.value), float(ms.value))

    def ldl_factor(self, block_dofs=0):
        """LDL^T factor of K_eff over diagonal blocks of block_dofs DOFs (0: one
        block, exact) for PC_LDL and SOLVE_DIRECT solves (asend_ldl_factor)."""
        info = LdlInfo()
        self._chk(_lib.asend_ldl_factor(self.h, _i64(block_dofs), ctypes.byref(info)), "asend_ldl_factor")
        return info.as_dict()

    def ldl_solve(self, b):
        b = _f64(b)
        x = np.zeros(self.n_dof)
        self._chk(_lib.asend_ldl_solve(self.h, _p(b), _p(x)), "asend_ldl_solve")
        return x

    def ldl_release(self):
        self._chk(_lib.asend_ldl_release(self.h), "asend_ldl_release")

    def element_sens_coords(self, u, lam):
        """(n_el, 8, 3): lam_e^T (dK_e/dX_{a,c}) u_e (brick8 shape sensitivities)."""
        u = _f64(u)
        lam = _f64(lam)
        out = np.zeros((self.n_el, 8, 3))
        self._chk(_lib.asend_element_sens_coords(self.h, _p(u), _p(lam), _p(out)), "asend_element_sens_coords")
        return out

    def element_centre_fields(self, u):
        """(n_el, 13): centre strains (6), stresses (6), element volume (brick8)."""
        u = _f64(u)
        out = np.zeros((self.n_el, 13))
        self._chk(_lib.asend_element_centre_fields(self.h, _p(u), _p(out)), "asend_element_centre_fields")
        return out

    def element_centre_adjoint(self, h):
        """sum over elements of B_c^T h_e (h: (n_el, 6) strain-space weights)."""
        h = _f64(np.asarray(h, dtype=np.float64).reshape(self.n_el, 6))
        y = np.zeros(self.n_dof)
        self._chk(_lib.asend_element_centre_adjoint(self.h, _p(h), _p(y)), "asend_element_centre_adjoint")
        return y

    def element_sens_material(self, u, lam):
        """(n_el, 9) partials of -lambda^T K_e u w.r.t. E1, E2, E3, nu12, nu13,
        nu23, G12, G13, G23 (asend_element_sens_material, brick8)."""
        u = _f64(u)
        lam = _f64(lam)
        out = np.zeros((self.n_el, 9))
        self._chk(_lib.asend_element_sens_material(self.h, _p(u), _p(lam), _p(out)), "asend_element_sens_material")
        return out

    def static_compliance(self, f, opts):
        f = _f64(f)
        u = np.zeros(self.n_dof)
        s = np.zeros(self.n_el)
        J = _dbl(0)
        st = SolveStats()
        self._chk(_lib.asend_static_compliance(self.h, _p(f), _p(u), _p(s), ctypes.byref(J), ctypes.byref(opts),
                                               ctypes.byref(st)), "asend_static_compliance")
        return u, s, float(J.value), st

    def nl_assemble(self, u):
        """Assemble the nonlinear tangent at u (with the penalty) as the
        context operator; returns the internal force r(u)."""
        u = _f64(u)
        r = np.zeros(self.n_dof)
        self._chk(_lib.asend_nl_assemble(self.h, _p(u), _p(r)), "asend_nl_assemble")
        return r

    def static_nonlinear(self, f, opts, load_steps=1, max_newton=50, du_tol=1e-12, u0=None, forcing=0.0):
        """Geometrically nonlinear static solve (Newton under the quadratic
        load ramp; forcing > 0 makes the Newton solves inexact); returns u and
        NlStats."""
        f = _f64(f)
        u = np.zeros(self.n_dof) if u0 is None else np.array(u0, dtype=np.float64)
        nl = NlOpts(int(load_steps), int(max_newton), float(du_tol), float(forcing))
        st = NlStats()
        self._chk(_lib.asend_static_nonlinear(self.h, _p(f), _p(u), ctypes.byref(nl), ctypes.byref(opts),
                                              ctypes.byref(st)), "asend_static_nonlinear")
        return u, st

    def brickim_internal(self, u=None):
        """brickIM bubble amplitudes (n_el x 9): of a linear solve reaching u,
        or, with u None, those held by the nonlinear path."""
        a = np.zeros((self.n_el, 9))
        self._chk(_lib.asend_brickim_internal(self.h, _p(None if u is None else _f64(u)), _p(a)),
                  "asend_brickim_internal")
        return a

    def brickim_set_internal(self, alpha):
        a = _f64(np.asarray(alpha, dtype=np.float64).reshape(self.n_el, 9))
        self._chk(_lib.asend_brickim_set_internal(self.h, _p(a)), "asend_brickim_set_internal")

    def bench_spmv(self, precision=32, n_iter=50):
        ms = _dbl(0)
        self._chk(_lib.asend_bench_spmv(self.h, ctypes.c_int(precision), ctypes.c_int(n_iter), ctypes.byref(ms)),
                  "asend_bench_spmv")
        return float(ms.value)

    def sync(self):
        self._chk(_lib.asend_sync(self.h), "asend_sync")

    def bench_copy(self, nbytes, n_iter=20):
        ms = _dbl(0)
        self._chk(_lib.asend_bench_copy(self.h, _i64(nbytes), ctypes.c_int(n_iter), ctypes.byref(ms)),
                  "asend_bench_copy")
        return float(ms.value)

    # ---- multigrid diagnostics ----
    def mg_info(self):
        n = ctypes.c_int(0)
        nd = np.zeros(16, dtype=np.int64)
        self._chk(_lib.asend_mg_info(self.h, ctypes.byref(n), _p(nd)), "asend_mg_info")
        return [int(v) for v in nd[:n.value]]

    def _mg_io(self, fn, level, x, n_out):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(n_out, dtype=np.float32)
        self._chk(fn(self.h, ctypes.c_int(level), _p(x), _p(y)), fn.__name__)
        return y

    def mg_apply(self, level, x):
        return self._mg_io(_lib.asend_mg_apply, level, x, self.mg_info()[level])

    def mg_restrict(self, level, x):
        return self._mg_io(_lib.asend_mg_restrict, level, x, self.mg_info()[level + 1])

    def mg_prolong(self, level, x):
        return self._mg_io(_lib.asend_mg_prolong, level, x, self.mg_info()[level])

    def mg_vcycle(self, b):
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.zeros(b.shape[0], dtype=np.float32)
        self._chk(_lib.asend_mg_vcycle(self.h, _p(b), _p(x)), "asend_mg_vcycle")
        return x

    def mg_coarse_solve(self, b):
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.zeros(b.shape[0], dtype=np.float32)
        self._chk(_lib.asend_mg_coarse_solve(self.h, _p(b), _p(x)), "asend_mg_coarse_solve")
        return x

    def newmark_set_solver(self, mode=NM_AUTO, ru_factor=1e-3):
        """Step solver of newmark_run (NM_MIXED or NM_FP64); ru_factor: the
        residual reduction between FP64 reliable updates of the mixed solver."""
        self._chk(_lib.asend_newmark_set_solver(self.h, ctypes.c_int(mode), _dbl(ru_factor)),
                  "asend_newmark_set_solver")

    def set_rayleigh_damping(self, cm=0.0, ck=0.0):
        """D = cm M + ck K for the next newmark_setup (rayDampCM / rayDampCK)."""
        self._chk(_lib.asend_set_rayleigh_damping(self.h, _dbl(cm), _dbl(ck)), "asend_set_rayleigh_damping")

    def set_material_damping(self, dmat, el_mat=None):
        """Material damping matrices (n_mat x 6 x 6, strain order 11, 22, 33,
        12, 13, 23) and each element's matrix (None: matrix 0) for the next
        newmark_setup (asend_set_material_damping, brick8)."""
        d = _f64(np.asarray(dmat, dtype=np.float64).reshape(-1, 36))
        idx = None if el_mat is None else np.ascontiguousarray(el_mat, dtype=np.int32)
        self._chk(_lib.asend_set_material_damping(self.h, _i64(d.shape[0]), _p(d), _p(idx)),
                  "asend_set_material_damping")

    def set_material_stiffness(self, cmat, el_mat=None):
        """General material stiffness matrices (n_mat x 6 x 6, strain order
        11, 22, 33, 12, 13, 23) and each element's matrix (None: matrix 0),
        asend_set_material_stiffness (brick8); cmat None removes them."""
        if cmat is None:
            self._chk(_lib.asend_set_material_stiffness(self.h, _i64(0), None, None), "asend_set_material_stiffness")
            return
        c = _f64(np.asarray(cmat, dtype=np.float64).reshape(-1, 36))
        idx = None if el_mat is None else np.ascontiguousarray(el_mat, dtype=np.int32)
        self._chk(_lib.asend_set_material_stiffness(self.h, _i64(c.shape[0]), _p(c), _p(idx)),
                  "asend_set_material_stiffness")

    def set_conductivity(self, kcond):
        """Heat conduction (asend_set_conductivity): n_el x 6 conductivities
        (k11 k22 k33 k12 k13 k23), or None to return to elasticity.  The
        temperature of node n is DOF dof_index[n, 0]."""
        k = None if kcond is None else _f64(np.asarray(kcond, dtype=np.float64).reshape(-1, 6))
        if k is not None and k.shape[0] != self.n_el:
            raise AsendError(f"conductivities for {k.shape[0]} elements, mesh has {self.n_el}")
        self._chk(_lib.asend_set_conductivity(self.h, _p(k)), "asend_set_conductivity")

    def newmark_setup(self, dt, beta=0.25, gamma=0.5):
        self._chk(_lib.asend_newmark_setup(self.h, _dbl(dt), _dbl(beta), _dbl(gamma)), "asend_newmark_setup")

    def newmark_run(self, n_steps, f_shape, load_fac, watch=(), u0=None, v0=None, a0=None, tol=1e-10,
                    max_it_step=50):
        """Returns (hist [n_steps x n_watch], u, v, a, NewmarkStats)."""
        f_shape = _f64(f_shape)
        load_fac = _f64(load_fac)
        watch = _i64a(np.asarray(watch, dtype=np.int64).reshape(-1))
        hist = np.zeros((n_steps, max(watch.size, 1)))
        u = np.zeros(self.n_dof)
        v = np.zeros(self.n_dof)
        a = np.zeros(self.n_dof)
        st = NewmarkStats()
        opt = [None if x is None else _f64(x) for x in (u0, v0, a0)]
        self._chk(_lib.asend_newmark_run(self.h, ctypes.c_int64(n_steps), _p(f_shape), _p(load_fac), _p(opt[0]),
                                         _p(opt[1]), _p(opt[2]), ctypes.c_int64(watch.size),
                                         _p(watch) if watch.size else None, _p(hist) if watch.size else None,
                                         _p(u), _p(v), _p(a), _dbl(tol), ctypes.c_int(max_it_step),
                                         ctypes.byref(st)), "asend_newmark_run")
        return hist[:, :watch.size], u, v, a, st

    def newmark_adjoint(self, n_steps, f_shape, load_fac, q, w, tol=1e-10, max_it_step=50):
        """Forward sweep from rest and the transient adjoint of
        L = sum_n w[n] q.u_n; returns (L, dL/dE per element, forward stats,
        adjoint stats)."""
        f_shape = _f64(f_shape)
        load_fac = _f64(load_fac)
        q = _f64(q)
        w = _f64(w)
        J = _dbl(0)
        dJ = np.zeros(self.n_el)
        fs, as_ = NewmarkStats(), NewmarkStats()
        self._chk(_lib.asend_newmark_adjoint(self.h, _i64(n_steps), _p(f_shape), _p(load_fac), _p(q), _p(w),
                                             _dbl(tol), ctypes.c_int(max_it_step), ctypes.byref(J), _p(dJ),
                                             ctypes.byref(fs), ctypes.byref(as_)), "asend_newmark_adjoint")
        return float(J.value), dJ, fs, as_

    def modal(self, n_modes, shift=0.0, opts=None, vectors=True):
        """Natural frequencies: (eigenvalues, frequencies in Hz, vectors
        [n_modes x n_dof] or None, ModalStats)."""
        if opts is None:
            opts = self.opts(SOLVE_PCG_FP64, PC_JACOBI, tol=1e-11, max_it=2000000)
        vals = np.zeros(n_modes)
        fr = np.zeros(n_modes)
        vecs = np.zeros((n_modes, self.n_dof)) if vectors else None
        st = ModalStats()
        self._chk(_lib.asend_modal(self.h, int(n_modes), _dbl(shift), ctypes.byref(opts), _p(vals), _p(fr),
                                   _p(vecs) if vectors else None, ctypes.byref(st)), "asend_modal")
        return vals, fr, vecs, st

    def eigen(self, kind, n_modes, u=None, shift=0.0, opts=None, vectors=True):
        """asend_eigen: kind 0 (frequency) or 1 (buckling) about the tangent at
        u (None: the linear stiffness); returns (eigenvalues, frequencies or
        load factors, vectors or None, ModalStats)."""
        if opts is None:
            opts = self.opts(SOLVE_PCG_FP64, PC_JACOBI, tol=1e-11, max_it=2000000)
        vals = np.zeros(n_modes)
        fac = np.zeros(n_modes)
        vecs = np.zeros((n_modes, self.n_dof)) if vectors else None
        st = ModalStats()
        self._chk(_lib.asend_eigen(self.h, int(kind), _p(None if u is None else _f64(u)), int(n_modes), _dbl(shift),
                                   ctypes.byref(opts), _p(vals), _p(fac), _p(vecs) if vectors else None,
                                   ctypes.byref(st)), "asend_eigen")
        return vals, fac, vecs, st
    def autotune(self, rhs=None, what=TUNE_SPMV | TUNE_SOLVER):
        t = Tuning()
        r = None if rhs is None else _f64(rhs)
        self._chk(_lib.asend_autotune(self.h, _p(r), ctypes.c_int(what), ctypes.byref(t)), "asend_autotune")
        return t

    def set_graphs(self, on=True):
        self._chk(_lib.asend_set_graphs(self.h, 1 if on else 0), "asend_set_graphs")

    def mg_set_smoother(self, kind=1, degree=2, lo_frac=0.1, hi_frac=1.15):
        """kind 0 = point Jacobi, 1 = 3x3 block Jacobi; takes effect at the next assembly."""
        self._chk(_lib.asend_mg_set_smoother(self.h, int(kind), int(degree), _dbl(lo_frac), _dbl(hi_frac)),
                  "asend_mg_set_smoother")

    # -------------------------------------------------- learned preconditioner --
    def mg_set_cycle_precision(self, bits=16):
        """Finest-level arithmetic inside the multigrid cycle: 16 (FP16
        row-scaled operator copy, the default) or 32."""
        self._chk(_lib.asend_mg_set_cycle_precision(self.h, ctypes.c_int(bits)), "asend_mg_set_cycle_precision")

    def learned_init(self, seed=1):
        self._chk(_lib.asend_learned_init(self.h, ctypes.c_uint64(seed)), "asend_learned_init")

    def learned_params(self):
        out = np.zeros(learned_num_params(), dtype=np.float32)
        self._chk(_lib.asend_learned_get_params(self.h, _p(out), out.size), "asend_learned_get_params")
        return out

    def learned_set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        self._chk(_lib.asend_learned_set_params(self.h, _p(a), a.size), "asend_learned_set_params")

    def learned_train(self, n_steps, lr=1e-3, seed=1, mode=1):
        """mode 0: Gaussian pairs, one-step loss; 1: solver-generated samples
        (V-cycle errors and V-cycle solutions), one-step loss; 2: end to end
        through one multigrid V-cycle; 3: as 1 with a Gaussian sample before
        every solver pair."""
        losses = np.zeros(max(n_steps, 1), dtype=np.float64)
        self._chk(_lib.asend_learned_train(self.h, int(n_steps), ctypes.c_float(lr), ctypes.c_uint64(seed),
                                           int(mode), _p(losses)), "asend_learned_train")
        return losses[:n_steps]

    def learned_eval(self, e, n_nodes):
        e = np.ascontiguousarray(e, dtype=np.float64)
        loss = ctypes.c_double(0.0)
        grad = np.zeros(learned_num_params(), dtype=np.float32)
        feat = np.zeros((n_nodes, 8), dtype=np.float32)
        out = np.zeros((n_nodes, 24), dtype=np.float32)
        self._chk(_lib.asend_learned_eval(self.h, _p(e), ctypes.byref(loss), _p(grad), _p(feat), _p(out)),
                  "asend_learned_eval")
        return loss.value, grad, feat, out

    def learned_vcycle_eval(self, e, keep_bounds=False):
        e = np.ascontiguousarray(e, dtype=np.float64)
        loss = ctypes.c_double(0.0)
        grad = np.zeros(learned_num_params(), dtype=np.float32)
        self._chk(_lib.asend_learned_vcycle_eval(self.h, _p(e), 1 if keep_bounds else 0, ctypes.byref(loss), _p(grad)),
                  "asend_learned_vcycle_eval")
        return loss.value, grad

    def learned_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        self._chk(_lib.asend_learned_apply(self.h, _p(r), _p(z)), "asend_learned_apply")
        return z

    def learned_edges(self, n_nodes):
        w = ctypes.c_int(0)
        self._chk(_lib.asend_learned_edges(self.h, None, None, ctypes.byref(w)), "asend_learned_edges")
        cnt = max(w.value, 1) * n_nodes
        enb = np.zeros(cnt, dtype=np.int32)
        erev = np.zeros(cnt, dtype=np.int32)
        self._chk(_lib.asend_learned_edges(self.h, _p(enb), _p(erev), ctypes.byref(w)), "asend_learned_edges")
        return enb.reshape(-1, n_nodes)[:w.value], erev.reshape(-1, n_nodes)[:w.value]


def learned_num_params():
    return int(_lib.asend_learned_num_params())



class Batch:
    """Lockstep multi-RHS solves over assembled variant contexts
    (asend_batch_create / asend_batch_solve)."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        arr = (ctypes.c_void_p * len(self.ctxs))(*[c.h for c in self.ctxs])
        h = ctypes.c_void_p()
        rc = _lib.asend_batch_create(arr, len(self.ctxs), ctypes.byref(h))
        if rc != 0:
            msg = _lib.asend_batch_error(h).decode() if h.value else "bad batch arguments"
            if h.value:
                _lib.asend_batch_destroy(h)
            raise AsendError(f"asend_batch_create failed ({rc}): {msg}")
        self.h = h

    def solve(self, rhs, tol=1e-8, inner_tol=1e-3, max_it=100000, check_every=16, verbose=False):
        """rhs: list (per variant) of (nrhs, n_dof) arrays; returns (list of
        solutions, BatchStats)."""
        nrhs = rhs[0].shape[0]
        rs = [np.ascontiguousarray(r, dtype=np.float64) for r in rhs]
        xs = [np.zeros_like(r) for r in rs]
        rp = (ctypes.c_void_p * len(rs))(*[r.ctypes.data for r in rs])
        xp = (ctypes.c_void_p * len(xs))(*[x.ctypes.data for x in xs])
        st = BatchStats()
        rc = _lib.asend_batch_solve(self.h, ctypes.c_int(nrhs), rp, xp, _dbl(tol), _dbl(inner_tol),
                                    ctypes.c_int64(max_it), ctypes.c_int(check_every), ctypes.c_int(1 if verbose else 0),
                                    ctypes.byref(st))
        if rc != 0:
            raise AsendError(f"asend_batch_solve failed ({rc}): {_lib.asend_batch_error(self.h).decode()}")
        return xs, st

    def close(self):
        if self.h:
            _lib.asend_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
