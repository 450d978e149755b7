"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import it; the product
package (asend3d_amd) never does.

Each wrapper names the reference routine its C++ function restates (paths
relative to the reference checkout, S/ = solverSrc/asend_run_job/src/).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_p = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_max_abs.restype = _dbl
        L.orc_ldl_max_bandwidth.restype = _i64
        L.orc_ldl_free.restype = None
        _lib = L
    return _lib


def set_threads(n):
    """OpenMP threads of jacobi_pcg, spmv_penalty, the element loop of
    assemble_solid and solid_sens (1 = serial, the default; results do not
    depend on the thread count)."""
    _chk(lib().orc_set_threads(ctypes.c_int(int(n))), "set_threads")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _chk(rc, name):
    if rc != 0:
        raise RuntimeError(f"oracle {name} failed with code {rc}")


def reorder_nodes(xyz, conn, dof_per_nd, constraints, block_dim):
    """Model::reorder_nodes (S/model_prep.rs:56-303).

    constraints: list of constraints, each a list of terms, each term an int
    array of node labels (its node set, in set order).
    Returns (ordered[int32], dof_index[n_nodes, 6] int64, n_dof).
    """
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    n_nodes = xyz.shape[0]
    n_el, nn = conn.shape
    cptr = [0]
    tptr = [0]
    nodes = []
    for c in constraints:
        for t in c:
            nodes.extend(int(v) for v in t)
            tptr.append(len(nodes))
        cptr.append(len(tptr) - 1)
    cptr = np.asarray(cptr, dtype=np.int64)
    tptr = np.asarray(tptr, dtype=np.int64)
    setn = np.asarray(nodes if nodes else [0], dtype=np.int32)
    ordered = np.zeros(n_nodes, dtype=np.int32)
    dof_index = np.zeros((n_nodes, 6), dtype=np.int64)
    ndof = _i64(0)
    rc = lib().orc_reorder_nodes(_i64(n_nodes), _ptr(xyz), _i64(n_el), ctypes.c_int(nn), ctypes.c_int(dof_per_nd),
                                 _ptr(conn), _i64(len(constraints)), _ptr(cptr), _ptr(tptr), _ptr(setn),
                                 _i64(block_dim), _ptr(ordered), _ptr(dof_index), ctypes.byref(ndof))
    _chk(rc, "reorder_nodes")
    return ordered, dof_index, int(ndof.value)


def solid_cmat(E, nu, G=None):
    """get_solid_stiff_dfd0 isotropic case (S/element_properties.rs:648-764)."""
    if G is None:
        G = E / (2.0 * (1.0 + nu))
    mod = np.array([E, E, E], dtype=np.float64)
    pr = np.array([nu, nu, nu], dtype=np.float64)
    sm = np.array([G, G, G], dtype=np.float64)
    C = np.zeros(36)
    _chk(lib().orc_solid_cmat(_ptr(mod), _ptr(pr), _ptr(sm), _ptr(C)), "solid_cmat")
    return C.reshape(6, 6)


def solid_element(etype, X, C, ue=None, ori=None):
    """get_ruk_dfd0 + condense_mat for brick8 (8) / brickIM (81).

    X: (8, 3) coordinates.  Returns (K[24,24], R[24])."""
    Xs = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T.reshape(-1))
    C = np.ascontiguousarray(C, dtype=np.float64)
    K = np.zeros(576)
    R = np.zeros(24)
    ue = None if ue is None else np.ascontiguousarray(ue, dtype=np.float64)
    ori = None if ori is None else np.ascontiguousarray(ori, dtype=np.float64)
    _chk(lib().orc_solid_element(ctypes.c_int(etype), _ptr(Xs), _ptr(ori), _ptr(C), _ptr(ue), _ptr(K), _ptr(R)),
         "solid_element")
    return K.reshape(24, 24), R


def solid_element_dE(etype, X, E, nu):
    Xs = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T.reshape(-1))
    K = np.zeros(576)
    dK = np.zeros(576)
    _chk(lib().orc_solid_element_dE(ctypes.c_int(etype), _ptr(Xs), None, _dbl(E), _dbl(nu), _ptr(K), _ptr(dK)),
         "solid_element_dE")
    return K.reshape(24, 24), dK.reshape(24, 24)


def solid_cmat_ortho(mod, pr, sm):
    """get_solid_stiff_dfd0 with the nine compliance parameters
    (S/element_properties.rs:728-759): mod = (E1, E2, E3), pr = (nu12, nu13,
    nu23), sm = (G12, G13, G23)."""
    mod, pr, sm = (np.ascontiguousarray(v, dtype=np.float64) for v in (mod, pr, sm))
    C = np.zeros(36)
    _chk(lib().orc_solid_cmat(_ptr(mod), _ptr(pr), _ptr(sm), _ptr(C)), "solid_cmat")
    return C.reshape(6, 6)


def solid_element_dmat(etype, X, mod, pr, sm, param):
    """(K, dK/dp) of one element, p = 0..8 indexing (E1, E2, E3, nu12, nu13,
    nu23, G12, G13, G23), by dual numbers through the QR inverse."""
    Xs = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T.reshape(-1))
    mod, pr, sm = (np.ascontiguousarray(v, dtype=np.float64) for v in (mod, pr, sm))
    K = np.zeros(576)
    dK = np.zeros(576)
    _chk(lib().orc_solid_element_dmat(ctypes.c_int(etype), _ptr(Xs), _ptr(mod), _ptr(pr), _ptr(sm),
                                      ctypes.c_int(param), _ptr(K), _ptr(dK)), "solid_element_dmat")
    return K.reshape(24, 24), dK.reshape(24, 24)


def solid_mass(etype, X, rho):
    Xs = np.ascontiguousarray(np.asarray(X, dtype=np.float64).T.reshape(-1))
    M = np.zeros(576)
    _chk(lib().orc_solid_mass(ctypes.c_int(etype), _ptr(Xs), _dbl(rho), _ptr(M)), "solid_mass")
    return M.reshape(24, 24)


class CSR:
    def __init__(self, rowptr, cols, vals):
        self.rowptr = rowptr
        self.cols = cols
        self.vals = vals
        self.n = rowptr.shape[0] - 1

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.cols, self.rowptr), shape=(self.n, self.n))


def assemble_solid(etype, xyz, conn, E, nu, dof_index, n_dof):
    """build_elastic_soln_load(true) scatter (S/model_analysis.rs:179-201)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    nnz = _i64(0)
    L = lib()
    args = (ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]), _ptr(conn), _ptr(E), _ptr(nu),
            _ptr(di), _i64(n_dof))
    _chk(L.orc_assemble_solid(*args, None, None, None, ctypes.byref(nnz)), "assemble(nnz)")
    rowptr = np.zeros(n_dof + 1, dtype=np.int64)
    cols = np.zeros(nnz.value, dtype=np.int32)
    vals = np.zeros(nnz.value, dtype=np.float64)
    _chk(L.orc_assemble_solid(*args, _ptr(rowptr), _ptr(cols), _ptr(vals), ctypes.byref(nnz)), "assemble")
    return CSR(rowptr, cols, vals)


def assemble_solid_mass(etype, xyz, conn, rho, dof_index, K):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    vals = np.zeros_like(K.vals)
    _chk(lib().orc_assemble_solid_mass(ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]),
                                       _ptr(conn), _ptr(rho), _ptr(di), _i64(K.n), _ptr(K.rowptr), _ptr(K.cols),
                                       _ptr(vals)), "assemble_mass")
    return CSR(K.rowptr, K.cols, vals)


def surface_traction(etype, xyz, conn, dof_index, n_dof, el_set, load, norm_dir, norm_tol_deg):
    """get_app_load_dfd0 surfaceTraction (S/element_equations.rs:1833-2056)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    es = np.ascontiguousarray(el_set, dtype=np.int32)
    ld = np.ascontiguousarray(load, dtype=np.float64)
    nd = np.ascontiguousarray(norm_dir, dtype=np.float64)
    f = np.zeros(n_dof)
    _chk(lib().orc_surface_traction(ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]),
                                    _ptr(conn), _ptr(di), _i64(es.shape[0]), _ptr(es), _ptr(ld), _ptr(nd),
                                    _dbl(norm_tol_deg), _ptr(f)), "surface_traction")
    return f


def surface_pressure(etype, xyz, conn, dof_index, n_dof, el_set, pressure, norm_dir, norm_tol_deg):
    """get_app_load_dfd0 surfacePressure (S/element_equations.rs:1970-2043)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    es = np.ascontiguousarray(el_set, dtype=np.int32)
    nd = np.ascontiguousarray(norm_dir, dtype=np.float64)
    f = np.zeros(n_dof)
    _chk(lib().orc_surface_pressure(ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]),
                                    _ptr(conn), _ptr(di), _i64(es.shape[0]), _ptr(es), _dbl(pressure), _ptr(nd),
                                    _dbl(norm_tol_deg), _ptr(f)), "surface_pressure")
    return f


BODY_FORCE, GRAVITATIONAL, CENTRIFUGAL = 0, 1, 2


def body_load(etype, kind, xyz, conn, dof_index, n_dof, el_set, rho=None, load=(0.0, 0.0, 0.0),
              center=(0.0, 0.0, 0.0), axis=(0.0, 0.0, 1.0), ang_vel=0.0):
    """get_app_load_dfd0 bodyForce / gravitational / centrifugal
    (S/element_equations.rs:1876-1969)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    es = np.ascontiguousarray(el_set, dtype=np.int32)
    rho = np.ones(conn.shape[0]) if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
    ld, ce, ax = (np.ascontiguousarray(v, dtype=np.float64) for v in (load, center, axis))
    f = np.zeros(n_dof)
    _chk(lib().orc_body_load(ctypes.c_int(etype), ctypes.c_int(kind), _i64(conn.shape[0]), _ptr(xyz), _ptr(conn),
                             _ptr(di), _i64(es.shape[0]), _ptr(es), _ptr(rho), _ptr(ld), _ptr(ce), _ptr(ax),
                             _dbl(ang_vel), _ptr(f)), "body_load")
    return f


def max_abs(K):
    """SparseMat::get_max_abs_val (S/list_ent_meth.rs:81-94)."""
    return lib().orc_max_abs(_i64(K.vals.shape[0]), _ptr(K.vals))


def spmv_penalty(K, pen, sf, x):
    pen = np.ascontiguousarray(pen, dtype=np.int64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(K.n)
    _chk(lib().orc_spmv_penalty(_i64(K.n), _ptr(K.rowptr), _ptr(K.cols), _ptr(K.vals), _i64(pen.shape[0]), _ptr(pen),
                                _dbl(sf), _ptr(x), _ptr(y)), "spmv_penalty")
    return y


class LDL:
    """LowerTriMat (S/lower_tri_mat_meth.rs:15-241) on K + sf C^T C."""

    def __init__(self, K, pen, sf, block_dofs):
        self.pen = np.ascontiguousarray(pen, dtype=np.int64)
        self.h = ctypes.c_void_p()
        _chk(lib().orc_ldl_build(_i64(K.n), _ptr(K.rowptr), _ptr(K.cols), _ptr(K.vals), _i64(self.pen.shape[0]),
                                 _ptr(self.pen), _dbl(sf), _i64(block_dofs), ctypes.byref(self.h)), "ldl_build")
        self.n = K.n

    def solve(self, rhs):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros(self.n)
        _chk(lib().orc_ldl_solve(self.h, _ptr(rhs), _ptr(x)), "ldl_solve")
        return x

    def pos_def(self):
        return bool(lib().orc_ldl_pos_def(self.h))

    def max_bandwidth(self):
        return int(lib().orc_ldl_max_bandwidth(self.h))

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().orc_ldl_free(self.h)
            self.h = None


def conj_grad_sparse(K, pen, sf, ldl, rhs, conv_tol, max_it):
    """conj_grad_sparse (S/matrix_functions.rs:285-362)."""
    pen = np.ascontiguousarray(pen, dtype=np.int64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(K.n)
    it = _i64(0)
    res = _dbl(0)
    _chk(lib().orc_conj_grad_sparse(_i64(K.n), _ptr(K.rowptr), _ptr(K.cols), _ptr(K.vals), _i64(pen.shape[0]),
                                    _ptr(pen), _dbl(sf), ldl.h, _ptr(rhs), _dbl(conv_tol), _i64(max_it), _ptr(x),
                                    ctypes.byref(it), ctypes.byref(res)), "conj_grad_sparse")
    return x, int(it.value), float(res.value)


def jacobi_pcg(K, pen, sf, rhs, rel_tol, max_it):
    pen = np.ascontiguousarray(pen, dtype=np.int64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    x = np.zeros(K.n)
    it = _i64(0)
    rr = _dbl(0)
    _chk(lib().orc_jacobi_pcg(_i64(K.n), _ptr(K.rowptr), _ptr(K.cols), _ptr(K.vals), _i64(pen.shape[0]), _ptr(pen),
                              _dbl(sf), _ptr(rhs), _dbl(rel_tol), _i64(max_it), _ptr(x), ctypes.byref(it),
                              ctypes.byref(rr)), "jacobi_pcg")
    return x, int(it.value), float(rr.value)


def solid_sens(etype, xyz, conn, E, nu, dof_index, u, lam):
    """Adjoint DV loop with dual-number dR/dd (S/model_analysis.rs:1379-1457, 1679-1701)."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    out = np.zeros(conn.shape[0])
    _chk(lib().orc_solid_sens(ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]), _ptr(conn),
                              _ptr(E), _ptr(nu), _ptr(di), _ptr(u), _ptr(lam), _ptr(out)), "solid_sens")
    return out


def dv_loop(etype, xyz, conn, E, nu, dof_index, n_dof, u, lam, load_el, n_dv):
    """The reference's design-variable loop for per-element modulus DVs
    0..n_dv-1 (d_relasticd_d + the DV loop of get_obj_gradient,
    S/model_analysis.rs:1379-1457, 1679-1701): its O(N_dof) sweeps per DV."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    le = np.ascontiguousarray(load_el, dtype=np.int32)
    out = np.zeros(n_dv)
    _chk(lib().orc_dv_loop(ctypes.c_int(etype), _i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]), _ptr(conn),
                           _ptr(E), _ptr(nu), _ptr(di), _i64(n_dof), _ptr(u), _ptr(lam), _i64(le.shape[0]), _ptr(le),
                           _i64(n_dv), _ptr(out)), "dv_loop")
    return out


def shell4_element(X, E, nu, t, rho):
    """Condensed 24x24 stiffness and 24x24 consistent mass of one shell4
    (X: 3 x 4 SoA coordinates)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    Ke = np.zeros(576)
    Me = np.zeros(576)
    _chk(lib().orc_shell4_element(_ptr(X), _dbl(E), _dbl(nu), _dbl(t), _dbl(rho), _ptr(Ke), _ptr(Me)), "shell4")
    return Ke.reshape(24, 24), Me.reshape(24, 24)


def assemble_shell(xyz, conn, E, nu, t, rho, dof_index, n_dof):
    """Shell4 stiffness and consistent mass (CSR pair) on the same pattern."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (E, nu, t, rho)]
    di = np.ascontiguousarray(dof_index, dtype=np.int64)
    nnz = ctypes.c_int64(0)
    rowptr = np.zeros(n_dof + 1, dtype=np.int64)
    _chk(lib().orc_assemble_shell(_i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]), _ptr(conn), *[_ptr(a) for a in args],
                                  _ptr(di), _i64(n_dof), _ptr(rowptr), None, None, None, ctypes.byref(nnz)),
         "assemble_shell")
    cols = np.zeros(nnz.value, dtype=np.int32)
    vals = np.zeros(nnz.value)
    mvals = np.zeros(nnz.value)
    _chk(lib().orc_assemble_shell(_i64(xyz.shape[0]), _ptr(xyz), _i64(conn.shape[0]), _ptr(conn), *[_ptr(a) for a in args],
                                  _ptr(di), _i64(n_dof), _ptr(rowptr), _ptr(cols), _ptr(vals), _ptr(mvals),
                                  ctypes.byref(nnz)), "assemble_shell")
    return CSR(rowptr, cols, vals), CSR(rowptr, cols, mvals)
