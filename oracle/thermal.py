"""CPU restatement of the reference's brick8 heat conduction statics (numpy).

TEST INFRASTRUCTURE ONLY: the checker of asend_set_conductivity +
asend_assemble, asend_surface_flux and the thermal job path; only tests/
import it.  Paths relative to the reference checkout, S/ =
solverSrc/asend_run_job/src/.

  conduction_element   Element::get_rtk_dfd0 (S/element_equations.rs:
                       1064-1123): sum over the 8 Gauss points of
                       w detJ G k G^T, G the 8 x 3 physical shape-function
                       gradients; k the 3 x 3 conductivity of
                       get_conductivity_dfd0 (S/element_properties.rs:
                       1715-1760) from the material's six values
                       (k11 k22 k33 k12 k13 k23).
  assemble_conduction  build_thermal_soln_load(true) scatter by node
                       (S/model_analysis.rs:203-227, get_rt_dfd0 :1206-1265).
  heat_rate            get_rtm_dfd0 with actual_props = false
                       (S/element_equations.rs:1125-1204): sum over the Gauss
                       points of w detJ N_a (N . T).
  body_heat_gen        bodyHeatGen of get_app_therm_load_dfd0
                       (S/element_equations.rs:2087-2111): heat_rate of a
                       field Q, rescaled to Q V.
  surface_flux         surfaceFlux of the same routine (:2113-2150),
                       including its running normal test (the dot products of
                       an element's surface faces add up across faces) and the
                       field kept across faces; the element adds the vector of
                       its last passing face.
  nodal loads          nodalHeatGen (S/model_analysis.rs:98-104).
  capacity_element     get_rtm_dfd0 with actual_props (rho c_p int N N^T).
  transient            the thermal Newmark loop (first-order, gamma).

Parity pins (tests/test_thermal_oracle.py): the unit-cube conduction matrix in
closed form (1/3, 0, -1/12, -1/12), the patch test of a linear field on a mesh
with moved interior nodes and an anisotropic conductivity (interior rows
vanish, boundary rows equal the face-flux integral), 1D closed-form
temperature profiles, box-volume and face-area sums, and the tensor-product
mass integrals that fix the flux distribution; the box capacity matrix from
the 1D mass integrals; uniform heating's exact discrete rise from a zero
starting rate; first-order convergence of the transient (its zero starting
rate) to the exact solution of the semi-discrete system (matrix
exponential).
"""
import numpy as np

_G = 0.5773502691896257
_SIGN = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
# hex face local nodes (S/element_meth.rs:537-590)
HEX_FACES = ((3, 2, 1, 0), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7))


def _gauss_points():
    pts = []
    for gp in range(8):
        pts.append(((_G if gp & 1 else -_G), (_G if gp & 2 else -_G), (_G if gp & 4 else -_G)))
    return np.array(pts)


_GP = _gauss_points()


def _shape(s):
    """N (8,) and dN/ds (8, 3) of the trilinear brick at s."""
    f = 1.0 + _SIGN * s[None, :]
    N = 0.125 * f[:, 0] * f[:, 1] * f[:, 2]
    dN = np.empty((8, 3))
    dN[:, 0] = 0.125 * _SIGN[:, 0] * f[:, 1] * f[:, 2]
    dN[:, 1] = 0.125 * f[:, 0] * _SIGN[:, 1] * f[:, 2]
    dN[:, 2] = 0.125 * f[:, 0] * f[:, 1] * _SIGN[:, 2]
    return N, dN


def kmat(k6):
    k = np.asarray(k6, dtype=np.float64)
    return np.array([[k[0], k[3], k[4]], [k[3], k[1], k[5]], [k[4], k[5], k[2]]])


def conduction_element(X, k6):
    """8 x 8 conduction matrix of the brick with nodes X (8 x 3)."""
    k = kmat(k6)
    Ke = np.zeros((8, 8))
    for s in _GP:
        _, dN = _shape(s)
        J = X.T @ dN  # J[i, c] = d x_i / d s_c
        det = np.linalg.det(J)
        G = dN @ np.linalg.inv(J)  # physical gradients, 8 x 3
        Ke += det * (G @ k @ G.T)
    return Ke


def heat_rate(X, T):
    """sum_gp w detJ N_a (N . T) (unit weights of the 2-point rule)."""
    out = np.zeros(8)
    vol = 0.0
    for s in _GP:
        N, dN = _shape(s)
        det = np.linalg.det(X.T @ dN)
        out += det * N * (N @ T)
        vol += det
    return out, vol


def assemble_conduction(xyz, conn, kcond):
    """Global conduction matrix in node numbering (scipy CSR, n_nodes^2)."""
    import scipy.sparse as sp
    n = xyz.shape[0]
    rows, cols, vals = [], [], []
    for e in range(conn.shape[0]):
        c = conn[e]
        Ke = conduction_element(xyz[c], kcond[e])
        rows.append(np.repeat(c, 8))
        cols.append(np.tile(c, 8))
        vals.append(Ke.reshape(-1))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def body_heat_gen(xyz, conn, el_set, Q):
    """bodyHeatGen: node vector of the uniform generation Q over el_set."""
    f = np.zeros(xyz.shape[0])
    for e in el_set:
        c = conn[e]
        v, vol = heat_rate(xyz[c], np.full(8, Q))
        f[c] += v * (Q * vol / v.sum())
    return f


def surface_faces(conn):
    """on_surf[e, k]: face k of element e has no twin (S/model_prep.rs:510-554)."""
    seen = {}
    for e in range(conn.shape[0]):
        for k, fl in enumerate(HEX_FACES):
            key = tuple(sorted(int(conn[e, i]) for i in fl))
            seen[key] = seen.get(key, 0) + 1
    on = np.zeros((conn.shape[0], 6), dtype=bool)
    for e in range(conn.shape[0]):
        for k, fl in enumerate(HEX_FACES):
            on[e, k] = seen[tuple(sorted(int(conn[e, i]) for i in fl))] == 1
    return on


def face_area_normal(X, fl):
    """Face::get_area_normal_dfd0: area and unit normal of a quad from the
    cross product of its diagonals."""
    v1 = X[fl[2]] - X[fl[0]]
    v2 = X[fl[3]] - X[fl[1]]
    n = np.cross(v1, v2)
    m = np.linalg.norm(n)
    return 0.5 * m, n / m


def surface_flux(xyz, conn, el_set, q, norm_dir, norm_tol_deg):
    """surfaceFlux: node vector of the flux q on el_set's faces."""
    nd = np.asarray(norm_dir, dtype=np.float64)
    nd = nd / np.linalg.norm(nd)
    cos_tol = np.cos(np.pi / 180.0 * norm_tol_deg)
    on = surface_faces(conn)
    f = np.zeros(xyz.shape[0])
    for e in el_set:
        c = conn[e]
        X = xyz[c]
        T = np.zeros(8)
        dp = 0.0
        last = None
        for k, fl in enumerate(HEX_FACES):
            if not on[e, k]:
                continue
            area, n = face_area_normal(X, fl)
            dp += nd @ n
            if not dp > cos_tol:
                continue
            T[list(fl)] = q
            v, _ = heat_rate(X, T)
            keep = np.zeros(8, dtype=bool)
            keep[list(fl)] = True
            v[~keep] = 0.0
            last = v * (q * area / v.sum())
        if last is not None:
            f[c] += last
    return f


def to_dofs(f_node, dof_index, n_dof):
    """Node vector onto the temperature DOFs dof_index[:, 0] of a 3-DOF-per-node numbering."""
    out = np.zeros(n_dof)
    out[np.asarray(dof_index).reshape(-1, 6)[:, 0]] = f_node
    return out


def capacity_element(X, rc):
    """rc int N N^T (get_rtm_dfd0 with actual_props: rc = density x specific heat)."""
    Ce = np.zeros((8, 8))
    for s in _GP:
        N, dN = _shape(s)
        Ce += rc * np.linalg.det(X.T @ dN) * np.outer(N, N)
    return Ce


def assemble_capacity(xyz, conn, rc):
    import scipy.sparse as sp
    n = xyz.shape[0]
    rows, cols, vals = [], [], []
    for e in range(conn.shape[0]):
        c = conn[e]
        rows.append(np.repeat(c, 8))
        cols.append(np.tile(c, 8))
        vals.append(capacity_element(xyz[c], rc[e]).reshape(-1))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def transient(K, C, pen, sf, t_pen, load_at, dt, gamma, sim_period):
    """The thermal Newmark loop of Model::solve (S/model_analysis.rs:596-603,
    solve_step :331-370, Node::update_tdot / advance_temp
    S/node_meth.rs:140-171), from T = 0, Tdot = 0: per step at time t the
    load of t minus K T minus C times the predicted rate
    -((1 - gamma) / gamma) Tdot plus the penalty load sf (T_c - T), solved
    with K + C / (gamma dt) + sf P for the increment; then
    Tdot = ((T_new - T) / dt - (1 - gamma) Tdot) / gamma.  Returns the
    temperatures and rates after each step."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = K.shape[0]
    P = sp.csr_matrix((np.full(len(pen), sf), (pen, pen)), shape=K.shape)
    A = spla.splu((K + C / (gamma * dt) + P).tocsc())
    T = np.zeros(n)
    Td = np.zeros(n)
    hist = []
    t = 0.0
    while t < sim_period:
        pred = -((1.0 - gamma) / gamma) * Td
        b = load_at(t) - K @ T - C @ pred - P @ T
        b[pen] += sf * t_pen
        dT = A.solve(b)
        Tn = T + dT
        Td = ((Tn - T) / dt - (1.0 - gamma) * Td) / gamma
        T = Tn
        hist.append((T.copy(), Td.copy()))
        t += dt
    return hist


def thermal_strain_load(xyz, conn, E, nu, alpha, temp):
    """The expansion term of get_ruk_dfd0 (S/element_equations.rs:272-298):
    node forces (n_nodes, 3) of sum_gp w detJ B^T C (alpha N.T), isotropic C."""
    f = np.zeros((xyz.shape[0], 3))
    for e in range(conn.shape[0]):
        c = conn[e]
        X = xyz[c]
        lam = E[e] * nu[e] / ((1 + nu[e]) * (1 - 2 * nu[e]))
        mu = E[e] / (2 * (1 + nu[e]))
        Cm = np.zeros((6, 6))
        Cm[:3, :3] = lam
        Cm[np.arange(3), np.arange(3)] += 2 * mu
        Cm[np.arange(3, 6), np.arange(3, 6)] = mu
        for s in _GP:
            N, dN = _shape(s)
            J = X.T @ dN
            G = dN @ np.linalg.inv(J)
            sv = Cm @ (np.asarray(alpha[e]) * (N @ temp[c]))
            S = np.array([[sv[0], sv[3], sv[4]], [sv[3], sv[1], sv[5]], [sv[4], sv[5], sv[2]]])
            f[c] += np.linalg.det(J) * (G @ S)
    return f
