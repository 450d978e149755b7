"""CPU oracle of the geometrically nonlinear brick8 static solve.

Test infrastructure only: the tests use it as the checker of the device
path; nothing in the product imports it.

Restates, in FP64 numpy:
- the Green-Lagrange strain of a solid element, E = (H + H^T + H^T H) / 2
  with H the displacement gradient (S/element_soln_fields.rs:1345-1461,
  engineering shears [E11, E22, E33, 2E12, 2E13, 2E23]);
- the element residual r_i = sum_gp w detJ sigma . dE/du_i with
  sigma = C E, and its tangent, the material part B(u)^T C B(u) plus the
  geometric part sigma . d2E/du_i du_j that is nonzero only between equal
  displacement components (S/element_equations.rs:206-418, nonlinear branch
  :319-331 and the B^T C B product :398-411);
- the Newton loop of one load step: assemble r(u) and the tangent, load
  fact f - r(u) + penalty load, solve, update, stop when max|du| <= 1e-12
  or after 50 iterations (S/model_analysis.rs:312-558, elastic branch
  :450-546);
- the quadratic load ramp of a static nonlinear solve,
  fact_s = 1 - ((L - 1 - s) / L)^2 for s = 0..L-1
  (S/model_analysis.rs:567-568, :636-643).
brickIM (type 81) adds three bubble modes 1 - s_k^2 per displacement
component whose natural derivatives -2 s_k use the Jacobian at the element
centroid (S/element_soln_fields.rs:161-225, :383-463); their 9 amplitudes
enter the displacement gradient like nodal values
(get_inst_disp_dfd0, S/element_soln_fields.rs:718-748).  Each Newton
iteration condenses them (condense_mat, update_external,
S/element_equations.rs:19-121): K_c = K_ee - K_ei K_ii^-1 K_ie and
r_c = r_e - K_ei K_ii^-1 r_i; after the solve the amplitudes move by
-K_ii^-1 (r_i + K_ie du_e) (update_internal, :123-194) and persist across
iterations and load steps.
The applied load is a dead load: the reference integrates surface
tractions on the undeformed faces (S/face_meth.rs:174-221 reads the node
coordinates, not the displacements).

Isotropic C (Lame lam, mu) in the global frame: the reference rotates the
linear part into the element's local frame (loc_ori) and keeps the quadratic
part H^T H frame-free, which for an isotropic material gives the same
energy and residual.

Pinned by tests/test_nonlinear_oracle.py without this module's formulas:
the residual is the finite-difference gradient of the separately written
strain energy (1/2) E:C:E, the tangent is the finite-difference Jacobian
of the residual, a finite rigid rotation gives zero strain and zero
residual, the tangent at u = 0 is the linear oracle's element matrix, and a
uniaxially loaded cube matches the closed-form St. Venant-Kirchhoff
nominal stress P = Y (1 + e) (e + e^2 / 2).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_G = 1.0 / np.sqrt(3.0)
# brick8 node signs (node t at (sx, sy, sz) of the reference cube)
_SIGN = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
# 2x2x2 Gauss points, weight 1 (S/element_meth.rs:175-191)
_GP = np.array([[(1 if g & 1 else -1) * _G, (1 if g & 2 else -1) * _G, (1 if g & 4 else -1) * _G]
                for g in range(8)])


def _dn_natural():
    """dN_t/dxi_c at the Gauss points: shape (gp, t, c)."""
    f = 1.0 + _GP[:, None, :] * _SIGN[None, :, :]        # (gp, t, c)
    d = np.empty((8, 8, 3))
    d[:, :, 0] = 0.125 * _SIGN[None, :, 0] * f[:, :, 1] * f[:, :, 2]
    d[:, :, 1] = 0.125 * f[:, :, 0] * _SIGN[None, :, 1] * f[:, :, 2]
    d[:, :, 2] = 0.125 * f[:, :, 0] * f[:, :, 1] * _SIGN[None, :, 2]
    return d


_DN = _dn_natural()


def lame(E, nu):
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    return E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


def _gradients(X, im=False):
    """Physical shape-function gradients G (ne, gp, t, 3) and detJ (ne, gp);
    with im, t runs over the 8 nodes and the 3 bubbles."""
    J = np.einsum('eti,gtc->egic', X, _DN)              # J_ic = dx_i / dxi_c
    det = np.linalg.det(J)
    Ji = np.linalg.inv(J)                                 # Ji_ci = dxi_c / dx_i
    G = np.einsum('gtk,egkc->egtc', _DN, Ji)
    if im:
        Jc = np.einsum('eti,tc->eic', X, 0.125 * _SIGN)   # Jacobian at the centroid
        Jci = np.linalg.inv(Jc)
        # bubble k: d(1 - s_k^2)/ds_k = -2 s_k at the Gauss point
        Gb = -2.0 * _GP[None, :, :, None] * Jci[:, None, :, :]   # (e, gp, k, c)
        G = np.concatenate([G, Gb], axis=2)
    return G, det


def _kinematics(X, U, E, nu):
    G, det = _gradients(X, im=U.shape[1] == 11)
    H = np.einsum('eti,egtj->egij', U, G)                 # H_ij = du_i / dx_j
    F = H + np.eye(3)
    Egl = 0.5 * (np.einsum('egki,egkj->egij', F, F) - np.eye(3))
    lam, mu = lame(E, nu)
    tr = np.trace(Egl, axis1=2, axis2=3)
    S = 2 * mu[:, None, None, None] * Egl + (lam[:, None] * tr)[:, :, None, None] * np.eye(3)
    return G, det, F, Egl, S, lam, mu


def element_energy(X, U, E, nu):
    """Strain energy sum_gp detJ (lam tr(E)^2 / 2 + mu E:E) per element,
    written from the Green-Lagrange tensor alone.  U: (ne, 8, 3), or
    (ne, 11, 3) for brickIM (nodal values, then the bubble amplitudes)."""
    G, det, F, Egl, S, lam, mu = _kinematics(X, U, E, nu)
    tr = np.trace(Egl, axis1=2, axis2=3)
    w = 0.5 * lam[:, None] * tr ** 2 + mu[:, None] * np.einsum('egij,egij->eg', Egl, Egl)
    return np.einsum('eg,eg->e', det, w)


def element_residual_tangent(X, U, E, nu, im=False):
    """Element residual (ne, 24) and tangent (ne, 24, 24), DOF 3 t + i;
    brickIM (im): (ne, 33) and (ne, 33, 33) over the 8 nodes and 3 bubbles.

    r_ti = sum_gp detJ (F S g_t)_i (first Piola stress on the gradient);
    K_ti,sk = sum_gp detJ [lam (F g_t)_i (F g_s)_k
                           + mu ((F F^T)_ik g_t.g_s + (F g_s)_i (F g_t)_k)
                           + delta_ik g_t^T S g_s].
    """
    nt = 11 if im else 8
    X = np.asarray(X, dtype=np.float64).reshape(-1, 8, 3)
    U = np.asarray(U, dtype=np.float64).reshape(-1, nt, 3)
    E = np.atleast_1d(np.asarray(E, dtype=np.float64))
    nu = np.broadcast_to(np.asarray(nu, dtype=np.float64), E.shape)
    G, det, F, Egl, S, lam, mu = _kinematics(X, U, E, nu)
    P = np.einsum('egij,egjl->egil', F, S)
    r = np.einsum('eg,egil,egtl->eti', det, P, G).reshape(-1, 3 * nt)
    FG = np.einsum('egil,egtl->egti', F, G)               # (F g_t)_i
    FFt = np.einsum('egil,egkl->egik', F, F)
    gg = np.einsum('egtl,egsl->egts', G, G)
    gSg = np.einsum('egtj,egjl,egsl->egts', G, S, G)
    K = (lam[:, None, None, None, None, None] * np.einsum('egti,egsk->egtisk', FG, FG)
         + mu[:, None, None, None, None, None] * (np.einsum('egik,egts->egtisk', FFt, gg)
                                                  + np.einsum('egsi,egtk->egtisk', FG, FG))
         + np.einsum('egts,ik->egtisk', gSg, np.eye(3)))
    K = np.einsum('eg,egtisk->etisk', det, K).reshape(-1, 3 * nt, 3 * nt)
    return r, K


def condense(r, K):
    """brickIM: condensed residual and tangent of (ne, 33) / (ne, 33, 33),
    and the maps W = K_ii^-1 K_ie, w = K_ii^-1 r_i of the internal update."""
    Kee, Kei, Kie, Kii = K[:, :24, :24], K[:, :24, 24:], K[:, 24:, :24], K[:, 24:, 24:]
    W = np.linalg.solve(Kii, Kie)
    w = np.linalg.solve(Kii, r[:, 24:, None])[:, :, 0]
    Kc = Kee - np.einsum('eik,ekj->eij', Kei, W)
    rc = r[:, :24] - np.einsum('eik,ek->ei', Kei, w)
    return rc, Kc, W, w


def _element_dofs(conn, dof_index):
    """DOF numbers (ne, 24) of the elements; dof_index holds the DOFs of a
    node in its first three columns (the reorder output has six)."""
    return np.asarray(dof_index)[:, :3][np.asarray(conn)].reshape(-1, 24)


def assemble(xyz, conn, E, nu, dof_index, n_dof, u, alpha=None):
    """Global internal residual r(u) (n_dof) and tangent (scipy CSR); with
    alpha (ne, 9, brickIM bubble amplitudes, 3 k + component) the condensed
    ones, and the internal-update maps (W, w) as a third item."""
    conn = np.asarray(conn)
    dofs = _element_dofs(conn, dof_index)
    X = np.asarray(xyz, dtype=np.float64)[conn]
    U = np.asarray(u, dtype=np.float64)[dofs].reshape(-1, 8, 3)
    maps = None
    if alpha is None:
        re, Ke = element_residual_tangent(X, U, E, nu)
    else:
        U = np.concatenate([U, np.asarray(alpha).reshape(-1, 3, 3)], axis=1)
        r33, K33 = element_residual_tangent(X, U, E, nu, im=True)
        re, Ke, W, w = condense(r33, K33)
        maps = (W, w)
    r = np.zeros(n_dof)
    np.add.at(r, dofs, re)
    rows = np.repeat(dofs, 24, axis=1).ravel()
    cols = np.tile(dofs, (1, 24)).ravel()
    K = sp.csr_matrix((Ke.ravel(), (rows, cols)), shape=(n_dof, n_dof))
    K.sum_duplicates()
    if maps is not None:
        return r, K, maps
    return r, K


def ramp_factors(load_steps):
    """Load factors of the static nonlinear ramp (S/model_analysis.rs:636-643)."""
    L = int(load_steps)
    return [1.0 - ((L - 1 - s) / L) ** 2 for s in range(L)]


def newton(xyz, conn, E, nu, dof_index, n_dof, f, pen_dofs=(), sf=None, load_steps=1, max_newton=50,
           du_tol=1e-12, fixed=None, u0=None, im=False, alpha=None):
    """Static nonlinear solve by the reference's Newton loop.

    Constraints: penalty sf * sum C^T C on pen_dofs (a DOF listed twice
    counts twice), sf = 1e5 max|K| of the first tangent when None
    (S/model_analysis.rs:255-259); or, with fixed = DOF list, exact
    elimination of those DOFs (for the closed-form checks).
    Returns u, the Newton iterations of each load step, the max|du| of
    every iteration and sf; with im (brickIM) also the bubble amplitudes
    (ne, 9), which start from alpha (zero by default).
    """
    u = np.zeros(n_dof) if u0 is None else np.array(u0, dtype=np.float64)
    pen = np.asarray(pen_dofs, dtype=np.int64)
    pdiag = np.zeros(n_dof)
    np.add.at(pdiag, pen, 1.0)
    free = None
    if fixed is not None:
        mask = np.ones(n_dof, bool)
        mask[np.asarray(fixed)] = False
        free = np.flatnonzero(mask)
    iters, steps = [], []
    if im:
        alpha = np.zeros((np.asarray(conn).shape[0], 9)) if alpha is None else np.array(alpha, dtype=np.float64)
        dofs = _element_dofs(conn, dof_index)
    for fact in ramp_factors(load_steps):
        it = 0
        while it < max_newton:
            if im:
                r, K, (W, w) = assemble(xyz, conn, E, nu, dof_index, n_dof, u, alpha)
            else:
                r, K = assemble(xyz, conn, E, nu, dof_index, n_dof, u)
            if sf is None and pen.size:
                sf = 1e5 * abs(K).max()
            rhs = fact * f - r
            if free is None:
                s = 0.0 if sf is None else sf
                rhs = rhs - s * pdiag * u
                A = (K + sp.diags(s * pdiag)).tocsc()
                du = spla.spsolve(A, rhs)
            else:
                du = np.zeros(n_dof)
                du[free] = spla.spsolve(K[free][:, free].tocsc(), rhs[free])
            u += du
            if im:
                alpha -= w + np.einsum('eij,ej->ei', W, du[dofs])
            it += 1
            dmax = float(np.abs(du).max())
            steps.append(dmax)
            if dmax <= du_tol:
                break
        iters.append(it)
    if im:
        return u, iters, steps, sf, alpha
    return u, iters, steps, sf
