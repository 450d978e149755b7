"""Pins the geometrically nonlinear brick8 oracle (oracle/nonlinear.py) with
checks that do not reuse its formulas: finite differences of the strain
energy and of the residual, a finite rigid rotation, the linear element at
u = 0, and the closed-form St. Venant-Kirchhoff uniaxial stretch.  CPU only."""
import numpy as np
import pytest
from scipy.optimize import brentq

import nonlinear as NL
import oracle as O
from asend3d_amd.meshgen import hex_block

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def distorted(seed=7):
    rng = np.random.default_rng(seed)
    return UNIT * np.array([2.0, 1.0, 0.5]) + 0.08 * rng.standard_normal((8, 3))


def lame_c(E, nu):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[np.arange(3), np.arange(3)] = lam + 2 * mu
    C[3, 3] = C[4, 4] = C[5, 5] = mu
    return C


def rotation(axis, deg):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    t = np.radians(deg)
    Kx = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * Kx + (1 - np.cos(t)) * Kx @ Kx


def test_ramp_factors():
    f = NL.ramp_factors(5)
    assert len(f) == 5 and f[-1] == 1.0
    assert f[0] == pytest.approx(9 / 25)          # (2L - 1) / L^2
    inc = np.diff([0.0] + f)
    assert np.all(inc > 0) and np.all(np.diff(inc) < 0)  # front-loaded steps
    assert NL.ramp_factors(1) == [1.0]


@pytest.mark.parametrize("seed", [1, 2])
def test_residual_is_energy_gradient(seed):
    X = distorted(seed)
    rng = np.random.default_rng(seed + 10)
    U = 0.06 * rng.standard_normal((8, 3))
    E, nu = 70e3, 0.31
    r, _ = NL.element_residual_tangent(X[None], U[None], E, nu)
    h = 1e-6
    fd = np.zeros(24)
    for k in range(24):
        Up, Um = U.copy().ravel(), U.copy().ravel()
        Up[k] += h
        Um[k] -= h
        fd[k] = (NL.element_energy(X[None], Up.reshape(1, 8, 3), [E], [nu])[0]
                 - NL.element_energy(X[None], Um.reshape(1, 8, 3), [E], [nu])[0]) / (2 * h)
    assert np.abs(r[0] - fd).max() <= 1e-6 * np.abs(r[0]).max()


@pytest.mark.parametrize("seed", [3, 4])
def test_tangent_is_residual_jacobian(seed):
    X = distorted(seed)
    rng = np.random.default_rng(seed + 20)
    U = 0.06 * rng.standard_normal((8, 3))
    E, nu = 70e3, 0.27
    _, K = NL.element_residual_tangent(X[None], U[None], E, nu)
    h = 1e-7
    fd = np.zeros((24, 24))
    for k in range(24):
        Up, Um = U.copy().ravel(), U.copy().ravel()
        Up[k] += h
        Um[k] -= h
        fd[:, k] = (NL.element_residual_tangent(X[None], Up.reshape(1, 8, 3), E, nu)[0][0]
                    - NL.element_residual_tangent(X[None], Um.reshape(1, 8, 3), E, nu)[0][0]) / (2 * h)
    scale = np.abs(K[0]).max()
    assert np.abs(K[0] - fd).max() <= 1e-6 * scale
    assert np.abs(K[0] - K[0].T).max() <= 1e-12 * scale


def test_rigid_rotation_is_strain_free():
    X = distorted(5)
    Q = rotation([1.0, 2.0, -0.5], 40.0)
    U = X @ Q.T - X + np.array([0.3, -0.1, 0.2])
    E, nu = 70e3, 0.3
    r, K = NL.element_residual_tangent(X[None], U[None], E, nu)
    Klin, _ = O.solid_element(8, X, lame_c(E, nu))
    lin = Klin @ U.ravel()
    assert np.abs(lin).max() > 1e3                      # the linear strain of a rotation is not zero
    assert np.abs(r[0]).max() <= 1e-10 * np.abs(lin).max()
    assert abs(NL.element_energy(X[None], U[None], [E], [nu])[0]) <= 1e-20 * np.abs(lin).max() ** 2


def test_tangent_at_zero_is_linear_element():
    X = distorted(6)
    E, nu = 200e9, 0.3
    r, K = NL.element_residual_tangent(X[None], np.zeros((1, 8, 3)), E, nu)
    Klin, _ = O.solid_element(8, X, lame_c(E, nu))
    assert np.abs(r).max() == 0.0
    assert np.abs(K[0] - Klin).max() <= 1e-12 * np.abs(Klin).max()


def dof_table(n_nodes):
    """Node-major DOF numbers in the reorder output's six-column layout."""
    di = np.full((n_nodes, 6), -1, dtype=np.int64)
    di[:, :3] = np.arange(3 * n_nodes).reshape(-1, 3)
    return di


def uniaxial_cube(n=2):
    m = hex_block(n, n, n, 1.0, 1.0, 1.0)
    di = dof_table(m.n_nodes)
    fixed = np.concatenate([di[np.isclose(m.xyz[:, c], 0.0), c] for c in range(3)])
    return m, di, fixed


def test_uniaxial_stretch_closed_form():
    # A cube on rollers (u_x = 0 on x = 0, u_y = 0 on y = 0, u_z = 0 on z = 0)
    # under a dead nominal traction t on x = 1 deforms homogeneously; with
    # S22 = S33 = 0 the Green strain is E22 = E33 = -nu E11 and the nominal
    # stress is P11 = F11 S11 = Y (1 + e) (e + e^2 / 2).
    m, di, fixed = uniaxial_cube(2)
    Y, nu = 70e3, 0.3
    t = 0.06 * Y
    tip = m.elements_at_x(m.nx - 1)
    f = O.surface_traction(8, m.xyz, m.conn, di, 3 * m.n_nodes, tip, (t, 0.0, 0.0), (1.0, 0.0, 0.0), 5.0)
    E = np.full(m.n_el, Y)
    nuv = np.full(m.n_el, nu)
    u, iters, steps, _ = NL.newton(m.xyz, m.conn, E, nuv, di, 3 * m.n_nodes, f, fixed=fixed, load_steps=2)
    e = brentq(lambda e: Y * (1 + e) * (e + 0.5 * e * e) - t, 0.0, 0.5)
    ey = np.sqrt(1.0 + 2.0 * (-nu * (e + 0.5 * e * e))) - 1.0
    ux = u[di[:, 0]]
    uy = u[di[:, 1]]
    assert np.allclose(ux, e * m.xyz[:, 0], rtol=0, atol=1e-10)
    assert np.allclose(uy, ey * m.xyz[:, 1], rtol=0, atol=1e-10)
    assert abs(e - t / Y) > 1e-3                        # far from the linear answer
    assert steps[-1] <= 1e-12 and max(iters) <= 8


def bent_cantilever(scale):
    m = hex_block(10, 2, 2, 1.0, 0.2, 0.2)
    di = dof_table(m.n_nodes)
    clamp = m.nodes_at_x(0)
    pen = np.concatenate([di[clamp, c] for c in range(3)])
    tip = m.elements_at_x(m.nx - 1)
    f = O.surface_traction(8, m.xyz, m.conn, di, 3 * m.n_nodes, tip, (0.0, 0.0, -scale), (1.0, 0.0, 0.0), 5.0)
    E = np.full(m.n_el, 200e9)
    nu = np.full(m.n_el, 0.3)
    return m, di, pen, f, E, nu


def test_newton_converges_quadratically_and_path_independent():
    m, di, pen, f, E, nu = bent_cantilever(2.0e8)   # tip deflection ~ 9% of the length
    n = 3 * m.n_nodes
    u1, it1, st1, sf = NL.newton(m.xyz, m.conn, E, nu, di, n, f, pen, load_steps=1)
    u3, it3, st3, _ = NL.newton(m.xyz, m.conn, E, nu, di, n, f, pen, sf=sf, load_steps=3)
    tip_w = u1[di[m.nodes_at_x(m.nx), 2]].mean()
    assert -0.15 < tip_w < -0.05
    assert np.abs(u1 - u3).max() <= 1e-9 * np.abs(u1).max()
    assert st1[-1] <= 1e-12 and it1[0] <= 12
    # quadratic convergence: somewhere the step size squares (ratio << previous)
    d = np.asarray(st1)
    assert np.any((d[1:] / d[:-1]) < 1e-3)


def test_small_load_linear_limit():
    out = []
    for s in (1e5, 1e4):
        m, di, pen, f, E, nu = bent_cantilever(s)
        n = 3 * m.n_nodes
        u, _, _, sf = NL.newton(m.xyz, m.conn, E, nu, di, n, f, pen)
        K = O.assemble_solid(8, m.xyz, m.conn, E, nu, di, n).to_scipy()
        import scipy.sparse as sp
        import scipy.sparse.linalg as spla
        pd = np.zeros(n)
        np.add.at(pd, pen, 1.0)
        ul = spla.spsolve((K + sp.diags(sf * pd)).tocsc(), f)
        out.append(np.abs(u - ul).max() / np.abs(ul).max())
    # the nonlinear correction is first order in the load
    assert 5.0 < out[0] / out[1] < 20.0
    assert out[0] < 1e-3


# ---------------------------------------------------------------- brickIM
def test_brickim_residual_tangent_energy_and_rotation():
    """brickIM: the residual over the 8 nodes and 3 bubbles is the gradient
    of the strain energy, the tangent the Jacobian of the residual, a finite
    rigid rotation (bubbles at rest) carries no force, and the condensed
    tangent at rest is the linear oracle's condensed element."""
    X = distorted(5)
    rng = np.random.default_rng(31)
    U = 0.05 * rng.standard_normal((11, 3))
    E, nu = 70e3, 0.3
    r, K = NL.element_residual_tangent(X[None], U[None], E, nu, im=True)
    h = 1e-6
    fd = np.zeros(33)
    for k in range(33):
        Up, Um = U.copy().ravel(), U.copy().ravel()
        Up[k] += h
        Um[k] -= h
        fd[k] = (NL.element_energy(X[None], Up.reshape(1, 11, 3), [E], [nu])[0]
                 - NL.element_energy(X[None], Um.reshape(1, 11, 3), [E], [nu])[0]) / (2 * h)
    assert np.abs(r[0] - fd).max() <= 1e-6 * np.abs(r[0]).max()
    h = 1e-7
    J = np.zeros((33, 33))
    for k in range(33):
        Up, Um = U.copy().ravel(), U.copy().ravel()
        Up[k] += h
        Um[k] -= h
        J[:, k] = (NL.element_residual_tangent(X[None], Up.reshape(1, 11, 3), E, nu, im=True)[0][0]
                   - NL.element_residual_tangent(X[None], Um.reshape(1, 11, 3), E, nu, im=True)[0][0]) / (2 * h)
    assert np.abs(K[0] - J).max() <= 1e-6 * np.abs(K[0]).max()
    R = rotation([1.0, -2.0, 0.5], 40.0)
    Urot = np.vstack([X @ R.T + np.array([0.3, -0.1, 0.2]) - X, np.zeros((3, 3))])
    rr, _ = NL.element_residual_tangent(X[None], Urot[None], E, nu, im=True)
    _, K0 = NL.element_residual_tangent(X[None], np.zeros((1, 11, 3)), E, nu, im=True)
    assert np.abs(rr).max() <= 1e-9 * np.abs(K0).max() * np.abs(Urot).max()
    _, Kc, _, _ = NL.condense(np.zeros((1, 33)), K0)
    Klin, _ = O.solid_element(81, X, lame_c(E, nu))
    assert np.abs(Kc[0] - Klin).max() <= 1e-10 * np.abs(Klin).max()


def hexbeam_moment(nx, ny):
    """hexBeam test 6 (E/testCases/hexBeam/tipMomentNonlinear.py): 10 x 1 x 0.1
    brickIM beam, E = 1e6, nu = 0, clamped at x = 0; at the tip, dead forces
    c (1, 0, 1) on the bottom edge nodes and -c (1, 0, 1) on the top edge
    nodes, c = 65.4498469 / (11 sqrt 2) for the reference's 11 nodes per
    edge (the total scaled to ny + 1 nodes), in 5 ramp steps."""
    m = hex_block(nx, ny, 1, 10.0, 1.0, 0.1)
    clamp = m.nodes_at_x(0)
    _, di, n_dof = O.reorder_nodes(m.xyz, m.conn, 3, [[clamp]] * 3, 2000000000)
    di = np.asarray(di).reshape(-1, 6)
    tip = m.nodes_at_x(nx)
    c = 65.4498469 / ((ny + 1) * np.sqrt(2.0))
    f = np.zeros(n_dof)
    for nd in tip:
        s = 1.0 if m.xyz[nd, 2] < 0.05 else -1.0
        f[di[nd, 0]] += s * c
        f[di[nd, 2]] += s * c
    pen = np.concatenate([di[clamp, k] for k in range(3)])
    E = np.full(m.n_el, 1e6)
    nu = np.zeros(m.n_el)
    return m, di, n_dof, f, pen, E, nu


def test_brickim_hexbeam_tip_moment_known_answer():
    """A 45 degree arc: tip displacement [-0.99684, 0, 3.729232]
    (E/testCases/hexBeam/notes.txt, test 6) on a 20 x 2 x 1 mesh, to 5e-4 of
    its length (the reference's 100 x 10 x 1 mesh gives [-0.996865, 0, 3.729197])."""
    m, di, n_dof, f, pen, E, nu = hexbeam_moment(20, 2)
    u, iters, steps, sf, alpha = NL.newton(m.xyz, m.conn, E, nu, di, n_dof, f, pen, load_steps=5, im=True)
    w = tip_axis_displacement(m, di, u)
    want = np.array([-0.99684, 0.0, 3.729232])
    assert np.abs(w - want).max() <= 5e-4 * np.linalg.norm(want), (w, iters)
    assert max(iters) < 50


def tip_axis_displacement(m, di, u):
    """Displacement of the tip's beam axis (z = h/2, y = b/2): the mean of the
    bottom and top tip nodes at y = b/2 (the known answer is the elastica's
    axis; a face node sits h/2 sin(theta) off it along x)."""
    tip = m.nodes_at_x(m.nx)
    at = tip[np.abs(m.xyz[tip, 1] - 0.5) < 1e-9]
    return np.mean([[u[di[nd, k]] for k in range(3)] for nd in at], axis=0)
