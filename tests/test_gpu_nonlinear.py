"""Device parity of the geometrically nonlinear brick8 static path (element
tangent and residual kernels, row gather, Newton loop under the load ramp)
against the CPU oracle (oracle/nonlinear.py) and the closed-form
St. Venant-Kirchhoff uniaxial stretch.  Every device call goes through the
C ABI."""

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.optimize import brentq

import nonlinear as NL
import oracle as O
from asend3d_amd import api, problems
from asend3d_amd.meshgen import hex_block
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def csr(ctx, n):
    rp, cols, vals = ctx.get_matrix(with_penalty=False)
    return sp.csr_matrix((vals, cols, rp), shape=(n, n))


def assert_operator_close(A, B, pen, sf):
    """A == B to 1e-12 of max|B| off the penalty diagonal; on it, to the
    rounding of removing sf (the download subtracts it from K + sf)."""
    d = np.zeros(A.shape[0])
    d[pen] = 1.0
    P = sp.diags(d)
    off = (A - B) - P @ (A - B) @ P
    assert abs(off).max() <= 1e-12 * abs(B).max()
    assert np.abs((A - B).diagonal()[pen]).max() <= 1e-14 * sf


@pytest.mark.parametrize("which", ["het-8x4x3", "het-8x4x3 distorted"])
def test_tangent_and_residual_match_oracle(ctx, which):
    gpu_case(f"nonlinear tangent {which}")
    p = problems.heterogeneous(8, 4, 3, seed=11)
    if "distorted" in which:
        p = problems.distorted(p, seed=12)
    problems.upload(ctx, p)
    rng = np.random.default_rng(13)
    # displacement gradients of order 0.1 (elements are 125 x 50 x 67 mm)
    u = rng.uniform(-8.0, 8.0, p.n_dof)
    r = ctx.nl_assemble(u)
    Kg = csr(ctx, p.n_dof)
    r_o, K_o = NL.assemble(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof, u)
    assert np.abs(r - r_o).max() <= 1e-12 * np.abs(r_o).max()
    sf = ctx.penalty_scale()
    assert_operator_close(Kg, K_o, p.pen_dofs, sf)
    # at u = 0 the tangent is the linear operator of asend_assemble
    r0 = ctx.nl_assemble(np.zeros(p.n_dof))
    K0 = csr(ctx, p.n_dof)
    ctx.assemble()
    Klin = csr(ctx, p.n_dof)
    assert np.abs(r0).max() == 0.0
    assert_operator_close(K0, Klin, p.pen_dofs, sf)


def bent(scale, nx=10, ny=2, nz=2):
    return problems.cantilever(nx, ny, nz, E=200e9, nu=0.3, traction=(0.0, 0.0, -scale))


def oracle_newton(p, sf=None, load_steps=1):
    return NL.newton(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof, p.f, p.pen_dofs, sf=sf,
                     load_steps=load_steps)


@pytest.mark.parametrize("load_steps", [1, 4])
def test_newton_matches_oracle(ctx, load_steps):
    gpu_case(f"nonlinear Newton cantilever 10x2x2 load_steps={load_steps}")
    p = bent(2.0e8)
    problems.upload(ctx, p)
    opts = api.Context.opts(tol=1e-12, max_outer=50)
    u, st = ctx.static_nonlinear(p.f, opts, load_steps=load_steps)
    u_o, iters, steps, sf = oracle_newton(p, load_steps=load_steps)
    assert abs(ctx.penalty_scale() - sf) <= 1e-12 * sf
    assert np.abs(u - u_o).max() <= 1e-9 * np.abs(u_o).max()
    tip = p.dof_index[p.mesh.nodes_at_x(p.mesh.nx), 2]
    assert -0.15 < u[tip].mean() < -0.05            # a large deflection
    assert st.load_steps == load_steps and st.unconverged_steps == 0 and st.unconverged_solves == 0
    assert st.last_du <= 1e-12
    assert st.newton_iterations <= sum(iters) + 2 * load_steps
    # the same final state whatever the ramp (the energy is path independent)
    u1, _ = ctx.static_nonlinear(p.f, opts, load_steps=1)
    assert np.abs(u1 - u).max() <= 1e-9 * np.abs(u).max()


def test_newton_with_multigrid_mixed_solves(ctx):
    gpu_case("nonlinear Newton cantilever 20x4x4, multigrid FP32 FCG + FP64 refresh")
    p = bent(2.0e8, 20, 4, 4)
    problems.upload(ctx, p)
    ctx.set_grid(20, 4, 4)
    opts = api.Context.opts(method=api.SOLVE_IR_MIXED, precond=api.PC_MG, tol=1e-11, inner_tol=1e-3,
                            max_outer=200, check_every=4)
    u, st = ctx.static_nonlinear(p.f, opts, load_steps=2)
    u_o, iters, steps, sf = oracle_newton(p, load_steps=2)
    assert np.abs(u - u_o).max() <= 1e-8 * np.abs(u_o).max()
    assert st.unconverged_steps == 0 and st.unconverged_solves == 0
    # inexact Newton: same fixed point, fewer Krylov iterations
    ui, sti = ctx.static_nonlinear(p.f, opts, load_steps=2, forcing=0.1)
    assert np.abs(ui - u_o).max() <= 1e-8 * np.abs(u_o).max()
    assert sti.unconverged_steps == 0 and sti.last_du <= 1e-12
    assert sti.linear_iterations < st.linear_iterations


def test_uniaxial_stretch_closed_form(ctx):
    # cube on rollers under a dead nominal traction: homogeneous stretch with
    # P11 = Y (1 + e) (e + e^2 / 2); the penalty rollers move by about
    # reaction / sf, which bounds the agreement with the closed form
    gpu_case("nonlinear uniaxial cube 3x3x3")
    m = hex_block(3, 3, 3, 1.0, 1.0, 1.0)
    sets = [m.xyz[:, c] == 0.0 for c in range(3)]
    cons = [[np.flatnonzero(s).astype(np.int32)] for s in sets]
    ordered, di, n_dof = api.reorder_nodes(m.xyz, m.conn, 3, cons, 24)
    pen = np.concatenate([di[np.flatnonzero(sets[c]), c] for c in range(3)]).astype(np.int64)
    Y, nu = 70e3, 0.3
    t = 0.06 * Y
    f = api.surface_traction(8, m.xyz, m.conn, di, n_dof, m.elements_at_x(m.nx - 1), (t, 0.0, 0.0),
                             (1.0, 0.0, 0.0), 5.0)
    E = np.full(m.n_el, Y)
    nuv = np.full(m.n_el, nu)
    ctx.upload_mesh(8, m.xyz, di, n_dof, m.conn, E, nuv)
    ctx.set_penalty(pen, 0.0)
    u, st = ctx.static_nonlinear(f, api.Context.opts(tol=1e-12, max_outer=50), load_steps=3)
    u_o, _, _, _ = NL.newton(m.xyz, m.conn, E, nuv, di, n_dof, f, pen, load_steps=3)
    assert np.abs(u - u_o).max() <= 1e-9 * np.abs(u_o).max()
    e = brentq(lambda e: Y * (1 + e) * (e + 0.5 * e * e) - t, 0.0, 0.5)
    ey = np.sqrt(1.0 - 2.0 * nu * (e + 0.5 * e * e)) - 1.0
    assert np.abs(u[di[:, 0]] - e * m.xyz[:, 0]).max() <= 1e-4 * e
    assert np.abs(u[di[:, 1]] - ey * m.xyz[:, 1]).max() <= 1e-4 * abs(ey)
    assert abs(e - t / Y) > 1e-3


def rotation(axis, deg):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    t = np.radians(deg)
    Kx = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * Kx + (1 - np.cos(t)) * Kx @ Kx


def test_full_size_rigid_rotation_and_equilibrium(ctx):
    # size-independent checks on the 1M-DOF block: a finite rigid rotation
    # carries no internal force, and the Newton solution is in equilibrium
    gpu_case("nonlinear 200x40x40: rigid rotation, equilibrium of a 2-step solve")
    p = problems.config2()
    problems.upload(ctx, p)
    X = p.mesh.xyz
    Q = rotation([0.3, 1.0, -0.2], 35.0)
    d = X @ Q.T - X
    u = np.zeros(p.n_dof)
    for c in range(3):
        u[p.dof_index[:, c]] = d[:, c]
    ctx.assemble()
    free = np.ones(p.n_dof, bool)
    free[p.pen_dofs] = False
    lin = np.abs(ctx.spmv(u)[free]).max()      # linear internal force of the rotation
    r = ctx.nl_assemble(u)
    assert np.abs(r).max() <= 1e-9 * lin
    ctx.set_grid(200, 40, 40)
    f = 60.0 * p.f
    opts = api.Context.opts(method=api.SOLVE_IR_MIXED, precond=api.PC_MG, tol=1e-10, inner_tol=1e-2,
                            max_outer=200, check_every=4)
    un, st = ctx.static_nonlinear(f, opts, load_steps=2, du_tol=1e-6, forcing=0.1)
    assert st.unconverged_steps == 0
    ri = ctx.nl_assemble(un)
    pd = np.zeros(p.n_dof)
    np.add.at(pd, p.pen_dofs, 1.0)
    res = f - ri - ctx.penalty_scale() * pd * un
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(f)
    tip = p.dof_index[p.mesh.nodes_at_x(200), 2]
    assert un[tip].mean() < -50.0                 # large deflection (mm, length 1000 mm)


def test_nonlinear_rejects_other_elements(ctx):
    gpu_case("nonlinear error path shell4")
    p = problems.shell_plate(4, 4, 1.0, 1.0, 0.01, 70e9, 0.3, 2700.0, clamp="x0")
    problems.upload_shell(ctx, p)
    f = np.zeros(p.n_dof)
    with pytest.raises(api.AsendError, match="brick8 and brickIM"):
        ctx.static_nonlinear(f, api.Context.opts())
    with pytest.raises(api.AsendError, match="brick8 and brickIM"):
        ctx.nl_assemble(np.zeros(p.n_dof))


# ---------------------------------------------------------------- brickIM
def brickim(which):
    import dataclasses
    p = problems.heterogeneous(8, 4, 3, seed=11)
    if "distorted" in which:
        p = problems.distorted(p, seed=12)
    return dataclasses.replace(p, etype=81)


@pytest.mark.parametrize("which", ["het-8x4x3", "het-8x4x3 distorted"])
def test_brickim_condensed_tangent_and_residual_match_oracle(ctx, which):
    gpu_case(f"nonlinear brickIM tangent {which}")
    p = brickim(which)
    problems.upload(ctx, p)
    rng = np.random.default_rng(17)
    u = rng.uniform(-8.0, 8.0, p.n_dof)
    r = ctx.nl_assemble(u)  # bubble amplitudes at zero after the upload
    Kg = csr(ctx, p.n_dof)
    alpha = np.zeros((p.mesh.n_el, 9))
    r_o, K_o, _ = NL.assemble(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof, u, alpha)
    assert np.abs(r - r_o).max() <= 1e-11 * np.abs(r_o).max()
    assert_operator_close(Kg, K_o, p.pen_dofs, ctx.penalty_scale())


@pytest.mark.parametrize("load_steps", [1, 3])
def test_brickim_newton_matches_oracle(ctx, load_steps):
    """Newton with per-iteration condensation and bubble updates: the same
    iterations and displacements as the oracle on a distorted mesh under a
    load that bends the cantilever by ~15% of its length."""
    gpu_case(f"nonlinear brickIM Newton load_steps={load_steps}")
    import dataclasses
    p = problems.distorted(problems.heterogeneous(10, 2, 2, seed=3, traction=(0.0, 0.0, -150.0)), seed=4)
    p = dataclasses.replace(p, etype=81)
    problems.upload(ctx, p)
    ctx.assemble()
    ctx.ldl_factor(0)  # each Newton solve by the exact factor of the new tangent
    # max|du| <= 1e-8 mm: the quadratic tail jumps from ~1e-5 to ~1e-12
    u, st = ctx.static_nonlinear(p.f, api.Context.opts(api.SOLVE_DIRECT, api.PC_JACOBI), load_steps=load_steps,
                                 du_tol=1e-8)
    u_o, iters, steps, sf, alpha = NL.newton(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof, p.f,
                                             p.pen_dofs, load_steps=load_steps, im=True, du_tol=1e-8)
    print(st.newton_iterations, iters, np.abs(u_o).max())
    assert abs(ctx.penalty_scale() - sf) <= 1e-12 * sf
    assert st.newton_iterations == sum(iters) and st.unconverged_steps == 0
    assert np.abs(u_o).max() > 0.1 * 1000.0  # a large deflection (length 1000 mm)
    assert np.abs(u - u_o).max() <= 1e-9 * np.abs(u_o).max()


@pytest.mark.parametrize("mesh", [(20, 2), (100, 10)])
def test_brickim_hexbeam_tip_moment(ctx, mesh):
    """hexBeam test 6 (E/testCases/hexBeam/notes.txt): a tip couple bends
    the 10 x 1 x 0.1 beam into a 45 degree arc, tip displacement
    [-0.99684, 0, 3.729232]; 100 x 10 x 1 is the reference's mesh (each
    Newton iteration solved by the device's exact LDL^T factor)."""
    import test_nonlinear_oracle as TNO
    nx, ny = mesh
    gpu_case(f"nonlinear brickIM hexBeam tip moment {nx}x{ny}x1")
    m, di, n_dof, f, pen, E, nu = TNO.hexbeam_moment(nx, ny)
    ctx.upload_mesh(81, m.xyz, di, n_dof, m.conn, E, nu)
    ctx.set_penalty(pen, 0.0)
    ctx.assemble()
    ctx.ldl_factor(0)
    u, st = ctx.static_nonlinear(f, api.Context.opts(api.SOLVE_DIRECT, api.PC_JACOBI), load_steps=5)
    w = TNO.tip_axis_displacement(m, di, u)
    want = np.array([-0.99684, 0.0, 3.729232])
    print(mesh, w, st.newton_iterations)
    assert st.unconverged_steps == 0
    assert np.abs(w - want).max() <= (5e-4 if nx == 20 else 5e-5) * np.linalg.norm(want)
    if nx == 20:
        u_o, iters, _, _, _ = NL.newton(m.xyz, m.conn, E, nu, di, n_dof, f, pen, load_steps=5, im=True)
        # the stop rule max|du| <= 1e-12 sits near the rounding of the solves
        assert abs(st.newton_iterations - sum(iters)) <= 2, (st.newton_iterations, iters)
        assert np.abs(u - u_o).max() <= 1e-8 * np.abs(u_o).max()
