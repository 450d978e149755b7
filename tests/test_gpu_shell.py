"""Device shell4 integration (asend_upload_mesh with ASEND_SHELL4) against the
oracle's restatement of the reference's quad shell: the assembled stiffness
and consistent mass on a distorted plate with per-element moduli, Poisson
ratios, thicknesses and densities; a static clamped-plate solve; the
reference's shellBeam known answers (examples/testCases/shellBeam/notes.txt);
Newmark stepping of a clamped plate under a half-sine centre load; and the
error paths of the shell upload."""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import dynamics
import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def distorted_plate(nx, ny, seed, clamp="edges"):
    """Flat plate with in-plane-perturbed interior nodes and per-element
    properties."""
    p = problems.shell_plate(nx, ny, 1.0, 0.8, 0.005, 70e9, 0.33, 2700.0, clamp=clamp)
    rng = np.random.default_rng(seed)
    xyz = p.mesh.xyz.copy()
    interior = np.setdiff1d(np.arange(p.mesh.n_nodes), p.mesh.boundary_nodes())
    h = min(1.0 / nx, 0.8 / ny)
    xyz[interior, :2] += rng.uniform(-0.2, 0.2, (interior.size, 2)) * h
    xyz[:, 2] = 0.25  # any plane normal to z
    n = p.mesh.n_el
    return dataclasses.replace(p, mesh=dataclasses.replace(p.mesh, xyz=xyz),
                               E=70e9 * np.exp(0.5 * rng.standard_normal(n)), nu=rng.uniform(0.2, 0.4, n),
                               t=rng.uniform(0.003, 0.008, n), rho=rng.uniform(2000.0, 8000.0, n))


def oracle_km(p):
    K, M = O.assemble_shell(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.t, p.rho, p.dof_index, p.n_dof)
    return K.to_scipy(), M.to_scipy()


def device_matrix(ctx, c_mass):
    ctx.assemble(c_mass)
    rp, cols, vals = ctx.get_matrix(with_penalty=False)
    return sp.csr_matrix((vals, cols, rp), shape=(ctx.n_dof, ctx.n_dof))


def penalised_solver(K, pen_dofs, extra=None):
    A = K if extra is None else K + extra
    sf = 1e5 * abs(A).max()
    pen = np.zeros(K.shape[0])
    np.add.at(pen, pen_dofs, sf)
    return spla.splu((A + sp.diags(pen)).tocsc()).solve


def test_shell_assembly_matches_oracle(ctx):
    gpu_case("shell4 assembly 9x7 distorted plate vs oracle")
    p = distorted_plate(9, 7, seed=3)
    K, M = oracle_km(p)
    problems.upload_shell(ctx, p)
    # FP64 roundoff: the condensation of the internal DOFs cancels digits
    # (measured 6e-12 of max|K| between the two orders of summation)
    Kd = device_matrix(ctx, 0.0)
    assert abs(Kd - K).max() <= 1e-10 * abs(K).max()
    c = abs(K).max() / abs(M).max()
    KMd = device_matrix(ctx, c)
    assert abs(KMd - (K + c * M)).max() <= 1e-10 * abs(K + c * M).max()
    Md = (KMd - Kd) / c
    # symmetric, and the translational mass is rho t area
    assert abs(Kd - Kd.T).max() <= 1e-10 * abs(Kd).max()
    di = np.asarray(p.dof_index).reshape(-1, 6)
    X = p.mesh.xyz[p.mesh.conn]
    area = 0.5 * np.abs(np.einsum("ei,ei->e", X[:, :, 0], np.roll(X[:, :, 1], -1, axis=1))
                        - np.einsum("ei,ei->e", X[:, :, 1], np.roll(X[:, :, 0], -1, axis=1)))
    ux = np.zeros(p.n_dof)
    ux[di[:, 0]] = 1.0
    assert np.isclose(ux @ (Md @ ux), np.sum(p.rho * p.t * area), rtol=1e-9)


def test_shell_static_plate_matches_oracle(ctx):
    gpu_case("shell4 static clamped plate 12x10 vs oracle")
    p = distorted_plate(12, 10, seed=5)
    K, _ = oracle_km(p)
    f = np.zeros(p.n_dof)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    f[di[:, 2]] = -1.0e3  # uniform transverse load
    f[di[:, 0]] = 2.0e2
    u_o = penalised_solver(K, p.pen_dofs)(f)
    problems.upload_shell(ctx, p)
    ctx.assemble(0.0)
    u, st = ctx.solve(f, api.Context.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-12, max_it=200000))
    assert st.converged, st.as_dict()
    assert np.abs(u - u_o).max() <= 1e-8 * np.abs(u_o).max()


def test_shell_beam_known_answers(ctx):
    gpu_case("shell4 shellBeam known answers")
    p = problems.shell_plate(10, 1, 10.0, 1.0, 0.1, 1.0e6, 0.0, 1.0, clamp="x0")
    di = np.asarray(p.dof_index).reshape(-1, 6)
    tip = p.mesh.nodes_at_x(10)
    problems.upload_shell(ctx, p)
    ctx.assemble(0.0)
    # the FP64 floor of the relative residual here is about eps max|K| |u| / |f| ~ 1e-10
    o = api.Context.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-9, max_it=200000)
    f = np.zeros(p.n_dof)
    f[di[tip, 2]] = 1.0  # Fz = 2: FL^3/(3EI) = 8
    u, st = ctx.solve(f, o)
    assert st.converged, st.as_dict()
    assert abs(u[di[tip, 2]].mean() - 8.0) <= 0.02 * 8.0
    K, _ = oracle_km(p)
    u_o = penalised_solver(K, p.pen_dofs)(f)
    err = np.abs(u - u_o).max() / np.abs(u_o).max()
    assert err <= 1e-6, err
    f = np.zeros(p.n_dof)
    f[di[tip, 0]] = 1.0  # Fx = 2: FL/(Ebh) = 2e-4
    u, st = ctx.solve(f, o)
    assert st.converged, st.as_dict()
    assert np.allclose(u[di[tip, 0]], 2e-4, rtol=1e-7), u[di[tip, 0]]


@pytest.mark.parametrize("mode", [api.NM_MIXED, api.NM_FP64])
def test_shell_newmark_matches_oracle(ctx, mode):
    gpu_case(f"shell4 newmark 16x16 clamped plate vs oracle, solver {mode}")
    p = problems.shell_plate(16, 16, 1.0, 1.0, 0.005, 70e9, 0.33, 2700.0)
    K, M = oracle_km(p)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    centre = p.mesh.node_id(8, 8)
    f = np.zeros(p.n_dof)
    f[di[centre, 2]] = 1.0e4
    dt, n, n_pulse = 1e-5, 80, 50
    fac = np.where(np.arange(n) < n_pulse, np.sin(np.pi * np.arange(n) / n_pulse), 0.0)
    watch = di[centre, :3]
    hist_o, u_o, v_o, a_o = dynamics.newmark(K, M, f, fac, dt, pen_dofs=p.pen_dofs, watch=watch)
    problems.upload_shell(ctx, p)
    ctx.newmark_setup(dt, 0.25, 0.5)
    try:
        ctx.newmark_set_solver(mode)
        hist, u, v, a, st = ctx.newmark_run(n, f, fac, watch=watch, tol=1e-12, max_it_step=4000)
    finally:
        ctx.newmark_set_solver(api.NM_AUTO)
    assert st.unconverged_steps == 0 and st.steps == n and st.step_solver == mode, st.as_dict()
    scale = np.abs(hist_o[:, 2]).max()
    assert np.abs(hist[:, 2] - hist_o[:, 2]).max() <= 1e-7 * scale
    for x, xo in ((u, u_o), (v, v_o)):
        assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()


def test_shell_mass_damped_newmark_matches_oracle(ctx):
    # mass-proportional Rayleigh damping on the plate (stiffness-proportional
    # damping of elements with condensed DOFs is not built: an error)
    gpu_case("shell4 newmark 12x12 with mass damping vs oracle")
    p = problems.shell_plate(12, 12, 1.0, 1.0, 0.005, 70e9, 0.33, 2700.0)
    K, M = oracle_km(p)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    centre = p.mesh.node_id(6, 6)
    f = np.zeros(p.n_dof)
    f[di[centre, 2]] = 1.0e4
    dt, n = 1e-5, 60
    fac = np.where(np.arange(n) < 40, np.sin(np.pi * np.arange(n) / 40.0), 0.0)
    watch = di[centre, :3]
    cm = 2000.0
    hist_o, u_o, v_o, a_o = dynamics.newmark(K, M, f, fac, dt, pen_dofs=p.pen_dofs, watch=watch, damp_cm=cm)
    problems.upload_shell(ctx, p)
    try:
        ctx.set_rayleigh_damping(0.0, 1e-6)
        with pytest.raises(api.AsendError, match="stiffness-proportional"):
            ctx.newmark_setup(dt, 0.25, 0.5)
        ctx.set_rayleigh_damping(cm, 0.0)
        ctx.newmark_setup(dt, 0.25, 0.5)
        hist, u, v, a, st = ctx.newmark_run(n, f, fac, watch=watch, tol=1e-12, max_it_step=4000)
    finally:
        ctx.set_rayleigh_damping(0.0, 0.0)
    assert st.unconverged_steps == 0 and st.steps == n, st.as_dict()
    scale = np.abs(hist_o[:, 2]).max()
    assert np.abs(hist[:, 2] - hist_o[:, 2]).max() <= 1e-7 * scale
    for x, xo in ((u, u_o), (v, v_o)):
        assert np.abs(x - xo).max() <= 1e-6 * np.abs(xo).max()
    hist_0, *_ = dynamics.newmark(K, M, f, fac, dt, pen_dofs=p.pen_dofs, watch=watch)
    assert np.abs(hist_0[:, 2] - hist_o[:, 2]).max() > 1e-3 * scale


def test_shell_error_paths(ctx):
    gpu_case("shell4 error paths")
    p = problems.shell_plate(4, 3, 1.0, 1.0, 0.01, 70e9, 0.3, 2700.0)
    ctx.upload_mesh(p.etype, p.mesh.xyz, p.dof_index, p.n_dof, p.mesh.conn, p.E, p.nu, p.rho)
    with pytest.raises(api.AsendError, match="thickness"):
        ctx.assemble(0.0)
    with pytest.raises(api.AsendError, match="positive"):
        ctx.set_thickness(np.zeros(p.mesh.n_el))
    ctx.set_thickness(p.t)
    # the penalty set of the previous (larger) mesh does not carry over
    ctx.assemble(0.0)
    assert ctx.penalty_scale() == 0.0
    with pytest.raises(api.AsendError, match="brick8"):
        ctx.set_grid(4, 3, 1)
    with pytest.raises(api.AsendError, match="brick8"):
        ctx.element_sens(np.zeros(p.n_dof), np.zeros(p.n_dof))
    bent = p.mesh.xyz.copy()
    bent[p.mesh.node_id(2, 1), 2] = 0.05
    with pytest.raises(api.AsendError, match="normal to z"):
        ctx.upload_mesh(p.etype, bent, p.dof_index, p.n_dof, p.mesh.conn, p.E, p.nu, p.rho)
