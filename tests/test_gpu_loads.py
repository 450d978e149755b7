"""Body and pressure loads through the device solve, against closed forms
the reference's own test notes give (examples/testCases/hexBeam/notes.txt):
the hexBeam bar (L = 10, b = 1, h = 0.1, E = 1e6, nu = 0, rho = 1, brick8
100 x 10 x 1) spun about the z axis through (0, 0.5, 0) at omega = 1
stretches by rho omega^2 L^3 / (3 E) = 3.3333e-4 at the tip (test 5); under
gravity and a tip pressure the clamp reactions of a stocky block balance the
applied totals."""
import numpy as np
import pytest

import oracle as O
from asend3d_amd import api, problems
from asend3d_amd.meshgen import hex_block
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def hexbeam(dims=(100, 10, 1), size=(10.0, 1.0, 0.1)):
    m = hex_block(*dims, *size)
    clamp = m.nodes_at_x(0)
    ordered, di, n_dof = api.reorder_nodes(m.xyz, m.conn, 3, [[clamp], [clamp], [clamp]],
                                           api.block_dim_auto(m.n_nodes))
    pen = np.concatenate([di[clamp, 0], di[clamp, 1], di[clamp, 2]]).astype(np.int64)
    return m, di, n_dof, clamp, pen


def solve(ctx, m, di, n_dof, pen, f, E=1e6, nu=0.0, rho=1.0):
    ctx.upload_mesh(8, m.xyz, di, n_dof, m.conn, np.full(m.n_el, E), np.full(m.n_el, nu), np.full(m.n_el, rho))
    ctx.set_penalty(pen, 0.0)
    ctx.assemble()
    u, st = ctx.solve(f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-12, max_it=400000))
    assert st.converged
    return u


def test_hexbeam_centrifugal_tip_stretch(ctx):
    gpu_case("hexBeam test 5: centrifugal load")
    m, di, n_dof, clamp, pen = hexbeam()
    every = np.arange(m.n_el, dtype=np.int32)
    f = api.body_load(8, api.CENTRIFUGAL, m.xyz, m.conn, di, n_dof, every, rho=np.ones(m.n_el),
                      center=[0.0, 0.5, 0.0], axis=[0.0, 0.0, 1.0], ang_vel=1.0)
    fo = O.body_load(8, O.CENTRIFUGAL, m.xyz, m.conn, di, n_dof, every, rho=np.ones(m.n_el),
                     center=[0.0, 0.5, 0.0], axis=[0.0, 0.0, 1.0], ang_vel=1.0)
    assert np.allclose(f, fo, rtol=1e-12, atol=1e-15)
    u = solve(ctx, m, di, n_dof, pen, f)
    tip = m.nodes_at_x(m.nx)
    ux = u[di[tip, 0]].mean()
    print(f"[loads] hexBeam centrifugal tip u_x {ux:.6e} (closed form {1000.0 / 3e6:.6e})", flush=True)
    assert abs(ux - 1000.0 / 3e6) <= 5e-3 * (1000.0 / 3e6)


def test_gravity_and_pressure_reactions(ctx):
    gpu_case("gravity + tip pressure: clamp reactions")
    m, di, n_dof, clamp, pen = hexbeam((20, 4, 4), (2.0, 0.4, 0.4))
    every = np.arange(m.n_el, dtype=np.int32)
    rho = np.linspace(1.0, 2.0, m.n_el)
    f = api.body_load(8, api.GRAVITATIONAL, m.xyz, m.conn, di, n_dof, every, rho=rho, load=[0.0, 0.0, -1.0])
    f = api.surface_pressure(8, m.xyz, m.conn, di, n_dof, m.elements_at_x(m.nx - 1), 0.5, [1.0, 0.0, 0.0], 5.0, f=f)
    ctx.upload_mesh(8, m.xyz, di, n_dof, m.conn, np.full(m.n_el, 1e6), np.full(m.n_el, 0.3), rho)
    ctx.set_penalty(pen, 0.0)
    ctx.assemble()
    u, st = ctx.solve(f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11, max_it=400000))
    assert st.converged, st.as_dict()
    # support reactions K u - f at the clamp (K without the penalty) balance
    # the whole applied load, including the share on the clamped nodes
    sf = ctx.penalty_scale()
    r = ctx.spmv(u)
    r[pen] -= sf * u[pen]
    reac = (r - f)[di[clamp, :3]].sum(axis=0)
    V = 2.0 * 0.4 * 0.4 / m.n_el
    total = np.array([-0.5 * 0.4 * 0.4, 0.0, -(rho * V).sum()])
    assert np.allclose(reac, -total, rtol=1e-7, atol=1e-9)
