"""Device brickIM (type 81: brick8 + 3 condensed bubble modes) against the
oracle's restatement: the condensed stiffness assembled on a regular mesh
(scaled reference blocks) and on a distorted mesh with per-element Poisson
ratios (per-element integration and condensation), the consistent mass,
element sensitivities against the oracle's dual-number loop, and a full
static compliance pipeline against a sparse direct solve."""
import dataclasses

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def im_problem(distort, seed=11):
    p = problems.cantilever(10, 3, 3, lx=1000.0, ly=300.0, lz=300.0, traction=(0.0, 0.0, -1.0), etype=81,
                            E=api.lognormal_moduli(seed, 90, 70e3, 0.8))
    if not distort:
        return p
    rng = np.random.default_rng(seed)
    m = p.mesh
    xyz = m.xyz.copy()
    h = 100.0
    inner = np.all((xyz > xyz.min(axis=0) + 1e-9) & (xyz < xyz.max(axis=0) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-0.2, 0.2, (int(inner.sum()), 3)) * h
    return dataclasses.replace(p, mesh=dataclasses.replace(m, xyz=xyz), nu=rng.uniform(0.2, 0.4, m.n_el))


def oracle_k(p):
    return O.assemble_solid(81, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)


def device_matrix(ctx, p, c_mass=0.0):
    ctx.assemble(c_mass)
    rp, cols, vals = ctx.get_matrix(with_penalty=False)
    return sp.csr_matrix((vals, cols, rp), shape=(p.n_dof, p.n_dof))


@pytest.mark.parametrize("distort", [False, True])
def test_brickim_assembly_and_mass_match_oracle(ctx, distort):
    gpu_case(f"brickIM assembly distort={distort}")
    p = im_problem(distort)
    rho = np.random.default_rng(2).uniform(1e-9, 8e-9, p.mesh.n_el)
    problems.upload(ctx, p, rho=rho)
    assert ctx.uniform_elements() == (not distort)
    K = oracle_k(p)
    Ks = K.to_scipy()
    Kd = device_matrix(ctx, p)
    # FP64 roundoff of the bubble condensation (QR in the oracle, Gauss-Jordan here)
    assert abs(Kd - Ks).max() <= 1e-11 * abs(Ks).max()
    M = O.assemble_solid_mass(81, p.mesh.xyz, p.mesh.conn, rho, p.dof_index, K).to_scipy()
    c = abs(Ks).max() / abs(M).max()
    KMd = device_matrix(ctx, p, c)
    ref = Ks + c * M
    assert abs(KMd - ref).max() <= 1e-11 * abs(ref).max()
    # the bubbles make the element softer than brick8 in bending
    K8 = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof).to_scipy()
    x = np.random.default_rng(4).standard_normal(p.n_dof)
    assert x @ (Kd @ x) < x @ (K8 @ x)


@pytest.mark.parametrize("distort", [False, True])
def test_brickim_sensitivities_match_dual_number_oracle(ctx, distort):
    gpu_case(f"brickIM dJ/dE distort={distort}")
    p = im_problem(distort)
    problems.upload(ctx, p)
    rng = np.random.default_rng(8)
    u = rng.standard_normal(p.n_dof)
    lam = rng.standard_normal(p.n_dof)
    s = ctx.element_sens(u, lam)
    s_ref = O.solid_sens(81, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, u, lam)
    assert np.abs(s - s_ref).max() <= 1e-10 * np.abs(s_ref).max()


def test_brickim_compliance_pipeline_matches_direct_solve(ctx):
    gpu_case("brickIM compliance pipeline")
    p = im_problem(True)
    problems.upload(ctx, p)
    Ks = oracle_k(p).to_scipy()
    sf = 1e5 * abs(Ks).max()
    pen = np.zeros(p.n_dof)
    np.add.at(pen, p.pen_dofs, sf)
    u_ref = spla.splu((Ks + sp.diags(pen)).tocsc()).solve(p.f)
    o = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_JACOBI, tol=1e-10, inner_tol=1e-4, max_it=400000, max_outer=60)
    u, dJdE, J, st = ctx.static_compliance(p.f, o)
    assert st.converged, st.as_dict()
    assert np.abs(u - u_ref).max() <= 1e-7 * np.abs(u_ref).max()
    assert abs(J - p.f @ u_ref) <= 1e-8 * abs(J)
    s_ref = O.solid_sens(81, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, u_ref, u_ref)
    assert np.abs(dJdE - s_ref).max() <= 1e-6 * np.abs(s_ref).max()
