"""Device parity of the static solve path against the CPU oracle.
Every device call goes through the C ABI (libasend_gpu.so)."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def oracle_system(p):
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    sf = 1e5 * O.max_abs(K)
    return K, sf


def small_problems():
    yield "C1", problems.config1()
    yield "het-14x5x4", problems.heterogeneous(14, 5, 4, seed=5)
    yield "het-14x5x4 distorted", problems.distorted(problems.heterogeneous(14, 5, 4, seed=5), seed=6)


@pytest.mark.parametrize("which", [0, 1, 2])
def test_assembly_matches_oracle(ctx, which):
    name, p = list(small_problems())[which]
    gpu_case(f"assembly {name}")
    problems.upload(ctx, p)
    assert ctx.uniform_elements() == (which < 2)
    ctx.assemble()
    K, sf = oracle_system(p)
    assert abs(ctx.penalty_scale() - sf) <= 1e-12 * sf
    rp, cols, vals = ctx.get_matrix(with_penalty=True)
    # identical sparsity (DOF numbering and connectivity are bit-exact)
    assert np.array_equal(rp, K.rowptr)
    assert np.array_equal(cols, K.cols)
    ref = K.vals.copy()
    diag_pos = []
    for d in p.pen_dofs:
        k = K.rowptr[d] + np.searchsorted(K.cols[K.rowptr[d]:K.rowptr[d + 1]], d)
        ref[k] += sf
        diag_pos.append(k)
    mask = np.ones(ref.size, bool)
    mask[diag_pos] = False
    scale = np.abs(K.vals).max()
    assert np.abs(vals[mask] - ref[mask]).max() <= 1e-12 * scale
    assert np.abs(vals[~mask] - ref[~mask]).max() <= 1e-14 * sf


@pytest.mark.parametrize("which", [1, 2])
def test_mass_assembly_matches_oracle(ctx, which):
    # K + c M with per-element densities, scaled reference blocks (which = 1)
    # and per-element integration (which = 2)
    name, p = list(small_problems())[which]
    gpu_case(f"mass assembly {name}")
    rho = np.random.default_rng(3).uniform(1e-9, 8e-9, p.mesh.n_el)
    problems.upload(ctx, p, rho=rho)
    K, _ = oracle_system(p)
    M = O.assemble_solid_mass(8, p.mesh.xyz, p.mesh.conn, rho, p.dof_index, K)
    Ks, Ms = K.to_scipy(), M.to_scipy()
    c = abs(Ks).max() / abs(Ms).max()
    ctx.assemble(c)
    rp, cols, vals = ctx.get_matrix(with_penalty=False)
    import scipy.sparse as sp
    A = sp.csr_matrix((vals, cols, rp), shape=(p.n_dof, p.n_dof))
    ref = Ks + c * Ms
    # FP64 rounding of two comparable terms summed in a different order
    # (measured 5e-12 of max|K + c M| on the scaled-reference path)
    assert abs(A - ref).max() <= 1e-11 * abs(ref).max()


def test_spmv_matches_oracle(ctx):
    name, p = list(small_problems())[1]
    gpu_case(f"spmv {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    K, sf = oracle_system(p)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(p.n_dof)
    y_ref = O.spmv_penalty(K, p.pen_dofs, ctx.penalty_scale(), x)
    y64 = ctx.spmv(x, 64)
    assert np.abs(y64 - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    y32 = ctx.spmv(x, 32)
    assert np.abs(y32 - y_ref).max() <= 1e-5 * np.abs(y_ref).max()


@pytest.mark.parametrize("which", [0, 1])
def test_fp64_pcg_matches_ldl(ctx, which):
    name, p = list(small_problems())[which]
    gpu_case(f"fp64 pcg {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    K, sf = oracle_system(p)
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    # FP64 attains ~1e-12 relative residual on this penalty-conditioned system
    # (the oracle's LDL^T solve itself lands at ~4e-12)
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11, max_it=200000))
    print(st.as_dict())
    assert st.converged and st.rel_residual <= 1e-11
    assert np.abs(u - u_ref).max() <= 1e-10 * np.abs(u_ref).max()


def test_mixed_ir_jacobi_reaches_fp64_residual(ctx):
    name, p = list(small_problems())[1]
    gpu_case(f"mixed IR {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    K, sf = oracle_system(p)
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_IR_MIXED, api.PC_JACOBI, tol=1e-10, inner_tol=1e-4, max_it=400000))
    assert st.converged and st.rel_residual <= 1e-10
    assert st.outer_iterations >= 2  # FP32 inner alone cannot reach 1e-10
    assert np.abs(u - u_ref).max() <= 1e-7 * np.abs(u_ref).max()
    r = p.f - O.spmv_penalty(K, p.pen_dofs, sf, u)
    assert np.linalg.norm(r) <= 1.01e-10 * np.linalg.norm(p.f)


def test_reference_stop_rule(ctx):
    # g^T M^-1 g <= conv_tol (absolute), S/matrix_functions.rs:328
    name, p = list(small_problems())[0]
    gpu_case(f"abs stop rule {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, stop_rule=api.STOP_REF_ABS, tol=1e-12,
                                    max_it=100000, check_every=1))
    assert st.converged and st.ref_residual <= 1e-12
    u1, st1 = ctx.solve(p.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, stop_rule=api.STOP_REF_ABS, tol=1e-30,
                                      max_it=3, check_every=1))
    assert st1.iterations == 3 and not st1.converged


@pytest.mark.parametrize("which", [1, 2])
def test_sensitivity_matches_dual_number_oracle(ctx, which):
    # which = 2: distorted elements (full Jacobians, so an index slip in J^-1
    # shows) with a per-element Poisson ratio
    name, p = list(small_problems())[which]
    gpu_case(f"sens {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    K, sf = oracle_system(p)
    u = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    rng = np.random.default_rng(4)
    lam = u + 0.3 * np.abs(u).max() * rng.standard_normal(u.shape)
    s_ref = O.solid_sens(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, u, lam)
    s = ctx.element_sens(u, lam)
    assert np.abs(s - s_ref).max() <= 1e-10 * np.abs(s_ref).max()


@pytest.mark.parametrize("which", [1, 2])
@pytest.mark.parametrize("method", [api.SOLVE_PCG_FP64, api.SOLVE_IR_MIXED])
def test_compliance_pipeline(ctx, method, which):
    name, p = list(small_problems())[which]
    gpu_case(f"compliance pipeline {name} method {method}")
    problems.upload(ctx, p)
    K, sf = oracle_system(p)
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    J_ref = p.f @ u_ref
    s_ref = O.solid_sens(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, u_ref, u_ref)
    tol = 1e-11 if method == api.SOLVE_PCG_FP64 else 1e-9
    u, s, J, st = ctx.static_compliance(p.f, ctx.opts(method, api.PC_JACOBI, tol=tol, max_it=400000))
    assert st.converged
    etol = 1e-10 if method == api.SOLVE_PCG_FP64 else 1e-6
    assert abs(J - J_ref) <= etol * abs(J_ref)
    assert np.abs(u - u_ref).max() <= etol * np.abs(u_ref).max()
    assert np.abs(s - s_ref).max() <= etol * np.abs(s_ref).max()
    assert st.device_ms > 0


def test_assembly_deterministic(ctx):
    name, p = list(small_problems())[1]
    gpu_case(f"determinism {name}")
    problems.upload(ctx, p)
    ctx.assemble()
    a = ctx.get_matrix()[2]
    ctx.assemble()
    b = ctx.get_matrix()[2]
    assert np.array_equal(a, b)
    u1, _ = ctx.solve(p.f, ctx.opts(tol=1e-10, max_it=100000))
    u2, _ = ctx.solve(p.f, ctx.opts(tol=1e-10, max_it=100000))
    assert np.array_equal(u1, u2)


def test_si_units_and_pivot_clamp(ctx):
    """SI units (m, Pa): lognormal moduli around 70 GPa plus one element at
    6e11 Pa, just under the nu = 0.3 threshold (6.35e11 Pa) of the
    reference's pivot-clamped compliance inverse, assemble to the oracle's K.
    An element at 5e12 Pa, where the oracle's (the reference's) clamped
    inverse is no longer the isotropic C, is refused at upload and at a
    moduli update, and the context keeps the mesh and moduli it held."""
    gpu_case("SI units and the compliance pivot clamp")
    p = problems.cantilever(14, 5, 4, E=api.lognormal_moduli(11, 280, 70e9, 0.5))
    p.E[7] = 6.0e11
    assert api.compliance_clamp_check(8, p.E, p.nu) == (0, -1)
    problems.upload(ctx, p)
    ctx.assemble()
    K, sf = oracle_system(p)
    rp, cols, vals = ctx.get_matrix(with_penalty=True)
    assert np.array_equal(rp, K.rowptr) and np.array_equal(cols, K.cols)
    # K to 1e-12 of max|K|; the penalised diagonals on the scale of sf
    diag = np.array([K.rowptr[d] + np.searchsorted(K.cols[K.rowptr[d]:K.rowptr[d + 1]], d) for d in p.pen_dofs])
    ref = K.vals.copy()
    ref[diag] += sf
    mask = np.ones(ref.size, bool)
    mask[diag] = False
    assert np.abs(vals[mask] - ref[mask]).max() <= 1e-12 * np.abs(K.vals).max()
    assert np.abs(vals[diag] - ref[diag]).max() <= 1e-14 * sf

    bad = p.E.copy()
    bad[3] = 5e12
    assert api.compliance_clamp_check(8, bad, p.nu) == (1, 3)
    C = O.solid_cmat(5e12, 0.3)
    assert not np.allclose(C, 5e12 * O.solid_cmat(1.0, 0.3), rtol=1e-3)
    with pytest.raises(api.AsendError, match="pivot clamp"):
        problems.upload(ctx, dataclasses.replace(p, E=bad))
    with pytest.raises(api.AsendError, match="pivot clamp"):
        ctx.update_moduli(bad)
    ctx.assemble()
    assert np.array_equal(ctx.get_matrix(with_penalty=True)[2], vals)


def test_material_sensitivities_match_oracle(ctx):
    """asend_element_sens_material on a distorted mesh with per-element E and
    nu: the partials of -lambda^T K_e u with respect to the nine compliance
    parameters (E1..3, nu12, nu13, nu23, G12, G13, G23) against the oracle's
    dual-number dK_e/dp through its QR inverse, and their isotropic sum
    E (dE1 + dE2 + dE3) + G (dG12 + dG13 + dG23) against E dJ/dE of
    asend_element_sens."""
    gpu_case("material-parameter sensitivities, distorted 6x3x2")
    p = problems.distorted(problems.heterogeneous(6, 3, 2, seed=8), seed=9)
    problems.upload(ctx, p)
    rng = np.random.default_rng(10)
    u = rng.standard_normal(p.n_dof)
    lam = rng.standard_normal(p.n_dof)
    dev = ctx.element_sens_material(u, lam)
    di = np.asarray(p.dof_index).reshape(-1, 6)[:, :3]
    ref = np.zeros_like(dev)
    for e in range(p.mesh.n_el):
        nodes = p.mesh.conn[e]
        E, nu = p.E[e], p.nu[e]
        G = E / (2.0 * (1.0 + nu))
        ue = u[di[nodes]].ravel()
        le = lam[di[nodes]].ravel()
        for q in range(9):
            _, dK = O.solid_element_dmat(8, p.mesh.xyz[nodes], [E] * 3, [nu] * 3, [G] * 3, q)
            ref[e, q] = -le @ dK @ ue
    for q in range(9):
        assert np.abs(dev[:, q] - ref[:, q]).max() <= 1e-10 * np.abs(ref[:, q]).max(), q
    G = p.E / (2.0 * (1.0 + p.nu))
    iso = p.E * dev[:, 0:3].sum(axis=1) + G * dev[:, 6:9].sum(axis=1)
    dJdE = ctx.element_sens(u, lam)
    assert np.abs(iso - p.E * dJdE).max() <= 1e-11 * np.abs(iso).max()


def test_shape_sensitivities_match_finite_differences(ctx):
    """lam_e^T (dK_e/dX) u_e of every element, node and coordinate (device dual
    numbers) against central differences of the oracle's element stiffness
    (get_ruk_dfd0 restated) on a distorted mesh with per-element moduli."""
    p = problems.distorted(problems.heterogeneous(4, 3, 2, seed=5), seed=8)
    gpu_case("shape sensitivities 4x3x2 distorted")
    problems.upload(ctx, p)
    ctx.assemble()
    rng = np.random.default_rng(12)
    u = rng.standard_normal(p.n_dof)
    lam = rng.standard_normal(p.n_dof)
    s = ctx.element_sens_coords(u, lam)
    m = p.mesh
    di = np.asarray(p.dof_index).reshape(-1, 6)[:, :3]
    worst = 0.0
    for e in range(m.n_el):
        X = m.xyz[m.conn[e]]
        C = O.solid_cmat(p.E[e], p.nu[e])
        ue = u[di[m.conn[e]]].reshape(-1)
        le = lam[di[m.conn[e]]].reshape(-1)
        h = 1e-6 * np.abs(X).max()
        fd = np.zeros((8, 3))
        for a in range(8):
            for c in range(3):
                Xp = X.copy()
                Xm = X.copy()
                Xp[a, c] += h
                Xm[a, c] -= h
                Kp, _ = O.solid_element(8, Xp, C)
                Km, _ = O.solid_element(8, Xm, C)
                fd[a, c] = le @ (Kp - Km) @ ue / (2.0 * h)
        worst = max(worst, np.abs(s[e] - fd).max() / np.abs(fd).max())
    assert worst <= 1e-6, worst


def test_orthotropic_materials_match_oracle(ctx):
    """General material matrices (asend_set_material_stiffness): two
    orthotropic materials from the reference's compliance inverse (the
    oracle's solid_cmat_ortho) on a distorted mesh against the oracle's
    element integration, scattered; the solve against a sparse direct solve;
    the isotropic-only paths refuse."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    gpu_case("orthotropic materials het-8x4x3 distorted")
    p = problems.distorted(problems.heterogeneous(8, 4, 3, seed=11), seed=12)
    C = np.array([O.solid_cmat_ortho((70e3, 20e3, 10e3), (0.3, 0.25, 0.2), (5e3, 4e3, 3e3)),
                  O.solid_cmat_ortho((30e3, 60e3, 45e3), (0.1, 0.35, 0.22), (9e3, 7e3, 11e3))])
    el_mat = np.random.default_rng(5).integers(0, 2, p.mesh.n_el).astype(np.int32)
    ones = np.ones(p.mesh.n_el)
    ctx.upload_mesh(8, p.mesh.xyz, p.dof_index, p.n_dof, p.mesh.conn, ones, 0.3 * ones)
    ctx.set_material_stiffness(C, el_mat)
    ctx.set_penalty(p.pen_dofs, 0.0)
    ctx.assemble()
    di = np.asarray(p.dof_index).reshape(-1, 6)[:, :3]
    rows, cols, vals = [], [], []
    for e in range(p.mesh.n_el):
        c = p.mesh.conn[e]
        Ke, _ = O.solid_element(8, p.mesh.xyz[c], C[el_mat[e]])
        d = di[c].reshape(-1)
        rows.append(np.repeat(d, 24))
        cols.append(np.tile(d, 24))
        vals.append(Ke.reshape(-1))
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(p.n_dof, p.n_dof))
    rp, cl, vl = ctx.get_matrix(with_penalty=False)
    Kd = sp.csr_matrix((vl, cl, rp), shape=(p.n_dof, p.n_dof))
    # each entry sums 81 products C_pq G_mn of both signs: 6e-12 of max|K| measured
    assert abs(Kd - K).max() <= 2e-11 * abs(K).max()
    sf = ctx.penalty_scale()
    assert abs(sf - 1e5 * abs(K).max()) <= 1e-12 * sf
    P = np.zeros(p.n_dof)
    np.add.at(P, p.pen_dofs, sf)
    u_ref = spla.spsolve((K + sp.diags(P)).tocsc(), p.f)
    ctx.ldl_factor(0)
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_DIRECT, api.PC_JACOBI))
    assert np.abs(u - u_ref).max() <= 1e-10 * np.abs(u_ref).max()
    with pytest.raises(api.AsendError, match="isotropic"):
        ctx.element_sens(u, u)
    with pytest.raises(api.AsendError, match="Jacobi or LDL"):
        ctx.solve(p.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_MG))
