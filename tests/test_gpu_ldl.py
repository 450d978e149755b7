"""Device LDL^T factor (the reference's LowerTriMat) against the oracle's
restatement of S/lower_tri_mat_meth.rs: block preconditioner solves, the exact
one-block factor behind the direct solver, and conj_grad_sparse preconditioned
by it (same iteration count under the reference's absolute stop rule).  Every
device call goes through the C ABI."""
import numpy as np
import pytest

import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def solid(name):
    if name == "C1":
        return problems.config1()
    if name == "het-12x6x6":
        return problems.heterogeneous(12, 6, 6, seed=11)
    if name == "het-12x6x6 distorted":
        return problems.distorted(problems.heterogeneous(12, 6, 6, seed=11), seed=12)
    raise KeyError(name)


def solid_system(ctx, p):
    problems.upload(ctx, p)
    ctx.assemble()
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    sf = 1e5 * O.max_abs(K)
    assert abs(ctx.penalty_scale() - sf) <= 1e-12 * sf
    return K, sf


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


# block sizes: one block (exact), the reference's 6 * solverBlockDim, a block
# shorter than a tile and one that is not a multiple of the tile edge
@pytest.mark.parametrize("name", ["C1", "het-12x6x6", "het-12x6x6 distorted"])
@pytest.mark.parametrize("blocks", ["one", "reference", "30", "200"])
def test_block_factor_solve_matches_oracle(ctx, name, blocks):
    p = solid(name)
    gpu_case(f"ldl factor {name} blocks={blocks}")
    K, sf = solid_system(ctx, p)
    bd = {"one": p.n_dof, "reference": 6 * p.block_dim, "30": 30, "200": 200}[blocks]
    info = ctx.ldl_factor(bd)
    ref = O.LDL(K, p.pen_dofs, sf, bd)
    assert info["max_bandwidth"] == ref.max_bandwidth()
    assert info["pos_def"] == int(ref.pos_def()) == 1
    assert info["n_blocks"] == (p.n_dof + min(bd, p.n_dof) - 1) // min(bd, p.n_dof)
    rng = np.random.default_rng(3)
    for b in (p.f, rng.standard_normal(p.n_dof)):
        x = ctx.ldl_solve(b)
        x_ref = ref.solve(b)
        assert rel_err(x, x_ref) <= 1e-10, (info, rel_err(x, x_ref))


def test_direct_solve_is_exact(ctx):
    p = solid("het-12x6x6 distorted")
    gpu_case("direct solve het-12x6x6 distorted")
    K, sf = solid_system(ctx, p)
    info = ctx.ldl_factor(0)
    assert info["n_blocks"] == 1 and info["block_dofs"] == p.n_dof
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_DIRECT, api.PC_JACOBI))
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    assert st.converged and st.iterations == 0
    # the penalty (1e5 max|K|) bounds the attainable residual near 1e-11
    assert st.rel_residual <= 1e-10, st.as_dict()
    assert rel_err(u, u_ref) <= 1e-10
    r = p.f - O.spmv_penalty(K, p.pen_dofs, sf, u)
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(p.f)


def test_direct_needs_one_block(ctx):
    p = solid("C1")
    gpu_case("direct solve refuses a block factor")
    solid_system(ctx, p)
    ctx.ldl_factor(6 * p.block_dim if 6 * p.block_dim < p.n_dof else 30)
    with pytest.raises(api.AsendError, match="one block"):
        ctx.solve(p.f, ctx.opts(api.SOLVE_DIRECT, api.PC_JACOBI))


@pytest.mark.parametrize("name", ["C1", "het-12x6x6", "het-12x6x6 distorted"])
def test_conj_grad_sparse_same_iterations(ctx, name):
    """PCG with the reference's block preconditioner and absolute stop rule:
    the iteration count of the oracle's conj_grad_sparse, the same answer."""
    p = solid(name)
    gpu_case(f"ldl pcg {name}")
    K, sf = solid_system(ctx, p)
    bd = min(6 * p.block_dim, 200)  # several blocks even on the small meshes
    ctx.ldl_factor(bd)
    ref = O.LDL(K, p.pen_dofs, sf, bd)
    scale = float(p.f @ ref.solve(p.f))  # g^T M^-1 g at the start
    for tol_rel in (1e-6, 1e-12):
        tol = tol_rel * scale
        x_ref, it_ref, res_ref = O.conj_grad_sparse(K, p.pen_dofs, sf, ref, p.f, tol, 100000)
        u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_LDL, stop_rule=api.STOP_REF_ABS, tol=tol,
                                        max_it=100000, check_every=1))
        print(name, tol_rel, it_ref, st.iterations, res_ref, st.ref_residual)
        assert st.converged and st.ref_residual <= tol
        assert abs(st.iterations - it_ref) <= 1, (st.iterations, it_ref)
        assert rel_err(u, x_ref) <= 1e-6 if tol_rel > 1e-9 else rel_err(u, x_ref) <= 1e-9


def test_factor_follows_reassembly(ctx):
    p = solid("het-12x6x6")
    gpu_case("ldl refactor after new moduli")
    _, sf = solid_system(ctx, p)
    ctx.ldl_factor(0)
    E2 = p.E * np.linspace(0.5, 2.0, p.mesh.n_el)
    ctx.update_moduli(E2)
    ctx.assemble()
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_DIRECT, api.PC_JACOBI))
    K2 = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, E2, p.nu, p.dof_index, p.n_dof)
    # the penalty scale stays the one of the first assembly (scale_const runs once)
    u_ref = O.LDL(K2, p.pen_dofs, sf, p.n_dof).solve(p.f)
    assert rel_err(u, u_ref) <= 1e-10


def test_shell_block_factor_matches_oracle(ctx):
    p = problems.shell_plate(10, 8, 1.0, 0.8, 0.01, 70e9, 0.33, 2700.0, clamp="x0")
    gpu_case("ldl factor shell4 10x8")
    problems.upload_shell(ctx, p)
    ctx.assemble()
    K, _ = O.assemble_shell(p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.t, p.rho, p.dof_index, p.n_dof)
    sf = ctx.penalty_scale()
    for bd in (p.n_dof, 6 * p.block_dim, 100):
        info = ctx.ldl_factor(bd)
        ref = O.LDL(K, p.pen_dofs, sf, bd)
        assert info["max_bandwidth"] == ref.max_bandwidth()
        b = np.random.default_rng(bd).standard_normal(p.n_dof)
        assert rel_err(ctx.ldl_solve(b), ref.solve(b)) <= 1e-9
