"""Device checks of the multigrid preconditioner: exact Galerkin coarse
operators (A_{l+1} = P^T A_l P with P^T the restriction), a symmetric positive
V-cycle, and MG-preconditioned mixed-precision solves that reach the FP64
residual the oracle's LDL^T solution sets."""
import numpy as np
import pytest

import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def setup(ctx, p, dims):
    problems.upload(ctx, p)
    ctx.set_grid(*dims)
    ctx.assemble()


@pytest.mark.parametrize("dims", [(14, 5, 4), (24, 6, 6)])
def test_galerkin_identity(ctx, dims):
    gpu_case(f"galerkin {dims}")
    p = problems.heterogeneous(*dims, seed=9)
    setup(ctx, p, dims)
    nd = ctx.mg_info()
    assert len(nd) >= 2 and nd[0] == p.n_dof
    rng = np.random.default_rng(2)
    for l in range(len(nd) - 1):
        xc = rng.standard_normal(nd[l + 1]).astype(np.float32)
        y1 = ctx.mg_apply(l + 1, xc).astype(np.float64)
        xf = ctx.mg_prolong(l, xc)
        y2 = ctx.mg_restrict(l, ctx.mg_apply(l, xf)).astype(np.float64)
        err = np.linalg.norm(y1 - y2) / np.linalg.norm(y2)
        assert err < 2e-5, (l, err)


@pytest.mark.parametrize("dims", [(4, 4, 4), (13, 3, 3), (40, 4, 4)])
def test_coarsest_inverse_is_exact(ctx, dims):
    # coarsest sizes 81, 216 and 567 DOF: the whole-matrix LDS, the blocked
    # panel and the streaming Gauss-Jordan kernels; the operator is read back
    # column by column through the level SpMV (FP32 values, exact products)
    gpu_case(f"coarsest inverse {dims}")
    p = problems.heterogeneous(*dims, seed=4)
    setup(ctx, p, dims)
    nd = ctx.mg_info()
    lc = len(nd) - 1
    n = nd[lc]
    A = np.zeros((n, n))
    e = np.zeros(n, dtype=np.float32)
    for j in range(n):
        e[j] = 1.0
        A[:, j] = ctx.mg_apply(lc, e)
        e[j] = 0.0
    b = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    x = ctx.mg_coarse_solve(b).astype(np.float64)
    x_ref = np.linalg.solve(A, b.astype(np.float64))
    assert np.abs(x - x_ref).max() <= 1e-5 * np.abs(x_ref).max()


def test_prolongation_interpolates_linear_fields(ctx):
    # P reproduces fields linear in the grid index exactly
    dims = (13, 5, 3)
    gpu_case(f"prolongation {dims}")
    p = problems.heterogeneous(*dims, seed=3)
    setup(ctx, p, dims)
    nd = ctx.mg_info()
    mx, my, mz = dims[0] + 1, dims[1] + 1, dims[2] + 1
    cx, cy, cz = (dims[0] + 1) // 2 + 1, (dims[1] + 1) // 2 + 1, (dims[2] + 1) // 2 + 1
    assert nd[1] == 3 * cx * cy * cz
    X = np.minimum(2 * np.arange(cx), dims[0])
    Y = np.minimum(2 * np.arange(cy), dims[1])
    Z = np.minimum(2 * np.arange(cz), dims[2])
    K, J, I = np.meshgrid(Z, Y, X, indexing="ij")
    xc = np.zeros((cx * cy * cz, 3), np.float32)
    xc[:, 0] = I.reshape(-1)
    xc[:, 1] = 2.0 * J.reshape(-1) + 1.0
    xc[:, 2] = -K.reshape(-1)
    xf = ctx.mg_prolong(0, xc.reshape(-1)).reshape(-1, 3)
    # fine row of grid node (i,j,k)
    di = p.dof_index
    k, j, i = np.meshgrid(np.arange(mz), np.arange(my), np.arange(mx), indexing="ij")
    g = (i + mx * (j + my * k)).reshape(-1)
    rows = di[g, 0] // 3
    np.testing.assert_allclose(xf[rows, 0], i.reshape(-1), atol=1e-5)
    np.testing.assert_allclose(xf[rows, 1], 2.0 * j.reshape(-1) + 1.0, atol=1e-5)
    np.testing.assert_allclose(xf[rows, 2], -k.reshape(-1), atol=1e-5)


def test_vcycle_symmetric_positive(ctx):
    dims = (24, 6, 6)
    gpu_case(f"vcycle {dims}")
    p = problems.heterogeneous(*dims, seed=4)
    setup(ctx, p, dims)
    rng = np.random.default_rng(5)
    b1 = rng.standard_normal(p.n_dof).astype(np.float32)
    b2 = rng.standard_normal(p.n_dof).astype(np.float32)
    x1 = ctx.mg_vcycle(b1).astype(np.float64)
    x2 = ctx.mg_vcycle(b2).astype(np.float64)
    s12, s21 = b2 @ x1, b1 @ x2
    assert abs(s12 - s21) <= 1e-3 * np.sqrt(abs((b1 @ x1) * (b2 @ x2)))
    assert b1 @ x1 > 0 and b2 @ x2 > 0


@pytest.mark.parametrize("dims", [(14, 5, 4), (40, 8, 8)])
def test_mg_mixed_solve_matches_ldl(ctx, dims):
    gpu_case(f"mg mixed solve {dims}")
    p = problems.heterogeneous(*dims, seed=6)
    setup(ctx, p, dims)
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    sf = 1e5 * O.max_abs(K)
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    u, st = ctx.solve(p.f, ctx.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=1e-10, inner_tol=1e-3, max_it=5000,
                                    check_every=2))
    _, stj = ctx.solve(p.f, ctx.opts(api.SOLVE_IR_MIXED, api.PC_JACOBI, tol=1e-10, inner_tol=1e-4,
                                     max_it=400000))
    print(st.as_dict(), stj.as_dict())
    assert st.converged and st.rel_residual <= 1e-10
    assert st.iterations * 5 < stj.iterations
    assert np.abs(u - u_ref).max() <= 1e-7 * np.abs(u_ref).max()
    r = p.f - O.spmv_penalty(K, p.pen_dofs, sf, u)
    assert np.linalg.norm(r) <= 1.05e-10 * np.linalg.norm(p.f)


def test_mg_compliance_pipeline_matches_oracle(ctx):
    dims = (20, 4, 4)
    gpu_case(f"mg compliance {dims}")
    p = problems.heterogeneous(*dims, seed=8)
    problems.upload(ctx, p)
    ctx.set_grid(*dims)
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    sf = 1e5 * O.max_abs(K)
    u_ref = O.LDL(K, p.pen_dofs, sf, p.n_dof).solve(p.f)
    s_ref = O.solid_sens(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, u_ref, u_ref)
    u, s, J, st = ctx.static_compliance(p.f, ctx.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=1e-10, inner_tol=1e-3,
                                                      max_it=5000, check_every=2))
    assert st.converged
    assert abs(J - p.f @ u_ref) <= 1e-8 * abs(J)
    assert np.abs(s - s_ref).max() <= 1e-7 * np.abs(s_ref).max()


def test_graph_replay_matches_eager_launches(ctx):
    # the FCG iteration blocks replayed from captured hipGraphs give bitwise the
    # results of launching the same kernels one by one, over several replays
    # with eager work in between (two pipelines, smoother changes)
    dims = (20, 6, 6)
    gpu_case(f"graph replay {dims}")
    p = problems.heterogeneous(*dims, seed=12)
    setup(ctx, p, dims)
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=1e-8, inner_tol=1e-3, max_it=100000, max_outer=40, check_every=3)
    res = {}
    for on in (False, True, True):
        ctx.set_graphs(on)
        out = []
        for kind in (1, 0, 1):
            ctx.mg_set_smoother(kind, 2, 0.1, 1.15)
            u, s, J, st = ctx.static_compliance(p.f, o)
            assert st.converged
            out.append((u.copy(), st.iterations))
        res.setdefault(on, []).append(out)
    ctx.set_graphs(True)
    ref = res[False][0]
    for run in res[True]:
        for (u1, it1), (u0, it0) in zip(run, ref):
            assert it1 == it0 and np.array_equal(u1, u0)


def test_learned_graph_replay_matches_eager_launches(ctx):
    # the learned pipeline captures the FCG blocks and the smoother bound's
    # power iteration; replays must equal eager launches bit for bit across
    # design updates (new moduli, new factor, same buffers) and across a
    # larger mesh in between (new hierarchy, reallocated factor storage)
    dims, big = (20, 6, 6), (30, 8, 8)
    gpu_case(f"learned graph replay {dims} / {big}")
    p = problems.heterogeneous(*dims, seed=12)
    q = problems.heterogeneous(*big, seed=13)
    E2 = p.E * np.exp(0.3 * np.random.default_rng(3).standard_normal(p.E.size))
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_LEARNED, tol=1e-8, inner_tol=1e-2, max_it=100000, max_outer=40,
                 check_every=2)
    ctx.learned_init(3)
    setup(ctx, p, dims)
    ctx.learned_train(40, lr=1e-3, seed=1, mode=1)
    res = {}
    try:
        for on in (False, True, True):
            ctx.set_graphs(on)
            out = []
            for mesh in ("p", "q", "p"):
                m, d = (p, dims) if mesh == "p" else (q, big)
                setup(ctx, m, d)
                for E in ((m.E, E2) if mesh == "p" else (m.E,)):
                    ctx.update_moduli(E)
                    u, s, J, st = ctx.static_compliance(m.f, o)
                    assert st.converged and st.fallback_used == 0, st.as_dict()
                    out.append((u.copy(), st.iterations, J))
            res.setdefault(on, []).append(out)
    finally:
        ctx.set_graphs(True)
    ref = res[False][0]
    for run in res[True]:
        for (u1, it1, J1), (u0, it0, J0) in zip(run, ref):
            assert it1 == it0 and J1 == J0 and np.array_equal(u1, u0)


def test_autotuner_choices_are_used(ctx):
    dims = (24, 6, 6)
    gpu_case(f"autotune {dims}")
    p = problems.heterogeneous(*dims, seed=21)
    setup(ctx, p, dims)
    t = ctx.autotune(p.f)
    assert t.spmv_waves_per_block in (1, 2, 4, 8)
    assert all(t.spmv_ms_fp32[k] > 0 for k in range(4))
    assert t.cheb_degree in (1, 2, 3) and t.inner_tol in (1e-1, 1e-2, 1e-3, 1e-4) and t.check_every in (1, 2, 4)
    assert t.n_trials == 4 + 36 and t.solve_ms > 0
    # a solve that passes 0 for the inner tolerance and check interval uses the tuned ones
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=1e-9, inner_tol=0.0, max_it=100000, max_outer=60,
                 check_every=0)
    u, s, J, st = ctx.static_compliance(p.f, o)
    assert st.converged
    ref = ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11, max_it=400000, check_every=32)
    u_ref = ctx.static_compliance(p.f, ref)[0]
    assert np.abs(u - u_ref).max() <= 1e-7 * np.abs(u_ref).max()
    ctx.mg_set_smoother(1, 2, 0.1, 1.15)
