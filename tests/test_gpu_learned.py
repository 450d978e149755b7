"""Device checks of the learned factored preconditioner against the CPU
checker oracle/mpnn.py: node features, head outputs, the training loss and
its parameter gradient (BF16-MFMA forward/backward), the apply z = L L^T r;
training reduces the loss and is deterministic; solves with the learned
preconditioner reach the FP64 solution, and a broken network falls back to
Jacobi PCG."""
import numpy as np
import pytest
import scipy.sparse as sp

import mpnn
from asend3d_amd import api, problems
from conftest import gpu_case, oracle_learned_graph

pytestmark = pytest.mark.gpu

GROUPS = [("W0", mpnn.OFF_W0, mpnn.OFF_B0), ("b0", mpnn.OFF_B0, mpnn.OFF_L1)] + \
    [(f"layer{l + 1}", mpnn.OFF_L1 + l * mpnn.LAYER_SIZE, mpnn.OFF_L1 + (l + 1) * mpnn.LAYER_SIZE)
     for l in range(mpnn.LAYERS)] + [("head", mpnn.OFF_WH, mpnn.OFF_WA), ("Wa", mpnn.OFF_WA, mpnn.NUM_PARAMS)]


def small_problem(dims=(6, 4, 3), seed=5):
    return problems.heterogeneous(*dims, seed=seed, lx=5.0 * dims[0], ly=5.0 * dims[1], lz=5.0 * dims[2])


def test_network_parity_with_oracle(ctx):
    gpu_case("learned parity 6x4x3")
    p = small_problem()
    problems.upload(ctx, p)
    ctx.assemble()
    ctx.learned_init(1)
    prm = mpnn.random_params(11).astype(np.float32)
    ctx.learned_set_params(prm)
    n = p.n_dof // 3
    g, _ = oracle_learned_graph(p)
    rng = np.random.default_rng(0)
    e = rng.standard_normal(p.n_dof)
    e[np.repeat(g.X[:, 4] > 0, 3)] = 0.0
    loss_d, grad_d, feat_d, out_d = ctx.learned_eval(e, n)
    # features: FP32 on the device, FP64 here
    assert np.abs(feat_d - g.X).max() < 1e-5
    # head outputs and loss: BF16 activations on both sides; a rounding flip
    # in one activation moves the outputs by ~1e-4 relative
    fw = mpnn.forward(g, prm.astype(np.float64))
    assert np.linalg.norm(out_d - fw["O"]) <= 2e-3 * np.linalg.norm(fw["O"])
    loss_o, grad_o = mpnn.loss_and_grad(g, prm.astype(np.float64), e)
    assert abs(loss_d - loss_o) <= 1e-4 * abs(loss_o)
    for name, a, b in GROUPS:
        err = np.linalg.norm(grad_d[a:b] - grad_o[a:b]) / np.linalg.norm(grad_o[a:b])
        assert err < 1e-2, (name, err)
    # the apply kernels, given the device's own head outputs
    r = rng.standard_normal(p.n_dof)
    z_d = ctx.learned_apply(r)
    Wa = mpnn.bf16(prm[mpnn.OFF_WA:].astype(np.float64)).reshape(9, 9)
    Ld, Le = mpnn.factor_blocks(g, out_d.astype(np.float64), Wa)
    L = mpnn.factor_matrix(g, {"Ld": Ld, "Le": Le})
    z_o = L @ (L.T @ r)
    assert np.linalg.norm(z_d - z_o) <= 1e-5 * np.linalg.norm(z_o)


def test_untrained_network_is_block_jacobi(ctx):
    gpu_case("learned untrained = block Jacobi")
    p = small_problem((8, 3, 3), seed=2)
    problems.upload(ctx, p)
    ctx.assemble()
    ctx.learned_init(3)
    rp, cols, vals = ctx.get_matrix(with_penalty=True)
    K = sp.csr_matrix((vals, cols, rp), shape=(p.n_dof, p.n_dof))
    n = p.n_dof // 3
    B = np.zeros((n, 3, 3))
    for a in range(3):
        for b in range(3):
            B[:, a, b] = np.asarray(K[3 * np.arange(n) + a, 3 * np.arange(n) + b]).ravel()
    Bs = 0.5 * (B + B.transpose(0, 2, 1))
    r = np.random.default_rng(1).standard_normal(p.n_dof)
    z = ctx.learned_apply(r)
    z_ref = np.einsum("iab,ib->ia", np.linalg.inv(Bs), r.reshape(-1, 3)).ravel()
    assert np.linalg.norm(z - z_ref) <= 1e-5 * np.linalg.norm(z_ref)


def test_training_reduces_loss_and_is_deterministic(ctx):
    dims = (24, 6, 6)
    gpu_case(f"learned training {dims}")
    p = small_problem(dims, seed=8)
    problems.upload(ctx, p)
    ctx.set_grid(*dims)
    ctx.assemble()
    # mode 3: the three sample kinds (Gaussian, V-cycle error, V-cycle
    # solution) have different loss levels; the Gaussian and V-cycle-error
    # losses must drop (observed: Gaussian 0.21 -> 0.08, V-cycle error
    # 0.133 -> 0.113)
    ctx.learned_init(5)
    losses = ctx.learned_train(240, lr=2e-3, seed=3, mode=3)
    assert np.all(np.isfinite(losses))
    for kind, factor in ((0, 0.7), (1, 0.93)):
        lk = losses[kind::3]
        assert lk[-10:].mean() < factor * lk[:10].mean(), (kind, lk[:10].mean(), lk[-10:].mean())
    # mode 1: solver-generated samples only (V-cycle errors on even steps,
    # V-cycle solutions on odd ones); the V-cycle-error loss must drop too
    ctx.learned_init(5)
    losses = ctx.learned_train(240, lr=2e-3, seed=3, mode=1)
    assert np.all(np.isfinite(losses))
    le = losses[0::2]
    assert le[-10:].mean() < 0.97 * le[:10].mean(), (le[:10].mean(), le[-10:].mean())
    w1 = ctx.learned_params()
    ctx.learned_init(5)
    ctx.learned_train(240, lr=2e-3, seed=3, mode=1)
    w2 = ctx.learned_params()
    assert np.array_equal(w1, w2)


def test_learned_solve_matches_fp64_and_falls_back(ctx):
    dims = (30, 6, 6)
    gpu_case(f"learned solve + fallback {dims}")
    p = small_problem(dims, seed=4)
    problems.upload(ctx, p)
    ctx.set_grid(*dims)
    ctx.assemble()
    ref = ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11, max_it=400000, check_every=32)
    u_ref, s_ref, J_ref, _ = ctx.static_compliance(p.f, ref)
    ctx.learned_init(9)
    ctx.learned_train(60, lr=1e-3, seed=1)
    for pc in (api.PC_LEARNED, api.PC_LEARNED_PCG):
        o = ctx.opts(api.SOLVE_IR_MIXED, pc, tol=1e-8, inner_tol=1e-3, max_it=100000, max_outer=40, check_every=4)
        u, s, J, st = ctx.static_compliance(p.f, o)
        assert st.converged and st.fallback_used == 0, (pc, st.as_dict())
        assert np.abs(u - u_ref).max() <= 1e-6 * np.abs(u_ref).max()
        assert abs(J - J_ref) <= 1e-8 * abs(J_ref)
    # solver tuning with the learned smoother times the same grid without
    # the network too and keeps the faster pipeline; a solve passing 0 for
    # the refresh factor and check interval takes the tuned ones, and runs
    # without the network (fallback_used 2) when that was the faster one
    ctx.assemble()
    tn = ctx.autotune(p.f, api.TUNE_SOLVER | api.TUNE_LEARNED)
    assert tn.n_trials == 72 and tn.inner_tol in (1e-1, 1e-2, 1e-3, 1e-4) and tn.check_every in (1, 2, 4)
    assert tn.solve_ms_learned > 0 and tn.solve_ms_no_network > 0
    assert tn.solve_ms == min(tn.solve_ms_learned, tn.solve_ms_no_network)
    assert tn.learned_active == (1 if tn.solve_ms_learned <= tn.solve_ms_no_network else 0)
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_LEARNED, tol=1e-8, inner_tol=0.0, max_it=100000, max_outer=40,
                 check_every=0)
    u, s, J, st = ctx.static_compliance(p.f, o)
    assert st.converged and st.fallback_used == (0 if tn.learned_active else 2), (tn.as_dict(), st.as_dict())
    assert np.abs(u - u_ref).max() <= 1e-6 * np.abs(u_ref).max()
    # a changed network clears the decision: the learned smoother runs again
    ctx.learned_set_params(ctx.learned_params())
    u, s, J, st = ctx.static_compliance(p.f, o)
    assert st.converged and st.fallback_used == 0, st.as_dict()
    fresh = api.Context(0)
    try:
        problems.upload(fresh, p)
        fresh.set_grid(*dims)
        fresh.assemble()
        with pytest.raises(api.AsendError, match="learned"):
            fresh.autotune(p.f, api.TUNE_SOLVER | api.TUNE_LEARNED)
    finally:
        fresh.close()
    # a learned-smoother solve on a context without the multigrid hierarchy
    # is refused with a message
    nogrid = api.Context(0)
    try:
        problems.upload(nogrid, p)
        nogrid.assemble()
        nogrid.learned_init(9)
        with pytest.raises(api.AsendError, match="asend_set_grid"):
            nogrid.static_compliance(p.f, o)
    finally:
        nogrid.close()
    # a network with non-finite weights must not take the solve down
    ctx.learned_set_params(np.full(api.learned_num_params(), np.nan, dtype=np.float32))
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_LEARNED, tol=1e-8, inner_tol=1e-3, max_it=400000, max_outer=40,
                 check_every=4)
    u, s, J, st = ctx.static_compliance(p.f, o)
    assert st.converged and st.fallback_used == 1, st.as_dict()
    assert np.abs(u - u_ref).max() <= 1e-6 * np.abs(u_ref).max()


def test_vcycle_training_gradient_matches_finite_differences(ctx):
    # end-to-end training differentiates through every use of P in one
    # V-cycle (Chebyshev bounds held fixed); check along directions of the
    # head and edge map, in which the loss is smooth
    dims = (24, 6, 6)
    gpu_case(f"learned V-cycle gradient FD {dims}")
    p = small_problem(dims, seed=3)
    problems.upload(ctx, p)
    ctx.set_grid(*dims)
    ctx.assemble()
    ctx.learned_init(2)
    prm = mpnn.random_params(5, head_scale=0.05).astype(np.float32)
    ctx.learned_set_params(prm)
    rng = np.random.default_rng(0)
    e = rng.standard_normal(p.n_dof)
    _, g0 = ctx.learned_vcycle_eval(e)
    h = 3e-3
    for _ in range(3):
        d = np.zeros(prm.size, np.float32)
        d[mpnn.OFF_WH:] = rng.standard_normal(prm.size - mpnn.OFF_WH).astype(np.float32)
        ctx.learned_set_params(prm + h * d)
        lp, _ = ctx.learned_vcycle_eval(e, keep_bounds=True)
        ctx.learned_set_params(prm - h * d)
        lm, _ = ctx.learned_vcycle_eval(e, keep_bounds=True)
        fd = (lp - lm) / (2 * h)
        # FP32 cycle: agreement to ~1e-2 relative (observed <= 6e-3)
        assert abs(fd - g0 @ d) <= 2e-2 * abs(fd), (fd, float(g0 @ d))
        ctx.learned_set_params(prm)
        ctx.learned_vcycle_eval(e)


def test_trained_network_cuts_pcg_iterations(ctx):
    # what the network gains over the method without it: the untrained
    # network is exactly block Jacobi (head and edge map zero); after training
    # on four other 40x8x8 lognormal meshes the learned L L^T must need fewer
    # PCG iterations on a held-out mesh (observed 1648 -> 1368)
    dims = (40, 8, 8)
    gpu_case(f"learned gain {dims}")
    o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_LEARNED_PCG, tol=1e-8, inner_tol=1e-3, max_it=400000, max_outer=40,
                 check_every=4)

    def setup(seed, sig=1.0):
        q = problems.heterogeneous(*dims, seed=seed, sigma_ln=sig, lx=200.0, ly=40.0, lz=40.0)
        problems.upload(ctx, q)
        ctx.set_grid(*dims)
        ctx.assemble()
        return q

    ctx.learned_init(7)
    q = setup(4242)
    u0, _, _, st0 = ctx.static_compliance(q.f, o)
    for m, sig in enumerate((0.6, 0.8, 1.0, 1.15)):
        setup(100 + m, sig)
        ctx.learned_train(150, lr=1e-3, seed=m + 1, mode=1)
    q = setup(4242)
    u1, _, _, st1 = ctx.static_compliance(q.f, o)
    assert st0.converged and st1.converged and st1.fallback_used == 0
    assert st1.iterations <= 0.92 * st0.iterations, (st0.iterations, st1.iterations)
    assert np.abs(u1 - u0).max() <= 1e-6 * np.abs(u0).max()
