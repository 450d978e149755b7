"""Allocation-free FP64 solve (asend_solve_noalloc) against asend_solve.

Both run the same kernels in the same order, so iterates, statistics and
solutions must agree bit for bit.  Two calls in a row, the second on a
smaller model (it reuses the scratch array the first left behind), and a
model beyond the scratch limit is refused with a clear error before any
device work."""
import numpy as np
import pytest

from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def _compare(ctx, p, tol):
    problems.upload(ctx, p)
    ctx.assemble()
    o = ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=tol, max_it=200000)
    u_ref, st_ref = ctx.solve(p.f, o)
    u, st = ctx.solve_noalloc(p.f, o)
    assert st_ref.converged and st.converged
    assert st.iterations == st_ref.iterations
    assert st.outer_iterations == st_ref.outer_iterations
    assert st.rel_residual == st_ref.rel_residual
    assert st.ref_residual == st_ref.ref_residual
    assert np.array_equal(u, u_ref)
    # and it is the solution: FP64 residual through the library's own SpMV
    r = p.f - ctx.spmv(u, 64)
    assert np.linalg.norm(r) <= 1.01 * tol * np.linalg.norm(p.f)


def test_noalloc_solve_equals_solve_on_two_calls(ctx):
    big = problems.heterogeneous(30, 8, 8, seed=11)    # 7,533 DOF
    small = problems.heterogeneous(10, 3, 3, seed=12)  # 528 DOF
    gpu_case(f"noalloc call 1: 30x8x8, {big.n_dof} DOF")
    _compare(ctx, big, 1e-10)
    gpu_case(f"noalloc call 2: 10x3x3, {small.n_dof} DOF")
    _compare(ctx, small, 1e-10)


def test_noalloc_refuses_models_beyond_its_scratch(ctx):
    p = problems.cantilever(100, 30, 28)  # 272,397 DOF > 262,144
    gpu_case(f"noalloc size limit: {p.n_dof} DOF")
    problems.upload(ctx, p)
    ctx.assemble()
    o = ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-8)
    with pytest.raises(api.AsendError, match="scratch limit"):
        ctx.solve_noalloc(p.f, o)
    o_mixed = ctx.opts(api.SOLVE_IR_MIXED, api.PC_JACOBI, tol=1e-8)
    with pytest.raises(api.AsendError, match="FP64 Jacobi PCG only"):
        ctx.solve_noalloc(p.f, o_mixed)
