"""Batched lockstep multi-RHS solves over variants of different sizes against
the single-system FP64 solve of each variant and right-hand side."""
import numpy as np
import pytest

from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def test_batch_matches_single_solves():
    gpu_case("batched 4 variants x 6 rhs")
    vs = problems.batch_variants(4, seed=3, min_el=300, max_el=3000, n_load=3, n_adjoint=3)
    ctxs = []
    for v in vs:
        c = api.Context(0)
        problems.upload(c, v.prob)
        c.assemble()
        ctxs.append(c)
    b = api.Batch(ctxs)
    xs, st = b.solve([v.rhs for v in vs], tol=1e-9, inner_tol=1e-3, check_every=8)
    assert st.unconverged == 0 and st.max_rel_residual <= 1e-9, st.as_dict()
    o = api.Context.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11, max_it=1000000, check_every=32)
    for c, v, x in zip(ctxs, vs, xs):
        for j in range(v.rhs.shape[0]):
            u = c.solve(v.rhs[j], o)[0]
            assert np.abs(x[j] - u).max() <= 1e-7 * np.abs(u).max(), j
    b.close()
    for c in ctxs:
        c.close()
