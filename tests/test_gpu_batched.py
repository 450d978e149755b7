"""Batched lockstep multi-RHS solves over design variants of different sizes
(the 64-variant workload: brick8 cantilevers with 10k-200k elements, aspect
5:1:1, lognormal moduli, tip tractions and tip-displacement adjoint
right-hand sides) against the oracle.

Small variants are compared with the oracle's direct LDL^T solve of the same
penalised system (the reference's banded factorisation,
S/lower_tri_mat_meth.rs, over one block spanning every DOF);
variants at the workload's largest size, where a direct solve would take
minutes, are held to the size-independent property the solver promises: the
relative residual of every system, measured with the oracle's own operator,
is at the requested tolerance."""
import os

import numpy as np
import pytest

import oracle as O
from asend3d_amd import api, problems
from conftest import gpu_case

pytestmark = pytest.mark.gpu


def _contexts(vs):
    ctxs = []
    for v in vs:
        c = api.Context(0)
        problems.upload(c, v.prob)
        c.assemble()
        ctxs.append(c)
    return ctxs


def _oracle_operator(p):
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    return K, 1e5 * O.max_abs(K)


def test_batch_matches_oracle_direct_solves():
    gpu_case("batched 4 variants x 6 rhs vs oracle LDL")
    vs = problems.batch_variants(4, seed=3, min_el=300, max_el=1500, n_load=3, n_adjoint=3)
    ctxs = _contexts(vs)
    b = api.Batch(ctxs)
    tol = 1e-9
    xs, st = b.solve([v.rhs for v in vs], tol=tol, inner_tol=1e-3, check_every=8)
    print(f"[batched small] {st.as_dict()}", flush=True)
    assert st.unconverged == 0 and st.max_rel_residual <= tol, st.as_dict()
    for v, x in zip(vs, xs):
        p = v.prob
        K, sf = _oracle_operator(p)
        # one block over all DOFs: the exact banded LDL^T factor
        ldl = O.LDL(K, p.pen_dofs, sf, p.n_dof)
        for j in range(v.rhs.shape[0]):
            u = ldl.solve(v.rhs[j])
            rel = np.linalg.norm(v.rhs[j] - O.spmv_penalty(K, p.pen_dofs, sf, x[j])) / np.linalg.norm(v.rhs[j])
            err = np.abs(x[j] - u).max() / np.abs(u).max()
            assert rel <= 1.05 * tol and err <= 1e-7, (p.mesh.n_el, j, rel, err)
    b.close()
    for c in ctxs:
        c.close()


def test_batch_full_size_variants_reach_tolerance():
    gpu_case("batched 3 full-size variants (150k-200k elements) x 4 rhs, oracle residuals")
    vs = problems.batch_variants(3, seed=11, min_el=150_000, max_el=200_000, n_load=2, n_adjoint=2)
    ctxs = _contexts(vs)
    b = api.Batch(ctxs)
    tol = 1e-8
    xs, st = b.solve([v.rhs for v in vs], tol=tol, inner_tol=1e-3, check_every=16)
    print(f"[batched full size] {st.as_dict()}", flush=True)
    assert st.unconverged == 0 and st.max_rel_residual <= tol, st.as_dict()
    # FP32 inner solves restarted from FP64 residuals: more than one
    # refinement step at this size
    assert st.outer_iterations >= 2
    O.set_threads(max(1, min(16, os.cpu_count() or 1)))
    try:
        for v, x in zip(vs, xs):
            p = v.prob
            K, sf = _oracle_operator(p)
            for j in range(v.rhs.shape[0]):
                rel = np.linalg.norm(v.rhs[j] - O.spmv_penalty(K, p.pen_dofs, sf, x[j])) / np.linalg.norm(v.rhs[j])
                print(f"[batched full size] {p.mesh.n_el} elements rhs {j}: oracle residual {rel:.3e}", flush=True)
                assert rel <= 1.05 * tol, (p.mesh.n_el, j, rel)
    finally:
        O.set_threads(1)
    b.close()
    for c in ctxs:
        c.close()


def test_batch_matches_single_solves():
    # the batch and the single-system FP64 path agree (both against the
    # same device operator)
    gpu_case("batched 3 variants x 4 rhs vs single solves")
    vs = problems.batch_variants(3, seed=5, min_el=300, max_el=3000, n_load=2, n_adjoint=2, sigma_ln=1.5)
    ctxs = _contexts(vs)
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
