"""Heat conduction statics on the device (asend_set_conductivity, then
asend_assemble and asend_solve) against the oracle's restatement of the
reference's thermal branch (oracle/thermal.py, S/element_equations.rs:
1064-1123, S/model_analysis.rs:331-370): the conduction operator on a mesh
with moved nodes and per-element anisotropic conductivities, the direct
solve against the oracle's LDL^T and the one-dimensional closed forms, and
conj_grad_sparse with the block LDL^T preconditioner at the oracle's
iteration count on the one-DOF-per-node system.  Every device call goes
through the C ABI."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import thermal as TH
from asend3d_amd import api
from asend3d_amd.meshgen import hex_block
from conftest import gpu_case

pytestmark = pytest.mark.gpu


class Thermal:
    """Mesh, conductivities, temperature constraints and the DOF numbering of
    a brick8 conduction problem (three DOFs per node, temperature first)."""

    def __init__(self, nx, ny, nz, lx, ly, lz, seed=0, distort=0.0, aniso=False, fixed=(), block_dim=24):
        m = hex_block(nx, ny, nz, lx, ly, lz)
        rng = np.random.default_rng(seed)
        xyz = m.xyz.copy()
        if distort:
            inner = np.all((xyz > 1e-9) & (xyz < np.array([lx, ly, lz]) - 1e-9), axis=1)
            h = np.array([lx / nx, ly / ny, lz / nz])
            xyz[inner] += rng.uniform(-distort, distort, (int(inner.sum()), 3)) * h
        if aniso:
            kc = np.column_stack([rng.uniform(1.0, 3.0, (m.n_el, 3)), rng.uniform(-0.3, 0.3, (m.n_el, 3))])
        else:
            kc = np.tile((3.0, 1.0, 2.0, 0.0, 0.0, 0.0), (m.n_el, 1))
        self.m, self.xyz, self.kc = m, xyz, kc
        self.fixed = [np.asarray(s, dtype=np.int32) for s, _ in fixed]
        self.fixed_T = [float(t) for _, t in fixed]
        self.block_dim = block_dim
        _, di, n_dof = api.reorder_nodes(xyz, m.conn, 3, [[s] for s in self.fixed], block_dim)
        self.di = np.asarray(di).reshape(-1, 6)
        self.n_dof = n_dof
        self.rank = self.di[:, 0] // 3  # the reference's sorted_rank
        self.cnodes = np.concatenate(self.fixed) if self.fixed else np.zeros(0, dtype=np.int32)
        self.cvals = np.concatenate([np.full(len(s), t) for s, t in zip(self.fixed, self.fixed_T)]) \
            if self.fixed else np.zeros(0)
        self.K = TH.assemble_conduction(xyz, m.conn, kc)  # node numbering

    def upload(self, ctx):
        ones = np.ones(self.m.n_el)
        ctx.upload_mesh(8, self.xyz, self.di, self.n_dof, self.m.conn, ones, 0.3 * ones)
        ctx.set_conductivity(self.kc)
        pen = np.concatenate([self.di[self.cnodes, c] for c in range(3)]).astype(np.int64)
        ctx.set_penalty(pen, 0.0)
        ctx.assemble()

    def rhs(self, f_node, sf):
        b = TH.to_dofs(f_node, self.di, self.n_dof)
        b[self.di[self.cnodes, 0]] += sf * self.cvals
        return b

    def scalar(self):
        """The conduction matrix in the reference's numbering (sorted rank)."""
        perm = np.argsort(self.rank)
        Ks = self.K[perm][:, perm].tocsr()
        Ks.sort_indices()
        return O.CSR(Ks.indptr.astype(np.int64), Ks.indices.astype(np.int32), Ks.data.copy())

    def to_scalar(self, f_node):
        out = np.zeros(self.m.n_nodes)
        out[self.rank] = f_node
        return out


def test_conduction_operator_matches_oracle(ctx):
    p = Thermal(5, 4, 3, 1.0, 0.8, 0.6, seed=3, distort=0.2, aniso=True)
    gpu_case("thermal operator 5x4x3 distorted anisotropic")
    p.upload(ctx)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(p.n_dof)
    y = ctx.spmv(x)
    want = np.zeros(p.n_dof)
    for c in range(3):
        want[p.di[:, c]] = p.K @ x[p.di[:, c]]
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()
    # no penalty in this problem: the scale is that of the conduction matrix
    rowptr, cols, vals = ctx.get_matrix(with_penalty=False)
    assert abs(np.abs(vals).max() - np.abs(p.K.data).max()) <= 1e-12 * np.abs(p.K.data).max()


def test_direct_solve_bar_closed_form(ctx):
    nx, L = 12, 2.0
    m = hex_block(nx, 2, 2, L, 0.4, 0.4)
    T0, T1, Q = 10.0, 30.0, 50.0
    p = Thermal(nx, 2, 2, L, 0.4, 0.4, fixed=((m.nodes_at_x(0), T0), (m.nodes_at_x(nx), T1)))
    gpu_case("thermal direct solve bar 12x2x2")
    p.upload(ctx)
    sf = ctx.penalty_scale()
    assert abs(sf - 1e5 * np.abs(p.K.data).max()) <= 1e-12 * sf
    f_dev = api.body_load(8, api.BODY_FORCE, p.xyz, m.conn, p.di, p.n_dof, np.arange(m.n_el, dtype=np.int32),
                          load=(Q, 0.0, 0.0))
    f_node = TH.body_heat_gen(p.xyz, m.conn, np.arange(m.n_el), Q)
    assert np.abs(f_dev - TH.to_dofs(f_node, p.di, p.n_dof)).max() <= 1e-13 * np.abs(f_node).max()
    b = f_dev.copy()
    b[p.di[p.cnodes, 0]] += sf * p.cvals
    info = ctx.ldl_factor(0)
    assert info["pos_def"] == 1
    u, st = ctx.solve(b, ctx.opts(api.SOLVE_DIRECT, api.PC_JACOBI))
    assert st.converged
    T = u[p.di[:, 0]]
    assert np.abs(u[p.di[:, 1:]]).max() == 0.0  # the inert copies carry no load
    Ks = p.scalar()
    ref = O.LDL(Ks, p.rank[p.cnodes], sf, Ks.n)
    T_ref = ref.solve(p.to_scalar(f_node) + _pen_load(p, sf))[p.rank]
    assert np.abs(T - T_ref).max() <= 1e-10 * np.abs(T_ref).max()
    x = p.xyz[:, 0]
    exact = T0 + (T1 - T0) * x / L + Q * x * (L - x) / (2 * 3.0)
    assert np.abs(T - exact).max() <= 1e-4 * np.abs(exact).max()
    # maxBandwidth of the three-DOF layout is three times the reference's
    assert info["max_bandwidth"] == 3 * ref.max_bandwidth()


def _pen_load(p, sf):
    out = np.zeros(p.m.n_nodes)
    out[p.rank[p.cnodes]] = sf * p.cvals
    return out


@pytest.mark.parametrize("distort", [0.0, 0.2])
def test_conj_grad_sparse_same_iterations(ctx, distort):
    """PCG preconditioned by the block LDL^T over 3 b DOFs against the
    oracle's conj_grad_sparse over blocks of b nodes on the one-DOF system:
    the same iterations (the copies stay zero), the same temperatures."""
    nx, ny, nz = 10, 4, 3
    m = hex_block(nx, ny, nz, 2.0, 0.8, 0.6)
    p = Thermal(nx, ny, nz, 2.0, 0.8, 0.6, seed=7, distort=distort, aniso=True,
                fixed=((m.nodes_at_x(0), 20.0),), block_dim=24)
    gpu_case(f"thermal pcg 10x4x3 distort={distort}")
    p.upload(ctx)
    sf = ctx.penalty_scale()
    tip = m.elements_at_x(nx - 1)
    f_node = TH.surface_flux(p.xyz, m.conn, tip, 4.0, (1.0, 0.0, 0.0), 5.0)
    f_node[m.nodes_at_x(nx // 2)] += 0.25  # nodalHeatGen
    f_dev = api.surface_flux(8, p.xyz, m.conn, p.di, p.n_dof, tip, 4.0, (1.0, 0.0, 0.0), 5.0)
    assert np.abs(f_dev - TH.to_dofs(TH.surface_flux(p.xyz, m.conn, tip, 4.0, (1.0, 0.0, 0.0), 5.0), p.di,
                                     p.n_dof)).max() <= 1e-13 * np.abs(f_dev).max()
    b = p.rhs(f_node, sf)
    bd = p.block_dim
    ctx.ldl_factor(3 * bd)
    Ks = p.scalar()
    ref = O.LDL(Ks, p.rank[p.cnodes], sf, bd)
    bs = p.to_scalar(f_node) + _pen_load(p, sf)
    scale = float(bs @ ref.solve(bs))
    for tol_rel in (1e-6, 1e-12):
        tol = tol_rel * scale
        x_ref, it_ref, _ = O.conj_grad_sparse(Ks, p.rank[p.cnodes], sf, ref, bs, tol, 100000)
        u, st = ctx.solve(b, ctx.opts(api.SOLVE_PCG_FP64, api.PC_LDL, stop_rule=api.STOP_REF_ABS, tol=tol,
                                      max_it=100000, check_every=1))
        print(distort, tol_rel, it_ref, st.iterations)
        assert st.converged and abs(st.iterations - it_ref) <= 1, (st.iterations, it_ref)
        T = u[p.di[:, 0]]
        err = np.abs(T - x_ref[p.rank]).max() / np.abs(x_ref).max()
        assert err <= (1e-6 if tol_rel > 1e-9 else 1e-9)


def test_conduction_refusals_and_return_to_elasticity(ctx):
    m = hex_block(4, 2, 2, 1.0, 0.3, 0.3)
    p = Thermal(4, 2, 2, 1.0, 0.3, 0.3, fixed=((m.nodes_at_x(0), 0.0),))
    gpu_case("thermal refusals")
    p.upload(ctx)
    b = np.zeros(p.n_dof)
    b[p.di[m.nodes_at_x(4), 0]] = 1.0
    with pytest.raises(api.AsendError, match="Jacobi or LDL"):
        ctx.solve(b, ctx.opts(api.SOLVE_PCG_FP64, api.PC_MG))
    with pytest.raises(api.AsendError, match="elastic operator"):
        ctx.element_sens(b, b)
    u, st = ctx.solve(b, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-12, max_it=10000, check_every=1))
    assert st.converged and u[p.di[:, 0]].max() > 0
    # back to the elastic operator, with its own penalty scale
    ctx.set_conductivity(None)
    ctx.assemble()
    E = np.ones(m.n_el)
    K = O.assemble_solid(8, p.xyz, m.conn, E, 0.3 * E, p.di, p.n_dof)
    sf = 1e5 * O.max_abs(K)
    assert abs(ctx.penalty_scale() - sf) <= 1e-12 * sf
    x = np.random.default_rng(0).standard_normal(p.n_dof)
    pen = np.concatenate([p.di[p.cnodes, c] for c in range(3)])
    want = O.spmv_penalty(K, pen, sf, x)
    assert np.abs(ctx.spmv(x) - want).max() <= 1e-12 * np.abs(want).max()
