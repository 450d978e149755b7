"""Static cantilever workloads of the solve path, built with the library's
host-side preparation (node reordering, traction loads, seeded moduli).

Clamp: one single-DOF penalty constraint per component on the x = 0 node set
(three constraints, as the reference input writes a clamp,
S/const_meth.rs:30-90).  Load: uniform surface traction on the x = lx face
elements with normDir (1,0,0) and normTol 5 degrees
(S/element_equations.rs:1833-2056).
"""
import dataclasses
from dataclasses import dataclass

import numpy as np

from . import api
from .meshgen import HexBlock, QuadGrid, hex_block, quad_grid

C2_SEED = 20261202


@dataclass
class StaticProblem:
    mesh: HexBlock
    etype: int
    E: np.ndarray
    nu: np.ndarray
    block_dim: int
    ordered: np.ndarray
    dof_index: np.ndarray
    n_dof: int
    clamp_nodes: np.ndarray
    pen_dofs: np.ndarray
    tip_elements: np.ndarray
    traction: tuple
    f: np.ndarray

    @property
    def constraints(self):
        return [[self.clamp_nodes], [self.clamp_nodes], [self.clamp_nodes]]


def cantilever(nx, ny, nz, lx=1.0, ly=0.2, lz=0.2, E=200e9, nu=0.3, traction=(0.0, 0.0, -1.0e6), etype=8,
               block_dim=None):
    m = hex_block(nx, ny, nz, lx, ly, lz)
    if np.isscalar(E):
        E = np.full(m.n_el, float(E))
    E = np.ascontiguousarray(E, dtype=np.float64)
    nu_a = np.full(m.n_el, float(nu)) if np.isscalar(nu) else np.ascontiguousarray(nu, dtype=np.float64)
    clamp = m.nodes_at_x(0)
    bd = api.block_dim_auto(m.n_nodes) if block_dim is None else block_dim
    constraints = [[clamp], [clamp], [clamp]]
    ordered, di, n_dof = api.reorder_nodes(m.xyz, m.conn, 3, constraints, bd)
    pen = np.concatenate([di[clamp, 0], di[clamp, 1], di[clamp, 2]]).astype(np.int64)
    tip = m.elements_at_x(nx - 1)
    f = api.surface_traction(etype, m.xyz, m.conn, di, n_dof, tip, traction, (1.0, 0.0, 0.0), 5.0)
    return StaticProblem(m, etype, E, nu_a, bd, ordered, di, n_dof, clamp, pen, tip, tuple(traction), f)


def config1(etype=8):
    """20x4x4 brick8 cantilever, 1.0 x 0.2 x 0.2 m, E = 200 GPa, nu = 0.3."""
    return cantilever(20, 4, 4, E=200e9, nu=0.3, etype=etype)


def heterogeneous(nx, ny, nz, seed=C2_SEED, median=70e3, sigma_ln=1.0, nu=0.3, lx=1000.0, ly=200.0, lz=200.0,
                  traction=(0.0, 0.0, -1.0), block_dim=None):
    """Lognormal per-element moduli, posed in mm-N-MPa (median 70e3 MPa =
    70 GPa, traction 1 MPa).  In these units 1/E stays far above the 1e-12
    pivot clamp of the reference's QR inverse (S/matrix_functions.rs:189-191),
    which in SI units would corrupt the material matrix of the stiff tail
    (E > 6.35e11 Pa at nu = 0.3) of a sigma_ln = 1 distribution; the upload
    refuses such moduli (api.compliance_clamp_check)."""
    n_el = nx * ny * nz
    E = api.lognormal_moduli(seed, n_el, median, sigma_ln)
    return cantilever(nx, ny, nz, lx, ly, lz, E=E, nu=nu, traction=traction, block_dim=block_dim)


def distorted(p, seed, amp=0.2, nu_range=(0.2, 0.4)):
    """The same problem on a mesh whose interior nodes move by up to amp
    times the smallest element edge in each direction (seeded), with a
    per-element Poisson ratio drawn from nu_range: the general, per-element
    integration path."""
    rng = np.random.default_rng(seed)
    m = p.mesh
    xyz = m.xyz.copy()
    h = min(m.lx / m.nx, m.ly / m.ny, m.lz / m.nz)
    inner = np.all((xyz > xyz.min(axis=0) + 1e-9) & (xyz < xyz.max(axis=0) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-amp, amp, (int(inner.sum()), 3)) * h
    return dataclasses.replace(p, mesh=dataclasses.replace(m, xyz=xyz), nu=rng.uniform(*nu_range, m.n_el))


def config2():
    """200x40x40 brick8 block (337,881 nodes, 1,013,643 DOF), lognormal E."""
    return heterogeneous(200, 40, 40)


def upload(ctx, prob, rho=None):
    ctx.upload_mesh(prob.etype, prob.mesh.xyz, prob.dof_index, prob.n_dof, prob.mesh.conn, prob.E, prob.nu, rho)
    ctx.set_penalty(prob.pen_dofs, 0.0)


def upload_shell(ctx, prob):
    """Shell mesh, thickness and clamps onto a device context."""
    ctx.upload_mesh(prob.etype, prob.mesh.xyz, prob.dof_index, prob.n_dof, prob.mesh.conn, prob.E, prob.nu, prob.rho)
    ctx.set_thickness(prob.t)
    ctx.set_penalty(prob.pen_dofs, 0.0)


@dataclass
class BatchVariant:
    """One design variant of a batch: the cantilever problem plus its load
    cases and adjoint right-hand sides (rows of `rhs`, n_rhs x n_dof)."""
    prob: StaticProblem
    rhs: np.ndarray
    n_load: int


def batch_variants(n_variants, seed=7, min_el=10_000, max_el=200_000, n_load=8, n_adjoint=8, sigma_ln=1.0,
                   progress=None):
    """Cantilever variants with aspect 5:1:1 and element counts drawn uniformly
    from [min_el, max_el] (nx = 5k, ny = nz = k), 5 mm cubes, lognormal moduli
    (seeded per variant).  Right-hand sides: n_load random unit tip-face
    tractions (1 MPa in a random direction) and n_adjoint adjoint right-hand
    sides of tip-displacement objectives (unit loads on random tip DOFs; for a
    compliance objective the adjoint right-hand side is the load case itself)."""
    rng = np.random.default_rng(seed)
    out = []
    for v in range(n_variants):
        target = rng.uniform(min_el, max_el)
        k = max(2, int(round((target / 5.0) ** (1.0 / 3.0))))
        nx, ny, nz = 5 * k, k, k
        E = api.lognormal_moduli(seed * 1000 + v, nx * ny * nz, 70e3, sigma_ln)
        p = cantilever(nx, ny, nz, 5.0 * nx, 5.0 * ny, 5.0 * nz, E=E, nu=0.3, traction=(0.0, 0.0, -1.0))
        rows = []
        for _ in range(n_load):
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            rows.append(api.surface_traction(p.etype, p.mesh.xyz, p.mesh.conn, p.dof_index, p.n_dof,
                                             p.tip_elements, tuple(d), (1.0, 0.0, 0.0), 5.0))
        tip_nodes = p.mesh.nodes_at_x(nx)
        di = np.asarray(p.dof_index).reshape(-1, 6)
        for _ in range(n_adjoint):
            nd = tip_nodes[rng.integers(len(tip_nodes))]
            r = np.zeros(p.n_dof)
            r[di[nd, rng.integers(3)]] = 1.0
            rows.append(r)
        out.append(BatchVariant(p, np.array(rows), n_load))
        if progress is not None:
            progress(v, p)
    return out


@dataclass
class ShellProblem:
    mesh: QuadGrid
    E: np.ndarray
    nu: np.ndarray
    t: np.ndarray
    rho: np.ndarray
    block_dim: int
    dof_index: np.ndarray
    n_dof: int
    clamp_nodes: np.ndarray
    pen_dofs: np.ndarray
    etype: int = 41


def shell_plate(nx, ny, lx, ly, t, E, nu, rho, clamp="edges"):
    """Flat shell4 plate in the xy plane; clamp = "edges" (all four edges, all
    six DOFs) or "x0" (the x = 0 edge)."""
    m = quad_grid(nx, ny, lx, ly)
    n_el = m.n_el
    clamp_nodes = m.boundary_nodes() if clamp == "edges" else m.nodes_at_x(0)
    bd = api.block_dim_auto(m.n_nodes)
    constraints = [[clamp_nodes]] * 6
    _, di, n_dof = api.reorder_nodes(m.xyz, m.conn, 6, constraints, bd)
    pen = np.concatenate([di[clamp_nodes, d] for d in range(6)]).astype(np.int64)
    full = lambda v: np.full(n_el, float(v)) if np.isscalar(v) else np.ascontiguousarray(v, dtype=np.float64)
    return ShellProblem(m, full(E), full(nu), full(t), full(rho), bd, di, n_dof, clamp_nodes, pen)


def config3():
    """1 x 1 m plate, 256 x 256 shell4, t = 5 mm, E = 70 GPa, nu = 0.33,
    rho = 2700 kg/m^3, edges clamped (66,049 nodes, 396,294 DOF)."""
    return shell_plate(256, 256, 1.0, 1.0, 0.005, 70e9, 0.33, 2700.0)
