"""Host-side product code: library exports, node reordering, traction loads,
seeded moduli.  CPU only (no device calls)."""
import os
import re

import numpy as np
import pytest

import oracle as O
from asend3d_amd import api
from asend3d_amd.meshgen import hex_block

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "asend_gpu.h")).read()
    declared = set(re.findall(r"\b(asend_[a-z0-9_]+)\s*\(", hdr))
    assert declared, "no declarations parsed"
    lib = api.lib()
    for name in sorted(declared):
        assert hasattr(lib, name), f"{name} declared but not exported"
    assert set(api.EXPORTED) <= declared
    assert "gfx950" in api.version()


def _fixture_cases():
    import json
    return json.load(open(os.path.join(ROOT, "tests", "golden", "reorder_hand_traced.json")))["cases"]


@pytest.mark.parametrize("case", _fixture_cases(), ids=lambda c: c["name"])
def test_reorder_matches_hand_traced_orderings(case):
    # the host reordering against orderings stepped by hand through the
    # reference routine (tests/golden/reorder_hand_traced.md)
    cons = [[np.asarray(t, dtype=np.int32) for t in con] for con in case["constraints"]]
    ordered, di, ndof = api.reorder_nodes(np.asarray(case["xyz"], float), np.asarray(case["conn"], dtype=np.int32),
                                          case["dof_per_nd"], cons, case["block_dim"])
    assert list(ordered) == case["ordered"]
    assert np.array_equal(di, np.asarray(case["dof_index"])) and ndof == case["n_dof"]


@pytest.mark.parametrize("dims,bd", [((20, 4, 4), None), ((13, 7, 3), None), ((9, 9, 9), 12), ((40, 8, 8), 30),
                                     ((6, 3, 2), 2000000000)])
def test_reorder_matches_oracle_bit_exact(dims, bd):
    m = hex_block(*dims, 1.0, 0.37, 0.21)
    clamp = m.nodes_at_x(0)
    cons = [[clamp], [clamp], [clamp]]
    bd = api.block_dim_auto(m.n_nodes) if bd is None else bd
    a = api.reorder_nodes(m.xyz, m.conn, 3, cons, bd)
    b = O.reorder_nodes(m.xyz, m.conn, 3, cons, bd)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_reorder_tie_breaks_and_multi_term_constraints():
    # uniform grid: many equal distances at restarts; a two-set tie constraint
    # couples nodes pairwise (S/model_prep.rs:31-41)
    m = hex_block(8, 6, 5, 1.0, 1.0, 1.0)
    s1 = m.nodes_at_x(0)
    s2 = m.nodes_at_x(m.nx)
    single = np.array([5], dtype=np.int32)
    cons = [[s1, s2], [single, s2[:1]], [s1]]
    for bd in (7, 18, 33):
        a = api.reorder_nodes(m.xyz, m.conn, 3, cons, bd)
        b = O.reorder_nodes(m.xyz, m.conn, 3, cons, bd)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_reorder_shell_dofs():
    # 6-DOF nodes number six DOFs per node in insertion order
    m = hex_block(4, 3, 2, 1.0, 1.0, 1.0)
    a = api.reorder_nodes(m.xyz, m.conn, 6, [], 10)
    b = O.reorder_nodes(m.xyz, m.conn, 6, [], 10)
    assert np.array_equal(a[1], b[1]) and a[2] == 6 * m.n_nodes


def test_block_dim_rule():
    assert api.block_dim_auto(525) == 24
    assert api.block_dim_auto(337881) == 582
    assert api.block_dim_auto(66049) == 258


@pytest.mark.parametrize("dims", [(20, 4, 4), (7, 3, 5)])
def test_traction_matches_oracle(dims):
    m = hex_block(*dims, 1.3, 0.2, 0.45)
    rng = np.random.default_rng(0)
    xyz = m.xyz + 0.02 * rng.standard_normal(m.xyz.shape) * np.array([1, 1, 1]) * (m.xyz[:, :1] < 1.2)
    clamp = m.nodes_at_x(0)
    _, di, nd = O.reorder_nodes(xyz, m.conn, 3, [[clamp]] * 3, 24)
    tip = m.elements_at_x(m.nx - 1)
    for load in ([0, 0, -1e6], [3.0, -1.0, 2.0]):
        a = api.surface_traction(8, xyz, m.conn, di, nd, tip, load, [1, 0, 0], 5.0)
        b = O.surface_traction(8, xyz, m.conn, di, nd, tip, load, [1, 0, 0], 5.0)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
    # all elements in the set: only faces facing +x within 5 degrees load
    every = np.arange(m.n_el, dtype=np.int32)
    a = api.surface_traction(8, xyz, m.conn, di, nd, every, [0, 0, -1.0], [1, 0, 0], 5.0)
    b = O.surface_traction(8, xyz, m.conn, di, nd, every, [0, 0, -1.0], [1, 0, 0], 5.0)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dims", [(6, 3, 2), (3, 1, 1)])
def test_thermal_loads_match_oracle(dims):
    """surfaceFlux (asend_surface_flux) and bodyHeatGen (asend_body_load kind
    0 with load (Q, 0, 0)) against the oracle's restatement, on a mesh with
    moved nodes, for flux directions that pass on one face and on several (the
    running normal test)."""
    import thermal as TH
    m = hex_block(*dims, 1.3, 0.4, 0.3)
    rng = np.random.default_rng(4)
    xyz = m.xyz + 0.01 * rng.standard_normal(m.xyz.shape)
    _, di, nd = O.reorder_nodes(xyz, m.conn, 3, [[m.nodes_at_x(0)]] * 3, 24)
    every = np.arange(m.n_el, dtype=np.int32)
    for ndir, tol in (([1, 0, 0], 5.0), ([-1, 0, 0], 5.0), ([0, 1, 0], 10.0), ([0, 0, -1], 45.0), ([1, 1, 0], 50.0)):
        a = api.surface_flux(8, xyz, m.conn, di, nd, every, -3.5, ndir, tol)
        b = TH.to_dofs(TH.surface_flux(xyz, m.conn, every, -3.5, ndir, tol), di, nd)
        assert np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1e-300), (ndir, tol)
        assert not a.reshape(-1, 3)[:, 1:].any()
    a = api.body_load(8, api.BODY_FORCE, xyz, m.conn, di, nd, every, load=(7.0, 0.0, 0.0))
    b = TH.to_dofs(TH.body_heat_gen(xyz, m.conn, every, 7.0), di, nd)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_thermal_strain_load_matches_oracle():
    import thermal as TH
    m = hex_block(4, 2, 3, 1.3, 0.4, 0.3)
    rng = np.random.default_rng(6)
    xyz = m.xyz + 0.01 * rng.standard_normal(m.xyz.shape)
    _, di, nd = O.reorder_nodes(xyz, m.conn, 3, [[m.nodes_at_x(0)]] * 3, 24)
    E = rng.uniform(50e3, 90e3, m.n_el)
    nu = rng.uniform(0.2, 0.35, m.n_el)
    al = rng.uniform(-1e-5, 3e-5, (m.n_el, 6))
    T = rng.uniform(-20.0, 120.0, m.n_nodes)
    a = api.thermal_strain_load(8, xyz, m.conn, di, nd, E, nu, al, T)
    b = TH.thermal_strain_load(xyz, m.conn, E, nu, al, T)
    d3 = np.asarray(di).reshape(-1, 6)[:, :3]
    assert np.abs(a[d3] - b).max() <= 1e-12 * np.abs(b).max()


def test_lognormal_moduli_seeded():
    a = api.lognormal_moduli(20261202, 200000, 70e3, 1.0)
    b = api.lognormal_moduli(20261202, 200000, 70e3, 1.0)
    assert np.array_equal(a, b)
    z = np.log(a / 70e3)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01
    c = api.lognormal_moduli(1, 10, 70e3, 1.0)
    assert not np.allclose(a[:10], c)


@pytest.mark.parametrize("etype", [8, 81])
def test_pressure_and_body_loads_match_oracle(etype):
    # distorted 6x3x2 block, per-element density; every load type on all
    # elements and on a subset
    m = hex_block(6, 3, 2, 1.3, 0.4, 0.3)
    rng = np.random.default_rng(4)
    inner = np.all((m.xyz > m.xyz.min(axis=0) + 1e-9) & (m.xyz < m.xyz.max(axis=0) - 1e-9), axis=1)
    xyz = m.xyz.copy()
    xyz[inner] += rng.uniform(-0.03, 0.03, (int(inner.sum()), 3))
    clamp = m.nodes_at_x(0)
    _, di, nd = O.reorder_nodes(xyz, m.conn, 3, [[clamp]] * 3, 24)
    rho = rng.uniform(500.0, 3000.0, m.n_el)
    every = np.arange(m.n_el, dtype=np.int32)
    some = every[1::2]  # odd labels: tip (x = 5) and bottom-layer elements among them
    for es in (every, some):
        for ndir in ([1.0, 0.0, 0.0], [0.0, 0.0, -2.0]):
            a = api.surface_pressure(etype, xyz, m.conn, di, nd, es, 2.5e5, ndir, 10.0)
            b = O.surface_pressure(etype, xyz, m.conn, di, nd, es, 2.5e5, ndir, 10.0)
            assert np.abs(b).max() > 0
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())
        for kind, kw in ((api.BODY_FORCE, {"load": [3.0, -2.0, 9.81]}),
                         (api.GRAVITATIONAL, {"rho": rho, "load": [0.0, 0.0, -9.81]}),
                         (api.CENTRIFUGAL, {"rho": rho, "center": [0.2, 0.1, -0.3], "axis": [0.0, 1.0, 1.0],
                                             "ang_vel": 30.0})):
            a = api.body_load(etype, kind, xyz, m.conn, di, nd, es, **kw)
            b = O.body_load(etype, kind, xyz, m.conn, di, nd, es, **kw)
            assert np.abs(b).max() > 0
            assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max()), kind
