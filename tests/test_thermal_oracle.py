"""Pins of the heat conduction oracle (oracle/thermal.py) by closed forms: the
box conduction matrix from one-dimensional integrals, the patch test of a
linear temperature field, box volume and face-area sums, the tensor-product
integrals behind the surfaceFlux distribution (including the reference's
running normal test), and one-dimensional temperature profiles."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import thermal as TH
from asend3d_amd.meshgen import hex_block


def _bits(a):
    s = TH._SIGN[a]
    return [int((s[d] + 1) // 2) for d in range(3)]


def _box_matrix(h, k6):
    """Conduction matrix of an axis-aligned box (edges h) from exact 1D
    integrals on [0, h_d]: S = int phi' phi', M = int phi phi, C = int phi' phi;
    int d_m N_a d_n N_b factorises over the three directions."""
    k = TH.kmat(k6)
    S = [np.array([[1.0, -1.0], [-1.0, 1.0]]) / hd for hd in h]
    M = [hd * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]) for hd in h]
    C = [np.array([[-0.5, -0.5], [0.5, 0.5]]) for _ in h]  # C[i, j] = int phi_i' phi_j
    K = np.zeros((8, 8))
    for a in range(8):
        ia = _bits(a)
        for b in range(8):
            ib = _bits(b)
            for m in range(3):
                for n in range(3):
                    v = 1.0
                    for d in range(3):
                        if d == m and d == n:
                            v *= S[d][ia[d], ib[d]]
                        elif d == m:
                            v *= C[d][ia[d], ib[d]]
                        elif d == n:
                            v *= C[d][ib[d], ia[d]]
                        else:
                            v *= M[d][ia[d], ib[d]]
                    K[a, b] += k[m, n] * v
    return K


def _box(h, origin=(0.0, 0.0, 0.0)):
    return np.array(origin) + TH._SIGN.clip(0, 1) * np.array(h)


def test_unit_cube_textbook_values():
    Ke = TH.conduction_element(_box((1.0, 1.0, 1.0)), (1, 1, 1, 0, 0, 0))
    for a in range(8):
        for b in range(8):
            nd = sum(x != y for x, y in zip(_bits(a), _bits(b)))
            want = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[nd]
            assert abs(Ke[a, b] - want) <= 1e-14


@pytest.mark.parametrize("k6", [(1, 1, 1, 0, 0, 0), (2.0, 0.5, 3.0, 0.3, -0.2, 0.1)])
def test_box_matrix_closed_form(k6):
    h = (0.7, 0.3, 1.9)
    Ke = TH.conduction_element(_box(h, (1.0, -2.0, 0.5)), k6)
    assert np.abs(Ke - _box_matrix(h, k6)).max() <= 1e-13 * np.abs(Ke).max()


def test_element_symmetric_psd_constant_null_space():
    rng = np.random.default_rng(5)
    X = _box((1.0, 0.8, 1.2)) + rng.uniform(-0.15, 0.15, (8, 3))
    Ke = TH.conduction_element(X, (2.0, 1.5, 1.0, 0.4, -0.3, 0.2))
    assert np.abs(Ke - Ke.T).max() <= 1e-14 * np.abs(Ke).max()
    assert np.abs(Ke @ np.ones(8)).max() <= 1e-13 * np.abs(Ke).max()
    w = np.linalg.eigvalsh(Ke)
    assert w[0] >= -1e-13 * w[-1] and w[1] > 1e-6 * w[-1]


def _grid_counts(nx, ny, nz):
    """For each node of hex_block: the number of boundary quads it touches
    on each of the six box planes (x0, x1, y0, y1, z0, z1)."""
    m = hex_block(nx, ny, nz, 1.0, 1.0, 1.0)
    idx = np.rint(m.xyz * np.array([nx, ny, nz])).astype(int)

    def touch(i, n):
        return 1 if i in (0, n) else 2

    out = np.zeros((m.n_nodes, 6))
    for r, (i, j, k) in enumerate(idx):
        if i == 0:
            out[r, 0] = touch(j, ny) * touch(k, nz)
        if i == nx:
            out[r, 1] = touch(j, ny) * touch(k, nz)
        if j == 0:
            out[r, 2] = touch(i, nx) * touch(k, nz)
        if j == ny:
            out[r, 3] = touch(i, nx) * touch(k, nz)
        if k == 0:
            out[r, 4] = touch(i, nx) * touch(j, ny)
        if k == nz:
            out[r, 5] = touch(i, nx) * touch(j, ny)
    return out


def test_patch_test_linear_field_anisotropic():
    nx, ny, nz = 4, 3, 3
    lx, ly, lz = 2.0, 1.0, 1.5
    m = hex_block(nx, ny, nz, lx, ly, lz)
    xyz = m.xyz.copy()
    rng = np.random.default_rng(9)
    inner = np.all((xyz > 1e-9) & (xyz < np.array([lx, ly, lz]) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-0.2, 0.2, (int(inner.sum()), 3)) * np.array([lx / nx, ly / ny, lz / nz])
    k6 = (2.0, 1.5, 1.0, 0.4, -0.3, 0.2)
    K = TH.assemble_conduction(xyz, m.conn, np.tile(k6, (m.n_el, 1)))
    grad = np.array([0.7, -1.3, 0.4])
    T = xyz @ grad + 5.0
    r = K @ T
    qv = TH.kmat(k6) @ grad  # flux vector k grad T
    cnt = _grid_counts(nx, ny, nz)
    area = np.array([ly * lz / (ny * nz)] * 2 + [lx * lz / (nx * nz)] * 2 + [lx * ly / (nx * ny)] * 2)
    normal = np.array([-qv[0], qv[0], -qv[1], qv[1], -qv[2], qv[2]])
    want = (cnt * area * normal).sum(axis=1) / 4.0
    assert np.abs(r - want).max() <= 1e-12 * np.abs(want).max()
    assert np.abs(r[inner]).max() <= 1e-12 * np.abs(want).max()


def test_body_heat_gen_box_and_distorted():
    m = hex_block(3, 2, 2, 1.5, 1.0, 1.0)
    f = TH.body_heat_gen(m.xyz, m.conn, np.arange(m.n_el), 4.0)
    V = 1.5 * 1.0 * 1.0 / m.n_el
    want = np.zeros(m.n_nodes)
    for e in range(m.n_el):
        want[m.conn[e]] += 4.0 * V / 8.0
    assert np.abs(f - want).max() <= 1e-14 * want.max()
    rng = np.random.default_rng(2)
    X = _box((1.0, 1.0, 1.0)) + rng.uniform(-0.1, 0.1, (8, 3))
    v, vol = TH.heat_rate(X, np.full(8, 1.0))
    assert abs(v.sum() - vol) <= 1e-14 * vol


def _mass_box(h):
    M1 = [hd * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]) for hd in h]
    M = np.zeros((8, 8))
    for a in range(8):
        for b in range(8):
            ia, ib = _bits(a), _bits(b)
            M[a, b] = np.prod([M1[d][ia[d], ib[d]] for d in range(3)])
    return M


def test_surface_flux_minus_x_face_is_area_weighted():
    nx, ny, nz = 3, 2, 2
    lx, ly, lz = 1.2, 0.6, 0.8
    m = hex_block(nx, ny, nz, lx, ly, lz)
    q = 2.5
    f = TH.surface_flux(m.xyz, m.conn, m.elements_at_x(0), q, (-1.0, 0.0, 0.0), 5.0)
    cnt = _grid_counts(nx, ny, nz)[:, 0]
    want = q * cnt * (ly * lz / (ny * nz)) / 4.0
    assert np.abs(f - want).max() <= 1e-14 * want.max()
    assert abs(f.sum() - q * ly * lz) <= 1e-13 * q * ly * lz


def test_surface_flux_running_normal_test():
    """With normDir +x, the face test adds n . normDir over the element's
    surface faces in face order (-z, +z, -y, +x, +y, -x): after the +x face
    passes, the +y face (dot 0) passes too, and the element's vector is the +y
    face's, computed with the field q on both faces' nodes."""
    nx, ny, nz = 3, 2, 2
    h = (0.4, 0.3, 0.5)
    m = hex_block(nx, ny, nz, nx * h[0], ny * h[1], nz * h[2])
    q = 1.7
    tip = m.elements_at_x(nx - 1)
    f = TH.surface_flux(m.xyz, m.conn, tip, q, (1.0, 0.0, 0.0), 5.0)
    M = _mass_box(h)
    on = TH.surface_faces(m.conn)
    want = np.zeros(m.n_nodes)
    for e in tip:
        c = m.conn[e]
        if on[e, 4]:  # +y face on the surface: it takes the load
            face, area = TH.HEX_FACES[4], h[0] * h[2]
            U = sorted(set(TH.HEX_FACES[3]) | set(face))
        else:
            face, area = TH.HEX_FACES[3], h[1] * h[2]
            U = list(face)
        v = np.array([M[a, U].sum() if a in face else 0.0 for a in range(8)])
        want[c] += v * (q * area / v.sum())
    assert np.abs(f - want).max() <= 1e-14 * np.abs(want).max()
    # the load leaves the +x face of the elements on the +y side
    assert abs(f.sum() - q * h[2] * nz * ((ny - 1) * h[1] + h[0])) <= 1e-13 * abs(f.sum())


def _penalised_solve(K, pen, T_pen, f):
    import scipy.sparse as sp
    sf = 1e5 * np.abs(K.data).max()
    P = sp.csr_matrix((np.full(len(pen), sf), (pen, pen)), shape=K.shape)
    b = f.copy()
    b[pen] += sf * T_pen
    return spla.spsolve((K + P).tocsc(), b)


def test_bar_profiles_closed_form():
    nx, L = 10, 2.0
    m = hex_block(nx, 2, 2, L, 0.4, 0.4)
    k11 = 3.0
    kc = np.tile((k11, 1.0, 2.0, 0.0, 0.0, 0.0), (m.n_el, 1))
    K = TH.assemble_conduction(m.xyz, m.conn, kc)
    x0, x1 = m.nodes_at_x(0), m.nodes_at_x(nx)
    x = m.xyz[:, 0]
    # fixed ends and uniform generation: T = T0 + (T1 - T0) x / L + Q x (L - x) / (2 k)
    T0, T1, Q = 10.0, 30.0, 50.0
    f = TH.body_heat_gen(m.xyz, m.conn, np.arange(m.n_el), Q)
    pen = np.concatenate([x0, x1])
    T = _penalised_solve(K, pen, np.concatenate([np.full(len(x0), T0), np.full(len(x1), T1)]), f)
    exact = T0 + (T1 - T0) * x / L + Q * x * (L - x) / (2 * k11)
    assert np.abs(T - exact).max() <= 1e-4 * np.abs(exact).max()
    # heat q into the x = 0 face, x = L held at 0: T = q (L - x) / k
    q = 6.0
    f = TH.surface_flux(m.xyz, m.conn, m.elements_at_x(0), q, (-1.0, 0.0, 0.0), 5.0)
    T = _penalised_solve(K, x1, np.zeros(len(x1)), f)
    exact = q * (L - x) / k11
    assert np.abs(T - exact).max() <= 1e-4 * np.abs(exact).max()


def test_capacity_box_closed_form():
    h = (0.7, 0.3, 1.9)
    Ce = TH.capacity_element(_box(h), 2.5)
    M1 = [hd * np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]]) for hd in h]
    for a in range(8):
        for b in range(8):
            ia, ib = _bits(a), _bits(b)
            want = 2.5 * np.prod([M1[d][ia[d], ib[d]] for d in range(3)])
            assert abs(Ce[a, b] - want) <= 1e-14 * 2.5 * np.prod(h)


def test_transient_uniform_heating_is_linear():
    """No constraint, uniform generation Q: K T = 0 for a uniform T, so
    rho c_p dT/dt = Q.  The reference starts from a zero rate, so the first
    step gains gamma dt Q / (rho c_p) and every later one dt Q / (rho c_p):
    T_n = (Q / rho c_p) dt (n - (1 - gamma)), rate Q / (rho c_p), at every
    node."""
    m = hex_block(3, 2, 2, 1.0, 0.5, 0.5)
    K = TH.assemble_conduction(m.xyz, m.conn, np.tile((2.0, 1.0, 1.5, 0.2, 0.0, 0.1), (m.n_el, 1)))
    C = TH.assemble_capacity(m.xyz, m.conn, np.full(m.n_el, 4.0))
    Q = 3.0
    f = TH.body_heat_gen(m.xyz, m.conn, np.arange(m.n_el), Q)
    for gamma in (0.5, 0.8):
        hist = TH.transient(K, C, np.zeros(0, dtype=int), 0.0, np.zeros(0), lambda t: f, 0.1, gamma, 0.5)
        for k, (T, Td) in enumerate(hist):
            assert np.abs(T - Q / 4.0 * 0.1 * (k + 1 - (1.0 - gamma))).max() <= 1e-12
            assert np.abs(Td - Q / 4.0).max() <= 1e-12


def test_transient_converges_to_semidiscrete_exact():
    """gamma = 1/2: the error against the exp(-A t) solution of
    C T' + (K + P) T = b halves with dt.  First order: the zero starting
    rate is inconsistent with C T'(0) = b, which costs dt/2 of rise in the
    first step (the trapezoidal rule alone would give 4)."""
    import scipy.linalg as sl
    import scipy.sparse as sp
    m = hex_block(4, 1, 1, 1.0, 0.25, 0.25)
    K = TH.assemble_conduction(m.xyz, m.conn, np.tile((1.0, 1.0, 1.0, 0.0, 0.0, 0.0), (m.n_el, 1)))
    C = TH.assemble_capacity(m.xyz, m.conn, np.ones(m.n_el))
    pen = m.nodes_at_x(0)
    sf = 1e3 * np.abs(K.data).max()  # a soft penalty keeps the exact solution smooth
    f = TH.body_heat_gen(m.xyz, m.conn, np.arange(m.n_el), 2.0)
    Kp = (K + sp.csr_matrix((np.full(len(pen), sf), (pen, pen)), shape=K.shape)).toarray()
    Cd = C.toarray()
    A = np.linalg.solve(Cd, Kp)
    Tinf = np.linalg.solve(Kp, f)
    t_end = 0.4
    exact = Tinf - sl.expm(-A * t_end) @ Tinf
    errs = []
    for dt in (0.02, 0.01):
        hist = TH.transient(K, C, pen, sf, np.zeros(len(pen)), lambda t: f, dt, 0.5, t_end - 1e-9)
        errs.append(np.abs(hist[-1][0] - exact).max())
    assert 1.8 <= errs[0] / errs[1] <= 2.3, errs
    # and the error is the first step's lost rise carried forward: below dt b / C
    assert errs[1] <= 0.01 * np.abs(exact).max()


def test_thermal_strain_load_is_free_expansion():
    """A uniform temperature T with expansion coefficients alpha expands a
    free body stress-free: u = T (a1 x + a4 y, a2 y + a6 z... ) is the linear
    field with engineering strains alpha T, so the load must equal K u for
    the oracle's own (separately pinned) stiffness, on a mesh with moved
    interior nodes."""
    import oracle as O
    m = hex_block(3, 2, 2, 1.5, 1.0, 0.8)
    xyz = m.xyz.copy()
    rng = np.random.default_rng(3)
    inner = np.all((xyz > 1e-9) & (xyz < np.array([1.5, 1.0, 0.8]) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-0.1, 0.1, (int(inner.sum()), 3))
    E = np.full(m.n_el, 70e3)
    nu = np.full(m.n_el, 0.3)
    al = np.array([1e-5, 2e-5, 3e-5, 4e-6, 0.0, 5e-6])
    T = np.full(m.n_nodes, 80.0)
    f = TH.thermal_strain_load(xyz, m.conn, E, nu, np.tile(al, (m.n_el, 1)), T)
    # u_x = e11 x + g12 y, u_y = e22 y + g23 z, u_z = e33 z gives strains (e11, e22, e33, g12, 0, g23)
    e = al * 80.0
    U = np.column_stack([e[0] * xyz[:, 0] + e[3] * xyz[:, 1], e[1] * xyz[:, 1] + e[5] * xyz[:, 2], e[2] * xyz[:, 2]])
    _, di6, nd = O.reorder_nodes(xyz, m.conn, 3, [], 1000)
    K = O.assemble_solid(8, xyz, m.conn, E, nu, di6, nd).to_scipy()
    di = np.asarray(di6).reshape(-1, 6)[:, :3]
    u = np.zeros(nd)
    u[di.reshape(-1)] = U.reshape(-1)
    Ku = K @ u
    assert np.abs(f - Ku[di]).max() <= 1e-10 * np.abs(Ku).max()
