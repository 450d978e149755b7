"""Pins the CPU restatement of the reference's Newmark stepping
(oracle/dynamics.py): the average-acceleration rule's closed-form free
vibration of one DOF, energy conservation of a free brick8 cantilever, and the
agreement of the literal first step with the predictor form from rest."""
import numpy as np
import scipy.sparse as sp

import dynamics
import oracle as O
from asend3d_amd import problems


def test_single_dof_free_vibration_closed_form():
    # trapezoidal rule: u_n = cos(n phi) with tan(phi / 2) = omega dt / 2
    m, k, dt = 2.0, 8.0e4, 1e-3
    omega = np.sqrt(k / m)
    n = 400
    hist, _, _, _ = dynamics.newmark(sp.csr_matrix([[k]]), sp.csr_matrix([[m]]), np.zeros(1), np.zeros(n), dt,
                                     u0=[1.0], v0=[0.0], a0=[-omega ** 2], watch=[0])
    phi = 2.0 * np.arctan(omega * dt / 2.0)
    exact = np.cos(phi * np.arange(1, n + 1))
    assert np.abs(hist[:, 0] - exact).max() < 1e-11
    # and the period elongation is what the rule predicts, not the exact 2 pi / omega
    assert abs(phi - omega * dt) > 1e-6


def _cantilever_km():
    p = problems.cantilever(6, 2, 2, E=70e9)
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof)
    M = O.assemble_solid_mass(8, p.mesh.xyz, p.mesh.conn, np.full(p.mesh.n_el, 2700.0), p.dof_index, K)
    return p, K.to_scipy(), M.to_scipy()


def test_energy_conserved_in_free_vibration():
    # a half-sine pulse for 20 steps, then free: 1/2 v^T M v + 1/2 u^T K u stays
    # constant under the average-acceleration rule (penalty springs included)
    p, K, M = _cantilever_km()
    fac = np.zeros(80)
    fac[:20] = np.sin(np.pi * np.arange(20) / 20.0)
    energies = []

    def record(step, u, v, a, Kp):
        if step >= 20:
            energies.append(0.5 * v @ (M @ v) + 0.5 * u @ (Kp @ u))

    dynamics.newmark(K, M, p.f, fac, 2e-5, pen_dofs=p.pen_dofs, on_step=record)
    energies = np.array(energies)
    assert energies[0] > 0
    assert np.abs(energies - energies[0]).max() <= 1e-9 * energies[0]


def test_reference_first_step_equals_predictor_from_rest():
    p, K, M = _cantilever_km()
    fac = np.sin(np.pi * np.arange(30) / 30.0)
    h1, *_ = dynamics.newmark(K, M, p.f, fac, 1e-5, pen_dofs=p.pen_dofs, watch=[p.n_dof - 1],
                              first_step="reference")
    h2, *_ = dynamics.newmark(K, M, p.f, fac, 1e-5, pen_dofs=p.pen_dofs, watch=[p.n_dof - 1],
                              first_step="predictor")
    assert np.array_equal(h1, h2)


def test_adjoint_matches_finite_differences():
    # L = sum_n w_n q.u_n (the tip deflection weighted over the history);
    # central differences of the forward run in single element moduli, with
    # the penalty scale held at the unperturbed value
    p = problems.cantilever(6, 2, 2, E=70e9)
    rng = np.random.default_rng(3)
    E = 70e9 * np.exp(0.4 * rng.standard_normal(p.mesh.n_el))
    rho = np.full(p.mesh.n_el, 2700.0)

    def km(Ev):
        K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, Ev, p.nu, p.dof_index, p.n_dof)
        M = O.assemble_solid_mass(8, p.mesh.xyz, p.mesh.conn, rho, p.dof_index, K)
        return K.to_scipy(), M.to_scipy()

    K, M = km(E)
    dt, n = 2e-5, 24
    fac = np.where(np.arange(n) < 12, np.sin(np.pi * np.arange(n) / 12.0), 0.0)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    q = np.zeros(p.n_dof)
    q[di[p.mesh.nodes_at_x(6), 2]] = 1.0
    w = 1.0 + 0.1 * np.arange(n)
    c1 = 1.0 / (dt * dt * (0.25 - 0.5))
    sf = 1e5 * abs(K - c1 * M).max()
    sens = lambda u, lam: O.solid_sens(8, p.mesh.xyz, p.mesh.conn, E, p.nu, p.dof_index, u, lam)
    L, dL = dynamics.newmark_adjoint(K, M, p.f, fac, dt, 0.25, 0.5, q, w, sens, pen_dofs=p.pen_dofs, sf=sf)
    for e in (0, 7, 23):
        h = 1e-5 * E[e]
        Lp, Lm = [], []
        for sgn, out in ((1.0, Lp), (-1.0, Lm)):
            Ev = E.copy()
            Ev[e] += sgn * h
            Kv, Mv = km(Ev)
            hist, *_ = dynamics.newmark(Kv, Mv, p.f, fac, dt, pen_dofs=p.pen_dofs, sf=sf, watch=np.flatnonzero(q))
            out.append(float(np.sum(w[:, None] * hist)))
        fd = (Lp[0] - Lm[0]) / (2 * h)
        assert abs(dL[e] - fd) <= 1e-5 * np.abs(dL).max(), (e, dL[e], fd)
    # the objective itself is the weighted history
    hist, *_ = dynamics.newmark(K, M, p.f, fac, dt, pen_dofs=p.pen_dofs, sf=sf, watch=np.flatnonzero(q))
    assert np.isclose(L, np.sum(w[:, None] * hist), rtol=1e-12)
