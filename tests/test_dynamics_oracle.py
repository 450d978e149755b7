"""Pins the CPU restatement of the reference's Newmark stepping
(oracle/dynamics.py): the average-acceleration rule's closed-form free
vibration of one DOF, the damped single DOF against the analytic decaying
oscillation, energy conservation of a free brick8 cantilever and the exact
discrete energy balance with Rayleigh damping, the agreement of the literal
first step with the predictor form from rest, and the transient adjoint
against finite differences with and without damping."""
import numpy as np
import pytest
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


def test_damped_single_dof_follows_the_analytic_decay():
    # m u'' + c u' + k u = 0 with c = cm m + ck k (zeta = 0.05), u(0) = 1,
    # u'(0) = 0: the trapezoidal rule converges to the analytic solution at
    # second order; a dropped or mis-signed damping term does not decay at
    # the analytic rate
    m, k = 2.0, 8.0e4
    omega = np.sqrt(k / m)
    cm, ck = 0.05 * omega, 0.05 / omega  # zeta = (cm / omega + ck omega) / 2
    zeta = 0.5 * (cm / omega + ck * omega)
    wd = omega * np.sqrt(1.0 - zeta ** 2)
    dt = 2e-5
    n = int(4 * 2 * np.pi / omega / dt)
    c = cm * m + ck * k
    hist, _, _, _ = dynamics.newmark(sp.csr_matrix([[k]]), sp.csr_matrix([[m]]), np.zeros(1), np.zeros(n), dt,
                                     u0=[1.0], v0=[0.0], a0=[-k / m], watch=[0], damp_cm=cm, damp_ck=ck)
    t = dt * np.arange(1, n + 1)
    exact = np.exp(-zeta * omega * t) * (np.cos(wd * t) + zeta * omega / wd * np.sin(wd * t))
    assert np.abs(hist[:, 0] - exact).max() < 1e-5
    assert abs(c - 2 * zeta * omega * m) < 1e-12 * c
    # the undamped run does not decay
    h0, *_ = dynamics.newmark(sp.csr_matrix([[k]]), sp.csr_matrix([[m]]), np.zeros(1), np.zeros(n), dt,
                              u0=[1.0], v0=[0.0], a0=[-k / m], watch=[0])
    assert np.abs(h0[:, 0] - exact).max() > 0.1


def test_rayleigh_energy_balance_is_exact():
    # free vibration from a consistent random state: the trapezoidal rule
    # gives E_{n+1} - E_n = -dt vbar^T D vbar exactly, vbar = (v_n + v_{n+1})/2
    p, K, M = _cantilever_km()
    cm, ck = 40.0, 2e-6
    pen = np.zeros(p.n_dof)
    dt = 2e-5
    c1 = 1.0 / (dt * dt * (0.25 - 0.5))
    sf = 1e5 * abs(K - c1 * M - c1 * dt * 0.5 * (cm * M + ck * K)).max()
    np.add.at(pen, p.pen_dofs, sf)
    Kp = K + sp.diags(pen)
    D = cm * M + ck * K
    rng = np.random.default_rng(5)
    u0 = 1e-6 * rng.standard_normal(p.n_dof)
    v0 = 1e-3 * rng.standard_normal(p.n_dof)
    u0[p.pen_dofs] = 0.0
    v0[p.pen_dofs] = 0.0
    a0 = sp.linalg.spsolve(M.tocsc(), -(Kp @ u0) - D @ v0)
    states = [(u0, v0)]
    dynamics.newmark(K, M, np.zeros(p.n_dof), np.zeros(60), dt, u0=u0, v0=v0, a0=a0, pen_dofs=p.pen_dofs, sf=sf,
                     on_step=lambda s, u, v, a, Kq: states.append((u.copy(), v.copy())), damp_cm=cm, damp_ck=ck)
    energy = [0.5 * v @ (M @ v) + 0.5 * u @ (Kp @ u) for u, v in states]
    for i in range(len(states) - 1):
        vb = 0.5 * (states[i][1] + states[i + 1][1])
        loss = dt * vb @ (D @ vb)
        assert loss > 0
        assert abs((energy[i + 1] - energy[i]) + loss) <= 1e-9 * energy[0]
    assert energy[-1] < 0.99 * energy[0]


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


@pytest.mark.parametrize("damp", [(0.0, 0.0, 0.25, 0.5), (300.0, 4e-6, 0.25, 0.5), (300.0, 4e-6, 0.25, 0.55)])
def test_adjoint_matches_finite_differences(damp):
    # L = sum_n w_n q.u_n (the tip deflection weighted over the history);
    # central differences of the forward run in single element moduli, with
    # the penalty scale held at the unperturbed value; with Rayleigh damping
    # the operator, the coupling terms and dR/dE all carry D (the D term of
    # the acceleration coupling, dt (1 + gamma / (2 (beta - gamma))), vanishes
    # whenever beta = gamma / 2: the third case keeps it)
    cm, ck, beta, gamma = damp
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
    c1 = 1.0 / (dt * dt * (beta - gamma))
    sf = 1e5 * abs(K - c1 * M - c1 * dt * gamma * (cm * M + ck * K)).max()
    sens = lambda u, lam: O.solid_sens(8, p.mesh.xyz, p.mesh.conn, E, p.nu, p.dof_index, u, lam)
    L, dL = dynamics.newmark_adjoint(K, M, p.f, fac, dt, beta, gamma, q, w, sens, pen_dofs=p.pen_dofs, sf=sf,
                                     damp_cm=cm, damp_ck=ck)
    for e in (0, 7, 23):
        h = 1e-5 * E[e]
        Lp, Lm = [], []
        for sgn, out in ((1.0, Lp), (-1.0, Lm)):
            Ev = E.copy()
            Ev[e] += sgn * h
            Kv, Mv = km(Ev)
            hist, *_ = dynamics.newmark(Kv, Mv, p.f, fac, dt, beta, gamma, pen_dofs=p.pen_dofs, sf=sf,
                                        watch=np.flatnonzero(q), damp_cm=cm, damp_ck=ck)
            out.append(float(np.sum(w[:, None] * hist)))
        fd = (Lp[0] - Lm[0]) / (2 * h)
        assert abs(dL[e] - fd) <= 1e-5 * np.abs(dL).max(), (e, dL[e], fd)
    # the objective itself is the weighted history
    hist, *_ = dynamics.newmark(K, M, p.f, fac, dt, beta, gamma, pen_dofs=p.pen_dofs, sf=sf,
                                watch=np.flatnonzero(q), damp_cm=cm, damp_ck=ck)
    assert np.isclose(L, np.sum(w[:, None] * hist), rtol=1e-12)
