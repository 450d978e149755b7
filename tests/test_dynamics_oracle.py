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
