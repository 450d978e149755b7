"""CPU restatement of the reference's Newmark time stepping (numpy/scipy, FP64).

TEST INFRASTRUCTURE ONLY: the checker of asend_newmark_run; only tests/ import
it.  Follows, per step of Model::solve's dynamic loop
(S/model_analysis.rs:560-615, S/ = solverSrc/asend_run_job/src/):

  solve_step(time)            (S/model_analysis.rs:312-558, elastic branch)
      rhs = f(time) - R(u) + penalty load, R(u) = K u + M a + D v with the
      node accelerations a and velocities v of the current state; one linear
      solve with K_eff = K + c_m M + c_m dt gamma D (+ sf C^T C), c_m = -c1
      (Element::get_ru_dfd0, S/element_equations.rs:948-1003); u += du;
      D = ray_damp_cm M + ray_damp_ck K is Rayleigh damping
      (Element::get_rud_dfd0, S/element_equations.rs:699-753; material
      damping matrices are not restated);
      update_vel_acc           (S/node_meth.rs:127-140)
          c1 = 1/(dt^2 (beta - gamma)), c2 = dt^2 (0.5 + beta - gamma)
          a = c1 (u_prev - u + dt v_prev + c2 a_prev)
          v = v_prev + dt ((1 - gamma) a_prev + gamma a)
  advance_disp + update_vel_acc (S/node_meth.rs:163-169) before the next step,
  which makes the next residual use a(u_n) = c1 (dt v_n + c2 a_n).
  time += dt after the step: the load of the step that produces u_{n+1} is the
  one at t_n.

first_step="reference" reproduces the reference literally: its first step
uses the node acceleration as initialised (a0), because update_vel_acc has
not run yet; first_step="predictor" uses a(u_0) like every later step (the
device path).  From rest (u0 = v0 = a0 = 0) the two coincide.

newmark_adjoint restates the backward sweep of Model::get_obj_gradient for a
dynamic elastic job with Rayleigh damping (S/model_analysis.rs:1563-1631):
for n = N..1, augmentd_ld_u (:936-1063, the M and D terms of :1033-1056)
adds step n+1's coupling to step n,
solve_for_adjoint (:1089-1100) forms the velocity and acceleration adjoints
and solves K_eff lam_n = dL/du_n, and the DV loop accumulates
-lam_n^T dR_n/dd with the state of step n (d_relasticd_d, :1379-1457):
dR/dE_e = dK/dE_e (u_n + ray_damp_ck v_n).  The
objective is a linear functional of the displacement history,
L = sum_n w_n q^T u_n (displacement terms, S/objective_meth.rs:618-663, with
exponent 1).

Pinning (tests/test_dynamics_oracle.py): single-DOF free vibration against
the closed form u_n = cos(n phi), tan(phi/2) = omega dt / 2 of the average
acceleration rule; the damped single DOF against the analytic decaying
oscillation; energy conservation of a free brick8 mesh, and with Rayleigh
damping the exact discrete energy balance E_{n+1} - E_n = -dt vbar^T D vbar
of the trapezoidal rule; the two first-step modes agree from rest;
newmark_adjoint against central finite differences of the forward run, with
and without damping (penalty scale held fixed, as the reference does not
differentiate it).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def newmark(K, M, f_shape, load_fac, dt, beta=0.25, gamma=0.5, u0=None, v0=None, a0=None, pen_dofs=(),
            sf=None, watch=(), first_step="predictor", on_step=None, damp_cm=0.0, damp_ck=0.0, damp_mat=None):
    """Run len(load_fac) steps.  K, M: FP64 sparse (no penalty); pen_dofs: DOFs
    clamped by a penalty sf (default 1e5 max|K_eff|, the reference scale
    taken on the first dynamic assembly); damp_cm, damp_ck: Rayleigh
    coefficients (job keys rayDampCM, rayDampCK); damp_mat: the assembled
    material damping matrix (sum of the elements' B^T D_m B integrals,
    get_rud_dfd0's material term), added to D.  on_step(n, u, v, a, Kp) is
    called after every step.  Returns (hist [n x len(watch)], u, v, a)."""
    K = sp.csr_matrix(K)
    M = sp.csr_matrix(M)
    n = K.shape[0]
    c1 = 1.0 / (dt * dt * (beta - gamma))
    c2 = dt * dt * (0.5 + beta - gamma)
    c_m = -c1
    D = (damp_cm * M + damp_ck * K).tocsr()
    if damp_mat is not None:
        D = (D + sp.csr_matrix(damp_mat)).tocsr()
    keff = (K + c_m * M + (c_m * dt * gamma) * D).tocsr()
    pen = np.zeros(n)
    if len(pen_dofs):
        if sf is None:
            sf = 1e5 * abs(keff).max()
        np.add.at(pen, np.asarray(pen_dofs, dtype=np.int64), sf)
    Kp = (K + sp.diags(pen)).tocsr()
    solve = spla.splu((keff + sp.diags(pen)).tocsc()).solve
    u = np.zeros(n) if u0 is None else np.array(u0, dtype=np.float64)
    v = np.zeros(n) if v0 is None else np.array(v0, dtype=np.float64)
    a = np.zeros(n) if a0 is None else np.array(a0, dtype=np.float64)
    watch = np.asarray(watch, dtype=np.int64)
    hist = np.zeros((len(load_fac), watch.size))
    for step, fac in enumerate(load_fac):
        if step == 0 and first_step == "reference":
            acc, vel = a, v
        else:
            acc = c1 * (dt * v + c2 * a)
            vel = v + dt * ((1.0 - gamma) * a + gamma * acc)
        rhs = fac * f_shape - Kp @ u - M @ acc - D @ vel
        du = solve(rhs)
        un = u + du
        an = c1 * (u - un + dt * v + c2 * a)
        v = v + dt * ((1.0 - gamma) * a + gamma * an)
        u, a = un, an
        hist[step] = u[watch]
        if on_step is not None:
            on_step(step, u, v, a, Kp)
    return hist, u, v, a


def newmark_adjoint(K, M, f_shape, load_fac, dt, beta, gamma, q, w, elem_sens, pen_dofs=(), sf=None,
                    damp_cm=0.0, damp_ck=0.0):
    """Forward Newmark run from rest, then the reference's backward adjoint
    sweep.  elem_sens(u, lam) -> per-element -lam^T (dK_e/dE_e) u.  Returns
    (L, dL/dE) with L = sum_n w[n] q.u_n."""
    K = sp.csr_matrix(K)
    M = sp.csr_matrix(M)
    n = K.shape[0]
    c1 = 1.0 / (dt * dt * (beta - gamma))
    c2 = dt * dt * (0.5 + beta - gamma)
    D = (damp_cm * M + damp_ck * K).tocsr()
    keff = (K - c1 * M - (c1 * dt * gamma) * D).tocsr()
    pen = np.zeros(n)
    if len(pen_dofs):
        if sf is None:
            sf = 1e5 * abs(keff).max()
        np.add.at(pen, np.asarray(pen_dofs, dtype=np.int64), sf)
    us, vs = [], []

    def keep(k, u, v, a, Kp):
        us.append(u.copy())
        vs.append(v.copy())

    newmark(K, M, f_shape, load_fac, dt, beta, gamma, pen_dofs=pen_dofs, sf=sf, on_step=keep, damp_cm=damp_cm,
            damp_ck=damp_ck)
    # partial derivatives of (a, v) of step n+1 with respect to (u, v, a) of
    # step n (augmentd_ld_u's d_ad_* and d_vd_*)
    d_vd_up = dt * gamma * c1
    d_vd_vp = 1.0 + dt * gamma * c1 * dt
    d_vd_ap = dt * (1.0 - gamma) + dt * gamma * c1 * c2
    q = np.asarray(q, dtype=np.float64)
    L = float(sum(w[k] * (q @ us[k]) for k in range(len(us))))
    solve = spla.splu((keff + sp.diags(pen)).tocsc()).solve
    lam = None
    a_adj = np.zeros(n)
    v_adj = np.zeros(n)
    dL = None
    for k in reversed(range(len(us))):
        gu = w[k] * q
        ga = np.zeros(n)
        gv = np.zeros(n)
        if lam is not None:
            t = a_adj - M @ lam
            dlam = D @ lam
            gu = gu + c1 * t - d_vd_up * dlam
            ga = c1 * c2 * t + dt * (1.0 - gamma) * v_adj - d_vd_ap * dlam
            gv = c1 * dt * t + v_adj - d_vd_vp * dlam
        v_adj = gv
        a_adj = ga + dt * gamma * v_adj
        lam = solve(gu - c1 * a_adj)
        s = elem_sens(us[k] + damp_ck * vs[k], lam)
        dL = s if dL is None else dL + s
    return L, dL
