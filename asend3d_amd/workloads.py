"""Measured workloads besides the headline static solve (bench.py):

  batched  -- 64 brick8 cantilever variants (10k-200k elements, aspect
              5:1:1, lognormal moduli), 8 load cases + 8 adjoint right-hand
              sides each: lockstep multi-RHS PCG over the whole batch against
              solving one system at a time with the same mixed-precision
              Jacobi PCG, and every system with the multigrid-preconditioned
              mixed solver.
  newmark  -- Newmark time stepping of a cantilever with a half-sine tip
              pulse, each step replayed as a hipGraph.
  adjoint  -- the newmark workload followed by the transient adjoint of the
              tip deflection summed over the history, with dL/dE for every
              element (one K_eff solve per step backwards).
  modal    -- the 10 lowest natural frequencies of the 1M-DOF heterogeneous
              hex8 block (200x40x40, lumped mass), shift-invert Lanczos whose
              operator applications are multigrid mixed-precision solves.
  shell    -- Newmark time stepping of a clamped shell4 plate (1 x 1 m,
              256 x 256 elements, t = 5 mm, aluminium) under a 1 ms half-sine
              10 kN centre load, dt = 1e-5 s, 1000 steps, hipGraph steps.
  nonlinear -- geometrically nonlinear static solve of the 1M-DOF
              heterogeneous hex8 block (200x40x40, lognormal E) under a 90 MPa
              tip traction (tip deflection ~10% of the length): Newton under
              the 5-step quadratic load ramp, each tangent solved by multigrid
              FP32 FCG with FP64 refreshes; then the same with inexact
              Newton solves (Eisenstat-Walker forcing).
  general  -- assembly of the 200x40x40 heterogeneous block on a distorted
              mesh (per-element integration and row gather).
  cycle    -- three multigrid mixed solves of the 200x40x40 block with graphs
              off (the timed step's solve kernels, for counter runs).
  reference -- the reference's own solvers on the device for the 200x40x40
              heterogeneous block: conj_grad_sparse preconditioned by the
              block LowerTriMat (blocks of 6 x solverBlockDim DOFs) to a 1e-8
              relative residual, and the direct solver (one LDL^T factor of
              the whole operator, numbering without reordering blocks).

    python -m asend3d_amd.workloads batched [--variants 64] [--sample 2]
    python -m asend3d_amd.workloads newmark [--nx 100 --ny 20 --nz 20 --steps 1000]
    python -m asend3d_amd.workloads shell [--nx 256 --ny 256 --steps 1000]
    python -m asend3d_amd.workloads adjoint [--nx 100 --ny 20 --nz 20 --steps 1000]
    python -m asend3d_amd.workloads modal [--modes 10]
    python -m asend3d_amd.workloads nonlinear [--nx 200 --ny 40 --nz 40 --load-steps 5]
    python -m asend3d_amd.workloads general [--nx 200 --ny 40 --nz 40]
    python -m asend3d_amd.workloads reference [--nx 200 --ny 40 --nz 40] [--no-direct]

Each prints one JSON line.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import api, problems


def log(msg):
    print(f"[workloads] {msg}", file=sys.stderr, flush=True)


def batched(n_variants=64, sample=2, seed=7, tol=1e-8):
    t0 = time.perf_counter()
    vs = problems.batch_variants(n_variants, seed=seed,
                                 progress=lambda v, p: (v % 16 == 15) and log(f"prepared {v + 1} variants"))
    n_el = [v.prob.mesh.n_el for v in vs]
    log(f"{n_variants} variants, {sum(n_el)} elements ({min(n_el)}..{max(n_el)}), "
        f"{sum(v.prob.n_dof for v in vs)} DOF, host prep {time.perf_counter() - t0:.1f} s")
    ctxs = []
    for k, v in enumerate(vs):
        c = api.Context(0)
        problems.upload(c, v.prob)
        c.assemble()
        ctxs.append(c)
        if k % 16 == 15:
            log(f"assembled {k + 1} variants")
    b = api.Batch(ctxs)
    rhs = [v.rhs for v in vs]
    xs, st = b.solve(rhs, tol=tol, inner_tol=1e-3, check_every=16, verbose=True)
    n_sys = sum(r.shape[0] for r in rhs)
    log(f"batched: {st.device_ms / 1e3:.3f} s for {n_sys} systems, {st.lockstep_iterations} lockstep iterations, "
        f"max rel residual {st.max_rel_residual:.2e}")
    # one system at a time, same algorithm (FP32 Jacobi PCG inside FP64 refinement), on a sample
    o = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_JACOBI, tol=tol, inner_tol=1e-3, max_it=2000000, max_outer=60,
                         check_every=16)
    single_ms, n_single, worst = 0.0, 0, 0.0
    for k, (c, v, x) in enumerate(zip(ctxs, vs, xs)):
        for j in range(min(sample, v.rhs.shape[0])):
            u, s = c.solve(v.rhs[j], o)
            single_ms += s.device_ms
            n_single += 1
            worst = max(worst, float(np.abs(x[j] - u).max() / max(np.abs(u).max(), 1e-300)))
        log(f"single solves: variant {k + 1}/{len(vs)}, {single_ms / 1e3:.1f} s so far")
    per_single = single_ms / n_single
    est_single_s = per_single * n_sys / 1e3
    # every system one at a time with the multigrid preconditioner (the
    # variants are structured blocks), hierarchy built once per variant
    om = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=tol, inner_tol=1e-2, max_it=2000000, max_outer=60,
                          check_every=2)
    mg_ms, mg_its, mg_worst_res, mg_worst_diff = 0.0, 0, 0.0, 0.0
    for k, (c, v, x) in enumerate(zip(ctxs, vs, xs)):
        m = v.prob.mesh
        c.set_grid(m.nx, m.ny, m.nz)
        for j in range(v.rhs.shape[0]):
            u, s = c.solve(v.rhs[j], om)
            mg_ms += s.device_ms
            mg_its += s.iterations
            mg_worst_res = max(mg_worst_res, s.rel_residual)
            mg_worst_diff = max(mg_worst_diff, float(np.abs(x[j] - u).max() / max(np.abs(u).max(), 1e-300)))
        if k % 16 == 15:
            log(f"multigrid solves: variant {k + 1}/{len(vs)}, {mg_ms / 1e3:.2f} s so far")
    b.close()
    for c in ctxs:
        c.close()
    return {"workload": "batched", "variants": n_variants, "systems": n_sys, "elements_total": int(sum(n_el)),
            "dof_total": int(sum(v.prob.n_dof for v in vs)), "batched_s": st.device_ms / 1e3,
            "lockstep_iterations": int(st.lockstep_iterations), "outer_iterations": int(st.outer_iterations),
            "max_rel_residual": st.max_rel_residual, "one_at_a_time_s_estimate": est_single_s,
            "one_at_a_time_sample": f"{n_single} systems ({sample} per variant), {single_ms / 1e3:.2f} s, scaled "
                                    f"by {n_sys}/{n_single}",
            "speedup": est_single_s / (st.device_ms / 1e3), "max_rel_diff_vs_single": worst,
            "multigrid_per_system": {"s": mg_ms / 1e3, "iterations": int(mg_its), "max_rel_residual": mg_worst_res,
                                     "max_rel_diff_vs_batched": mg_worst_diff,
                                     "speedup_vs_batched": (st.device_ms / 1e3) / (mg_ms / 1e3)},
            "tol": tol}


def newmark(nx=100, ny=20, nz=20, steps=1000, dt=1e-5, graphs=True, tol=1e-8, max_it=400, solver="auto", ru_factor=1e-3):
    from .problems import cantilever
    p = cantilever(nx, ny, nz, lx=1.0, ly=0.2, lz=0.2, E=70e9, nu=0.33, traction=(0.0, 0.0, -1.0e6))
    ctx = api.Context(0)
    problems.upload(ctx, p, rho=np.full(p.mesh.n_el, 2700.0))
    ctx.set_graphs(graphs)
    ctx.newmark_setup(dt, 0.25, 0.5)
    ctx.newmark_set_solver({"mixed": api.NM_MIXED, "fp64": api.NM_FP64, "auto": api.NM_AUTO}[solver], ru_factor)
    n_pulse = int(round(1e-3 / dt))
    fac = np.where(np.arange(steps) < n_pulse, np.sin(np.pi * np.arange(steps) / n_pulse), 0.0)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    tip = p.mesh.nodes_at_x(nx)
    watch = di[tip, 2]
    ctx.newmark_run(min(steps, 20), p.f, fac, watch=watch, tol=tol, max_it_step=max_it)  # warm-up
    hist, u, v, a, st = ctx.newmark_run(steps, p.f, fac, watch=watch, tol=tol, max_it_step=max_it)
    ctx.close()
    return {"workload": "newmark", "mesh": f"{nx}x{ny}x{nz} brick8", "dof": int(p.n_dof), "steps": steps,
            "dt": dt, "device_s": st.device_ms / 1e3, "us_per_step": st.device_ms * 1e3 / steps,
            "dof_steps_per_s": p.n_dof * steps / (st.device_ms / 1e3), "pcg_iterations": int(st.total_iterations),
            "max_step_iterations": int(st.max_step_iterations), "unconverged_steps": int(st.unconverged_steps),
            "graph": bool(st.graph_used), "step_solver": "mixed" if st.step_solver == api.NM_MIXED else "fp64", "ru_factor": ru_factor if st.step_solver == api.NM_MIXED else None, "step_tol": tol, "max_it_step": max_it,
            "tip_w_min": float(hist.mean(axis=1).min())}


def adjoint(nx=100, ny=20, nz=20, steps=1000, dt=1e-5, graphs=True, tol=1e-8, max_it=400, solver="auto", ru_factor=1e-3):
    from .problems import cantilever
    p = cantilever(nx, ny, nz, lx=1.0, ly=0.2, lz=0.2, E=70e9, nu=0.33, traction=(0.0, 0.0, -1.0e6))
    ctx = api.Context(0)
    problems.upload(ctx, p, rho=np.full(p.mesh.n_el, 2700.0))
    ctx.set_graphs(graphs)
    ctx.newmark_setup(dt, 0.25, 0.5)
    ctx.newmark_set_solver({"mixed": api.NM_MIXED, "fp64": api.NM_FP64, "auto": api.NM_AUTO}[solver], ru_factor)
    n_pulse = int(round(1e-3 / dt))
    fac = np.where(np.arange(steps) < n_pulse, np.sin(np.pi * np.arange(steps) / n_pulse), 0.0)
    di = np.asarray(p.dof_index).reshape(-1, 6)
    q = np.zeros(p.n_dof)
    q[di[p.mesh.nodes_at_x(nx), 2]] = 1.0
    w = np.full(steps, dt)
    log(f"adjoint {nx}x{ny}x{nz}: {p.n_dof} DOF, {steps} steps, history {steps * p.n_dof * 8 / 2**30:.2f} GiB")
    L, dL, fst, ast = ctx.newmark_adjoint(steps, p.f, fac, q, w, tol=tol, max_it_step=max_it)
    ctx.close()
    return {"workload": "newmark-adjoint", "mesh": f"{nx}x{ny}x{nz} brick8", "dof": int(p.n_dof), "steps": steps,
            "dt": dt, "objective": "sum_n dt * (sum of tip w)_n", "L": L,
            "forward_s": fst.device_ms / 1e3, "adjoint_s": ast.device_ms / 1e3, "forward_step_solver": "mixed" if fst.step_solver == api.NM_MIXED else "fp64",
            "forward_pcg_iterations": int(fst.total_iterations), "adjoint_pcg_iterations": int(ast.total_iterations),
            "unconverged_steps": int(fst.unconverged_steps + ast.unconverged_steps),
            "dLdE_abs_max": float(np.abs(dL).max()), "dLdE_sum_E": float(np.sum(dL * p.E)), "step_tol": tol}


def modal(n_modes=10, nx=200, ny=40, nz=40):
    p = problems.heterogeneous(nx, ny, nz)
    ctx = api.Context(0)
    problems.upload(ctx, p, rho=np.full(p.mesh.n_el, 2.7e-9))
    ctx.set_grid(nx, ny, nz)
    o = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=1e-8, inner_tol=1e-2, max_it=200000, max_outer=60,
                         check_every=4)
    log(f"modal {nx}x{ny}x{nz}: {p.n_dof} DOF, {n_modes} modes")
    ctx.modal(2, opts=o, vectors=False)  # warm-up (hierarchy, graphs)
    vals, fr, _, st = ctx.modal(n_modes, opts=o, vectors=True)
    ctx.close()
    return {"workload": "modal", "mesh": f"{nx}x{ny}x{nz} brick8, lognormal E", "dof": int(p.n_dof),
            "modes": n_modes, "freqs_hz": [float(f) for f in fr], "device_s": st.device_ms / 1e3,
            "lanczos_steps": int(st.lanczos_steps), "solve_iterations": int(st.solve_iterations),
            "max_ritz_residual": float(st.max_ritz_residual), "unconverged_solves": int(st.unconverged_solves),
            "solve_tol": 1e-8}


def nonlinear(nx=200, ny=40, nz=40, load_steps=5, traction=90.0, du_tol=1e-6, tol=1e-10, graphs=True):
    # mm-N-MPa: du_tol 1e-6 mm is ~1e-8 of the tip deflection
    p = problems.heterogeneous(nx, ny, nz, traction=(0.0, 0.0, -traction))
    ctx = api.Context(0)
    problems.upload(ctx, p)
    ctx.set_grid(nx, ny, nz)
    ctx.set_graphs(graphs)
    o = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=tol, inner_tol=1e-2, max_it=200000, max_outer=200,
                         check_every=4)
    log(f"nonlinear {nx}x{ny}x{nz}: {p.n_dof} DOF, {load_steps} load steps, traction {traction} MPa")
    ctx.assemble()
    lin, lst = ctx.solve(p.f, o)
    t0 = time.perf_counter()
    u, st = ctx.static_nonlinear(p.f, o, load_steps=load_steps, du_tol=du_tol)
    wall = time.perf_counter() - t0
    tip = p.dof_index[p.mesh.nodes_at_x(nx), 2]
    out = {"workload": "nonlinear", "mesh": f"{nx}x{ny}x{nz} brick8, lognormal E", "dof": int(p.n_dof),
           "traction_mpa": traction, "load_steps": load_steps, "du_tol_mm": du_tol, "solve_tol": tol,
           "tip_w_mm": float(u[tip].mean()), "tip_w_linear_mm": float(lin[tip].mean()),
           "device_s": st.device_ms / 1e3, "wall_s": wall, "linear_solve_s": lst.device_ms / 1e3,
           **{k: v for k, v in st.as_dict().items() if k != "device_ms"}}
    out["ms_per_newton"] = st.device_ms / max(1, st.newton_iterations)
    # inexact Newton (Eisenstat-Walker forcing, first solves to 0.1)
    ui, sti = ctx.static_nonlinear(p.f, o, load_steps=load_steps, du_tol=du_tol, forcing=0.1)
    out["inexact_newton"] = {"forcing": 0.1, "device_s": sti.device_ms / 1e3,
                             "newton_iterations": int(sti.newton_iterations),
                             "linear_iterations": int(sti.linear_iterations),
                             "unconverged_steps": int(sti.unconverged_steps),
                             "max_abs_du_vs_exact_mm": float(np.abs(ui - u).max())}
    ctx.close()
    return out


def general(nx=200, ny=40, nz=40, reps=10, seed=6):
    """Assembly of the heterogeneous block on a distorted mesh (interior
    nodes moved by up to 0.2 of an element edge, per-element Poisson ratio):
    every element integrated once, then the row gather into the operator."""
    p = problems.distorted(problems.heterogeneous(nx, ny, nz), seed)
    ctx = api.Context(0)
    problems.upload(ctx, p)
    assert not ctx.uniform_elements()
    ctx.assemble()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.assemble()
    ctx.sync()
    wall = (time.perf_counter() - t0) / reps
    ctx.close()
    return {"workload": "general", "mesh": f"{nx}x{ny}x{nz} brick8, distorted, lognormal E, per-element nu",
            "elements": int(p.mesh.n_el), "dof": int(p.n_dof), "assemblies": reps,
            "wall_ms_per_assembly": wall * 1e3}


def shell(nx=256, ny=256, steps=1000, dt=1e-5, graphs=True, tol=1e-8, max_it=2000, load=1.0e4, solver="auto", ru_factor=1e-3):
    p = problems.shell_plate(nx, ny, 1.0, 1.0, 0.005, 70e9, 0.33, 2700.0)
    ctx = api.Context(0)
    problems.upload_shell(ctx, p)
    ctx.set_graphs(graphs)
    t0 = time.perf_counter()
    ctx.newmark_setup(dt, 0.25, 0.5)
    ctx.newmark_set_solver({"mixed": api.NM_MIXED, "fp64": api.NM_FP64, "auto": api.NM_AUTO}[solver], ru_factor)
    log(f"shell {nx}x{ny}: {p.n_dof} DOF, assembly of M and K + c_m M {time.perf_counter() - t0:.2f} s")
    di = np.asarray(p.dof_index).reshape(-1, 6)
    centre = p.mesh.node_id(nx // 2, ny // 2)
    f = np.zeros(p.n_dof)
    f[di[centre, 2]] = load
    n_pulse = int(round(1e-3 / dt))
    fac = np.where(np.arange(steps) < n_pulse, np.sin(np.pi * np.arange(steps) / n_pulse), 0.0)
    watch = di[centre, :3]
    ctx.newmark_run(min(steps, 20), f, fac, watch=watch, tol=tol, max_it_step=max_it)  # warm-up
    log("warm-up done")
    hist, u, v, a, st = ctx.newmark_run(steps, f, fac, watch=watch, tol=tol, max_it_step=max_it)
    ctx.close()
    w = hist[:, 2]
    return {"workload": "shell", "mesh": f"{nx}x{ny} shell4 plate, edges clamped", "dof": int(p.n_dof),
            "steps": steps, "dt": dt, "load": f"{load} N half-sine over {n_pulse} steps at the centre",
            "device_s": st.device_ms / 1e3, "us_per_step": st.device_ms * 1e3 / steps,
            "dof_steps_per_s": p.n_dof * steps / (st.device_ms / 1e3), "pcg_iterations": int(st.total_iterations),
            "max_step_iterations": int(st.max_step_iterations), "unconverged_steps": int(st.unconverged_steps),
            "graph": bool(st.graph_used), "step_solver": "mixed" if st.step_solver == api.NM_MIXED else "fp64", "ru_factor": ru_factor if st.step_solver == api.NM_MIXED else None, "step_tol": tol, "max_it_step": max_it,
            "centre_w_min": float(w.min()), "centre_w_max": float(w.max()),
            "centre_w_min_step": int(np.argmin(w))}


def cycle(nx=200, ny=40, nz=40, solves=3, tol=1e-8):
    """Multigrid mixed solves of the heterogeneous block with graphs off: the
    solve kernels of the bench's timed step without the rest of the bench,
    for hardware-counter runs."""
    p = problems.heterogeneous(nx, ny, nz)
    ctx = api.Context(0)
    ctx.set_graphs(False)
    problems.upload(ctx, p)
    ctx.assemble()
    ctx.set_grid(nx, ny, nz)
    o = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=tol, inner_tol=1e-2, check_every=2)
    runs = []
    for _ in range(solves):
        _, st = ctx.solve(p.f, o)
        runs.append({"ms": st.device_ms, "iterations": int(st.iterations), "rel_residual": st.rel_residual})
    ctx.close()
    return {"workload": "cycle", "mesh": [nx, ny, nz], "solves": runs}


def nonlinear_brickim(nx=100, ny=20, nz=20, load_steps=2, traction=60.0, du_tol=1e-6, tol=1e-10, graphs=True):
    """Geometrically nonlinear brickIM cantilever (lognormal E, mm-N-MPa):
    condensed Newton tangents and bubble updates on the device, each Newton
    system solved by FP64 Jacobi PCG (the multigrid hierarchy is built for
    brick8)."""
    import dataclasses
    p = dataclasses.replace(problems.heterogeneous(nx, ny, nz, traction=(0.0, 0.0, -traction)), etype=81)
    ctx = api.Context(0)
    problems.upload(ctx, p)
    ctx.set_graphs(graphs)
    o = api.Context.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=tol, max_it=2000000, check_every=16)
    log(f"nonlinear brickIM {nx}x{ny}x{nz}: {p.n_dof} DOF, {load_steps} load steps, traction {traction} MPa")
    ctx.assemble()
    t0 = time.perf_counter()
    u, st = ctx.static_nonlinear(p.f, o, load_steps=load_steps, du_tol=du_tol)
    wall = time.perf_counter() - t0
    tip = p.dof_index[p.mesh.nodes_at_x(nx), 2]
    out = {"workload": "nonlinear_brickim", "mesh": f"{nx}x{ny}x{nz} brickIM, lognormal E", "dof": int(p.n_dof),
           "elements": int(p.mesh.n_el), "traction_mpa": traction, "load_steps": load_steps,
           "tip_w_mm": float(u[tip].mean()), "device_s": st.device_ms / 1e3, "wall_s": wall,
           **{k: v for k, v in st.as_dict().items() if k != "device_ms"}}
    ctx.close()
    return out


def reference(nx=200, ny=40, nz=40, tol=1e-8, direct=True, graphs=True):
    out = {"workload": "reference", "mesh": [nx, ny, nz], "tol": tol}
    p = problems.heterogeneous(nx, ny, nz)
    ctx = api.Context(0)
    ctx.set_graphs(graphs)
    problems.upload(ctx, p)
    ctx.assemble()
    log(f"{p.n_dof} DOF, solverBlockDim {p.block_dim}")
    o_fast = api.Context.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=tol, inner_tol=1e-2)
    ctx.set_grid(nx, ny, nz)
    u_fast, st_fast = ctx.solve(p.f, o_fast)
    t0 = time.perf_counter()
    info = ctx.ldl_factor(6 * p.block_dim)
    t_fac = time.perf_counter() - t0
    log(f"block factor: {info}")
    o = api.Context.opts(api.SOLVE_PCG_FP64, api.PC_LDL, tol=tol, max_it=100000, check_every=16)
    u, st = ctx.solve(p.f, o)
    log(f"block LDL^T PCG: {st.iterations} iterations, {st.device_ms / 1e3:.3f} s, residual {st.rel_residual:.2e}")
    out["block"] = {"ldl": info, "factor_call_s": t_fac, "iterations": int(st.iterations),
                    "solve_s": st.device_ms / 1e3, "rel_residual": st.rel_residual,
                    "max_rel_diff_vs_multigrid": float(np.abs(u - u_fast).max() / np.abs(u_fast).max()),
                    "multigrid_mixed_s": st_fast.device_ms / 1e3}
    ctx.ldl_release()
    ctx.close()
    if direct:
        pd = problems.heterogeneous(nx, ny, nz, block_dim=2000000000)
        c2 = api.Context(0)
        problems.upload(c2, pd)
        c2.assemble()
        t0 = time.perf_counter()
        info2 = c2.ldl_factor(0)
        t_fac2 = time.perf_counter() - t0
        log(f"direct factor: {info2}")
        u2, st2 = c2.solve(pd.f, api.Context.opts(api.SOLVE_DIRECT, api.PC_JACOBI))
        log(f"direct solve: {st2.device_ms / 1e3:.3f} s, residual {st2.rel_residual:.2e}")
        di_b = np.asarray(p.dof_index).reshape(-1, 6)[:, :3]
        di_d = np.asarray(pd.dof_index).reshape(-1, 6)[:, :3]
        out["direct"] = {"ldl": info2, "factor_call_s": t_fac2, "solve_s": st2.device_ms / 1e3,
                         "rel_residual": st2.rel_residual,
                         "max_rel_diff_vs_block_pcg": float(np.abs(u2[di_d] - u[di_b]).max() / np.abs(u[di_b]).max())}
        c2.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["batched", "newmark", "shell", "adjoint", "modal", "nonlinear",
                                      "nonlinear_brickim", "general", "reference", "cycle"])
    ap.add_argument("--load-steps", type=int, default=5)
    ap.add_argument("--modes", type=int, default=10)
    ap.add_argument("--variants", type=int, default=64)
    ap.add_argument("--sample", type=int, default=2)
    ap.add_argument("--nx", type=int, default=None, help="default 100 (newmark), 256 (shell)")
    ap.add_argument("--ny", type=int, default=None, help="default 20 (newmark), 256 (shell)")
    ap.add_argument("--nz", type=int, default=None, help="default 20 (newmark), 40 (nonlinear)")
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--no-graphs", action="store_true")
    ap.add_argument("--no-direct", action="store_true", help="reference: skip the direct solver")
    ap.add_argument("--solver", choices=["auto", "mixed", "fp64"], default="auto", help="Newmark step solver")
    ap.add_argument("--ru-factor", type=float, default=1e-3,
                    help="mixed step solver: residual reduction between FP64 reliable updates")
    a = ap.parse_args()
    if a.which == "batched":
        out = batched(a.variants, a.sample)
    elif a.which == "modal":
        out = modal(a.modes)
    elif a.which == "cycle":
        out = cycle(a.nx or 200, a.ny or 40, a.nz or 40)
    elif a.which == "reference":
        out = reference(a.nx or 200, a.ny or 40, a.nz or 40, direct=not a.no_direct, graphs=not a.no_graphs)
    elif a.which == "general":
        out = general(a.nx or 200, a.ny or 40, a.nz or 40)
    elif a.which == "nonlinear_brickim":
        out = nonlinear_brickim(a.nx or 100, a.ny or 20, a.nz or 20, load_steps=2, graphs=not a.no_graphs)
    elif a.which == "nonlinear":
        out = nonlinear(a.nx or 200, a.ny or 40, a.nz or 40, a.load_steps, graphs=not a.no_graphs)
    elif a.which == "adjoint":
        out = adjoint(a.nx or 100, a.ny or 20, a.nz or 20, a.steps, graphs=not a.no_graphs, solver=a.solver, ru_factor=a.ru_factor)
    elif a.which == "shell":
        out = shell(a.nx or 256, a.ny or 256, a.steps, graphs=not a.no_graphs, solver=a.solver, ru_factor=a.ru_factor)
    else:
        out = newmark(a.nx or 100, a.ny or 20, a.nz or 20, a.steps, graphs=not a.no_graphs, solver=a.solver, ru_factor=a.ru_factor)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
