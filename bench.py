"""Benchmark of the MI355X solve path on the 1M-DOF heterogeneous hex8 static
solve + compliance adjoint (200x40x40 brick8, lognormal E, seeded).

One step = one pass of the hot path on resident data: element integration and
coloured assembly, penalty scale, solve to ||f - K u|| / ||f|| <= 1e-8 (FP64),
compliance, adjoint (lambda = u for compliance) and dJ/dE for all elements.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nx 200 --ny 40 --nz 40]

Prints one JSON line.  `value` is the mean time-to-solution of the learned +
mixed-precision path: FP32 flexible PCG inside FP64 refinement, preconditioned
by a multigrid V-cycle whose finest Chebyshev smoother is the learned factor
L L^T.  The network is first trained on the GPU (8 lognormal 110x22x22 meshes,
250 steps each = 2000 solver-generated samples; time reported as
`train_s`, outside `value`), then inferred inside every timed step.  The
autotuner times the pipeline with and without the network and keeps the
faster; when that is the one without it, the network steps aside
(`fallback_used` 2, `ablation.learned_active` false) and the step runs the
same multigrid with its block-Jacobi smoother.  The plain
FP64 Jacobi PCG pipeline on the same GPU is timed alongside (`fp64_jacobi_s`,
`speedup_vs_fp64_jacobi`), and so is the same multigrid without the network
(`mg_no_network_s`).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(key, name="sell3_spmv_c2_pmc.md"):
    """Calibrated HBM bytes per launch of a kernel on the 200x40x40 operator,
    from a committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE scaled by
    a copy kernel of known size); None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return int(line.split(":")[1])
    return None


# Iterations of the reference's conj_grad_sparse with its block-banded LDL^T
# preconditioner (blocks of 6 x 582 = 3,492 DOFs) to a 1e-8 relative residual
# on the 200x40x40 block: counted by running that algorithm on the device over
# the full workload (reference_algorithm_on_device; 1,203).  Without that
# run (--no-reference) the count is extrapolated from the oracle's counts on
# 25- and 50-element-long slices of the same cross-section and block size
# (247 and 395 iterations) by the power law through those two points,
# its ~ nx^0.68, which gives 1,010 at nx = 200.
REF_ITS_25, REF_ITS_50 = 247, 395


def reference_iterations(nx):
    return REF_ITS_25 * (nx / 25.0) ** (math.log(REF_ITS_50 / REF_ITS_25) / math.log(2.0))


def reference_algorithm_on_device(ctx, prob, tol):
    """The reference's own solver run on the device over the full workload:
    conj_grad_sparse preconditioned by the block LowerTriMat (blocks of
    6 x solverBlockDim DOFs), stopped at the metric's relative residual.  Its
    iteration count is the one the reference's serial run would take (the
    device restatement takes the oracle's iteration counts on small meshes)."""
    from asend3d_amd import api
    info = ctx.ldl_factor(6 * prob.block_dim)
    _, st = ctx.solve(prob.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_LDL, tol=tol, max_it=100000, check_every=16))
    ctx.ldl_release()
    return {"iterations": int(st.iterations), "factor_s": info["factor_ms"] / 1e3, "pcg_s": st.device_ms / 1e3,
            "rel_residual": st.rel_residual, "block_dofs": int(info["block_dofs"]),
            "factor_bytes": int(info["bytes"]), "max_bandwidth": int(info["max_bandwidth"])}


def cpu_baseline(prob, u, fp64_iters, nx, ny, nz, ref_its=None):
    """The reference's own algorithm, restated by the oracle and run serially
    as the reference runs, timed on bounded samples of the workload and
    extrapolated (each scaling stated in `sample`):
      - element loop and assembly on an 8-element-long slice, x element ratio;
      - block-banded LDL^T factor (blocks of 6 x solver_block_dim DOFs, the
        workload's size) on that slice, x block ratio;
      - conj_grad_sparse iteration (SpMV, penalty, block solves) on that slice,
        x DOF ratio, x the iteration count of the same preconditioned PCG run
        on the device over the full workload (ref_its; reference_iterations()
        when that run is skipped);
      - the design-variable loop (per-element modulus DVs) on the full
        workload: 300 DVs timed, x N_el / 300.
    The Jacobi PCG of the device baseline, restated on the same slice, is
    reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from asend3d_amd import api, problems
    sx = max(2, nx // 25)
    sl = problems.heterogeneous(sx, ny, nz, lx=1000.0 * sx / nx)
    block_dofs = 6 * api.block_dim_auto(prob.mesh.n_nodes)
    t0 = time.perf_counter()
    K = O.assemble_solid(8, sl.mesh.xyz, sl.mesh.conn, sl.E, sl.nu, sl.dof_index, sl.n_dof)
    t_asm = time.perf_counter() - t0
    sf = 1e5 * O.max_abs(K)
    t0 = time.perf_counter()
    ldl = O.LDL(K, sl.pen_dofs, sf, block_dofs)
    t_fac = time.perf_counter() - t0
    n_it = 8
    t0 = time.perf_counter()
    O.conj_grad_sparse(K, sl.pen_dofs, sf, ldl, sl.f, 0.0, n_it)
    t_it = (time.perf_counter() - t0) / n_it
    n_jac = 300
    t0 = time.perf_counter()
    O.jacobi_pcg(K, sl.pen_dofs, sf, sl.f, 1e-30, n_jac)
    t_jac = (time.perf_counter() - t0) / n_jac
    n_dv = 300
    t0 = time.perf_counter()
    O.dv_loop(8, prob.mesh.xyz, prob.mesh.conn, prob.E, prob.nu, prob.dof_index, prob.n_dof, u, u, prob.tip_elements,
              n_dv)
    t_dv = (time.perf_counter() - t0) / n_dv
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    r_el = prob.mesh.n_el / sl.mesh.n_el
    r_dof = prob.n_dof / sl.n_dof
    blocks_full = -(-prob.n_dof // block_dofs)
    blocks_sl = -(-sl.n_dof // block_dofs)
    its = ref_its if ref_its else reference_iterations(nx)
    its_note = (f"the count of the same block-LDL^T PCG run on the device over the full workload to the same "
                f"residual" if ref_its else
                f"power law through {REF_ITS_25} and {REF_ITS_50} measured on 25- and 50-long slices")
    est = t_asm * r_el + t_fac * blocks_full / blocks_sl + t_it * r_dof * its + t_dv * prob.mesh.n_el
    est_jac = t_asm * r_el + t_jac * r_dof * fp64_iters + t_dv * prob.mesh.n_el
    return {"value": est, "unit": "s", "cores": 1, "kind": "port",
            "sample": (f"reference algorithm restated by the oracle, serial: assembly {t_asm:.2f} s on a {sx}x{ny}x{nz} "
                       f"slice ({sl.n_dof} DOF) x{r_el:.0f}; block-banded LDL^T ({block_dofs}-DOF blocks) factor "
                       f"{t_fac:.2f} s for {blocks_sl} blocks x{blocks_full / blocks_sl:.1f}; conj_grad_sparse "
                       f"{t_it:.3f} s per iteration on the slice x{r_dof:.1f} DOF ratio x {its:.0f} iterations "
                       f"({its_note}); "
                       f"design-variable loop {t_dv * 1e3:.2f} ms per DV (300 DVs timed "
                       f"on the full mesh) x {prob.mesh.n_el} DVs; host {cpu_model}, {os.cpu_count()} hardware "
                       f"threads; with Jacobi PCG instead ({fp64_iters} iterations, {t_jac * 1e3:.1f} ms each on the "
                       f"slice): {est_jac:.0f} s"),
            "jacobi_variant_s": est_jac}


def train_network(ctx, n_meshes, steps):
    """GPU training of the learned preconditioner on lognormal 110x22x22 brick8
    meshes (53,240 elements each, sigma_ln cycling 0.6..1.15, i.e. moduli
    contrast up to ~1e3 at +-3 sigma), `steps` solver-generated samples per
    mesh."""
    from asend3d_amd import api, problems
    ctx.learned_init(7)
    t0 = time.perf_counter()
    first, last = [], []
    for m in range(n_meshes):
        q = problems.heterogeneous(110, 22, 22, seed=1000 + m, sigma_ln=(0.6, 0.8, 1.0, 1.15)[m % 4], lx=550.0,
                                   ly=110.0, lz=110.0)
        problems.upload(ctx, q)
        ctx.set_grid(110, 22, 22)
        ctx.assemble()
        L = ctx.learned_train(steps, lr=1e-3, seed=m + 1, mode=1)
        first.append(float(np.mean(L[0:30:3])))
        last.append(float(np.mean(L[-30::3])))
    ctx.sync()
    t = time.perf_counter() - t0
    log(f"trained the network on {n_meshes} meshes x {steps} steps in {t:.1f} s "
        f"(Gaussian-sample loss {first[0]:.3f} -> {last[-1]:.3f})")
    return {"meshes": n_meshes, "steps_per_mesh": steps, "samples": n_meshes * steps, "elements_per_mesh": 53240,
            "train_s": t, "loss_first": first[0], "loss_last": last[-1]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=200)
    ap.add_argument("--ny", type=int, default=40)
    ap.add_argument("--nz", type=int, default=40)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the device run of the reference's block-LDL^T PCG (iteration count extrapolated)")
    ap.add_argument("--fast-precond", default="learned", choices=["jacobi", "mg", "learned"])
    ap.add_argument("--train-meshes", type=int, default=8)
    ap.add_argument("--train-steps", type=int, default=250)
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the captured iteration blocks eagerly (kernel-trace profiling runs)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="assemble, then time only the SpMV and copy kernels (short run for counter profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding is a design "
              "note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    from asend3d_amd import api, problems

    t0 = time.perf_counter()
    prob = problems.heterogeneous(args.nx, args.ny, args.nz)
    log(f"prepared {args.nx}x{args.ny}x{args.nz}: {prob.n_dof} DOF, {prob.mesh.n_el} elements "
        f"({time.perf_counter() - t0:.1f} s host prep)")
    ctx = api.Context(0)
    if args.no_graphs:
        ctx.set_graphs(False)
    problems.upload(ctx, prob)
    if args.kernels_only:
        ctx.assemble()
        ms32 = ctx.bench_spmv(32, 20)
        ms64 = ctx.bench_spmv(64, 20)
        copy_bytes = 1 << 30
        msc = ctx.bench_copy(copy_bytes, 20)
        print(json.dumps({"spmv32_ms": ms32, "spmv64_ms": ms64, "copy_bytes": copy_bytes, "copy_ms": msc,
                          "copy_gbs": 2 * copy_bytes / (msc * 1e-3) / 1e9}), flush=True)
        ctx.close()
        return
    pc = {"jacobi": api.PC_JACOBI, "mg": api.PC_MG, "learned": api.PC_LEARNED}[args.fast_precond]
    train = None
    if pc == api.PC_LEARNED:
        train = train_network(ctx, args.train_meshes, args.train_steps)
        problems.upload(ctx, prob)
    if pc != api.PC_JACOBI:
        ctx.set_grid(args.nx, args.ny, args.nz)
    tuning = None
    if pc != api.PC_JACOBI:
        # project autotuner: SpMV workgroup geometry and the Chebyshev degree /
        # refresh factor / check interval of the solve that will run, with the
        # learned smoother when it is the fast path (setup, not timed)
        ctx.assemble()
        t0 = time.perf_counter()
        what = api.TUNE_SPMV | api.TUNE_SOLVER | (api.TUNE_LEARNED if pc == api.PC_LEARNED else 0)
        tn = ctx.autotune(prob.f, what)
        tuning = tn.as_dict()
        tuning["tune_s"] = time.perf_counter() - t0
        log(f"autotuned in {tuning['tune_s']:.1f} s: {tn.spmv_waves_per_block} waves/workgroup, Chebyshev degree "
            f"{tn.cheb_degree}, inner tol {tn.inner_tol:g}, check every {tn.check_every}"
            + (f", learned smoother {'kept' if tn.learned_active else 'stepped aside'}" if pc == api.PC_LEARNED else ""))
    fast = ctx.opts(api.SOLVE_IR_MIXED, pc, tol=args.tol, inner_tol=1e-4 if pc == api.PC_JACOBI else 0.0,
                    max_it=2000000, max_outer=40, check_every=32 if pc == api.PC_JACOBI else 0)
    base = ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=args.tol, max_it=2000000, check_every=32)

    # reference solution: FP64 PCG to 1e-11 (about the attainable FP64 residual here)
    u_ref, _, J_ref, st_ref = ctx.static_compliance(prob.f, ctx.opts(api.SOLVE_PCG_FP64, api.PC_JACOBI, tol=1e-11,
                                                                     max_it=4000000, check_every=64))
    log(f"FP64 reference to 1e-11: {st_ref.iterations} it, {st_ref.device_ms / 1e3:.2f} s")

    for _ in range(args.warmup):
        ctx.static_compliance(prob.f, fast)
    ctx.sync()
    times = []
    wall0 = time.perf_counter()
    for k in range(args.steps):
        u, s, J, st = ctx.static_compliance(prob.f, fast)
        times.append(st.device_ms / 1e3)
        log(f"step {k}: fast path {st.device_ms / 1e3:.3f} s, {st.iterations} inner it, "
            f"{st.outer_iterations} outer, relres {st.rel_residual:.2e}")
    ctx.sync()
    wall = time.perf_counter() - wall0
    t_fast = float(np.mean(times))
    err_u = float(np.abs(u - u_ref).max() / np.abs(u_ref).max())
    err_J = abs(J - J_ref) / abs(J_ref)

    ub, sb, Jb, stb = ctx.static_compliance(prob.f, base)
    t_base = stb.device_ms / 1e3
    log(f"FP64 Jacobi PCG: {t_base:.3f} s, {stb.iterations} it, relres {stb.rel_residual:.2e}")
    err_s = float(np.abs(s - sb).max() / np.abs(sb).max())
    ablation = None
    if pc == api.PC_LEARNED:
        # the same pipeline without the network (block-Jacobi Chebyshev
        # smoother) at the tuned settings, and each pipeline's best tuned solve
        mg_o = ctx.opts(api.SOLVE_IR_MIXED, api.PC_MG, tol=args.tol, inner_tol=0.0, max_it=2000000, max_outer=40,
                        check_every=0)
        ctx.static_compliance(prob.f, mg_o)
        _, _, _, stm = ctx.static_compliance(prob.f, mg_o)
        ablation = {"mg_no_network_s": stm.device_ms / 1e3, "mg_no_network_iterations": int(stm.iterations),
                    "learned_active": bool(tuning["learned_active"]),
                    "tuned_solve_ms_learned": tuning["solve_ms_learned"],
                    "tuned_solve_ms_no_network": tuning["solve_ms_no_network"]}
        log(f"multigrid without the network: {stm.device_ms / 1e3:.3f} s, {stm.iterations} it; tuned solves: "
            f"learned {tuning['solve_ms_learned']:.2f} ms, no network {tuning['solve_ms_no_network']:.2f} ms, "
            f"network {'kept' if tuning['learned_active'] else 'stepped aside'}")

    # roofline of the dominant kernel: FP32 block-row SpMV, HIP events on the context stream
    import ctypes
    nnz = ctypes.c_int64(0)
    api.lib().asend_get_matrix(ctx.h, None, None, None, ctypes.byref(nnz), 1)
    nnzb = nnz.value // 9
    n_rows = prob.n_dof // 3
    spmv32_bytes = nnzb * (9 * 4 + 4) + (n_rows + 1) * 4 + prob.n_dof * 4 * 2
    spmv64_bytes = nnzb * (9 * 8 + 4) + (n_rows + 1) * 4 + prob.n_dof * 8 * 2
    ms32 = ctx.bench_spmv(32, 100)
    ms64 = ctx.bench_spmv(64, 100)
    gbs32 = spmv32_bytes / (ms32 * 1e-3) / 1e9
    gbs64 = spmv64_bytes / (ms64 * 1e-3) / 1e9
    copy_bytes = 1 << 30
    ms_copy = ctx.bench_copy(copy_bytes, 20)
    copy_gbs = 2 * copy_bytes / (ms_copy * 1e-3) / 1e9
    log(f"SpMV fp32 {ms32 * 1e3:.1f} us ({gbs32:.0f} GB/s), fp64 {ms64 * 1e3:.1f} us ({gbs64:.0f} GB/s); "
        f"copy {copy_gbs:.0f} GB/s")
    # the timed step's leading kernel, timed inside the pipeline: the
    # multigrid cycle's FP16 row-scaled SpMV with residual epilogue (y = f - K x:
    # FP16 values and int32 columns once, x gathered, f, the FP32 row scale
    # and y once each), HIP events around each of its launches on the context
    # stream during one more step run with graphs off
    cyc = None
    if pc != api.PC_JACOBI:
        ctx.set_graphs(False)
        ctx.cycle_spmv_timing(True)
        ctx.static_compliance(prob.f, fast)
        n_cyc, ms_cyc = ctx.cycle_spmv_timing(False)
        ctx.set_graphs(True)
        if n_cyc:
            cyc_bytes = nnzb * (9 * 2 + 4) + (n_rows + 1) * 4 + prob.n_dof * 4 * 4
            cyc_ms = ms_cyc / n_cyc
            cyc = {"launches": n_cyc, "avg_ms": cyc_ms, "bytes_per_launch": cyc_bytes,
                   "achieved": cyc_bytes / (cyc_ms * 1e-3) / 1e9}
            log(f"cycle SpMV in the pipeline: {n_cyc} launches, {cyc_ms * 1e3:.2f} us each, "
                f"{cyc['achieved']:.0f} GB/s")

    c2 = (args.nx, args.ny, args.nz) == (200, 40, 40)
    cpu = None
    ref_alg = None
    if not args.no_cpu:
        if not args.no_reference:
            ref_alg = reference_algorithm_on_device(ctx, prob, args.tol)
            log(f"reference algorithm on the device: {ref_alg}")
        cpu = cpu_baseline(prob, u, stb.iterations, args.nx, args.ny, args.nz,
                           ref_its=ref_alg["iterations"] if ref_alg else None)
        log(f"cpu baseline estimate {cpu['value']:.1f} s")

    out = {
        "metric": "time-to-1e-8-residual, 1M-DOF hex8: learned+mixed PCG vs plain FP64 PCG",
        "value": t_fast,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_fast * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined",
        "data": "synthetic (seeded lognormal moduli, structured mesh)",
        "config": {"workload": f"hex8 cantilever {args.nx}x{args.ny}x{args.nz} static + compliance adjoint",
                   "n_dof": prob.n_dof, "n_elements": prob.mesh.n_el, "tol": args.tol,
                   "fast_precond": {api.PC_JACOBI: "jacobi", api.PC_MG: "mg", api.PC_LEARNED: "learned"}[pc],
                   "cycle_storage": "fp16 row-scaled finest operator and learned factor (fp32 arithmetic)",
                   "training": train,
                   "autotune": tuning,
                   "parallelism": "single GPU"},
        "fp64_jacobi_s": t_base,
        "speedup_vs_fp64_jacobi": t_base / t_fast,
        "fast_iterations": int(st.iterations),
        "fallback_used": int(st.fallback_used),
        "ablation": ablation,
        "fast_outer": int(st.outer_iterations),
        "fp64_iterations": int(stb.iterations),
        "rel_residual": st.rel_residual,
        "u_rel_err_vs_fp64_1e-11": err_u,
        "compliance_rel_err": err_J,
        "dJdE_rel_err_vs_fp64_path": err_s,
        "dof_per_s": prob.n_dof / t_fast,
        "wall_s_timed": wall,
        "roofline": ({"bound": "hbm", "achieved": cyc["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": cyc["achieved"] / HBM_PEAK_GBS,
                      "traffic": measured_traffic("traffic_bytes_per_launch_fp16_cycle", "sell3h_cycle_c2_pmc.md")
                      if c2 else None,
                      "kernel": "k_sell3h<3> FP16 cycle SpMV, y = f - K x (timed inside the pipeline)",
                      "bytes_per_launch": cyc["bytes_per_launch"], "avg_ms": cyc["avg_ms"],
                      "launches_per_step": cyc["launches"]} if cyc else
                     {"bound": "hbm", "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": gbs32 / HBM_PEAK_GBS,
                      "traffic": measured_traffic("traffic_bytes_per_launch_fp32") if c2 else None,
                      "kernel": "k_sell3<float,0> sliced-ELL SpMV", "bytes_per_launch": spmv32_bytes,
                      "avg_ms": ms32}),
        "spmv_fp32_loop": {"achieved": gbs32, "frac": gbs32 / HBM_PEAK_GBS, "avg_ms": ms32,
                           "bytes_per_launch": spmv32_bytes,
                           "traffic": measured_traffic("traffic_bytes_per_launch_fp32") if c2 else None},
        "spmv_fp64": {"achieved": gbs64, "frac": gbs64 / HBM_PEAK_GBS, "avg_ms": ms64,
                      "bytes_per_launch": spmv64_bytes},
        "copy_achievable_gbs": copy_gbs,
        "spmv32_frac_of_copy": gbs32 / copy_gbs,
        "cpu_baseline": cpu,
        "reference_algorithm_gpu": ref_alg,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
