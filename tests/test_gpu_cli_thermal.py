"""The job-file driver on a heat conduction job ("thermal: yes" beside the
elastic field, no thermal expansion): a brick8 block with an anisotropic
conductivity, temperature constraints at both ends (one written with
coefficient -1), bodyHeatGen, surfaceFlux (where the reference's running
normal test moves the corner element's flux onto its -x face), nodalHeatGen
active only at the first of two static load times.  Checked against the
oracle (oracle/thermal.py): node temperatures and reactionHeatGen at both load
times (the state files and writeNodeResults' timeSteps), the elastic
displacements of the same solve, the thermal maxBandwidth, and for the
iterative method the conj_grad_sparse iteration count of both thermal steps
on the reference's one-DOF-per-node system."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
import thermal as TH
from asend3d_amd.meshgen import hex_block
from conftest import gpu_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "asend3d_amd", "asend_run_job")
K6 = (2.0, 1.5, 1.0, 0.3, -0.2, 0.1)
Q_BODY, Q_FLUX, Q_NODE = 3.0, 2.0, 0.5
T_MIN, T_MAX = 20.0, 5.0


def dump(path, data):
    import yaml
    path.write_text(yaml.dump(data, sort_keys=False, width=200).replace("'", "").replace('"', ""))


def mesh():
    m = hex_block(8, 3, 2, 2.0, 0.6, 0.4)
    xyz = m.xyz.copy()
    rng = np.random.default_rng(21)
    inner = np.all((xyz > 1e-9) & (xyz < np.array([2.0, 0.6, 0.4]) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-0.05, 0.05, (int(inner.sum()), 3))
    return m, xyz


def write_job(work, m, xyz, method, block_dim=None, conv_tol=None):
    y_max_els = [e for e in range(m.n_el) if xyz[m.conn[e], 1].max() > 0.6 - 1e-9]
    mid = [int(i) for i in m.nodes_at_x(4)]
    model = {
        "nodes": [str([i] + [float(v) for v in xyz[i]]) for i in range(m.n_nodes)],
        "elements": [{"type": "brick8",
                      "connectivity": [str([e] + [int(v) for v in m.conn[e]]) for e in range(m.n_el)]}],
        "sets": {"node": {"xMin": [int(i) for i in m.nodes_at_x(0)], "xMax": [int(i) for i in m.nodes_at_x(m.nx)],
                          "mid": mid},
                 "element": {"yMaxEls": y_max_els, "tipEls": [int(e) for e in m.elements_at_x(m.nx - 1)]}},
        "sections": [{"type": "solid", "material": "mat", "elementSet": "all"}],
        "materials": {"mat": {"density": 1.0,
                              "elastic": {"E": str([1e6] * 3), "nu": str([0.25] * 3), "G": str([4e5] * 3)},
                              "thermal": {"conductivity": str(list(K6))}}},
    }
    dump(work / "model.yaml", model)
    cons = [{"type": "displacement", "terms": [{"nodeSet": "xMin", "dof": d, "coef": 1.0}], "rhs": 0.0}
            for d in (1, 2, 3)]
    cons += [{"type": "temperature", "terms": [{"nodeSet": "xMin", "coef": 1.0}], "rhs": T_MIN},
             {"type": "temperature", "terms": [{"nodeSet": "xMax", "coef": -1.0}], "rhs": -T_MAX}]
    dump(work / "constraints.yaml", {"constraints": cons})
    loads = [{"type": "surfaceTraction", "elementSet": "tipEls", "load": str([0.0, 0.0, -100.0]),
              "normDir": str([1.0, 0.0, 0.0]), "normTolerance": 5.0},
             {"type": "bodyHeatGen", "elementSet": "all", "load": str([Q_BODY])},
             {"type": "surfaceFlux", "elementSet": "yMaxEls", "load": str([Q_FLUX]),
              "normDir": str([0.0, 1.0, 0.0]), "normTolerance": 5.0},
             {"type": "nodalHeatGen", "nodeSet": "mid", "load": str([Q_NODE]), "activeTime": str([0.0, 0.5])}]
    dump(work / "loads.yaml", {"loads": loads})
    (work / "results").mkdir()
    solve = {"command": "solve", "thermal": "yes", "solverMethod": method, "solnHistDir": "results",
             "staticLoadTime": str([0.0, 1.0])}
    if block_dim:
        solve["solverBlockDim"] = block_dim
        solve["convergenceTol"] = conv_tol
    job = {"jobCommands": [
        {"command": "readModelInput", "fileName": "model.yaml"},
        {"command": "readConstraints", "fileName": "constraints.yaml"},
        {"command": "readLoads", "fileName": "loads.yaml"},
        solve,
        {"command": "writeNodeResults", "fileName": "results/nodes.csv", "nodeSet": "all",
         "fields": ["temperature", "reactionHeatGen", "displacement"]},
        {"command": "writeNodeResults", "fileName": "results/first.csv", "nodeSet": "all",
         "fields": ["temperature", "reactionHeatGen"], "timeSteps": 0},
    ]}
    dump(work / "job.yaml", job)
    return np.array(y_max_els), np.array(mid)


def oracle_problem(m, xyz, y_max_els, mid, bd):
    kc = np.tile(K6, (m.n_el, 1))
    K = TH.assemble_conduction(xyz, m.conn, kc)
    sf = 1e5 * np.abs(K.data).max()
    f1 = TH.body_heat_gen(xyz, m.conn, np.arange(m.n_el), Q_BODY)
    f1 += TH.surface_flux(xyz, m.conn, y_max_els, Q_FLUX, (0.0, 1.0, 0.0), 5.0)
    f0 = f1.copy()
    f0[mid] += Q_NODE
    xmin, xmax = m.nodes_at_x(0), m.nodes_at_x(m.nx)
    pen = np.concatenate([xmin, xmax])
    cr = np.concatenate([np.full(len(xmin), T_MIN), np.full(len(xmax), T_MAX)])  # coefficient x rhs
    # the reference's numbering: displacement constraints, then temperature ones
    _, di, _ = O.reorder_nodes(xyz, m.conn, 3, [[xmin]] * 3 + [[xmin], [xmax]], bd)
    rank = np.asarray(di).reshape(-1, 6)[:, 0] // 3
    return K, sf, f0, f1, pen, cr, rank


def penalised(K, sf, pen):
    return (K + sp.csr_matrix((np.full(len(pen), sf), (pen, pen)), shape=K.shape)).tocsr()


def read_nodes(path):
    import csv
    with open(path) as fh:
        rows = list(csv.reader(fh))
    return rows[0], np.array([[float(x) for x in r if x != ""] for r in rows[1:]])


def run(work):
    r = subprocess.run([CLI, "job.yaml"], cwd=work, capture_output=True, text=True, timeout=300)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cli_thermal_job_direct(tmp_path):
    gpu_case("asend_run_job thermal + elastic, direct")
    m, xyz = mesh()
    work = tmp_path / "direct"
    work.mkdir()
    y_max_els, mid = write_job(work, m, xyz, "direct")
    out = run(work)
    K, sf, f0, f1, pen, cr, rank = oracle_problem(m, xyz, y_max_els, mid, 2000000000)
    Kp = penalised(K, sf, pen)
    T0 = spla.spsolve(Kp.tocsc(), f0 + _scatter(pen, sf * cr, m.n_nodes))
    T1 = spla.spsolve(Kp.tocsc(), f1 + _scatter(pen, sf * cr, m.n_nodes))
    head, last = read_nodes(work / "results" / "nodes.csv")
    assert head[:3] == ["node", "T", "RHG"] and head[3] == "U1" and head[-1] == "time"
    _, first = read_nodes(work / "results" / "first_timestep0.csv")
    for T_ref, rows in ((T1, last), (T0, first)):
        T, rhg = rows[:, 1], rows[:, 2]
        assert np.abs(T - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
        want = K @ T_ref
        assert np.abs(rhg - want).max() <= 1e-7 * np.abs(want).max()
        # K 1 = 0: the reactions at the held nodes balance the heat put in
        assert abs(rhg.sum()) <= 1e-8 * np.abs(rhg).sum()
    assert np.abs(first[:, -1]).max() == 0.0  # time of staticLoadTime 0
    # the held temperatures, to the penalty's accuracy
    assert np.abs(last[m.nodes_at_x(0), 1] - T_MIN).max() <= 1e-4 * T_MIN
    assert np.abs(last[m.nodes_at_x(m.nx), 1] - T_MAX).max() <= 1e-3 * T_MAX
    # the state file of each load time: temperature and rate, then the elastic state
    raw = np.fromfile(work / "results" / "solnTStep0.out", dtype=">f8").reshape(m.n_nodes, 2 + 9)
    assert np.abs(raw[:, 0] - first[:, 1]).max() <= 1e-12 * np.abs(T0).max() and not raw[:, 1].any()
    # the thermal matrix's bandwidth is printed first
    bw = [int(line.split(":")[1]) for line in out.splitlines() if line.startswith("maxBandwidth:")]
    Ks = _scalar(K, rank)
    ref = O.LDL(Ks, rank[pen], sf, Ks.n)
    assert bw[0] == ref.max_bandwidth()
    # the elastic field of the same solve
    U = last[:, 3:9]
    assert np.abs(U[:, 2]).max() > 0 and np.abs(U[m.nodes_at_x(0)]).max() <= 1e-6 * np.abs(U).max()


def _scatter(idx, vals, n):
    out = np.zeros(n)
    np.add.at(out, idx, vals)
    return out


def _scalar(K, rank):
    perm = np.argsort(rank)
    Ks = K[perm][:, perm].tocsr()
    Ks.sort_indices()
    return O.CSR(Ks.indptr.astype(np.int64), Ks.indices.astype(np.int32), Ks.data.copy())


def test_cli_thermal_job_iterative(tmp_path):
    gpu_case("asend_run_job thermal + elastic, iterative")
    m, xyz = mesh()
    bd = 20
    y_max_els = np.array([e for e in range(m.n_el) if xyz[m.conn[e], 1].max() > 0.6 - 1e-9])
    mid = m.nodes_at_x(4)
    K, sf, f0, f1, pen, cr, rank = oracle_problem(m, xyz, y_max_els, mid, bd)
    Ks = _scalar(K, rank)
    ref = O.LDL(Ks, rank[pen], sf, bd)
    b0 = _scatter(rank, f0, m.n_nodes) + _scatter(rank[pen], sf * cr, m.n_nodes)
    tol = 1e-10 * float(b0 @ ref.solve(b0))
    work = tmp_path / "iterative"
    work.mkdir()
    write_job(work, m, xyz, "iterative", block_dim=bd, conv_tol=float(f"{tol:.6e}"))
    out = run(work)
    tol = float(f"{tol:.6e}")
    # step 0 from T = 0; step 1 solves for the increment from the step-0 temperatures
    x0, it0, _ = O.conj_grad_sparse(Ks, rank[pen], sf, ref, b0, tol, 100000)
    b1 = _scatter(rank, f1, m.n_nodes) + _scatter(rank[pen], sf * cr, m.n_nodes) - O.spmv_penalty(Ks, rank[pen], sf,
                                                                                                   x0)
    x1, it1, _ = O.conj_grad_sparse(Ks, rank[pen], sf, ref, b1, tol, 100000)
    its = [int(line.split(":")[1]) for line in out.splitlines() if line.startswith("Total CG iterations:")]
    # per load time: the thermal solve, then the elastic one
    assert len(its) == 4, its
    assert abs(its[0] - it0) <= 1 and abs(its[2] - it1) <= 1, (its, it0, it1)
    _, last = read_nodes(work / "results" / "nodes.csv")
    T_ref = (x0 + x1)[rank]
    assert np.abs(last[:, 1] - T_ref).max() <= 1e-7 * np.abs(T_ref).max()


def test_cli_transient_thermal_job(tmp_path):
    """thermal: yes, elastic: no, dynamic: yes (the thermal Newmark loop,
    gamma 0.6): every state file and the final temperatures, rates and
    reactionHeatGen (K T + C Tdot) against the oracle's restatement of the
    loop, with the penalty scale of the first matrix K + C / (gamma dt)."""
    gpu_case("asend_run_job transient heat conduction")
    m, xyz = mesh()
    work = tmp_path / "transient"
    work.mkdir()
    y_max_els, mid = write_job(work, m, xyz, "direct")
    model = (work / "model.yaml").read_text().replace("density: 1.0", "density: 2.0")
    model = model.replace("thermal:\n      conductivity", "thermal:\n      specHeat: 900.0\n      conductivity")
    assert "specHeat: 900.0" in model
    (work / "model.yaml").write_text(model)
    import yaml
    job = yaml.safe_load((work / "job.yaml").read_text())
    cmds = job["jobCommands"]
    solve = [c for c in cmds if c["command"] == "solve"][0]
    solve.pop("staticLoadTime")
    # (safe_load reads yes/no as booleans: write the job's words back)
    solve.update({"thermal": "yes", "elastic": "no", "dynamic": "yes", "timeStep": 0.05, "newmarkGamma": 0.6,
                  "simPeriod": 0.5})
    for c in cmds:
        if c["command"] == "writeNodeResults":
            c["fields"] = ["temperature", "tdot", "reactionHeatGen"]
    dump(work / "job.yaml", job)
    out = run(work)
    # time accumulates 0.05 ten times to 0.49999999999999994 < simPeriod: an
    # eleventh step, in the reference's loop as here
    assert "transient heat conduction: 11 steps" in out
    K, _, f0, f1, pen, cr, rank = oracle_problem(m, xyz, y_max_els, mid, 2000000000)
    C = TH.assemble_capacity(xyz, m.conn, np.full(m.n_el, 1800.0))
    dt, g = 0.05, 0.6
    sf = 1e5 * np.abs((K + C / (g * dt)).data).max()
    load_at = lambda t: f0 if t <= 0.5 else f1  # nodalHeatGen active over [0, 0.5]
    hist = TH.transient(K, C, pen, sf, cr, load_at, dt, g, 0.5)
    assert len(hist) == 11
    for step in (1, 4, 11):
        raw = np.fromfile(work / "results" / f"solnTStep{step}.out", dtype=">f8").reshape(m.n_nodes, 2)
        T_ref, Td_ref = hist[step - 1]
        assert np.abs(raw[:, 0] - T_ref).max() <= 1e-9 * np.abs(T_ref).max(), step
        assert np.abs(raw[:, 1] - Td_ref).max() <= 1e-8 * np.abs(Td_ref).max(), step
    head, last = read_nodes(work / "results" / "nodes.csv")
    assert head[:4] == ["node", "T", "TDOT", "RHG"]
    T_ref, Td_ref = hist[-1]
    assert np.abs(last[:, 1] - T_ref).max() <= 1e-9 * np.abs(T_ref).max()
    want = K @ T_ref + C @ Td_ref
    assert np.abs(last[:, 3] - want).max() <= 1e-7 * np.abs(want).max()


def test_cli_thermal_expansion_free_expansion(tmp_path):
    """thermal + elastic with thermal expansion: every node held at T = 80,
    rollers on the x = 0, y = 0 and z = 0 faces, no loads.  The body expands
    stress-free, u = 80 (a1 x, a2 y, a3 z) exactly (a patch test on a mesh
    with moved interior nodes), and the reactions vanish."""
    gpu_case("asend_run_job thermal expansion, free expansion")
    m = hex_block(4, 3, 3, 1.0, 0.8, 0.6)
    xyz = m.xyz.copy()
    rng = np.random.default_rng(8)
    inner = np.all((xyz > 1e-9) & (xyz < np.array([1.0, 0.8, 0.6]) - 1e-9), axis=1)
    xyz[inner] += rng.uniform(-0.04, 0.04, (int(inner.sum()), 3))
    al = (1e-5, 2e-5, 3e-5, 0.0, 0.0, 0.0)
    work = tmp_path / "expansion"
    work.mkdir()
    near = lambda k: [int(i) for i in np.flatnonzero(np.abs(xyz[:, k]) < 1e-9)]
    model = {
        "nodes": [str([i] + [float(v) for v in xyz[i]]) for i in range(m.n_nodes)],
        "elements": [{"type": "brick8",
                      "connectivity": [str([e] + [int(v) for v in m.conn[e]]) for e in range(m.n_el)]}],
        "sets": {"node": {"x0": near(0), "y0": near(1), "z0": near(2)}},
        "sections": [{"type": "solid", "material": "mat", "elementSet": "all"}],
        "materials": {"mat": {"density": 1.0,
                              "elastic": {"E": str([70e3] * 3), "nu": str([0.3] * 3), "G": str([70e3 / 2.6] * 3)},
                              "thermal": {"conductivity": str([2.0, 2.0, 2.0, 0.0, 0.0, 0.0]),
                                          "expansion": str(list(al))}}},
        "constraints": [{"type": "displacement", "terms": [{"nodeSet": s, "dof": d, "coef": 1.0}], "rhs": 0.0}
                        for s, d in (("x0", 1), ("y0", 2), ("z0", 3))]
        + [{"type": "temperature", "terms": [{"nodeSet": "all", "coef": 1.0}], "rhs": 80.0}],
    }
    dump(work / "model.yaml", model)
    (work / "results").mkdir()
    dump(work / "job.yaml", {"jobCommands": [
        {"command": "readModelInput", "fileName": "model.yaml"},
        {"command": "solve", "thermal": "yes", "solverMethod": "direct", "solnHistDir": "results"},
        {"command": "writeNodeResults", "fileName": "results/nodes.csv", "nodeSet": "all",
         "fields": ["temperature", "displacement", "reactionForce"]}]})
    run(work)
    head, rows = read_nodes(work / "results" / "nodes.csv")
    assert np.abs(rows[:, 1] - 80.0).max() <= 1e-9 * 80.0
    U = rows[:, 2:5]
    want = 80.0 * xyz * np.array(al[:3])
    assert np.abs(U - want).max() <= 1e-9 * np.abs(want).max()
    RF = rows[:, 8:11]
    # stress-free: the reactions are rounding next to the expansion load's scale (E alpha T h^2 ~ 2e-2)
    assert np.abs(RF).max() <= 1e-8
