"""Writes the C1 job fixture in tests/golden/c1_job/: the 20x4x4 brick8
cantilever (1.0 x 0.2 x 0.2 m, E = 200 GPa, nu = 0.3, rho = 7850) clamped
at x = 0 with a -1 MPa z traction on the tip face, as the reference's job
files describe it: model, constraints, loads, design variables, objective
and three jobs (static with solverMethod direct and iterative; dynamic under a
traction pulse, with state files and time-step node results).

The files follow the layout of the reference's Python writers
(yaml.dump(..., sort_keys=False) of the same nested keys, quotes removed;
P/model/Model.py:467-480, P/job/ASenDJob.py:223-233,
P/objective/Objective.py:41-51, P/variables/DesignVariables.py:31-40), which
the solver's line reader (S/model_input.rs:34-119) consumes.  Node labels
run x fastest, then y, then z (P/meshing/Mesh3D.py:111-136), as
asend3d_amd.meshgen.hex_block numbers them.

    python tests/golden/make_c1_job.py
"""
import os
import sys

import yaml

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from asend3d_amd.meshgen import hex_block  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_job")


def dump(path, data, width=None):
    kw = {"sort_keys": False}
    if width:
        kw["width"] = width
    s = yaml.dump(data, **kw).replace("'", "").replace('"', "")
    with open(os.path.join(OUT, path), "w") as fh:
        fh.write(s)


def main():
    os.makedirs(OUT, exist_ok=True)
    m = hex_block(20, 4, 4, 1.0, 0.2, 0.2)
    E, nu = 200e9, 0.3
    G = E / (2.0 * (1.0 + nu))
    nodes = [str([i] + [float(v) for v in m.xyz[i]]) for i in range(m.n_nodes)]
    conn = [str([e] + [int(v) for v in m.conn[e]]) for e in range(m.n_el)]
    x_min = [int(i) for i in m.nodes_at_x(0)]
    x_max = [int(i) for i in m.nodes_at_x(m.nx)]
    tip_els = [int(e) for e in m.elements_at_x(m.nx - 1)]
    model = {
        "nodes": nodes,
        "elements": [{"type": "brick8", "connectivity": conn}],
        "sets": {"node": {"xMin": x_min, "xMax": x_max}, "element": {"tipEls": tip_els}},
        "sections": [{"type": "solid", "material": "steel", "elementSet": "all"}],
        "materials": {"steel": {"density": 7850.0,
                                "elastic": {"E": str([E, E, E]), "nu": str([nu, nu, nu]), "G": str([G, G, G])}}},
    }
    dump("c1_model.yaml", model, width=200)
    cons = {"constraints": [{"type": "displacement", "terms": [{"nodeSet": "xMin", "dof": d, "coef": 1.0}],
                             "rhs": 0.0} for d in (1, 2, 3)]}
    dump("c1_constraints.yaml", cons, width=200)
    loads = {"loads": [{"type": "surfaceTraction", "activeTime": str([0.0, 1e100]), "normDir": str([1.0, 0.0, 0.0]),
                        "normTolerance": 5.0, "load": str([0.0, 0.0, -1.0e6]), "elementSet": "tipEls"}]}
    dump("c1_loads.yaml", loads, width=200)
    dvs = []
    for cat, comp, es, coef in (("modulus", 1, "all", 1.0), ("modulus", 2, "all", 1.0), ("modulus", 3, "all", 1.0),
                                ("shearModulus", 1, "all", 1.0), ("shearModulus", 3, "all", 1.0),
                                ("poissonRatio", 1, "all", 1.0), ("poissonRatio", 3, "all", 1.0),
                                ("modulus", 1, "5", 1.0), ("modulus", 3, "tipEls", 2.0)):
        dvs.append({"category": cat, "component": comp, "elementSet": es, "activeTime": str(0.0),
                    "coefficients": coef})
    dump("c1_dvars.yaml", {"designVariables": dvs})
    terms = [{"category": "displacement", "operator": "powerNorm", "component": 3, "nodeSet": "xMax",
              "activeTime": str([0.0, 1e100]), "coefficient": 1.0, "exponent": 1.0, "targetValue": 0.0},
             {"category": "displacement", "operator": "powerNorm", "component": 1, "nodeSet": "xMax",
              "activeTime": str([0.0, 1e100]), "coefficient": 0.5, "exponent": 2.0, "targetValue": 0.0}]
    dump("c1_objective.yaml", {"objectiveTerms": terms})
    for name, method, extra in (("c1_job.yaml", "direct", {}),
                                ("c1_job_iterative.yaml", "iterative", {"maxIterations": 200000,
                                                                         "convergenceTol": 1e-24})):
        solve = {"command": "solve", "staticLoadTime": [0.0], "loadRampSteps": 1, "timeStep": 1.0,
                 "newmarkBeta": 0.25, "newmarkGamma": 0.5, "simPeriod": 1.0, "saveSolnHist": "yes",
                 "solnHistDir": "results" if method == "direct" else "results_iterative",
                 "solverMethod": method, "solverBlockDim": 2000000000}
        solve.update(extra)
        tag = "" if method == "direct" else "_iterative"
        job = {"jobCommands": [
            {"command": "readModelInput", "fileName": "c1_model.yaml"},
            {"command": "readConstraints", "fileName": "c1_constraints.yaml"},
            {"command": "readLoads", "fileName": "c1_loads.yaml"},
            {"command": "readDesignVarInput", "fileName": "c1_dvars.yaml"},
            {"command": "readObjectiveInput", "fileName": "c1_objective.yaml"},
            solve,
            {"command": "calcObjGradient"},
            {"command": "writeNodeResults", "fileName": f"results{tag}/nodeResults.csv", "nodeSet": "all",
             "fields": ["displacement", "reactionForce"]},
            {"command": "writeObjective", "fileName": f"results{tag}/objectiveResults.csv",
             "include": ["category", "nodeSet"]},
        ]}
        dump(name, job)
    # dynamic job: the same cantilever under a tip traction pulse (active for
    # the steps that start at t <= 9e-4 s), Rayleigh damping, 2e-4 s steps over
    # 4 ms, state files in results_dynamic/ and node results of five steps
    pulse = {"loads": [{"type": "surfaceTraction", "activeTime": str([0.0, 9e-4]), "normDir": str([1.0, 0.0, 0.0]),
                        "normTolerance": 5.0, "load": str([0.0, 0.0, -1.0e6]), "elementSet": "tipEls"}]}
    dump("c1_loads_pulse.yaml", pulse, width=200)
    solve = {"command": "solve", "elastic": "yes", "dynamic": "yes", "timeStep": 2e-4, "newmarkBeta": 0.25,
             "newmarkGamma": 0.5, "simPeriod": 4e-3, "rayDampCM": 5.0, "rayDampCK": 1e-5, "saveSolnHist": "yes",
             "solnHistDir": "results_dynamic", "solverMethod": "direct"}
    job = {"jobCommands": [
        {"command": "readModelInput", "fileName": "c1_model.yaml"},
        {"command": "readConstraints", "fileName": "c1_constraints.yaml"},
        {"command": "readLoads", "fileName": "c1_loads_pulse.yaml"},
        solve,
        {"command": "writeNodeResults", "fileName": "results_dynamic/nodeResults.csv", "nodeSet": "all",
         "fields": ["displacement", "velocity", "acceleration", "reactionForce"], "timeSteps": "0 21 5"},
    ]}
    dump("c1_job_dynamic.yaml", job)


if __name__ == "__main__":
    main()
