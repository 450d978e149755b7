"""Writes reorder_hand_traced.json: inputs and hand-traced expected node
orderings of the reference's node reordering (S/model_prep.rs:20-54, 56-258;
get_dist S/matrix_functions.rs:33-41).

The expected orderings below were derived by stepping the reference routine
by hand, line by line, on each input; the derivations are written out in
reorder_hand_traced.md beside this file.  Nothing here runs the oracle or the
product: the meshes come from plain coordinate arithmetic and the orderings
are literals.

    python tests/golden/make_reorder_cases.py
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def hex_grid(nx, ny, nz):
    """Unit-cell brick8 grid: nodes x-fastest, brick local node order."""
    nid = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)
    xyz = [[float(i), float(j), float(k)] for k in range(nz + 1) for j in range(ny + 1) for i in range(nx + 1)]
    conn = []
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                conn.append([nid(ex, ey, ez), nid(ex + 1, ey, ez), nid(ex + 1, ey + 1, ez), nid(ex, ey + 1, ez),
                             nid(ex, ey, ez + 1), nid(ex + 1, ey, ez + 1), nid(ex + 1, ey + 1, ez + 1),
                             nid(ex, ey + 1, ez + 1)])
    return xyz, conn


def two_cubes(gap_x):
    """Two disjoint unit cubes, the second shifted by gap_x along x."""
    cube = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
    xyz = [[float(c) for c in p] for p in cube] + [[p[0] + gap_x, float(p[1]), float(p[2])] for p in cube]
    return xyz, [list(range(8)), list(range(8, 16))]


def quad_strip():
    """2x1 shell4 strip in z = 0: nodes x-fastest, quad order counter-clockwise."""
    xyz = [[float(i), float(j), 0.0] for j in range(2) for i in range(3)]
    return xyz, [[0, 1, 4, 3], [1, 2, 5, 4]]


def cases():
    out = []
    xyz, conn = hex_grid(2, 1, 1)
    out.append(dict(name="hex2x1x1_bd3", xyz=xyz, conn=conn, dof_per_nd=3, block_dim=3, constraints=[],
                    ordered=[0, 1, 4, 3, 9, 6, 7, 10, 11, 2, 5, 8]))
    clamp = [0, 3, 6, 9]
    out.append(dict(name="hex2x1x1_clamp_bd3", xyz=xyz, conn=conn, dof_per_nd=3, block_dim=3,
                    constraints=[[clamp], [clamp], [clamp]],
                    ordered=[2, 1, 5, 4, 3, 0, 6, 7, 8, 11, 10, 9]))
    xyz3, conn3 = hex_grid(3, 2, 1)
    out.append(dict(name="hex3x2x1_bd5", xyz=xyz3, conn=conn3, dof_per_nd=3, block_dim=5, constraints=[],
                    ordered=[0, 1, 5, 4, 12, 13, 14, 2, 6, 18, 17, 3, 7, 15, 19, 11, 10, 23, 22, 9, 21, 8, 16, 20]))
    xyzc, connc = two_cubes(3.0)
    out.append(dict(name="two_cubes_bd100", xyz=xyzc, conn=connc, dof_per_nd=3, block_dim=100, constraints=[],
                    ordered=[0, 1, 2, 3, 4, 5, 6, 7, 15, 8, 9, 10, 11, 12, 13, 14]))
    out.append(dict(name="two_cubes_tie_bd100", xyz=xyzc, conn=connc, dof_per_nd=3, block_dim=100,
                    constraints=[[[1, 2], [8, 11]]],
                    ordered=[0, 1, 2, 3, 4, 5, 6, 7, 8, 11, 9, 10, 12, 13, 14, 15]))
    xyzq, connq = quad_strip()
    edge = [0, 3]
    out.append(dict(name="shell2x1_clamp_bd2", xyz=xyzq, conn=connq, dof_per_nd=6, block_dim=2,
                    constraints=[[edge]] * 6, ordered=[2, 1, 5, 4, 0, 3]))
    for c in out:
        n = len(c["xyz"])
        dpn = c["dof_per_nd"]
        di = [[-1] * 6 for _ in range(n)]
        for k, node in enumerate(c["ordered"]):
            for d in range(dpn):
                di[node][d] = dpn * k + d
        c["dof_index"] = di
        c["n_dof"] = dpn * n
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "reorder_hand_traced.json"), "w") as fh:
        json.dump({"source": "S/model_prep.rs:20-54,56-258 stepped by hand (reorder_hand_traced.md)",
                   "cases": cases()}, fh, indent=1)
