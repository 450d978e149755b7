import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
for p in (ROOT, ORACLE):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    if not os.path.exists(os.path.join(ORACLE, "liboracle.so")):
        subprocess.run(["make", "-s", "-C", ORACLE], check=True)


def gpu_case(msg):
    """Print and flush the name of a GPU case before it runs."""
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="session")
def ctx():
    from asend3d_amd import api
    c = api.Context(0)
    yield c
    c.close()


def oracle_learned_graph(p):
    """Oracle graph of the learned preconditioner for a brick8 problem: the
    oracle's FP64 assembly with the reference penalty (1e5 max|K| on every
    clamped DOF, per occurrence), connectivity in DOF-block ranks."""
    import numpy as np
    import oracle as O
    import mpnn
    K = O.assemble_solid(8, p.mesh.xyz, p.mesh.conn, p.E, p.nu, p.dof_index, p.n_dof).to_scipy().tolil()
    sf = 1e5 * abs(K.tocsr()).max()
    diag = np.zeros(p.n_dof)
    np.add.at(diag, np.asarray(p.pen_dofs, dtype=np.int64), sf)
    K = (K.tocsr() + __import__("scipy.sparse", fromlist=["diags"]).diags(diag)).tocsr()
    rank = np.asarray(p.dof_index).reshape(-1, 6)[:, 0] // 3
    conn_r = rank[np.asarray(p.mesh.conn).reshape(-1, 8)]
    clamped = np.zeros(p.n_dof // 3, dtype=bool)
    clamped[np.asarray(p.pen_dofs, dtype=np.int64) // 3] = True
    return mpnn.Graph(K, conn_r, p.E, clamped), sf
