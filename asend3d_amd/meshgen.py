"""Synthetic structured meshes for the solve-path workloads.

Node and element label order follows the reference's swept-mesh generator:
nodes x-fastest, then y, then z (P/meshing/Mesh3D.py:111-136); hex nodes in the
brick8 local order (-1,-1,-1), (1,-1,-1), (1,1,-1), (-1,1,-1), then the same at
z = +1 (S/element_meth.rs:207-231).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class HexBlock:
    nx: int
    ny: int
    nz: int
    lx: float
    ly: float
    lz: float
    xyz: np.ndarray = field(repr=False)
    conn: np.ndarray = field(repr=False)

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_el(self):
        return self.conn.shape[0]

    def node_id(self, i, j, k):
        return i + (self.nx + 1) * (j + (self.ny + 1) * k)

    def nodes_at_x(self, i):
        """Node labels of the plane x-index i, in label order."""
        j, k = np.meshgrid(np.arange(self.ny + 1), np.arange(self.nz + 1), indexing="xy")
        ids = self.node_id(i, j.reshape(-1), k.reshape(-1))
        return np.sort(ids).astype(np.int32)

    def elements_at_x(self, ex):
        ey, ez = np.meshgrid(np.arange(self.ny), np.arange(self.nz), indexing="xy")
        ids = ex + self.nx * (ey.reshape(-1) + self.ny * ez.reshape(-1))
        return np.sort(ids).astype(np.int32)


def hex_block(nx, ny, nz, lx, ly, lz):
    """Structured brick8 block [0,lx]x[0,ly]x[0,lz] with nx*ny*nz elements."""
    xs = np.arange(nx + 1, dtype=np.float64) * (lx / nx)
    ys = np.arange(ny + 1, dtype=np.float64) * (ly / ny)
    zs = np.arange(nz + 1, dtype=np.float64) * (lz / nz)
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    xyz = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], axis=1)
    ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ex = ex.reshape(-1)
    ey = ey.reshape(-1)
    ez = ez.reshape(-1)

    def nid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    conn = np.stack([nid(ex, ey, ez), nid(ex + 1, ey, ez), nid(ex + 1, ey + 1, ez), nid(ex, ey + 1, ez),
                     nid(ex, ey, ez + 1), nid(ex + 1, ey, ez + 1), nid(ex + 1, ey + 1, ez + 1),
                     nid(ex, ey + 1, ez + 1)], axis=1).astype(np.int32)
    return HexBlock(nx, ny, nz, lx, ly, lz, np.ascontiguousarray(xyz), np.ascontiguousarray(conn))


def block_dim_auto(n_nodes):
    """Iterative-solver block size: smallest multiple of 6 whose square reaches
    the node count (S/model_prep.rs:888-902)."""
    b = 6
    while b * b < n_nodes:
        b += 6
    return b


@dataclass
class QuadGrid:
    nx: int
    ny: int
    lx: float
    ly: float
    xyz: np.ndarray = field(repr=False)
    conn: np.ndarray = field(repr=False)

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_el(self):
        return self.conn.shape[0]

    def node_id(self, i, j):
        return i + (self.nx + 1) * j

    def boundary_nodes(self):
        i = np.arange(self.nx + 1)
        j = np.arange(self.ny + 1)
        ids = np.concatenate([self.node_id(i, 0), self.node_id(i, self.ny), self.node_id(0, j),
                              self.node_id(self.nx, j)])
        return np.unique(ids).astype(np.int32)

    def nodes_at_x(self, i):
        return np.sort(self.node_id(i, np.arange(self.ny + 1))).astype(np.int32)


def quad_grid(nx, ny, lx, ly):
    """Structured shell4 grid on [0,lx]x[0,ly] in the plane z = 0, nodes
    x-fastest, element nodes counter-clockwise from (-1,-1) (the quad4 order of
    S/element_soln_fields.rs:194-213)."""
    xs = np.arange(nx + 1, dtype=np.float64) * (lx / nx)
    ys = np.arange(ny + 1, dtype=np.float64) * (ly / ny)
    Y, X = np.meshgrid(ys, xs, indexing="ij")
    xyz = np.stack([X.reshape(-1), Y.reshape(-1), np.zeros(X.size)], axis=1)
    ey, ex = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ex = ex.reshape(-1)
    ey = ey.reshape(-1)
    n0 = ex + (nx + 1) * ey
    conn = np.stack([n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)], axis=1).astype(np.int32)
    return QuadGrid(nx, ny, lx, ly, xyz, conn)
