"""CPU checker of the learned factored preconditioner (numpy, FP64).

TEST INFRASTRUCTURE ONLY: this module is the checker of the device network in
asend3d_amd/csrc/learned.hip.  Only tests/ import it; the product package never
does.  The learned preconditioner has no counterpart in the reference (it is
new to this build), so parity is between the device kernels and this
restatement of the same mathematics:

  features   x_i  (8): mean/min/max of log E - mean(log E) over the elements of
             node i, (#neighbours)/26, clamp flag, tanh(log(B11/B00)),
             tanh(log(B22/B00)), (|B01|+|B02|+|B12|)/tr(B), B = diagonal block
  layer 0    h = bf16(relu(W0 x + b0))
  layers 1-3 a = bf16(mean_{j in N(i)} h_j);  h = bf16(relu(Wcat [h; a] + b))
  head       o = Wh h + bh  (24 outputs)
  factor     C = chol(B), Ci = C^-1;  Ahat_ij = Ci_i A_ij Ci_j^T
             Lam_ii = [[e^o0,0,0],[o3,e^o1,0],[o4,o5,e^o2]]  (o0..o2 clamped to +-10)
             Lam_ij = o_i[6:15] + o_j[15:24] + Wa vec(Ahat_ij)    (j a mesh-edge neighbour)
             L_ij = Ci_i^T Lam_ij,  P = L L^T
  loss       r = A e, z = P r:  min_w ||e - w z||_A^2 / ||e||_A^2
                                  = 1 - (z.r)^2 / ((z.A z)(e.A e))

All weights enter as BF16-rounded values, biases as FP32, exactly as on the
device; `bf16=False` switches every rounding off, which is what the
finite-difference test of the backward pass uses.

Pinning (tests/test_learned_oracle.py): the backward pass against central
finite differences along random directions in FP64; the untrained network
(head and Wa zero) reduces P to the inverse of the 3x3 diagonal blocks, checked
against numpy.linalg.inv; P is symmetric positive definite for random
parameters.
"""
import numpy as np
import scipy.sparse as sp

FEAT, HID, LAYERS, HEAD = 8, 64, 3, 24
OFF_W0 = 0
OFF_B0 = OFF_W0 + HID * FEAT
LAYER_SIZE = HID * 2 * HID + HID
OFF_L1 = OFF_B0 + HID
OFF_WH = OFF_L1 + LAYERS * LAYER_SIZE
OFF_BH = OFF_WH + HEAD * HID
OFF_WA = OFF_BH + HEAD
NUM_PARAMS = OFF_WA + 81
EXP_CLAMP = 10.0

HEX_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    x32 = np.ascontiguousarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


class Graph:
    """Mesh graph in rank (DOF-block) order.

    K: FP64 operator (scipy, 3n x 3n, penalty folded); conn_r: hex8
    connectivity in ranks (n_el x 8); E: element moduli; clamped: per-node flag.
    """

    def __init__(self, K, conn_r, E, clamped):
        K = sp.csr_matrix(K)
        n = K.shape[0] // 3
        self.n = n
        self.K = K
        conn_r = np.asarray(conn_r, dtype=np.int64)
        E = np.asarray(E, dtype=np.float64)
        n_el = conn_r.shape[0]
        # node -> elements, node neighbours (sharing an element)
        rows = np.repeat(np.arange(n_el), 8)
        n2e = sp.csr_matrix((np.ones(8 * n_el), (conn_r.ravel(), rows)), shape=(n, n_el))
        adj = (n2e @ n2e.T).tocsr()
        adj.data[:] = 1.0
        adj.setdiag(0)
        adj.eliminate_zeros()
        self.deg = np.asarray(adj.sum(axis=1)).ravel()
        self.adj = adj
        self.mean_adj = sp.diags(1.0 / np.maximum(self.deg, 1)) @ adj
        # features
        le = np.log(E) - np.log(E).mean()
        n2e_b = n2e.tocsr()
        cnt = np.diff(n2e_b.indptr)
        s = n2e_b @ le
        mn = np.full(n, np.inf)
        mx = np.full(n, -np.inf)
        for nd in range(n):
            v = le[n2e_b.indices[n2e_b.indptr[nd]:n2e_b.indptr[nd + 1]]]
            if v.size:
                mn[nd] = v.min()
                mx[nd] = v.max()
        B = np.zeros((n, 3, 3))
        for a in range(3):
            for b in range(3):
                B[:, a, b] = np.asarray(K[3 * np.arange(n) + a, 3 * np.arange(n) + b]).ravel()
        self.B = B
        X = np.zeros((n, FEAT))
        X[:, 0] = np.where(cnt > 0, s / np.maximum(cnt, 1), 0.0)
        X[:, 1] = np.where(cnt > 0, mn, 0.0)
        X[:, 2] = np.where(cnt > 0, mx, 0.0)
        X[:, 3] = self.deg / 26.0
        X[:, 4] = np.asarray(clamped, dtype=np.float64)
        X[:, 5] = np.tanh(np.log(B[:, 1, 1] / B[:, 0, 0]))
        X[:, 6] = np.tanh(np.log(B[:, 2, 2] / B[:, 0, 0]))
        X[:, 7] = (np.abs(B[:, 0, 1]) + np.abs(B[:, 0, 2]) + np.abs(B[:, 1, 2])) / np.trace(B, axis1=1, axis2=2)
        self.X = X
        # mesh-edge neighbours (directed both ways)
        pairs = set()
        for a, b in HEX_EDGES:
            for u, v in zip(conn_r[:, a], conn_r[:, b]):
                if u != v:
                    pairs.add((int(u), int(v)))
                    pairs.add((int(v), int(u)))
        ed = np.array(sorted(pairs), dtype=np.int64).reshape(-1, 2)
        self.esrc, self.edst = ed[:, 0], ed[:, 1]
        # Cholesky scaling and scaled edge blocks
        Bs = 0.5 * (B + B.transpose(0, 2, 1))
        C = np.linalg.cholesky(Bs)
        self.Ci = np.linalg.inv(C)
        Aij = np.zeros((len(self.esrc), 3, 3))
        for a in range(3):
            for b in range(3):
                Aij[:, a, b] = np.asarray(K[3 * self.esrc + a, 3 * self.edst + b]).ravel()
        self.Ahat = np.einsum('eik,ekl,ejl->eij', self.Ci[self.esrc], Aij, self.Ci[self.edst])


def _weights(params, use_bf16):
    p = np.asarray(params, dtype=np.float64)
    q = bf16(p) if use_bf16 else p.copy()
    W0 = q[OFF_W0:OFF_B0].reshape(HID, FEAT)
    b0 = p[OFF_B0:OFF_L1]
    Wl, bl = [], []
    for l in range(LAYERS):
        off = OFF_L1 + l * LAYER_SIZE
        Wl.append(q[off:off + HID * 2 * HID].reshape(HID, 2 * HID))
        bl.append(p[off + HID * 2 * HID:off + LAYER_SIZE])
    Wh = q[OFF_WH:OFF_BH].reshape(HEAD, HID)
    bh = p[OFF_BH:OFF_WA]
    Wa = q[OFF_WA:NUM_PARAMS].reshape(9, 9)
    return W0, b0, Wl, bl, Wh, bh, Wa


def forward(g, params, use_bf16=True):
    """Network forward and factor.  Returns a dict with the activations, the
    head outputs O (n x 24) and the factor blocks (Ld: n x 3 x 3, Le: edges)."""
    rnd = bf16 if use_bf16 else (lambda v: v)
    W0, b0, Wl, bl, Wh, bh, Wa = _weights(params, use_bf16)
    H = [rnd(np.maximum(g.X @ W0.T + b0, 0.0))]
    A = []
    for l in range(LAYERS):
        a = rnd(g.mean_adj @ H[l])
        A.append(a)
        H.append(rnd(np.maximum(np.concatenate([H[l], a], axis=1) @ Wl[l].T + bl[l], 0.0)))
    O = H[-1] @ Wh.T + bh
    Ld, Le = factor_blocks(g, O, Wa)
    return {"H": H, "A": A, "O": O, "Ld": Ld, "Le": Le, "W": (W0, b0, Wl, bl, Wh, bh, Wa)}


def factor_blocks(g, O, Wa):
    """Factor blocks from the head outputs O (n x 24) and the edge map Wa (9 x 9):
    node blocks Ld (n x 3 x 3) and edge blocks Le (one per directed edge)."""
    oc = np.clip(O[:, :3], -EXP_CLAMP, EXP_CLAMP)
    lam = np.zeros((g.n, 3, 3))
    lam[:, 0, 0] = np.exp(oc[:, 0])
    lam[:, 1, 1] = np.exp(oc[:, 1])
    lam[:, 2, 2] = np.exp(oc[:, 2])
    lam[:, 1, 0] = O[:, 3]
    lam[:, 2, 0] = O[:, 4]
    lam[:, 2, 1] = O[:, 5]
    lam_e = (O[g.esrc, 6:15] + O[g.edst, 15:24] + g.Ahat.reshape(-1, 9) @ np.asarray(Wa).T).reshape(-1, 3, 3)
    CiT = g.Ci.transpose(0, 2, 1)
    return CiT @ lam, CiT[g.esrc] @ lam_e


def factor_matrix(g, fw):
    """L as a scipy sparse matrix (3n x 3n)."""
    n = g.n
    rows, cols, vals = [], [], []
    for (bi, bj, blk) in ((np.arange(n), np.arange(n), fw["Ld"]), (g.esrc, g.edst, fw["Le"])):
        for a in range(3):
            for b in range(3):
                rows.append(3 * bi + a)
                cols.append(3 * bj + b)
                vals.append(blk[:, a, b])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * n, 3 * n))


def apply(g, params, r, use_bf16=True):
    """z = L L^T r."""
    L = factor_matrix(g, forward(g, params, use_bf16))
    return L @ (L.T @ r)


def loss_and_grad(g, params, e, use_bf16=True):
    """Loss and gradient w.r.t. the flat parameter vector (FP64 arithmetic,
    with the device's BF16 rounding points when use_bf16)."""
    rnd = bf16 if use_bf16 else (lambda v: v)
    fw = forward(g, params, use_bf16)
    W0, b0, Wl, bl, Wh, bh, Wa = fw["W"]
    H, A, O = fw["H"], fw["A"], fw["O"]
    L = factor_matrix(g, fw)
    K = g.K
    r = K @ e
    t = L.T @ r
    z = L @ t
    q = K @ z
    c, a, b = float(e @ r), float(z @ r), float(z @ q)
    loss = 1.0 - a * a / (b * c)
    alpha = -2.0 * a / (b * c)       # dloss/da
    beta = 2.0 * a * a / (b * b * c)  # -2 dloss/db
    u = L.T @ q
    # dloss/dL = g1 t^T + g2 u^T
    W3 = (2.0 * alpha * r + beta * q).reshape(-1, 3)
    R3 = (beta * r).reshape(-1, 3)
    T3, U3 = t.reshape(-1, 3), u.reshape(-1, 3)
    # dL blocks, chained through L = Ci^T Lam
    dLd = np.einsum('ia,ib->iab', W3, T3) + np.einsum('ia,ib->iab', R3, U3)
    dLe = np.einsum('ea,eb->eab', W3[g.esrc], T3[g.edst]) + np.einsum('ea,eb->eab', R3[g.esrc], U3[g.edst])
    dlam_d = g.Ci @ dLd
    dlam_e = g.Ci[g.esrc] @ dLe
    dO = np.zeros_like(O)
    live = (O[:, :3] > -EXP_CLAMP) & (O[:, :3] < EXP_CLAMP)
    for k in range(3):
        dO[:, k] = np.where(live[:, k], dlam_d[:, k, k] * np.exp(np.clip(O[:, k], -EXP_CLAMP, EXP_CLAMP)), 0.0)
    dO[:, 3] = dlam_d[:, 1, 0]
    dO[:, 4] = dlam_d[:, 2, 0]
    dO[:, 5] = dlam_d[:, 2, 1]
    de = dlam_e.reshape(-1, 9)
    np.add.at(dO[:, 6:15], g.esrc, de)
    np.add.at(dO[:, 15:24], g.edst, de)
    grad = np.zeros(NUM_PARAMS)
    grad[OFF_WA:NUM_PARAMS] = (de.T @ g.Ahat.reshape(-1, 9)).ravel()
    grad[OFF_WH:OFF_BH] = (dO.T @ H[-1]).ravel()
    grad[OFF_BH:OFF_WA] = dO.sum(axis=0)
    dH = dO @ Wh
    for l in range(LAYERS - 1, -1, -1):
        dz = rnd(np.where(H[l + 1] > 0, dH, 0.0))
        Xin = np.concatenate([H[l], A[l]], axis=1)
        off = OFF_L1 + l * LAYER_SIZE
        grad[off:off + HID * 2 * HID] = (dz.T @ Xin).ravel()
        grad[off + HID * 2 * HID:off + LAYER_SIZE] = dz.sum(axis=0)
        dX = dz @ Wl[l]
        dH = dX[:, :HID] + g.mean_adj.T @ dX[:, HID:]
    dz0 = np.where(H[0] > 0, dH, 0.0)
    grad[OFF_W0:OFF_B0] = (dz0.T @ g.X).ravel()
    grad[OFF_B0:OFF_L1] = dz0.sum(axis=0)
    return loss, grad


def loss_only(g, params, e, use_bf16=True):
    L = factor_matrix(g, forward(g, params, use_bf16))
    r = g.K @ e
    z = L @ (L.T @ r)
    c, a, b = float(e @ r), float(z @ r), float(z @ (g.K @ z))
    return 1.0 - a * a / (b * c)


def random_params(seed, head_scale=0.05):
    """Glorot-uniform layers plus a small random head and edge map, so that
    every gradient block is non-zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NUM_PARAMS)
    p[OFF_W0:OFF_B0] = rng.uniform(-1, 1, HID * FEAT) * np.sqrt(6.0 / (FEAT + HID))
    p[OFF_B0:OFF_L1] = rng.uniform(-0.1, 0.1, HID)
    for l in range(LAYERS):
        off = OFF_L1 + l * LAYER_SIZE
        p[off:off + HID * 2 * HID] = rng.uniform(-1, 1, HID * 2 * HID) * np.sqrt(6.0 / (3 * HID))
        p[off + HID * 2 * HID:off + LAYER_SIZE] = rng.uniform(-0.1, 0.1, HID)
    p[OFF_WH:OFF_BH] = rng.uniform(-1, 1, HEAD * HID) * head_scale
    p[OFF_BH:OFF_WA] = rng.uniform(-1, 1, HEAD) * head_scale
    p[OFF_WA:NUM_PARAMS] = rng.uniform(-1, 1, 81) * head_scale
    return p
