"""Pins the CPU checker of the learned preconditioner (oracle/mpnn.py): its
backward pass against central finite differences (FP64, every parameter
group), the untrained network against the exact inverse of the 3x3 diagonal
blocks, symmetry and positive definiteness of P = L L^T, and the BF16
rounding against hand-computed bit patterns."""
import numpy as np
import pytest

import mpnn
from asend3d_amd import problems
from conftest import oracle_learned_graph


@pytest.fixture(scope="module")
def small():
    p = problems.heterogeneous(4, 3, 2, seed=5, lx=20.0, ly=15.0, lz=10.0)
    g, _ = oracle_learned_graph(p)
    rng = np.random.default_rng(1)
    e = rng.standard_normal(p.n_dof)
    e[np.repeat(g.X[:, 4] > 0, 3)] = 0.0
    return p, g, e


def test_bf16_rounding():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1.0
    vals = np.array([one, one + ulp / 2, one + 3 * ulp / 2, one + ulp / 4, -2.5, 3.0e38, 1e-40])
    got = mpnn.bf16(vals)
    # ties go to even: 1 + ulp/2 -> 1, 1 + 3ulp/2 -> 1 + 2ulp; below half rounds down
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == one + 2 * ulp and got[3] == 1.0
    assert got[4] == -2.5
    assert np.isclose(got[5], 3.0e38, rtol=2 ** -8)
    # FP32 subnormal: BF16 spacing there is 2^-133, and 1e-40 = 1.09 * 2^-133
    assert got[6] == 2.0 ** -133


@pytest.mark.parametrize("group", ["W0", "b0", "layers", "head", "Wa", "all"])
def test_backward_matches_finite_differences(small, group):
    _, g, e = small
    prm = mpnn.random_params(11)
    loss, grad = mpnn.loss_and_grad(g, prm, e, use_bf16=False)
    lo, hi = {"W0": (mpnn.OFF_W0, mpnn.OFF_B0), "b0": (mpnn.OFF_B0, mpnn.OFF_L1),
              "layers": (mpnn.OFF_L1, mpnn.OFF_WH), "head": (mpnn.OFF_WH, mpnn.OFF_WA),
              "Wa": (mpnn.OFF_WA, mpnn.NUM_PARAMS), "all": (0, mpnn.NUM_PARAMS)}[group]
    # fixed seeds (str hashes vary per process): a random direction can put a
    # ReLU kink inside +-h, which spoils the central difference
    rng = np.random.default_rng({"W0": 11, "b0": 12, "layers": 13, "head": 14, "Wa": 15, "all": 16}[group])
    for _ in range(3):
        d = np.zeros(mpnn.NUM_PARAMS)
        d[lo:hi] = rng.standard_normal(hi - lo)
        h = 1e-6
        fd = (mpnn.loss_only(g, prm + h * d, e, False) - mpnn.loss_only(g, prm - h * d, e, False)) / (2 * h)
        an = grad @ d
        assert abs(an) > 1e-8, "gradient block is identically zero"
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3), (group, fd, an)


def test_untrained_network_is_block_jacobi(small):
    p, g, _ = small
    prm = mpnn.random_params(4)
    prm[mpnn.OFF_WH:] = 0.0  # head and edge map zero
    r = np.random.default_rng(3).standard_normal(p.n_dof)
    z = mpnn.apply(g, prm, r)
    Bs = 0.5 * (g.B + g.B.transpose(0, 2, 1))
    z_ref = np.einsum("iab,ib->ia", np.linalg.inv(Bs), r.reshape(-1, 3)).ravel()
    assert np.abs(z - z_ref).max() <= 1e-13 * np.abs(z_ref).max()


def test_preconditioner_is_spd(small):
    p, g, _ = small
    L = mpnn.factor_matrix(g, mpnn.forward(g, mpnn.random_params(7, head_scale=0.3)))
    P = (L @ L.T).toarray()
    assert np.abs(P - P.T).max() <= 1e-12 * np.abs(P).max()
    assert np.linalg.eigvalsh(P).min() > 0.0


def test_edge_lists_are_symmetric(small):
    _, g, _ = small
    fwd = set(zip(g.esrc.tolist(), g.edst.tolist()))
    assert all((j, i) in fwd for i, j in fwd)
    # interior nodes of a structured hex grid have 6 edge neighbours
    cnt = np.bincount(g.esrc, minlength=g.n)
    assert cnt.max() == 6 and cnt.min() == 3
