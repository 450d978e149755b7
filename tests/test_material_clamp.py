"""Host-side check of isotropic materials against the pivot clamp of the
reference's compliance inversion (q_rfactor, S/matrix_functions.rs:167-215:
pivots below 1e-12 become 1e-12 before each Givens rotation; the last column
is never checked).  CPU only.

Pinned three ways: by the closed-form thresholds of the two matrices (the
smallest pivot the QR checks), by the oracle's own clamped inverse (the
check must fire exactly where the oracle's stiffness stops scaling with E),
and by per-element bookkeeping (count, first label)."""
import numpy as np
import pytest

import oracle as O
from asend3d_amd import api

NUS = [0.0, 0.1, 0.2, 0.3, 0.45, 0.49]


def solid_threshold(nu):
    # 6x6 compliance [[S_n, 0], [0, 2(1+nu)/E I]], S_n = (1/E)[[1,-nu,-nu],...]:
    # checks see 1/E (column 0, first rotation), then for column 2 (against
    # the zero rows 3..5) the last pivot |r22| of the QR of S_n, which is
    # det(S_n) / |c0 x c1| = (1+nu)(1-2nu) / sqrt(2nu^2 + (1-nu)^2) / E;
    # column 1 checks (1-nu^2)/sqrt(1+nu^2)/E and |r11|/E, columns 3, 4 see
    # 2(1+nu)/E, and column 5 is not checked.
    r22 = (1 + nu) * (1 - 2 * nu) / np.sqrt(2 * nu * nu + (1 - nu) ** 2)
    c1 = (1 - nu * nu) / np.sqrt(1 + nu * nu)
    return 1e12 * min(1.0, r22, c1)


def shell_threshold(nu):
    # 3x3 plane-stress compliance: column 0 checks 1/E and sqrt(1+nu^2)/E,
    # column 1 checks (1-nu^2)/sqrt(1+nu^2)/E, column 2 is not checked
    return 1e12 * min(1.0, (1 - nu * nu) / np.sqrt(1 + nu * nu))


@pytest.mark.parametrize("nu", NUS)
def test_threshold_closed_form(nu):
    for etype, thr in ((8, solid_threshold(nu)), (81, solid_threshold(nu)), (api.SHELL4, shell_threshold(nu))):
        assert api.compliance_clamp_check(etype, [thr * (1 - 1e-9)], nu) == (0, -1)
        assert api.compliance_clamp_check(etype, [thr * (1 + 1e-9)], nu) == (1, 0)


def test_threshold_nu_zero_is_exact():
    # nu = 0: the compliance is diagonal, every check sees 1/E or 2/E, so the
    # clamp fires exactly when the double 1/E is below 1e-12
    for E in (1e12, np.nextafter(1e12, 0), np.nextafter(1e12, 2e12), 1.0000000001e12, 0.9999999999e12):
        expect = int(1.0 / E < 1e-12)
        assert api.compliance_clamp_check(8, [E], 0.0)[0] == expect
        assert api.compliance_clamp_check(api.SHELL4, [E], 0.0)[0] == expect


def _oracle_solid_scales(E, nu):
    # without the clamp C(E) = E C(1) to rounding
    return np.allclose(O.solid_cmat(E, nu), E * O.solid_cmat(1.0, nu), rtol=1e-11, atol=1e-11 * E)


def _oracle_shell_scales(E, nu):
    X = np.array([[0.0, 1.0, 1.0, 0.0], [0.0, 0.0, 1.0, 1.0], [0.0, 0.0, 0.0, 0.0]])
    K1, _ = O.shell4_element(X, 1.0, nu, 0.01, 0.0)
    K, _ = O.shell4_element(X, E, nu, 0.01, 0.0)
    return np.allclose(K, E * K1, rtol=1e-10, atol=1e-10 * E * np.abs(K1).max())


@pytest.mark.parametrize("nu", NUS)
def test_check_agrees_with_oracle_inverse(nu):
    # log sweep across the threshold: the check fires exactly where the
    # oracle's (clamped) stiffness stops being E times the unit-modulus one
    for E in np.geomspace(1e10, 1e14, 41) * 1.0137:
        fires = api.compliance_clamp_check(8, [E], nu)[0] == 1
        assert fires == (not _oracle_solid_scales(E, nu)), (E, nu)
        fires = api.compliance_clamp_check(api.SHELL4, [E], nu)[0] == 1
        assert fires == (not _oracle_shell_scales(E, nu)), (E, nu)


def test_counts_and_first_label():
    rng = np.random.default_rng(3)
    E = rng.uniform(1e9, 1.5e11, 1000)  # the nu = 0.45 threshold is 1.72e11
    nu = rng.uniform(0.0, 0.45, 1000)
    assert api.compliance_clamp_check(8, E, nu) == (0, -1)
    bad = [17, 400, 999]
    E[bad] = 5e12
    assert api.compliance_clamp_check(8, E, nu) == (3, 17)
    # mm-N-MPa moduli of any realistic material are far below the threshold
    assert api.compliance_clamp_check(8, E * 1e-6, nu) == (0, -1)
    assert api.compliance_clamp_check(8, [], []) == (0, -1)
    with pytest.raises(api.AsendError):
        api.compliance_clamp_check(3, E, nu)
