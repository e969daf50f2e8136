"""Pins of the reionization oracle (oracle/reion_oracle.py): the tanh step in y = (1 + z)^(3/2)
integrates in closed form against a matter-only background, the root solve inverts the optical
depth, and the limits of the ionisation fraction."""
import numpy as np
import pytest

from oracle import reion_oracle as R


@pytest.mark.parametrize("zre", [6.0, 7.7, 12.0, 30.0])
def test_matter_only_closed_form(zre):
    H0, akthom = 67.4 / 2.99792458e5, 2.3038921e-9 * 0.752 * 0.0493 * 67.4 ** 2  # H0 [1/Mpc], n_H sigma_T a^2 [1/Mpc]
    dtauda = lambda a: a ** -0.5 / H0
    got = R.added_depth(zre, 0.0, akthom, dtauda, lambda la: 0.0 * la)
    ref = R.matter_only_depth(zre, akthom, H0)
    assert abs(got / ref - 1.0) < 1e-10
    # a dropped 1/a (integrand in a instead of ln a) or a wrong power of y fails this
    bad = R.added_depth(zre, 0.0, akthom, lambda a: a ** -1.5 / H0, lambda la: 0.0 * la)
    assert abs(bad / ref - 1.0) > 1e-2


def test_root_solve_inverts_the_depth():
    H0, akthom, fHe = 67.4 / 2.99792458e5, 2.3038921e-9 * 0.752 * 0.0493 * 67.4 ** 2, 0.0823
    dtauda = lambda a: a ** -0.5 / H0
    xrec = lambda la: 2e-4 + 0.0 * la
    for tau in (0.03, 0.054, 0.09):
        z = R.solve_zre(tau, fHe, akthom, dtauda, xrec)
        assert abs(R.added_depth(z, fHe, akthom, dtauda, xrec) - tau) < 1e-12


def test_limits():
    fHe = 0.0823
    # today: hydrogen and both helium ionisations complete (up to tanh(7) for the second)
    x = 2e-4 + R.dxe(1.0, 2e-4, 7.7, fHe)
    assert abs(x - (1.0 + 2.0 * fHe)) < 1e-6
    # well before reionization: nothing added
    assert abs(R.dxe(1.0 / 101.0, 2e-4, 7.7, fHe)) < 1e-300
    # at z_re the hydrogen step is half way
    yre, dy, fre = R.reion_params(7.7, fHe)
    assert abs(R.dxe(1.0 / 8.7, 0.0, 7.7, 0.0) - 0.5) < 1e-12


def test_a_derivative_by_differences():
    fHe = 0.0823
    for a in (0.08, 1.0 / 8.7, 0.2, 0.5):
        d, _ = R.dxe_da(a, 2e-4, 7.7, fHe)
        h = 1e-6 * a
        num = (R.dxe(a + h, 2e-4, 7.7, fHe) - R.dxe(a - h, 2e-4, 7.7, fHe)) / (2 * h)
        assert abs(d - num) < 1e-7 * max(1.0, abs(num))
