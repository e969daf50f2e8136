"""Pins the oracle's large-scale-structure functions (oracle/lss_oracle.py) where the
mathematics fixes them; the reference has no tests or fixtures for these functions.  CPU only."""
import numpy as np
import pytest
import scipy.integrate
import scipy.signal
import scipy.special

from oracle import lss_oracle as L

PARAM = dict(A_s=2.1e-9, n_s=0.965, k_p=0.05)


@pytest.mark.parametrize("w", [5, 8, 39, 101])
@pytest.mark.parametrize("p", [2, 3, 4])
def test_savgol_coefficients_match_scipy(w, p):
    """util.py:291-328 builds the taps scipy.signal.savgol_coeffs documents (deriv 0, 'conv' use).
    The oracle's exact taps agree with scipy's floating-point least squares to the latter's
    rounding, which grows with the unscaled Vandermonde matrix's conditioning: 1e-15 for short
    windows, 3e-14 at 39 points, 5e-12 at 39 points order 4, 8e-10 at 101 points order 4."""
    c = L.savgol_coeffs(w, p)
    tol = {4: 1e-9}.get(p, 1e-12) if w >= 39 else 1e-13
    np.testing.assert_allclose(c, scipy.signal.savgol_coeffs(w, p), rtol=0, atol=tol * np.abs(c).max())
    assert abs(c.sum() - 1.0) < 1e-14  # a constant passes unchanged


def test_savgol_filter_reproduces_cubics():
    """A cubic filter returns a cubic unchanged wherever the window fits inside the data; near the
    ends the 'same' convolution sees zeros."""
    x = np.linspace(-2, 3, 200)
    y = 0.3 * x ** 3 - x ** 2 + 2 * x - 5
    out = L.savgol_filter(y, 21, 3)
    np.testing.assert_allclose(out[10:-10], y[10:-10], rtol=1e-12, atol=1e-12)
    assert abs(out[0] - y[0]) > 1e-3


def test_kaiser_legendre_projection_gives_the_multipoles():
    """(2l + 1)/2 int P(k, mu) L_l(mu) d mu of the Kaiser spectrum (degree 4 in mu, so 5-point
    Gauss-Legendre is exact) equals power_multipoles for l = 0, 2, 4, for any delta and theta."""
    rng = np.random.default_rng(0)
    k = np.geomspace(1e-3, 1.0, 17)
    y = rng.standard_normal((17, 20))
    b = 1.7
    x, wq = np.polynomial.legendre.leggauss(5)
    amp = np.sqrt(2 * np.pi ** 2 * PARAM["A_s"] * (k / PARAM["k_p"]) ** (PARAM["n_s"] - 1) * k ** -3.0)
    pkmu = (b * amp[:, None] * y[:, 4:5] - x[None, :] ** 2 * amp[:, None] * y[:, 5:6]) ** 2
    proj = [(2 * l + 1) / 2 * (pkmu * scipy.special.eval_legendre(l, x)[None, :] * wq[None, :]).sum(axis=1) for l in (0, 2, 4)]
    for got, ref in zip(L.power_multipoles(y, k, b, PARAM), proj):
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    P, mu = L.power_kaiser(y, k, b, 7, PARAM)
    np.testing.assert_allclose(mu, np.linspace(-1, 1, 7))
    np.testing.assert_allclose(P[:, 3], (b * amp * y[:, 4]) ** 2, rtol=1e-14)  # mu = 0: the tracer's own spectrum


def test_smoothed_power_keeps_a_power_law_and_its_ends():
    """log P linear in log k is a polynomial of degree 1 in the window index: the cubic filter keeps
    it; the end points keep the unsmoothed P (first w // 2, last (w + 1) // 2)."""
    k = np.geomspace(1e-4, 1.0, 300)
    rng = np.random.default_rng(1)
    y = np.ones((300, 20))
    y[:, 4] = k ** 1.3 * (1 + 0.02 * rng.standard_normal(300))
    w = L.smoothing_window(k, 0.5)
    assert w % 2 == 1
    ps = L.get_power_smoothed(k, y, 0.5, 4, PARAM)
    p = L.get_power(k, y, 4, PARAM)
    np.testing.assert_array_equal(ps[:w // 2], p[:w // 2])
    np.testing.assert_array_equal(ps[-((w + 1) // 2):], p[-((w + 1) // 2):])
    y[:, 4] = k ** 1.3
    ps = L.get_power_smoothed(k, y, 0.5, 4, PARAM)
    np.testing.assert_allclose(ps, L.get_power(k, y, 4, PARAM), rtol=1e-12)


def _xi_quad(P, r, ell):
    return np.array([scipy.integrate.quad(lambda kk: kk ** 2 * P(kk) * scipy.special.spherical_jn(ell, kk * x), 0, 6.0, limit=4000,
                                          epsabs=0, epsrel=1e-12)[0] for x in r]) / (2 * np.pi ** 2)


@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
def test_fftlog_correlation_function_converges_to_quadrature():
    """xi_l = Re(i^l) int k^2 P(k) j_l(k r) dk / (2 pi^2) for P = k^0.5 exp(-(8 k)^2): the FFTLog
    result approaches direct quadrature at first order in 1/N (the r = 2 pi / k assignment is half
    a grid step off the transform's own radii) -- 8.7e-4 at N = 4096, 2.2e-4 at N = 16384 for the
    monopole -- for l = 0, 2, 4.  A wrong phase, Gamma argument or normalisation does not
    converge."""
    P = lambda k: k ** 0.5 * np.exp(-(8.0 * k) ** 2)
    errs = {}
    for N in (4096, 16384):
        k = np.geomspace(1e-5, 1e3, N)
        for ell in (0, 2, 4):
            xi, r = L.get_xi_from_P(k, P(k), ell)
            idx = np.where((r > 2) & (r < 100))[0][::N // 64]
            ref = np.real(1j ** ell) * _xi_quad(P, r[idx], ell)
            errs[N, ell] = np.abs(xi[idx] - ref).max() / np.abs(ref).max()
    print({key: f"{v:.2e}" for key, v in errs.items()})
    for ell in (0, 2, 4):
        assert errs[4096, ell] < 1e-3 and errs[16384, ell] < 2.6e-4
        assert 3.5 < errs[4096, ell] / errs[16384, ell] < 4.5
