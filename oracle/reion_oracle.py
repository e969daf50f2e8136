"""Oracle (test infrastructure only; never imported by the product path): the reionization
extension of the thermal history, restated in numpy/scipy.

The reference's thermal histories have no reionization (its 9-parameter sampling lists tau_reion,
but nothing reads it).  The extension follows the tanh model of Lewis (2008, Phys. Rev. D 78,
023002) as CAMB uses it:

  x_e = x_rec + (f_re - x_rec) F_H + f_He F_He,
  F_H = (1 + tanh((y_re - y) / dy)) / 2,  y = (1 + z)^(3/2),  y_re = (1 + z_re)^(3/2),
  dy = (3/2) (1 + z_re)^(1/2) dz,  dz = 0.5,  f_re = 1 + f_He,
  F_He = (1 + tanh((3.5 - z) / 0.5)) / 2  (the second helium ionisation),

with z_re chosen so that the optical depth added to the recombination history,
akthom int_{1/151}^{1} (x_e - x_rec) / a^2 dtau/da da, equals the requested tau_reion.

Pinned by: the closed form of that integral for a matter-only background (dtau/da = a^(-1/2) / H0,
x_rec = 0, f_He = 0): tau = (2/3) (akthom / H0) (dy / 2) [u1 + ln(2 cosh u1)], u1 = (y_re - 1) / dy,
since the tanh step in y integrates exactly (tests/test_reion_oracle.py).
"""
import numpy as np
from scipy import integrate, optimize
from scipy.interpolate import CubicSpline

Z_HE, DZ_HE, DZ = 3.5, 0.5, 0.5
LOGA_MIN = -np.log(151.0)


def reion_params(zre, fHe):
    yre = (1.0 + zre) ** 1.5
    dy = 1.5 * np.sqrt(1.0 + zre) * DZ
    return yre, dy, 1.0 + fHe


def dxe(a, xrec, zre, fHe):
    """x_e - x_rec of the tanh model at scale factor(s) a."""
    yre, dy, fre = reion_params(zre, fHe)
    a = np.asarray(a, dtype=np.float64)
    y = a ** -1.5
    z = 1.0 / a - 1.0
    FH = 0.5 * (1.0 + np.tanh((yre - y) / dy))
    FHe = 0.5 * (1.0 + np.tanh((Z_HE - z) / DZ_HE))
    return (fre - xrec) * FH + fHe * FHe


def dxe_da(a, xrec, zre, fHe):
    """d/da of dxe at fixed x_rec, and the hydrogen fraction F_H (the table's dx_e/dtau is
    x_rec' (1 - F_H) + dxe_da / (dtau/da))."""
    yre, dy, fre = reion_params(zre, fHe)
    a = np.asarray(a, dtype=np.float64)
    th = np.tanh((yre - a ** -1.5) / dy)
    te = np.tanh((Z_HE - (1.0 / a - 1.0)) / DZ_HE)
    d = (fre - xrec) * 0.5 * (1.0 - th * th) * 1.5 * a ** -2.5 / dy + fHe * 0.5 * (1.0 - te * te) / (DZ_HE * a * a)
    return d, 0.5 * (1.0 + th)


def added_depth(zre, fHe, akthom, dtauda, xrec_of_loga):
    """akthom int (x_e - x_rec) / a^2 dtau/da da over z in [0, 150], by adaptive quadrature in ln a
    (split at the transitions so that quad resolves them)."""
    def f(la):
        a = np.exp(la)
        return dxe(a, xrec_of_loga(la), zre, fHe) * akthom / a * dtauda(a)

    pts = sorted({-np.log1p(zre), -np.log1p(Z_HE), LOGA_MIN, 0.0})
    pts = [p for p in pts if LOGA_MIN <= p <= 0.0]
    return sum(integrate.quad(f, lo, hi, limit=400, epsabs=0.0, epsrel=1e-12)[0] for lo, hi in zip(pts[:-1], pts[1:]))


def solve_zre(tau_reion, fHe, akthom, dtauda, xrec_of_loga):
    """z_re in [0, 100] whose added optical depth is tau_reion (Brent)."""
    g = lambda z: added_depth(z, fHe, akthom, dtauda, xrec_of_loga) - tau_reion
    return optimize.brentq(g, 0.0, 100.0, xtol=1e-13, rtol=1e-13)


def xrec_spline(loga, xe):
    """The recombination x_e as the natural cubic spline in ln a of its table (as the GPU builds it)."""
    s = CubicSpline(np.asarray(loga), np.asarray(xe), bc_type="natural")
    return lambda la: s(la)


def matter_only_depth(zre, akthom, H0):
    """Closed form of the hydrogen part for dtau/da = a^(-1/2) / H0, x_rec = 0, f_He = 0."""
    yre, dy, _ = reion_params(zre, 0.0)
    u1 = (yre - 1.0) / dy
    return (2.0 / 3.0) * (akthom / H0) * 0.5 * dy * (u1 + np.log(2.0 * np.cosh(u1)))
