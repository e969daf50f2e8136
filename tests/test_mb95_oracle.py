"""Pins the oracle's MB95 thermal history (oracle/mb95_oracle.py, a restatement of
thermodynamics_mb95.py) where physics and numerics fix it; the reference has no MB95 tests or
fixtures.  CPU only."""
import numpy as np
import pytest

from conftest import disco_notebook
from oracle import mb95_oracle as M


@pytest.fixture(scope="module")
def setup(oracle):
    P = disco_notebook()
    bg = oracle.Background(P, 1024)
    sc = dict(P, **{s: bg.scalar(s) for s in ["grhom", "grhog", "grhor", "OmegaDE", "adotrad", "taumin", "taumax"]})
    lr = lambda x: float(bg.spline("logrhonu_of_loga", np.array([x]))[0])
    hist = {N: M.thermal_history(sc, lr, N, sc["taumin"], sc["taumax"]) for N in (1024, 4096, 16384)}
    return P, bg, sc, hist


def test_helium_saha_fixed_point():
    """At the fixed point of the helium iteration x1 xe / (1 - x1 - x2) = r1 (singly ionised,
    checked where helium is partly neutral) and x2 xe / x1 = r2 (doubly ionised, checked where it
    is partly singly ionised), with xe = x0 + c (x1 + 2 x2)."""
    YHe, H0, Ob = 0.248, 67.742, 0.0488911
    c = 0.25 * YHe / (1 - YHe)
    for T, a, x0, eq in [(8e3, 3.4e-4, 0.999, 1), (6.5e3, 4.2e-4, 0.99, 1), (2e4, 1.4e-4, 1.0, 2), (1.6e4, 1.7e-4, 1.0, 2)]:
        x1, x2 = M.ionHe(T, a, x0, 0.0, 1.0, YHe, H0, Ob, n_iter=200)
        r1, r2 = M.saha_rhs(T, a, YHe, H0, Ob)
        xe = x0 + c * (x1 + 2 * x2)
        if eq == 1:
            assert 1 - x1 - x2 > 1e-4 and abs(x1 * xe / (1 - x1 - x2) / r1 - 1) < 1e-10
        else:
            assert x1 > 1e-4 and abs(x2 * xe / x1 / r2 - 1) < 1e-12


def test_baryon_temperature_limits(setup):
    """Tightly coupled early (T_b = T_cmb / a while Thomson cooling dominates) and adiabatic late
    (T_b ~ a^-2, slowed slightly by the residual Compton heating)."""
    P, bg, sc, hist = setup
    h = hist[16384]
    early = h["a"] < 1e-5
    assert np.abs(h["tb"][early] * h["a"][early] / P["Tcmb"] - 1).max() < 1e-8
    i, j = np.searchsorted(h["a"], [0.05, 0.1])
    assert -2.02 < np.log(h["tb"][j] / h["tb"][i]) / np.log(h["a"][j] / h["a"][i]) < -1.9
    # fully ionised hydrogen and doubly ionised helium at the start
    assert h["xe"][0] == pytest.approx(1 + 0.5 * P["YHe"] / (1 - P["YHe"]), rel=1e-14)


def test_convergence_in_the_number_of_points(setup):
    """The scan converges: x_e through recombination changes by at least four times less from
    4096 to 16384 points than from 1024 to 4096 (measured 6.8..13)."""
    _, _, _, hist = setup
    for z in (1100, 1000, 800):
        v = [np.interp(-np.log(1 + z), np.log(hist[N]["a"]), hist[N]["xe"]) for N in (1024, 4096, 16384)]
        assert abs(v[0] - v[1]) > 4 * abs(v[1] - v[2]), z


def test_recombination_against_recfast(setup):
    """MB95 is the three-level (Peebles) model; RECFAST adds the fudged case-B rates.  Their x_e
    agree to 4% through recombination (z = 1000..1400: 1.7..3.5% measured) and to 15% through
    freeze-out (z = 100..800); dropping the Peebles factor or the helium Saha step moves MB95 by
    far more."""
    _, bg, _, hist = setup
    h = hist[16384]
    ar, xr = bg.array("aexp"), bg.array("xe")
    for z, tol in [(1400, 0.04), (1200, 0.04), (1100, 0.04), (1000, 0.04), (800, 0.15), (500, 0.15), (100, 0.15)]:
        a = np.log(1 / (1 + z))
        m, r = np.interp(a, np.log(h["a"]), h["xe"]), np.interp(a, np.log(ar), xr)
        assert abs(m / r - 1) < tol, (z, m, r)


def test_grid_ends_at_the_reference_tau_max(setup):
    """The scan runs to taumax = taumin + Romberg of dtau/da over [1e-9, 1.01] (background.py:
    181-187), whose 1025 linear points put the radiation-matter transition inside the first panel:
    taumax exceeds tau(1.01) by 9%, so MB95's last point lies near a = 1.38 (the splines then
    cover a = 1 with room to spare)."""
    _, bg, sc, hist = setup
    t101 = float(bg.spline("tau_of_a", np.array([1.01]))[0])
    assert 1.05 < (sc["taumax"] - sc["taumin"]) / t101 < 1.12
    assert 1.3 < hist[16384]["a"][-1] < 1.45 and hist[16384]["tau"][-1] == pytest.approx(sc["taumax"], rel=1e-12)
