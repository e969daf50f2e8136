"""Pins the oracle's CMB chain against the reference's own fixtures where they can pin it, and
its dark-energy terms against the mathematics where no fixture exists (w_a != 0).

The reference's CMB fixtures (tests/test_cmb/test_benchmark_*.csv) come from its batched
shared-step run of DISCO-Notebook (128 log-spaced k in 1e-4..1, 320 output times, rtol = atol =
1e-4, batches of 16).  They predate the radiation-streaming (RSA) and ultra-relativistic-fluid
(UFA) switches and the spline-derivative ISW term: with the switches off and S1 = e^-tau_c
(eta' + alpha'') the oracle reproduces the states, metric and polarisation columns at k0 and
k63 to 1e-7..2e-6.  CPU only.
"""
import numpy as np
import pytest
import scipy.integrate

from conftest import disco_notebook
from oracle import cmb_oracle as C

LM = dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3)
LMAX = [11, 11, 11, 8, 3]


@pytest.fixture(scope="module")
def fixture_run(oracle):
    """The fixtures' configuration with RSA and UFA off, the oracle's batched solver."""
    P = dict(disco_notebook(), use_rsa=False, use_ur_fluid=False)
    bg = oracle.Background(P, 1024)
    aexp = np.concatenate([oracle.geomspace(3e-4, 5e-3, 256, endpoint=False), oracle.geomspace(5e-3, 1.0, 64)])
    k = oracle.geomspace(1e-4, 1.0, 128)
    y, tau, st, na, nr = bg.evolve(k, aexp, **LM, rtol=1e-4, atol=1e-4, return_full=True, batch_size=16)
    assert (st == 0).all()
    yp = np.stack([np.stack([bg.rhs(LMAX, tau[j], y[i, j], k[i]) for j in range(tau.size)]) for i in range(k.size)])
    q, w = oracle.nu_bins(3)
    sc = dict(P, **{s: bg.scalar(s) for s in ["grhom", "grhog", "grhor", "OmegaDE", "amnu"]})
    bgd = dict(sc, rhonu=np.exp(bg.spline("logrhonu_of_loga", np.log(aexp))), tau_tab=bg.array("tau"), opac=bg.array("opac"),
               gvis=bg.array("gvis"), gvisprime=bg.array("gvisprime"), gvispprime=bg.array("gvispprime"),
               optical_depth=bg.array("optical_depth"), q=q, w=w)
    tau0 = float(bg.spline("tau_of_a", np.array([1.0]))[0])
    return dict(P=P, bg=bg, aexp=aexp, k=k, y=y, yp=yp, tau=tau, bgd=bgd, tau0=tau0)


def _err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_states_metric_and_polarisation_match_fixtures(fixture_run, golden):
    """tests/test_cmb/test_benchmark_{extract_perturbations, compute_metric_perturbations,
    compute_polarization_terms}.csv at k0 and k63 to 1e-5 (measured 1e-10..2e-6).  k127 sits in the
    last batch, whose step sequence parts from the reference's at a rounding-level accept decision
    (~160 steps): checked at the reference's own regression tolerance 0.2."""
    r, g = fixture_run, golden("cmb_chain")
    pt = C.extract_perturbations(r["y"], r["yp"], 11, 11, 11)
    o = C.sources_from_states(r["y"], r["yp"], r["k"], r["aexp"], r["tau"], r["bgd"], 11, 11, 11, 3, isw="algebraic")
    cols = {"deltag": pt["deltag"], "deltac": pt["deltac"], "thetag": pt["thetag"], "deltab": pt["deltab"], "deltar": pt["deltar"],
            "alpha": o["metric"]["alpha"], "polarization_term": o["pol"]["polarization_term"]}
    for name, arr in cols.items():
        for ik in (0, 63, 127):
            key = f"{name}_k{ik}"
            if key not in g:
                continue
            e = _err(arr[ik], g[key])
            print(f"  {key:24s} {e:.2e}")
            assert e < (1e-5 if ik < 127 else 0.2), key
    assert _err(o["metric"]["alphaprime"][0], g["alphaprime_k0"]) < 1e-5


def test_theta_ell_and_cell_with_algebraic_isw(fixture_run, golden):
    """Theta_l(k_min) and C_l (l <= 100) of tests/test_cmb/test_benchmark_compute_{theta_ell, Cell}.csv.
    With S1 = e^-tau_c (eta' + alpha'') Theta_l(k_min) matches to 2.5e-7 for l >= 10 and 1.2e-6 at
    l = 5; the monopole and dipole (l = 0, 1: 8.8e-3, 1.1e-3) and l = 2 (3.2e-4) carry the rest.
    C_l agrees to a median 8.6e-4 (max 3.1e-3): C_l at l <= 100 is governed by the modes of
    k ~ 1e-3..3e-2 (batches 2-4), of which only k63 has fixtures (reproduced to 5e-9); the oracle's
    own C_l moves by <= 2.5e-7 under a 1e-13 change of its inputs, so this is not step-sequence
    noise of the run; its source lies in modes no fixture covers.  Theta_l(k_max) (k127, the
    last batch) to the reference's 0.2.  The spline-derivative S1 of the current reference does not
    reproduce these fixtures (C_l off by ~0.1): they predate it."""
    r, g = fixture_run, golden("cmb_chain")
    res = {}
    for isw in ("algebraic", "spline"):
        o = C.sources_from_states(r["y"], r["yp"], r["k"], r["aexp"], r["tau"], r["bgd"], 11, 11, 11, 3, isw=isw)
        th, kk = C.compute_theta_ell(100, r["k"], r["tau"], o["S"], r["tau0"])
        res[isw] = (th, C.compute_Cell(th, kk, r["P"]["n_s"], r["P"]["k_p"]))
    th, cl = res["algebraic"]
    ref = g["theta_ell_k_min"]
    d = np.abs(th[0] - ref) / np.abs(ref).max()
    print(f"  Theta_l(k_min): l >= 10 {d[10:].max():.2e}, l = 5..9 {d[5:10].max():.2e}, l = 2 {d[2]:.2e}, l = 0, 1 {d[0]:.2e} {d[1]:.2e}")
    assert d[10:].max() < 1e-6 and d[5:10].max() < 5e-6 and d[2] < 1e-3 and d[:2].max() < 2e-2
    rel = np.abs(cl[2:] / g["Cell"][2:] - 1)
    print(f"  C_l: median {np.median(rel):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 2e-3 and rel.max() < 5e-3
    assert _err(th[127], g["theta_ell_k_max"]) < 0.2
    rel_s = np.abs(res["spline"][1][2:] / g["Cell"][2:] - 1)
    assert np.median(rel_s) > 0.05  # the fixtures were not made with the spline-derivative ISW term


def test_neutrino_fixture_columns_are_not_from_the_alpha_run(fixture_run, golden):
    """test_benchmark_compute_neutrino_perturbations.csv (delta rho_nu at k0, k63, k127, delta p_nu at
    k0) disagrees with the oracle by up to a few percent (ratio 0.98..1.06 at k0), while alpha --
    built from delta rho_nu through h' -- matches its own fixture to 7e-10.  Putting the fixture's
    delta rho_nu into the Einstein equations instead moves alpha at k0 by orders of magnitude more
    than that: the neutrino columns were written by another run of the reference than the metric
    columns, so they are checked only at the reference's own regression tolerance (0.2 at k0)."""
    r, g = fixture_run, golden("cmb_chain")
    y, yp, k, aexp = r["y"][:1], r["yp"][:1], r["k"][:1], r["aexp"]
    o = C.sources_from_states(y, yp, k, aexp, r["tau"], r["bgd"], 11, 11, 11, 3, isw="algebraic")
    own = _err(o["metric"]["alpha"][0], g["alpha_k0"])
    pt = C.extract_perturbations(y, yp, 11, 11, 11)
    bq = C.background_quantities(aexp, r["bgd"], r["bgd"]["rhonu"])
    nu = dict(o["nu"])
    nu["drhonu"] = g["drhonu_k0"][None, :]
    swapped = _err(C.metric_perturbations(pt, nu, bq, r["bgd"], k, aexp)["alpha"][0], g["alpha_k0"])
    print(f"  alpha(k0) vs fixture: own delta rho_nu {own:.1e}, fixture's delta rho_nu {swapped:.1e}")
    assert own < 1e-8 and swapped > 1e3 * own
    assert _err(o["nu"]["drhonu"][0], g["drhonu_k0"]) < 0.2 and _err(o["nu"]["dpnu"][0], g["dpnu_k0"]) < 0.2


# ---------------------------------------------------------------------------------------------
# w0wa dark energy: no reference fixture has wa != 0
# ---------------------------------------------------------------------------------------------
def test_conformal_time_with_cpl_dark_energy(oracle):
    """tau(a) - tau(a_0) against quadrature of dtau/da with rho_DE = a^(-3(1 + w0 + wa)) e^(3 (a - 1) wa)."""
    P = dict(disco_notebook(), Nmnu=0.0, mnu=0.0, Neff=3.046, w_DE_0=-0.95, w_DE_a=0.1)
    bg = oracle.Background(P, 256)
    a, tau = bg.array("aexp"), bg.array("tau")
    grhom, grhog, grhor, ode = bg.scalar("grhom"), bg.scalar("grhog"), bg.scalar("grhor"), bg.scalar("OmegaDE")
    w0, wa = P["w_DE_0"], P["w_DE_a"]

    def dtauda(x):
        rde = x ** (-3 * (1 + w0 + wa)) * np.exp(3 * (x - 1) * wa)
        return np.sqrt(3.0 / (grhom * P["Omegam"] * x + grhog + grhor * P["Neff"] + grhom * ode * rde * x ** 4))

    for i in [5, 60, 150, 230, 255]:
        ref = scipy.integrate.quad(dtauda, a[0], a[i], epsrel=1e-13, limit=500)[0]
        assert abs((tau[i] - tau[0]) - ref) <= 1e-9 * ref, i
    # and the wa term matters at this level: the same integral with wa = 0 misses by far more
    ref0 = scipy.integrate.quad(lambda x: np.sqrt(3.0 / (grhom * P["Omegam"] * x + grhog + grhor * P["Neff"] +
                                                        grhom * ode * x ** (-3 * (1 + w0)) * x ** 4)), a[0], a[-1], epsrel=1e-13, limit=500)[0]
    assert abs((tau[-1] - tau[0]) - ref0) > 1e-5 * ref0


def test_adiabatic_sound_speed_of_the_fluid(oracle):
    """c_a^2 = p_Q' / rho_Q' for rho_Q ~ a^(-3(1 + w0 + wa)) e^(3 (a - 1) wa), p_Q = w rho_Q, against the
    value the right-hand side uses.  The delta_Q row with only theta_Q = 1 set (h' = 0 then) is
    -(1 + w) - 9 (1 + w)(c_s^2 - c_a^2) aH^2 / k^2 (BL10 3.5), so c_a^2 can be read back from it;
    the reference regularises 1 / (1 + w) by +1e-6 (perturbations.py:521), which the comparison
    with the exact ratio resolves at w = -0.95 (a relative 2e-5 effect)."""
    P = dict(disco_notebook(), w_DE_0=-0.95, w_DE_a=0.1)
    bg = oracle.Background(P, 256)
    n = oracle.nvar(*LMAX)
    k = 0.01
    w0, wa = P["w_DE_0"], P["w_DE_a"]
    for a in [1e-3, 0.1, 0.5, 0.9]:
        tau = float(bg.spline("tau_of_a", np.array([a]))[0])
        y = np.zeros(n)
        y[0] = a
        y[n - 1] = 1.0
        f = bg.rhs(LMAX, tau, y, k)
        wq = w0 + wa * (1 - a)
        rhonu = np.exp(bg.spline("logrhonu_of_loga", np.array([np.log(a)]))[0])
        sc = dict(P, **{s: bg.scalar(s) for s in ["grhom", "grhog", "grhor", "OmegaDE"]})
        aH = C.aprimeoa(a, sc, rhonu)
        ca2_rhs = 1.0 - (-(f[n - 2] + (1 + wq)) / (9 * (1 + wq) * aH ** 2 / k ** 2))
        # exact: rho' = -3 aH (1 + w) rho, w' = -wa aH a  =>  c_a^2 = w - w' / (3 aH (1 + w))
        ca2_exact = wq - (-wa * aH * a) / (3 * aH * (1 + wq))
        ca2_ref = wq - (-wa * aH * a) / 3 / ((1 + wq) + 1e-6) / aH
        # numerical p' / rho' from the closed forms, by central differences in a
        rho = lambda x: x ** (-3 * (1 + w0 + wa)) * np.exp(3 * (x - 1) * wa)
        p = lambda x: (w0 + wa * (1 - x)) * rho(x)
        h = 1e-5 * a
        ca2_num = (p(a + h) - p(a - h)) / (rho(a + h) - rho(a - h))
        print(f"  a {a}: rhs {ca2_rhs:.12f} reference form {ca2_ref:.12f} exact {ca2_exact:.12f} numerical {ca2_num:.12f}")
        assert abs(ca2_num - ca2_exact) < 1e-8
        assert abs(ca2_rhs - ca2_ref) < 1e-9 * max(1.0, abs(ca2_ref))
        assert abs(ca2_ref - ca2_exact) > 1e-7 * abs(ca2_exact - wq)  # the +1e-6 is there, and resolved


@pytest.mark.parametrize("w0,wa", [(-0.95, 0.1), (-1.1, -0.1)])
def test_superhorizon_fluid_attractor(oracle, w0, wa):
    """Outside the horizon in the radiation era the c_s^2 = 1 fluid follows the attractor
    delta_Q / delta_c = (1 + w)(4 - 3 c_s^2) / (4 - 6 w + 3 c_s^2) (BL10 eqs. 3.5-3.7 with
    delta_c = -(k tau)^2 C / 4), w = w0 + wa (1 - a) ~ w0 + wa at a << 1.  The reference's initial
    delta_Q (perturbations.py:732) has the opposite sign to BL10 eq. 3.7 and CLASS, so the solution
    first relaxes onto the attractor; from a = 3e-7 the ratio minus its linear-in-a matter
    correction, extrapolated to a -> 0, must sit on it.  A dropped or mis-signed (c_s^2 - c_a^2),
    (1 + w) or h' term of the fluid equations moves the attractor by O(1)."""
    P = dict(disco_notebook(), w_DE_0=w0, w_DE_a=wa, Nmnu=0.0, mnu=0.0, Neff=3.046)
    bg = oracle.Background(P, 1024)
    aexp = np.concatenate([[1e-9], np.geomspace(3e-7, 3e-6, 6)])
    y, tau, st, _, _ = bg.evolve(np.array([1e-4]), aexp, lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=8, nqmax=3, rtol=1e-9, atol=1e-14,
                                 return_full=True)
    assert (st == 0).all() and (1e-4 * tau < 1e-2).all()
    n = y.shape[-1]
    w = w0 + wa * (1 - aexp)
    pred = (1 + w) / (4 - 6 * w + 3)
    r = y[0, :, n - 2] / y[0, :, 3]
    assert r[0] / pred[0] < 0  # the reference's initial sign
    slope, at0 = np.polyfit(aexp[1:], r[1:] / pred[1:], 1)
    print(f"  w0 {w0} wa {wa}: attractor at a -> 0 {at0 - 1:+.1e}, matter correction {slope:.0f} a")
    assert abs(at0 - 1) < 2e-3
    assert np.abs(r[1:] / pred[1:] - 1).max() < 2e-2
