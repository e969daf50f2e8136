"""Pins the CPU oracle against the reference's golden data, closed forms and invariants.

Each check is chosen so that a plausible restatement error (a dropped term, a wrong index,
a wrong tableau entry, a wrong controller exponent) fails it.
"""
import numpy as np
import pytest
import scipy.integrate
import scipy.interpolate
import scipy.special

from conftest import disco_notebook

LMAX_DEFAULT = [11, 11, 11, 8, 3]


# ---------------------------------------------------------------------------------------
# background (reference fixture: tests/test_cmb/test_benchmark_evolve_background.csv)
# ---------------------------------------------------------------------------------------
def test_background_matches_reference_fixture(oracle, golden):
    g = golden("evolve_background")
    bg = oracle.Background(disco_notebook(), int(g["num_thermo"]))
    tau = bg.spline("tau_of_a", g["a"])
    exp_xe = np.exp(bg.spline("xe_of_loga", np.log(g["a"])))
    np.testing.assert_allclose(tau, g["tau"], rtol=1e-12)
    np.testing.assert_allclose(exp_xe, g["exp_xe"], rtol=1e-10)  # a=1e-3 sits inside recombination
    np.testing.assert_allclose(bg.scalar("grhom"), g["grhom"], rtol=1e-14)
    assert bg.scalar("recfast_maxstep_hits") == 0


def test_background_quantities_fixture(oracle, golden):
    """compute_background_quantities at the 320 output a of tests/test_cmb.py:150-153."""
    g = golden("background_quantities")
    bg = oracle.Background(disco_notebook(), 1024)
    a = np.concatenate([oracle.geomspace(3e-4, 5e-3, 256, endpoint=False), oracle.geomspace(5e-3, 1.0, 64)])
    rhonu = np.exp(bg.spline("logrhonu_of_loga", np.log(a)))
    np.testing.assert_allclose(rhonu, g["rhonu"], rtol=1e-12)
    P = disco_notebook()
    rho_Q = a ** (-3 * (1 + P["w_DE_0"] + P["w_DE_a"])) * np.exp(3 * (a - 1) * P["w_DE_a"])
    np.testing.assert_allclose(rho_Q, g["rho_Q"], rtol=1e-12)


def test_conformal_time_against_quadrature(oracle):
    """tau(a_i) - tau(a_0) is the integral of dtau/da (background.py:72-81) between the knots."""
    P = dict(disco_notebook(), Nmnu=0.0, mnu=0.0, Neff=3.046)
    bg = oracle.Background(P, 256)
    a = bg.array("aexp")
    tau = bg.array("tau")
    grhom, grhog, grhor = bg.scalar("grhom"), bg.scalar("grhog"), bg.scalar("grhor")
    ode = bg.scalar("OmegaDE")
    rhonu = np.exp(bg.spline("logrhonu_of_loga", [0.0]))[0]

    def dtauda(x):  # massless neutrinos: rhonu is constant
        g2 = grhom * P["Omegam"] * x + grhog + grhor * P["Neff"] + grhom * ode * x ** (-3 * (1 + P["w_DE_0"])) * x ** 4
        return np.sqrt(3.0 / g2)

    for i in [5, 40, 120, 200, 255]:
        ref = scipy.integrate.quad(dtauda, a[0], a[i], epsrel=1e-13, limit=500)[0]
        assert abs((tau[i] - tau[0]) - ref) <= 1e-9 * ref
    assert rhonu == pytest.approx(1.0, rel=2e-3)  # Fermi-Dirac normalisation with 8 GL nodes


def test_taumin_quirk(oracle):
    """tau[0] = taumin = amin/adotrad is attached to the first *end* point a_1 (recfast :453,489)."""
    bg = oracle.Background(disco_notebook(), 256)
    assert bg.array("tau")[0] == bg.scalar("taumin")
    assert bg.array("aexp")[0] > 1e-9 * 2.0


def test_recfast_v010_outside_recombination(oracle, golden):
    """DISCO-EB v0.1.0 x_e(a): agrees away from recombination; around it the v0.1.0 file was made by an
    older thermal-history code (the current reference's own fixture is matched to 1e-10 above)."""
    g = golden("recfast_v010")
    P = dict(disco_notebook(), Neff=2.046)
    bg = oracle.Background(P, 1024)
    a = g["a"]
    xe = bg.spline("xe_of_tau", bg.spline("tau_of_a", a))
    early = (a > 1e-8) & (a < 1e-4)  # fully ionised H and He: x_e = 1 + 2 f_He
    np.testing.assert_allclose(xe[early], g["xe"][early], rtol=2e-5)
    late = (a > 0.02) & (a < 1.0)  # freeze-out tail: the older code differs by ~1%
    np.testing.assert_allclose(xe[late], g["xe"][late], rtol=0.02)


# ---------------------------------------------------------------------------------------
# numerics building blocks
# ---------------------------------------------------------------------------------------
def test_spline_vs_scipy(oracle):
    rng = np.random.default_rng(4)
    x = np.sort(rng.uniform(0, 10, 40))
    y = np.sin(x) + 0.1 * x ** 2
    cs = scipy.interpolate.CubicSpline(x, y, bc_type="natural")
    xq = np.linspace(x[0], x[-1], 301)
    np.testing.assert_allclose(oracle.spline_eval(x, y, xq, 0), cs(xq), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(oracle.spline_eval(x, y, xq, 1), cs(xq, 1), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(oracle.spline_eval(x, y, xq, 2), cs(xq, 2), rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(oracle.spline_eval(x, y, xq, 3), cs.integrate(x[0], xq) if False else
                               np.array([cs.integrate(x[0], v) for v in xq]), rtol=1e-11, atol=1e-11)
    # the reference extrapolates with the end cubics (index clipped to [0, n-2])
    xe = np.array([-1.0, 11.0])
    np.testing.assert_allclose(oracle.spline_eval(x, y, xe, 0), cs(xe, extrapolate=True), rtol=1e-10)
    # uniform index path on a uniform grid gives identical results
    xu = np.linspace(0, 5, 30)
    yu = np.exp(-xu)
    np.testing.assert_allclose(oracle.spline_eval(xu, yu, xq / 2, 0, uniform=True), oracle.spline_eval(xu, yu, xq / 2, 0),
                               rtol=1e-13)


def test_gauss_laguerre_bins(oracle):
    """nq=8 Golub-Welsch nodes (util.py:83-124) and the Fermi-Dirac weights of background.py:40-45."""
    q, w = oracle.nu_bins(8)
    xr, wr = scipy.special.roots_genlaguerre(8, 1.0)
    np.testing.assert_allclose(q, xr, rtol=1e-12)
    fdc = 5.682196976983475
    np.testing.assert_allclose(w, wr * xr ** 3 / (1 + np.exp(-xr)) / xr / fdc, rtol=1e-10)
    # the CAMB 3-point table reweighted by -0.25 dlnf/dlnq
    q3, w3 = oracle.nu_bins(3)
    np.testing.assert_allclose(q3, [0.913201, 3.37517, 7.79184])
    dl = -q3 / (1 + np.exp(-q3))
    np.testing.assert_allclose(w3 * fdc * (-0.25 * dl), [0.0687359, 3.31435, 2.29911], rtol=1e-13)


def test_romberg(oracle):
    """romb divmax=6 is R(6,6): exact for polynomials up to degree 13, spectral for smooth f."""
    assert oracle.romb_test(2, 0.0, 1.7) == pytest.approx(1.7 ** 12 / 12, rel=1e-13)
    assert oracle.romb_test(0, 0.3, 1.1) == pytest.approx(np.exp(1.1) - np.exp(0.3), rel=1e-14)
    assert oracle.romb_test(1, 0.0, 1.0) == pytest.approx(np.pi / 4, rel=1e-12)


# ---------------------------------------------------------------------------------------
# perturbations
# ---------------------------------------------------------------------------------------
def _batched_fixture_run(oracle, golden, use_approx, batches):
    g = golden("evolve_perturbations_batched")
    P = disco_notebook()
    if not use_approx:
        P.update(use_rsa=False, use_ur_fluid=False)
    bg = oracle.Background(P, 1024)
    aout = np.concatenate([oracle.geomspace(3e-4, 5e-3, 256, endpoint=False), oracle.geomspace(5e-3, 1.0, 64)])
    k = oracle.geomspace(1e-4, 1.0, 64)
    np.testing.assert_allclose(k, g["kmodes"], rtol=1e-12)
    out = {}
    for b in batches:
        kb = k[16 * b:16 * (b + 1)]
        y, tau, st, na, nr = bg.evolve(kb, aout, rtol=1e-4, atol=1e-4, return_full=True, batch_size=16)
        assert (st == 0).all()
        out[b] = (y, tau, kb)
    return g, out


def _scaled_err(mine, ref):
    err = np.abs(mine - ref) / (np.abs(ref).max(axis=0) + 1e-300)
    err[:, 4] = 0.0  # theta_c is identically zero up to rounding in synchronous gauge
    return err


def test_batched_state_fixture_without_approximations(oracle, golden):
    """The reference's full-state fixture predates the RSA/UFA switches; with both off the batched
    Rodas5 + PI controller + linear SaveAt interpolation reproduce it (modes 0, 31, 63 of 64)."""
    g, out = _batched_fixture_run(oracle, golden, False, [0, 1, 3])
    assert _scaled_err(out[0][0][0], g["y_k0"]).max() < 1e-8
    assert _scaled_err(out[1][0][15], g["y_k31"]).max() < 1e-6
    assert _scaled_err(out[3][0][15], g["y_k63"]).max() < 5e-4  # 553 accepted steps at rtol 1e-4


def test_batched_state_fixture_with_approximations_before_trigger(oracle, golden):
    """With the default RSA/UFA on, the state agrees with the fixture until k*tau first exceeds the
    RSA trigger (80) for mode 31."""
    g, out = _batched_fixture_run(oracle, golden, True, [0, 1])
    y, tau, kb = out[1]
    m = kb[15] * tau < 79.0
    assert m.sum() > 300
    assert _scaled_err(y[15][m], g["y_k31"][m]).max() < 1e-6
    assert _scaled_err(out[0][0][0], g["y_k0"]).max() < 1e-8


def _state_at(oracle, bg, k, a_out, lmax):
    y, tau, st, na, nr = bg.evolve([k], [a_out], lmaxg=lmax[0], lmaxgp=lmax[1], lmaxr=lmax[2], lmaxnu=lmax[3],
                                   nqmax=lmax[4], rtol=1e-5, atol=1e-8, return_full=True)
    return y[0, 0], tau[0]


@pytest.mark.parametrize("k,a_out", [(0.05, 1e-3), (0.5, 2e-3), (0.3, 0.5)])
def test_jacobian_against_finite_differences(oracle, k, a_out):
    P = dict(disco_notebook(), w_DE_0=-0.95, w_DE_a=0.1)
    bg = oracle.Background(P, 256)
    lmax = [12, 12, 12, 17, 8]
    y, tau = _state_at(oracle, bg, k, a_out, lmax)
    J, dT = bg.jacobian(lmax, tau, y, k)
    n = y.size
    for j in range(n):
        h = 1e-4 * max(abs(y[j]), 1e-6) if j == 0 else 1.0
        yp = y.copy(); yp[j] += h
        ym = y.copy(); ym[j] -= h
        fd = (bg.rhs(lmax, tau, yp, k) - bg.rhs(lmax, tau, ym, k)) / (2 * h)
        scale = np.abs(J).max(axis=1) + 1e-300
        assert (np.abs(fd - J[:, j]) / scale).max() < 1e-6, j
    ht = 1e-6 * tau
    fdt = (bg.rhs(lmax, tau + ht, y, k) - bg.rhs(lmax, tau - ht, y, k)) / (2 * ht)
    assert np.abs(fdt - dT).max() <= 1e-6 * (np.abs(dT).max() + 1e-300)


def test_rhs_is_linear_in_state(oracle):
    """f(y + c e_j) - f(y) is linear in c for every j >= 1 (the RHS is linear in all but a)."""
    bg = oracle.Background(disco_notebook(), 256)
    lmax = [12, 12, 12, 8, 3]
    k = 0.2
    y, tau = _state_at(oracle, bg, k, 5e-3, lmax)
    f0 = bg.rhs(lmax, tau, y, k)
    for j in range(1, y.size, 5):
        e = np.zeros_like(y); e[j] = 1.0
        d1 = bg.rhs(lmax, tau, y + 1.0 * e, k) - f0
        d3 = bg.rhs(lmax, tau, y + 3.0 * e, k) - f0
        np.testing.assert_allclose(d3, 3 * d1, rtol=1e-9, atol=1e-9 * (np.abs(d1).max() + 1e-300))


def test_rodas5_local_order(oracle):
    """One Rodas5 step on the perturbation RHS: local error ~ dt^6 (order 5), against a Radau
    reference.  A wrong tableau entry, dT term or Jacobian column breaks the ratio."""
    P = dict(disco_notebook(), w_DE_0=-0.95, w_DE_a=0.1)
    bg = oracle.Background(P, 256)
    lmax = [8, 8, 8, 6, 3]
    k = 0.05
    y0, t0 = _state_at(oracle, bg, k, 2e-3, lmax)
    errs = []
    for dt in [t0 * 0.02, t0 * 0.01]:
        y1, _ = bg.rodas5_step(lmax, k, t0, dt, y0)
        sol = scipy.integrate.solve_ivp(lambda t, y: bg.rhs(lmax, t, y, k), (t0, t0 + dt), y0, method="Radau",
                                        rtol=1e-13, atol=1e-16, jac=lambda t, y: bg.jacobian(lmax, t, y, k)[0])
        ref = sol.y[:, -1]
        errs.append(np.abs(y1 - ref)[[2, 3, 5, 7]].max() / np.abs(ref[[2, 3, 5, 7]]).max())
    ratio = errs[0] / errs[1]
    assert ratio > 2 ** 4.5, (errs, ratio)


def test_initial_conditions_invariants(oracle):
    bg = oracle.Background(disco_notebook(), 256)
    lmax = LMAX_DEFAULT
    for k in [1e-4, 1e-2, 1.0]:
        t = 0.99 * bg.start_time(k)
        y = bg.ics(lmax, t, k)
        assert y[2] == pytest.approx(-1.0, abs=1e-3)  # superhorizon eta -> -1 (curvature_ini)
        assert y[3] == pytest.approx(0.75 * y[7], rel=1e-14)  # delta_c = 3/4 delta_gamma
        assert y[4] == 0.0  # theta_c = 0
        assert y[0] == pytest.approx(bg.spline("a_of_tau", [t])[0], rel=1e-14)


def test_start_time_rule(oracle):
    """determine_starting_time: large k starts when tau_H/tau_k = 0.07, small k when tau_c/tau_H = 4e-4."""
    bg = oracle.Background(disco_notebook(), 256)
    taumin = bg.scalar("taumin")
    tau1 = bg.spline("tau_of_a", [0.1])[0]
    ts = np.array([bg.start_time(k) for k in [1e-5, 1e-3, 1e-1, 10.0]])
    assert np.all(ts >= taumin) and np.all(ts <= tau1)
    assert np.all(np.diff(ts) <= 0)  # larger k starts earlier
    # bisection over log tau with 7 halvings: grid spacing (log tau1 - log taumin)/128
    step = (np.log(tau1) - np.log(taumin)) / 128
    frac = (np.log(ts) - np.log(taumin)) / step
    np.testing.assert_allclose(frac, np.round(frac - 0.5) + 0.5, atol=1e-9)


def test_power_spectrum_vs_class(oracle, golden):
    """End-to-end physics: P_bc(k, a=0.01) against CLASS (reference test bound 0.5%)."""
    g = golden("class_pk")
    m = (g["k"] >= 2e-3) & (g["k"] <= 0.5)
    k = g["k"][m][::5]
    bg = oracle.Background(disco_notebook(), 1024)
    y, tau, st, na, nr = bg.evolve(k, [0.01], lmaxg=31, lmaxgp=31, lmaxr=31, lmaxnu=31, nqmax=5, rtol=1e-5, atol=1e-5)
    assert (st == 0).all()
    Pk = oracle.get_power(k, y[:, -1, :], 6, disco_notebook())
    np.testing.assert_allclose(Pk, g["Pkbc"][m][::5], rtol=0.005)


def test_max_steps_reports_failure(oracle):
    bg = oracle.Background(disco_notebook(), 256)
    y, tau, st, na, nr = bg.evolve([0.1], [1.0], max_steps=10)
    assert st[0] == 1 and np.isnan(y).all()


def test_batch_size_must_divide(oracle):
    bg = oracle.Background(disco_notebook(), 256)
    with pytest.raises(ValueError):
        bg.evolve(oracle.geomspace(1e-3, 1e-1, 10), [1.0], batch_size=4)
