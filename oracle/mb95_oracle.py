"""CPU FP64 restatement of the MB95 thermal history (src/discoeb/thermodynamics_mb95.py): the
sequential scan over conformal time log-uniform in [taumin, taumax] with the semi-implicit
hydrogen step, helium Saha iterations and Thomson-coupled baryon temperature.  TEST
INFRASTRUCTURE: only tests/ use it, as the checker of the HIP path.

Pinning (tests/test_mb95_oracle.py): the Saha equations at the helium iteration's fixed point,
the tightly coupled and the adiabatic limits of T_b, first-order convergence of x_e in the number
of points, and the recombination history against the oracle's RECFAST (MB95 is the cruder
three-level model: the two agree to a few percent in x_e through recombination).
"""
import numpy as np

from oracle.cmb_oracle import aprimeoa


def ionize(tempb, a, adot, dtau, xe, YHe, H0, Omegab):
    """thermodynamics_mb95.py:6-44."""
    iswitch, tion, beta0, dec2g = 0.5, 1.5789e5, 43.082, 8.468e14
    alpha0 = 2.3866e-6 * (1.0 - YHe) * Omegab * H0 ** 2
    crec = 8.0138e-26 * (1.0 - YHe) * Omegab * H0 ** 2
    phi2 = max(0.448 * np.log(tion / tempb), 0.0)
    alpha = alpha0 / np.sqrt(tempb) * phi2 / a ** 3
    beta = tempb * phi2 * np.exp(beta0 - tion / tempb)
    if tempb <= 200.0:
        cpeebles = 1.0
    else:
        cp1 = crec * dec2g * (1.0 - xe) / (a * adot)
        cp2 = crec * tempb * phi2 * np.exp(beta0 - 0.25 * tion / tempb) * (1.0 - xe) / (a * adot)
        cpeebles = (1.0 + cp1) / (1.0 + cp1 + cp2)
    aa = a * dtau * alpha * cpeebles
    bb = a * dtau * beta * cpeebles
    b1 = 1.0 + iswitch * bb
    bbxe = bb + xe - (1.0 - iswitch) * (bb * xe + aa * xe * xe)
    rat = iswitch * aa * bbxe / (b1 * b1)
    return bbxe / b1 * (1.0 - rat) if rat < 5e-5 else b1 / (2.0 * iswitch * aa) * (np.sqrt(4.0 * rat + 1.0) - 1.0)


def saha_rhs(tempb, a, YHe, H0, Omegab):
    """r1, r2 of the helium Saha equations (thermodynamics_mb95.py:53-63)."""
    b0 = 2.150e24 / ((1.0 - YHe) * Omegab * H0 ** 2)
    b = b0 * a ** 3 * tempb * np.sqrt(tempb)
    return 4.0 * b * np.exp(-2.855e5 / tempb), b * np.exp(-6.313e5 / tempb)


def ionHe(tempb, a, x0, x1, x2, YHe, H0, Omegab, n_iter=6):
    """thermodynamics_mb95.py:47-89: n_iter fixed-point iterations."""
    r1, r2 = saha_rhs(tempb, a, YHe, H0, Omegab)
    c = 0.25 * YHe / (1.0 - YHe)
    for _ in range(n_iter):
        xe = x0 + c * (x1 + 2.0 * x2)
        den = r1 * r2 + xe * r1 + xe * xe
        x2, x1 = r1 * r2 / den, xe * r1 / den
    return x1, x2


def thermal_history(sc, logrhonu, N, taumin, taumax):
    """compute_thermo (thermodynamics_mb95.py:92-212).  sc: background scalars (Tcmb, YHe, H0,
    Omegab, adotrad and those of aprimeoa); logrhonu(log a): the neutrino density spline.
    Returns dict of arrays a, tau, tb, xHII, xe, xHeII, xHeIII, cs2 (N each)."""
    Tcmb, YHe, H0, Omegab = sc["Tcmb"], sc["YHe"], sc["H0"], sc["Omegab"]
    thomc0 = 5.0577e-8 * Tcmb ** 4
    dlntau = np.log(taumax / taumin) / (N - 1)
    barssc_raw = 9.1820e-14
    fx = 0.25 * YHe / (1.0 - YHe)
    adot, tau = sc["adotrad"], taumin
    a = adot * tau
    tb, xHII, xHeII, xHeIII = Tcmb / a, 1.0, 0.0, 1.0
    xe = xHII + fx * (xHeII + 2.0 * xHeIII)
    cs2 = 4.0 / 3.0 * barssc_raw * (1.0 - 0.75 * YHe + (1.0 - YHe) * xe) * tb
    out = {k: np.empty(N) for k in ("a", "tau", "tb", "xHII", "xe", "xHeII", "xHeIII", "cs2")}
    for i in range(N):
        for k, v in (("a", a), ("tau", tau), ("tb", tb), ("xHII", xHII), ("xe", xe), ("xHeII", xHeII), ("xHeIII", xHeIII), ("cs2", cs2)):
            out[k][i] = v
        if i == N - 1:
            break
        new_tau = taumin * np.exp((i + 1) * dlntau)
        dtau = new_tau - tau
        new_a = a + adot * dtau
        new_adot = aprimeoa(new_a, sc, np.exp(logrhonu(np.log(new_a)))) * new_a
        new_a = a + 2 * dtau / (1.0 / adot + 1.0 / new_adot)
        tg0 = Tcmb / a
        ahalf, adothalf = 0.5 * (a + new_a), 0.5 * (adot + new_adot)
        fe = (1.0 - YHe) * xe / (1.0 - 0.75 * YHe + (1.0 - YHe) * xe)
        thomc = thomc0 * fe / adothalf / ahalf ** 3
        if fe < 1e-3:
            a2t = (a - new_a) * Tcmb + a ** 2 * (tb - tg0) + (0.5 * (a - new_a) ** 2 * Tcmb + a ** 2 * (a - new_a) * (tb - tg0)) * thomc
        else:
            etc = np.exp(-thomc * (new_a - a))
            a2t = a ** 2 * (tb - tg0) * etc - Tcmb / thomc * (1.0 - etc)
        new_tb = Tcmb / new_a + a2t / new_a ** 2
        tbhalf = 0.5 * (tb + new_tb)
        new_xHII = ionize(tbhalf, ahalf, adothalf, dtau, xHII, YHe, H0, Omegab)
        xHeII, xHeIII = ionHe(tb, new_a, new_xHII, xHeII, xHeIII, YHe, H0, Omegab)
        new_xe = new_xHII + fx * (xHeII + 2.0 * xHeIII)
        dtbdla = -2.0 * new_tb - thomc * a2t / new_a
        cs2 = barssc_raw * (1.0 - 0.75 * YHe + (1.0 - YHe) * new_xe) * new_tb * (1.0 - dtbdla / new_tb / 3.0)
        a, adot, tau, tb, xHII, xe = new_a, new_adot, new_tau, new_tb, new_xHII, new_xe
    return out
