"""Pins the oracle's discrete adjoint of the frozen-mesh replay (the checker of the HIP adjoint).

The adjoint (oracle/deb_oracle.cpp adjoint_replay) reverses each Rodas5 step exactly; the map
it differentiates is the replay of a fixed accepted-step mesh (fractions of each mode's range).
Central differences of that same map over shifted cosmologies, at a relative step small enough
that no approximation switch or thermal-history step decision moves inside the stencil, must
agree with it; a dropped term of the reverse sweep (the Jacobian's dependence on the state, the
step size's dependence on the mesh ends, the output interpolation weights) shows up far above
that level.  The parts are checked separately where the mathematics fixes them.
"""
import math

import numpy as np
import pytest

H = 0.674
BASE = dict(Omegam=(0.120 + 0.0224) / H ** 2, Omegab=0.0224 / H ** 2, H0=100 * H, Tcmb=2.7255, YHe=0.248, Neff=2.046, Nmnu=1.0,
            mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)
THETA = np.array([0.120, 0.0224, H, 0.965, math.log(21.0), 0.06, -0.95, 0.1, 0.054])
ODE = (0, 1, 2, 5, 6, 7)
LMAX = [8, 8, 8, 8, 3]


def theta_to_param(th):
    oc, ob, h, ns, lnA, mnu, w0, wa, _ = th
    p = dict(BASE)
    p.update(Omegam=(oc + ob) / h ** 2, Omegab=ob / h ** 2, H0=100 * h, n_s=ns, A_s=math.exp(lnA) / 1e10, mnu=mnu, w_DE_0=w0, w_DE_a=wa)
    return p


@pytest.fixture(scope="module")
def setup(oracle):
    rel = 1e-6
    bg = oracle.Background(theta_to_param(THETA), 256)
    aout = np.array([0.5, 1.0])
    tau = bg.spline("tau_of_a", aout)
    dth = np.array([rel * max(abs(THETA[j]), 1e-2) for j in ODE])
    sh = []
    for i, j in enumerate(ODE):
        for sg in (1.0, -1.0):
            t = THETA.copy()
            t[j] += sg * dth[i]
            sh.append(oracle.Background(theta_to_param(t), 256))
    tsh = np.array([b.spline("tau_of_a", aout) for b in sh])
    kk = np.array([1e-3, 0.05, 0.3])
    meshes = [bg.mesh(k, tau, LMAX, 1e-5, 1e-10)[1] for k in kk]
    lens = np.array([len(s) for s in meshes])
    M = np.ones((kk.size, lens.max()))
    for i, s in enumerate(meshes):
        M[i, :len(s)] = s
    rng = np.random.default_rng(0)
    gbar = np.zeros((kk.size, aout.size, 20))
    for f in (0, 4, 9, 12):  # eta, delta_m, theta_c / aH, delta_gamma
        gbar[:, :, f] = rng.standard_normal((kk.size, aout.size))
    return dict(bg=bg, sh=sh, dth=dth, tau=tau, tsh=tsh, kk=kk, M=M, lens=lens, gbar=gbar)


def test_adjoint_matches_frozen_mesh_differences(setup):
    s = setup
    g, parts, lam0, tb = s["bg"].adjoint(s["sh"], s["dth"], s["kk"], s["tau"], s["tsh"], s["M"], s["lens"], s["gbar"], LMAX)
    fd = np.zeros_like(g)
    for i in range(6):
        yp = s["sh"][2 * i].replay_norm(s["kk"], s["tsh"][2 * i], s["M"], s["lens"], LMAX)
        ym = s["sh"][2 * i + 1].replay_norm(s["kk"], s["tsh"][2 * i + 1], s["M"], s["lens"], LMAX)
        fd[:, i] = ((yp - ym) * s["gbar"]).sum(axis=(1, 2)) / (2 * s["dth"][i])
    err = np.abs(g - fd) / np.abs(fd).max(axis=1, keepdims=True)
    print("adjoint vs differences, scaled by each mode's largest entry:", err.max(axis=1))
    assert err.max() < 2e-6
    np.testing.assert_allclose(parts.sum(axis=1), g, rtol=1e-12, atol=0)


def test_adjoint_initial_state_cotangent(setup, oracle):
    """lam0 = d<gbar, outputs>/d y(t0) with the mesh and background fixed: compare with central
    differences of the replay started from a perturbed initial state along a random direction."""
    s = setup
    bg = s["bg"]
    g, parts, lam0, tb = bg.adjoint(s["sh"], s["dth"], s["kk"][:1], s["tau"], s["tsh"], s["M"][:1], s["lens"][:1],
                                    s["gbar"][:1], LMAX)
    k = s["kk"][0]
    t0 = bg._lib.orc_start_tau(bg._h, float(k), s["tau"].ctypes.data_as(oracle._dp), s["tau"].size)
    t1 = s["tau"].max()
    m = s["M"][0, :s["lens"][0]]
    steps = np.stack([np.concatenate([[t0], t0 + m[:-1] * (t1 - t0)]), np.concatenate([t0 + m[:-1] * (t1 - t0), [t1]])], axis=1)
    y0 = bg.ics(LMAX, t0, k)
    v = np.random.default_rng(1).standard_normal(y0.size) * np.abs(y0).clip(1e-3)
    v[0] = 0.0  # the scale factor moves the output times' background too: keep this check linear
    e = 1e-3

    def outputs(y_init):
        lib = bg._lib
        lm, lp = oracle._i(LMAX)
        n = y_init.size
        yout = np.empty((s["tau"].size, 20))
        # one step at a time from a given state (Rodas5 steps and the same interpolation/outputs)
        y = y_init.copy()
        isave = 0
        for (ta, tb_) in steps:
            y1, _ = bg.rodas5_step(LMAX, k, ta, tb_ - ta, y)
            while isave < s["tau"].size and s["tau"][isave] <= tb_:
                th = (s["tau"][isave] - ta) / (tb_ - ta)
                yi = y + th * (y1 - y)
                yout[isave] = bg.convert(LMAX, k, yi)
                isave += 1
            y = y1
        del lib, lp, n
        return yout

    # RSA is off at these (k, tau): the outputs are the converted interpolants
    fp, fm = outputs(y0 + e * v), outputs(y0 - e * v)
    fd = ((fp - fm) * s["gbar"][0]).sum() / (2 * e)
    assert abs(fd - lam0[0] @ v) <= 1e-8 * abs(fd)
