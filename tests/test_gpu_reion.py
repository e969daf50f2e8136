"""Reionization (an extension: the reference's thermal history has none; tau_reion > 0 in the
parameters switches it on): the tanh model on the RECFAST history, z_re solved in the background
kernel.  Checked against the oracle restatement (oracle/reion_oracle.py, pinned by a closed form)
on the GPU's own recombination history and background, by the optical depth the tables carry, by
the damping of the temperature spectrum by e^(-2 tau) on small scales, and by the gradient
column it gives.
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from oracle import reion_oracle as R
from test_gpu_parity import lcdm_param

pytestmark = pytest.mark.gpu
TAU = 0.054


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _np(x):
    return x.detach().cpu().numpy() if torch.is_tensor(x) else np.asarray(x)


def test_history_against_the_oracle(deb):
    print("case: reionization history, LCDM, tau_reion 0.054, num_thermo 1024", flush=True)
    P = lcdm_param()
    p0 = deb.evolve_background(param=dict(P), num_thermo=1024)
    p1 = deb.evolve_background(param=dict(P, tau_reion=TAU), num_thermo=1024)
    a = _np(p0["aexp"])
    assert np.array_equal(_np(p0["tau"]), _np(p1["tau"]))  # the background is untouched
    xrec = _np(p0["xe"])
    # dtau/da of the background (massless neutrinos: no neutrino spline), header scalars of the GPU tables
    g = {k: p0[k] for k in ("grhom", "grhog", "grhor", "OmegaDE")}

    def dtauda(x):
        grho2 = g["grhom"] * P["Omegam"] * x + (g["grhog"] + g["grhor"] * P["Neff"]) + g["grhom"] * g["OmegaDE"] * x ** 4
        return np.sqrt(3.0 / grho2)

    zo = R.solve_zre(TAU, p0["fHe"], p0["akthom"], dtauda, R.xrec_spline(np.log(a), xrec))
    print(f"  z_re GPU {p1['z_reion']:.10f}, oracle {zo:.10f}; added depth {p1['tau_reion_achieved']:.12f}")
    assert abs(p1["z_reion"] - zo) < 1e-7
    assert abs(p1["tau_reion_achieved"] - TAU) < 1e-10
    xo = xrec + R.dxe(a, xrec, p1["z_reion"], p0["fHe"])
    err = np.abs(_np(p1["xe"]) - xo).max()
    print(f"  x_e table vs oracle max abs {err:.1e}; x_e today {_np(p1['xe'])[-1]:.8f} (1 + 2 f_He = {1 + 2 * p0['fHe']:.8f})")
    assert err < 1e-13
    assert abs(_np(p1["xe"])[np.argmin(np.abs(a - 1.0))] - (1.0 + 2.0 * p0["fHe"])) < 1e-5
    # the optical depth the tables carry (opacity spline integral) rises by tau_reion
    d = p1["optical_depth_today"] - p0["optical_depth_today"]
    print(f"  optical depth today: increase {d:.6f}")
    assert abs(d / TAU - 1.0) < 2e-3
    # the conformal-time derivative row: x_rec' (1 - F_H) + (d/da at fixed x_rec) / (dtau/da)
    d, FH = R.dxe_da(a, xrec, p1["z_reion"], p0["fHe"])
    xpo = _np(p0["xeprime_recf"]) * (1.0 - FH) + d / dtauda(a)
    xp = _np(p1["xeprime_recf"])
    errp = np.abs(xp - xpo).max() / np.abs(xpo).max()
    print(f"  dx_e/dtau row vs oracle: max scaled {errp:.1e}")
    assert errp < 1e-12


def test_spectrum_damping_and_polarisation_bump(deb):
    """C_l^TT with reionization over C_l^TT without: e^(-2 tau) on scales inside the horizon at
    reionization (l 200..1500, median within 1%); C_l^EE gains the reionization bump at l <= 6."""
    print("case: C_l TT/EE with and without reionization, ellmax 1500", flush=True)
    kw = dict(ellmax=1500, nmodes=128, solver="per_mode", lmaxg=24, lmaxgp=24, lmaxr=12)
    c0, _ = deb.compute_Cell_spectra_from_cosmo_params(lcdm_param(), spectra=("TT", "EE"), **kw)
    c1, _ = deb.compute_Cell_spectra_from_cosmo_params(dict(lcdm_param(), tau_reion=TAU), spectra=("TT", "EE"), **kw)
    r = _np(c1["TT"])[200:1501] / _np(c0["TT"])[200:1501]
    med = float(np.median(r / np.exp(-2 * TAU) - 1.0))
    print(f"  TT ratio / e^-2tau - 1: median {med:.2e}, range {r.min() / np.exp(-2 * TAU) - 1:.2e}..{r.max() / np.exp(-2 * TAU) - 1:.2e}")
    assert abs(med) < 1e-2
    bump = _np(c1["EE"])[2:7] / _np(c0["EE"])[2:7]
    print(f"  EE ratio at l = 2..6: {bump}")
    assert bump.min() > 10.0


def _weighted(deb, base, k, aexp, w, ellmax, nfft):
    """L = sum_l w_l C_l^TT by the adaptive chain (fused sources, per-mode steps)."""
    p = deb.evolve_background(param=dict(base), num_thermo=1024)
    kk = torch.as_tensor(k, dtype=torch.float64, device="cuda")
    src = deb.compute_sources_fused(p, kk, aexp, lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
    tau0 = float(p["tau_of_a_spline"].evaluate(np.array([1.0])).cpu()[0])
    th, kt = deb.compute_theta_ell(ellmax, kk, src["tau"], src["S"], tau0, n_fftlog=nfft)
    return float((w * _np(deb.compute_Cell(th, kt, p["n_s"], p["k_p"]))[2:]).sum())


def test_tau_reion_gradient_column(deb):
    """The tau_reion column of the weighted-TT gradient with reionization switched on (a frozen-mesh
    difference at a step of 1.6e-6) against a difference of the adaptive chain at a step of 0.005,
    on output times that resolve the reionization visibility (750 points below z = 600, the CMB
    workload's split).  The adaptive chain's step sequences move with tau_reion, which puts noise
    of about 1e-3 of L into each value (a few percent of the quotient at this step); its truncation
    error is 1e-5.  Without reionization in the base the column is zero."""
    print("case: cell_vjp tau_reion column, 48 k, 970 tau, ellmax 400", flush=True)
    base = dict(lcdm_param(), tau_reion=TAU)
    theta = deb.param_to_theta(base, tau_reion=TAU)
    k = np.geomspace(1e-4, 0.3, 48)
    aexp = deb.cmb_aexp_out(32, 64, 128, 750)
    w = np.random.default_rng(1).uniform(0.0, 1.0, 399)
    L, gr = deb.cell_vjp(theta, base, k, w, ellmax=400, aexp_out=aexp, n_fftlog=4096)
    g8 = float(gr[8])
    d = 0.005
    Lp = _weighted(deb, dict(base, tau_reion=TAU + d), k, aexp, w, 400, 4096)
    Lm = _weighted(deb, dict(base, tau_reion=TAU - d), k, aexp, w, 400, 4096)
    fd = (Lp - Lm) / (2 * d)
    print(f"  L {L:.6e}, d L / d tau_reion {g8:.6e}; adaptive chain difference {fd:.6e}; ratio to -2 L {g8 / (-2 * L):.4f}")
    assert abs(g8 / fd - 1.0) < 3e-2
    assert g8 < 0.0
    _, g0 = deb.cell_vjp(theta, lcdm_param(), k, w, ellmax=400, aexp_out=aexp, n_fftlog=4096)
    assert float(g0[8]) == 0.0
