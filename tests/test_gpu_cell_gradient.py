"""Gradient of a weighted TT spectrum, L = sum_l w_l C_l^TT, with respect to the nine sampling
parameters (cell_vjp): frozen-mesh central differences of the whole chain (background, fused
sources replayed on the base solve's accepted steps, line of sight, C_l) for the six parameters
that enter the perturbations, the exact n_s column, zero A_s and tau_reion columns.  Each stage
of the chain is checked against the oracle elsewhere; here: the replayed chain reproduces the
adaptive one, the n_s column equals a difference quotient in n_s, and the difference columns
converge in the step size.
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from test_gpu_parity import lcdm_param

pytestmark = pytest.mark.gpu

LM = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
ELLMAX, NFFT = 400, 4096


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


@pytest.fixture(scope="module")
def setup(deb):
    base = lcdm_param()
    theta = deb.param_to_theta(base)
    k = np.geomspace(1e-4, 0.3, 48)
    aexp = deb.cmb_aexp_out(32, 64, 128, 96)
    w = np.random.default_rng(1).uniform(0.0, 1.0, ELLMAX - 1)
    kw = dict(ellmax=ELLMAX, aexp_out=aexp, n_fftlog=NFFT, **LM)
    print("case: cell_vjp, LCDM, 48 k in 1e-4..0.3, 320 tau, lmax 24, ellmax 400", flush=True)
    L0, g, cl = deb.cell_vjp(theta, base, k, w, return_cell=True, **kw)
    return dict(base=base, theta=theta, k=k, aexp=aexp, w=w, kw=kw, L0=L0, g=g.cpu().numpy(), cl=cl)


def _L(deb, S, theta, replay_of=None):
    """L at theta by the adaptive chain (the driver's stages, per-mode solver)."""
    p = deb.evolve_background(param=deb.theta_to_param(theta, S["base"]), num_thermo=1024)
    k = torch.as_tensor(S["k"], dtype=torch.float64, device="cuda")
    src = deb.compute_sources_fused(p, k, S["aexp"], rtol=1e-4, atol=1e-4, **LM)
    tau0 = float(p["tau_of_a_spline"].evaluate(np.array([1.0])).cpu()[0])
    th, kk = deb.compute_theta_ell(ELLMAX, k, src["tau"], src["S"], tau0, n_fftlog=NFFT)
    cl = deb.compute_Cell(th, kk, p["n_s"], p["k_p"]).cpu().numpy()
    return float((S["w"] * cl[2:]).sum()), th, kk, p


def test_value_and_fixed_columns(deb, setup):
    S = setup
    L_ad, th, kk, p = _L(deb, S, S["theta"])
    print(f"  L = {S['L0']:.6e}; adaptive chain {L_ad:.6e}; grad {S['g']}")
    assert abs(S["L0"] / L_ad - 1.0) < 1e-13
    assert S["g"][4] == 0.0 and S["g"][8] == 0.0
    assert np.isfinite(S["g"]).all()
    # n_s: the exact column against a difference quotient in n_s on the same Theta_l (C_l is
    # analytic in n_s; the quotient's error is (d^2 / 6) <ln^3(k / k_p)> ~ 1e-9 at d = 1e-4)
    d = 1e-4
    clp = deb.compute_Cell(th, kk, p["n_s"] + d, p["k_p"]).cpu().numpy()
    clm = deb.compute_Cell(th, kk, p["n_s"] - d, p["k_p"]).cpu().numpy()
    fd = float((S["w"] * (clp - clm)[2:]).sum()) / (2 * d)
    print(f"  n_s column {S['g'][3]:.8e}, difference quotient {fd:.8e}")
    assert abs(S["g"][3] / fd - 1.0) < 1e-7


def test_replay_reproduces_the_adaptive_chain(deb, setup):
    """The gradient's base: the frozen mesh replayed at theta itself gives the adaptive chain's C_l."""
    S = setup
    base, theta = S["base"], S["theta"]
    p = deb.evolve_background(param=deb.theta_to_param(theta, base), num_thermo=1024)
    k = torch.as_tensor(S["k"], dtype=torch.float64, device="cuda")
    src = deb.compute_sources_fused(p, k, S["aexp"], step_log=True, **LM)
    mesh, mlen = deb.grad._accepted_mesh(src["step_log"], src["n_accept"], src["tau"].max().expand(k.numel()))
    rep = deb.compute_sources_fused(p, k, S["aexp"], replay=(mesh, mlen), **LM)
    d = float((rep["S"] - src["S"]).abs().max() / src["S"].abs().max())
    print(f"  replayed sources vs adaptive: max scaled difference {d:.1e}")
    assert d < 1e-12


def test_step_convergence(deb, setup):
    """Difference columns at two relative steps agree (the frozen-mesh map is smooth away from the
    approximation switches)."""
    S = setup
    _, g2 = deb.cell_vjp(S["theta"], S["base"], S["k"], S["w"], rel_step=1e-5, **S["kw"])
    g2 = g2.cpu().numpy()
    idx = [0, 1, 2, 6, 7]  # omega_c, omega_b, h, w0, wa (m_nu: no massive neutrino here, the column is exactly 0)
    rel = np.abs(g2[idx] / S["g"][idx] - 1.0)
    print(f"  columns at steps 3e-5 and 1e-5: relative differences {rel}")
    assert rel.max() < 1e-3
    assert S["g"][5] == 0.0 and g2[5] == 0.0
