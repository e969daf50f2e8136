"""Large-scale-structure observables on the GPU against the oracle (oracle/lss_oracle.py) on
identical inputs: the solver's own delta_m and theta_m / aH at a = 1 for DISCO-Notebook on 512
log-spaced k in 1e-4..1/Mpc."""
import numpy as np
import pytest
import torch

from conftest import disco_notebook, load_package
from oracle import lss_oracle as L

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


@pytest.fixture(scope="module")
def run(deb):
    p = deb.evolve_background(param=disco_notebook(), num_thermo=256)
    y, k, p = deb.evolve_perturbations(param=p, aexp_out=[1.0], kmin=1e-4, kmax=1.0, num_k=512, k_sampling_method="log")
    par = dict(A_s=float(p["A_s"]), n_s=float(p["n_s"]), k_p=float(p.get("k_p", 0.05)))
    return y[:, 0, :].contiguous(), k, par


def _np(t):
    return t.detach().cpu().numpy()


def _close(got, ref, tol=1e-12):
    err = np.abs(_np(got) - ref).max() / np.abs(ref).max()
    assert err < tol, err
    return err


@pytest.mark.parametrize("w,p", [(5, 2), (8, 3), (39, 3), (101, 4)])
def test_savgol_filter(deb, run, w, p):
    """Against the oracle's exact (rational) taps: the GPU solves the scaled normal equations
    (5e-15 from exact at window 101, order 4, where a float lstsq on the unscaled Vandermonde
    matrix, the reference's and scipy's route, is 8e-10 off)."""
    print(f"case: Savitzky-Golay window {w} order {p} on log P", flush=True)
    y, k, par = run
    lp = np.log(L.get_power(_np(k), _np(y), 4, par))
    _close(deb.savgol_filter(y=torch.as_tensor(lp, device="cuda"), window_length=w, polyorder=p), L.savgol_filter(lp, w, p))


@pytest.mark.parametrize("dlogk", [0.2, 0.5])
def test_power_smoothed(deb, run, dlogk):
    print(f"case: smoothed P(k), dlogk {dlogk}", flush=True)
    y, k, par = run
    for idx in (4, 5):
        _close(deb.get_power_smoothed(k=k, y=y, dlogk=dlogk, idx=idx, param=par), L.get_power_smoothed(_np(k), _np(y), dlogk, idx, par))


@pytest.mark.parametrize("mu_sampling", [True, False])
@pytest.mark.parametrize("smooth", [None, 0.5])
def test_kaiser_and_multipoles(deb, run, mu_sampling, smooth):
    print(f"case: Kaiser P(k, mu), mu_sampling {mu_sampling}, smoothing {smooth}", flush=True)
    y, k, par = run
    P, mu = deb.power_Kaiser(y=y, kmodes=k, bias=1.8, mu_sampling=mu_sampling, smooth_dlogk=smooth, nmu=11, param=par)
    Po, muo = L.power_kaiser(_np(y), _np(k), 1.8, 11, par, mu_sampling=mu_sampling, smooth_dlogk=smooth)
    np.testing.assert_allclose(_np(mu), muo, rtol=0, atol=1e-15)
    _close(P, Po)
    for got, ref in zip(deb.power_multipoles(y=y, kmodes=k, b=1.8, param=par), L.power_multipoles(_np(y), _np(k), 1.8, par)):
        _close(got, ref)


@pytest.mark.parametrize("ell", [0, 2, 4])
def test_correlation_function(deb, run, ell):
    """Direct transforms on the GPU against numpy's FFT in the oracle: 1e-11 of the largest |xi|."""
    print(f"case: xi_{ell}(r) by FFTLog, N = 512", flush=True)
    y, k, par = run
    pk = deb.get_power(k=k, y=y, idx=4, param=par)
    xi, r = deb.get_xi_from_P(k=k, Pk=pk, N=512, ell=ell)
    xio, ro = L.get_xi_from_P(_np(k), _np(pk), ell)
    np.testing.assert_allclose(_np(r), ro, rtol=1e-15)
    print(f"  max scaled err {_close(xi, xio, 1e-11):.1e}")


def test_argument_errors(deb, run):
    y, k, par = run
    with pytest.raises(ValueError):
        deb.get_xi_from_P(k=k[:-1], Pk=k[:-1], N=511)
    with pytest.raises(ValueError):
        deb.get_power_smoothed(k=k, y=y, dlogk=0.02, idx=4, param=par)  # a one-point window
    with pytest.raises(ValueError):
        deb.savgol_filter(y=k, window_length=3, polyorder=3)
