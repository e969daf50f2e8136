"""The reference's per-stage CMB functions on the GPU (cmb.py:170-633): each stage against the
oracle's on identical inputs, the stages chained against the fused source kernel, and the
chain against the reference's own per-stage fixtures in the configuration they were made in."""
import numpy as np
import pytest
import torch

from conftest import disco_notebook, load_package
from oracle import cmb_oracle as C
from test_gpu_cmb import LM, _bgd, _cmb_grid, _np

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _stages(deb, p, y, yp, k, aexp, tau, lm):
    pt = deb.extract_perturbations(y, yp, lm["lmaxg"], lm["lmaxgp"], lm["lmaxr"])
    vis = deb.compute_visibility_functions(tau, p)
    nu = deb.compute_neutrino_perturbations(y, yp, aexp, p, lm["nqmax"], pt["iq0"])
    bq = deb.compute_background_quantities(aexp, p)
    metric = deb.compute_metric_perturbations(pt, nu, bq, p, k, aexp)
    src = deb.compute_source_function(pt, metric, vis, y, yp, k, lm["lmaxg"], lm["lmaxgp"], tau)
    return pt, vis, nu, metric, src


@pytest.fixture(scope="module")
def small(deb):
    p = deb.evolve_background(param=disco_notebook(), num_thermo=1024)
    aexp = deb.perturbations._geomspace(1e-3, 1.0, 90)
    y, k, p = deb.evolve_perturbations(param=p, aexp_out=aexp, kmin=1e-3, kmax=0.3, num_k=6, rtol=1e-4, atol=1e-4, return_full=True,
                                       k_sampling_method="log", **LM)
    tau = p["tau_out"]
    yp = deb.compute_time_derivatives(y, tau, k, p, *LM.values())
    return p, aexp, y, yp, k, tau


def _err(got, ref):
    ref = np.asarray(ref)
    return np.abs(_np(got) - ref).max() / max(np.abs(ref).max(), 1e-300)


def test_each_stage_matches_oracle(deb, oracle, small):
    print("case: stages on 6 k x 90 tau (DISCO-Notebook, massive nu), against the oracle's stages", flush=True)
    p, aexp, y, yp, k, tau = small
    pt, vis, nu, metric, src = _stages(deb, p, y, yp, k, aexp, tau, LM)
    o = C.sources_from_states(_np(y), _np(yp), _np(k), aexp, _np(tau), _bgd(deb, oracle, p, aexp), 11, 11, 11, 3)
    checks = [(nu[n], o["nu"][n]) for n in ("drhonu", "dpnu", "fnu", "shearnu", "rho_nu_prime", "shear_nu_prime")]
    checks += [(metric[n], o["metric"][n]) for n in ("hprime", "etaprime", "alpha", "alphaprime", "alphapprime")]
    checks += [(src["polarization_terms"][n], o["pol"][n]) for n in ("polarization_term", "polarization_termprime", "polarization_termpprime")]
    checks += [(src[n], o[n]) for n in ("S1", "S2", "S3", "S4", "S")]
    worst = max(_err(g, r) for g, r in checks)
    print(f"  worst stage output vs oracle {worst:.1e}")
    assert worst < 1e-12


def test_stages_compose_to_the_fused_kernel(deb, small):
    """compute_source_function over the stages equals compute_source_function_from_states (one fused
    kernel over the same point functions) to rounding of the background inputs (the stages take
    aH, rho_nu, rho_Q, w_Q from compute_neutrino_perturbations and compute_background_quantities)."""
    print("case: stage chain vs fused source kernel", flush=True)
    p, aexp, y, yp, k, tau = small
    _, _, _, metric, src = _stages(deb, p, y, yp, k, aexp, tau, LM)
    fused = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, **LM)
    for name in ("S1", "S2", "S3", "S4", "S"):
        assert _err(src[name], _np(fused[name])) < 1e-13, name
    assert _err(metric["alpha"], _np(fused["alpha"])) < 1e-13 and _err(metric["alphaprime"], _np(fused["alphaprime"])) < 1e-13


def test_stages_reproduce_the_reference_fixtures(deb, golden):
    """The reference's per-stage fixtures (tests/test_cmb/test_benchmark_compute_{metric_perturbations,
    polarization_terms, source_term_*}.csv, extract_perturbations) through the stage functions, in
    the configuration they were made in (batched, 16 per batch, radiation-streaming and
    ultra-relativistic-fluid approximations off): k0 and k63 to 1e-5."""
    print("case: per-stage functions on the reference fixture configuration (128 log k x 320 tau, batches of 16, RSA/UFA off)", flush=True)
    p = deb.evolve_background(param=dict(disco_notebook(), use_rsa=False, use_ur_fluid=False), num_thermo=1024)
    aexp = _cmb_grid(deb)
    y, k, p = deb.evolve_perturbations_batched(param=p, aexp_out=aexp, kmin=1e-4, kmax=1.0, num_k=128, rtol=1e-4, atol=1e-4,
                                               return_full=True, batch_size=16, k_sampling_method="log", **LM)
    tau = p["tau_out"]
    yp = deb.compute_time_derivatives(y, tau, k, p, *LM.values())
    pt, vis, nu, metric, src = _stages(deb, p, y, yp, k, aexp, tau, LM)
    g = golden("cmb_chain")
    got = {"deltag": pt["deltag"], "deltac": pt["deltac"], "thetag": pt["thetag"], "alpha": metric["alpha"],
           "polarization_term": src["polarization_terms"]["polarization_term"], "source_sw": src["S2"], "source_doppler": src["S3"],
           "source_polarization": src["S4"]}
    worst = 0.0
    for name, arr in got.items():
        for ik in (0, 63):
            key = f"{name}_k{ik}"
            if key in g:
                e = _err(arr[ik], g[key])
                worst = max(worst, e)
                assert e < 1e-5, key
    assert _err(metric["alphaprime"][0], g["alphaprime_k0"]) < 1e-5
    print(f"  worst fixture column {worst:.1e}")
