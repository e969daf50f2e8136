"""CMB chain on the GPU: sources, line of sight, C_l, D_l.

Each stage is compared with the oracle (oracle/cmb_oracle.py) on identical inputs -- the GPU's
own background tables, states and time derivatives -- so the comparisons are at rounding level.
The reference's own fixtures came from its batched shared-step run with log-spaced k: the
batched solver reproduces their sources at k0 and k63 to the transfer tolerance; the per-mode
solver (a different step sequence) is checked at the reference's regression tolerance.
"""
import ctypes

import numpy as np
import pytest
import torch

from conftest import disco_notebook, load_package
from oracle import cmb_oracle as C
from test_gpu_parity import oracle_on_gpu_tables

pytestmark = pytest.mark.gpu
LM = dict(lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3)


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


def _bgd(deb, oracle, p, aexp):
    """Background pieces for the oracle chain, taken from the GPU's own tables."""
    bg = oracle_on_gpu_tables(deb, oracle, p)
    q, w = oracle.nu_bins(p.get("nqmax", 3))
    d = {k: float(p[k]) for k in ["grhom", "grhog", "grhor", "OmegaDE", "amnu"]}
    d.update({k: float(p[k]) for k in ["Omegam", "Omegab", "Neff", "Nmnu", "w_DE_0", "w_DE_a"]})
    d["Omegak"] = float(p.get("Omegak", 0.0))
    d.update(rhonu=np.exp(bg.spline("logrhonu_of_loga", np.log(aexp))), q=q, w=w)
    for name, key in [("tau", "tau_tab"), ("opac", "opac"), ("gvis", "gvis"), ("gvisprime", "gvisprime"), ("gvispprime", "gvispprime"),
                      ("optical_depth", "optical_depth")]:
        d[key] = _np(p[name])
    return d


def test_contract_matches_numpy(deb):
    print("case: deb_contract m,n,k = (150,133,80) (3,300,16) (257,129,512)", flush=True)
    L = deb._lib.lib()
    rng = np.random.default_rng(1)
    for m, n, k in [(150, 133, 80), (3, 300, 16), (257, 129, 512)]:
        A, B = rng.standard_normal((m, k)), rng.standard_normal((n, k))
        Ad, Bd = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
        Cd = torch.empty((m, n), dtype=torch.float64, device="cuda")
        deb._lib.check(L.deb_contract(deb._lib.ptr(Ad), deb._lib.ptr(Bd), m, n, k, deb._lib.ptr(Cd), deb._lib.stream()))
        ref = A @ B.T
        err = np.abs(_np(Cd) - ref).max() / np.abs(ref).max()
        print(f"  {m}x{n}x{k} max scaled err {err:.2e}")
        assert err < 1e-13
    assert L.deb_contract(None, None, 4, 4, 17, None, None) < 0


def _cmb_grid(deb):
    a = np.concatenate([deb.perturbations._geomspace(3e-4, 5e-3, 256, endpoint=False), deb.perturbations._geomspace(5e-3, 1.0, 64)])
    return a


def test_visibility_matches_oracle_and_fixture(deb, oracle, golden):
    print("case: visibility DISCO-Notebook num_thermo=1024, 320 output times", flush=True)
    p = deb.evolve_background(param=disco_notebook(), num_thermo=1024)
    aexp = _cmb_grid(deb)
    tau = p["tau_of_a_spline"].evaluate(aexp)
    v = deb.compute_visibility_functions(tau, p, aexp=aexp)
    d = _bgd(deb, oracle, p, aexp)
    o = C.visibility_functions(_np(tau), d["tau_tab"], d["opac"], d["gvis"], d["gvisprime"], d["gvispprime"], d["optical_depth"])
    o["aprimeoa"] = C.aprimeoa(aexp, d, d["rhonu"])
    o["rhonu"] = d["rhonu"]
    for name in ["opac", "opacprime", "gvis", "gvisprime", "gvispprime", "aprimeoa", "rhonu"]:
        err = np.abs(_np(v[name]) - o[name]).max() / np.abs(o[name]).max()
        print(f"  {name:10s} max scaled err {err:.2e}")
        assert err < 1e-11, name
    assert np.abs(np.exp(-_np(v["optical_depth"])) - np.exp(-o["optical_depth"])).max() < 1e-12
    g = golden("visibility_functions")
    for name in ["gvis", "gvisprime"]:
        ref = g[name]
        err = np.abs(_np(v[name]) - ref) / np.maximum(np.abs(ref), 1e-6 * np.abs(ref).max())
        assert err.max() < 1e-5, name
    assert np.abs(np.exp(-_np(v["optical_depth"])) - np.exp(-g["optical_depth"])).max() < 1e-6


@pytest.fixture(scope="module")
def small_cmb_run(deb):
    """DISCO-Notebook, 6 log-spaced k in 1e-3..0.3, 90 output times 1e-3..1, full states."""
    p = deb.evolve_background(param=disco_notebook(), num_thermo=1024)
    aexp = deb.perturbations._geomspace(1e-3, 1.0, 90)
    y, k, p = deb.evolve_perturbations(param=p, aexp_out=aexp, kmin=1e-3, kmax=0.3, num_k=6, rtol=1e-4, atol=1e-4, return_full=True,
                                       k_sampling_method="log", **LM)
    tau = p["tau_out"]
    yp = deb.compute_time_derivatives(y, tau, k, p, *LM.values())
    return p, aexp, y, yp, k, tau


@pytest.mark.parametrize("isw", ["spline", "algebraic"])
def test_sources_match_oracle(deb, oracle, small_cmb_run, isw):
    print(f"case: sources 6 k x 90 tau (DISCO-Notebook, massive nu), isw {isw}", flush=True)
    p, aexp, y, yp, k, tau = small_cmb_run
    src = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, isw=isw, **LM)
    d = _bgd(deb, oracle, p, aexp)
    o = C.sources_from_states(_np(y), _np(yp), _np(k), aexp, _np(tau), d, 11, 11, 11, 3, isw=isw)
    o["alpha"], o["alphaprime"] = o["metric"]["alpha"], o["metric"]["alphaprime"]
    for name in ["S1", "S2", "S3", "S4", "S", "alpha", "alphaprime"]:
        ref = o[name]
        err = np.abs(_np(src[name]) - ref).max() / np.abs(ref).max()
        print(f"  {name:10s} max scaled err {err:.2e}")
        assert err < 1e-9, name


@pytest.mark.parametrize("n_k_dense", [0, 40])
def test_theta_ell_cell_dell_match_oracle(deb, golden, small_cmb_run, n_k_dense):
    print(f"case: line of sight ellmax=300 n_fftlog=4096 n_k_dense={n_k_dense}", flush=True)
    p, aexp, y, yp, k, tau = small_cmb_run
    src = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, **LM)
    tau0 = float(p["tau0"])
    th, kk = deb.compute_theta_ell(300, k, tau, src["S"], tau0, n_fftlog=4096, n_k_dense=n_k_dense)
    tho, kko = C.compute_theta_ell(300, _np(k), _np(tau), _np(src["S"]), tau0, n_fftlog=4096, n_k_dense=n_k_dense)
    assert np.allclose(_np(kk), kko, rtol=1e-14, atol=0)
    err = np.abs(_np(th) - tho).max() / np.abs(tho).max()
    print(f"  Theta_l max scaled err {err:.2e}")
    assert err < 1e-10
    cl = deb.compute_Cell(th, kk, p["n_s"], p["k_p"])
    clo = C.compute_Cell(_np(th), _np(kk), p["n_s"], p["k_p"])
    assert np.abs(_np(cl) - clo).max() / np.abs(clo).max() < 1e-12
    ell, dl = deb.compute_Dell(torch.ones(2501, dtype=torch.float64, device="cuda"), 2.1064e-09, 2.7255, ellmax=2500)
    g = golden("compute_dell")
    np.testing.assert_array_equal(_np(ell), g["ell"])
    np.testing.assert_allclose(_np(dl), g["Dell"], rtol=1e-14)


FIXTURE_PLANES = [("alpha", "alpha"), ("source_sw", "S2"), ("source_doppler", "S3"), ("source_polarization", "S4")]


def _fixture_chain(deb, golden, batched, approximations=True, isw="spline"):
    param = disco_notebook() if approximations else dict(disco_notebook(), use_rsa=False, use_ur_fluid=False)
    p = deb.evolve_background(param=param, num_thermo=1024)
    aexp = _cmb_grid(deb)
    kw = dict(param=p, aexp_out=aexp, kmin=1e-4, kmax=1.0, num_k=128, rtol=1e-4, atol=1e-4, return_full=True, k_sampling_method="log",
              **LM)
    y, k, p = deb.evolve_perturbations_batched(batch_size=16, **kw) if batched else deb.evolve_perturbations(**kw)
    tau = p["tau_out"]
    yp = deb.compute_time_derivatives(y, tau, k, p, *LM.values())
    src = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, isw=isw, **LM)
    th, kk = deb.compute_theta_ell(100, k, tau, src["S"], float(p["tau0"]))
    cl = _np(deb.compute_Cell(th, kk, p["n_s"], p["k_p"]))
    g = golden("cmb_chain")
    rel = np.abs(cl[2:] / g["Cell"][2:] - 1)
    print(f"  C_l (l = 2..100) vs reference: median rel {np.median(rel):.2e}, max rel {rel.max():.2e}")
    errs = {}
    for name, plane in FIXTURE_PLANES:
        for ik in (0, 63, 127):
            ref = g[f"{name}_k{ik}"]
            errs[f"{name}_k{ik}"] = np.abs(_np(src[plane][ik]) - ref).max() / np.abs(ref).max()
    errs["alphaprime_k0"] = np.abs(_np(src["alphaprime"][0]) - g["alphaprime_k0"]).max() / np.abs(g["alphaprime_k0"]).max()
    for name, ik in (("theta_ell_k_min", 0), ("theta_ell_k_max", 127)):
        errs[name] = np.abs(_np(th[ik]) - g[name]).max() / np.abs(g[name]).max()
    pt = C.extract_perturbations(_np(y), _np(yp), 11, 11, 11)
    for name in ("deltag", "deltac", "thetag", "deltab", "deltar"):
        for ik in (0, 63, 127):
            if f"{name}_k{ik}" in g:
                errs[f"{name}_k{ik}"] = np.abs(pt[name][ik] - g[f"{name}_k{ik}"]).max() / np.abs(g[f"{name}_k{ik}"]).max()
    for name, e in errs.items():
        print(f"  {name}: max scaled diff {e:.2e}")
    errs["theta_ell_k_min_by_l"] = np.abs(_np(th[0]) - g["theta_ell_k_min"]) / np.abs(g["theta_ell_k_min"]).max()
    return rel, errs


def test_chain_against_reference_fixtures(deb, golden):
    """The reference's CMB test configuration (tests/test_cmb.py:93-240): DISCO-Notebook,
    num_thermo 1024, 128 k in 1e-4..1, 320 output times, rtol = atol = 1e-4, full states;
    Theta_l and C_l for l <= 100 (n_fftlog 16384, no densification), with the per-mode solver.
    The fixtures were made with log-spaced k (their k63 reproduces only with that sampling) by
    the reference's batched shared-step solver; per-mode steps are a different step sequence, so
    this is a regression check at the reference's own tolerance (rtol 0.2), not parity."""
    print("case: reference CMB fixture configuration (128 log k x 320 tau), ellmax 100, per-mode steps", flush=True)
    rel, errs = _fixture_chain(deb, golden, batched=False)
    assert np.median(rel) < 0.2
    for name, e in errs.items():  # at k0 a few coarse per-mode steps interpolate far from the fixture's
        if name.endswith("_k63") or name.endswith("_k127"):
            assert e < 0.2, name


def test_batched_chain_against_reference_fixtures(deb, golden):
    """The same configuration through the batched shared-step solver (16 modes per batch), the
    solver the fixtures were made with, with the default approximations on: the sources at k0
    and k63, and alpha, alpha' at k0, agree to the 1e-5 transfer tolerance (measured 2e-9..8e-7).
    alpha at k63 (1.5e-4) and the states there differ by more because the fixtures were made
    before the radiation-streaming approximation existed: with it switched off (the next test)
    alpha(k63) agrees to 1.5e-8.  At k127 (the last batch, ~160 accepted steps) the step
    sequences part at a rounding-level accept decision: checked at the reference's 0.2.
    Theta_l and C_l are reported only here (the spline-derivative ISW term postdates them)."""
    print("case: reference CMB fixture configuration (128 log k x 320 tau), ellmax 100, batches of 16", flush=True)
    rel, errs = _fixture_chain(deb, golden, batched=True)
    for name, e in errs.items():
        if name == "alpha_k63" or name.startswith("theta_ell") or name.split("_k")[0] in ("deltag", "deltac", "thetag", "deltab", "deltar"):
            continue
        elif name.endswith("_k0") or name.endswith("_k63"):
            assert e < 1e-5, name
        elif name.endswith("_k127"):
            assert e < 0.2, name
    assert errs["alpha_k63"] < 1e-3


def test_batched_chain_as_the_fixtures_were_made(deb, golden):
    """The fixtures' own configuration: batched solver, radiation-streaming and ultra-relativistic-
    fluid approximations off (they predate both), ISW term S1 = e^-tau_c (eta' + alpha'') (the
    algebraic form the reference's comment at cmb.py:406-408 describes; the spline derivative of
    cmb.py:437-449 came later).  States, alpha and the sources at k0 and k63 to 1e-5 (the oracle
    measures 1e-10..2e-6); Theta_l(k_min) for l >= 10 to 1e-6 (oracle 3e-7), l = 5..9 to 5e-6,
    the lowest three multipoles to 2e-2; C_l (l = 2..100) median < 2e-3, max < 5e-3 (oracle 8.6e-4 /
    3.1e-3 -- the remainder comes from modes k ~ 1e-3..3e-2 that no fixture covers)."""
    print("case: reference CMB fixture configuration, RSA and UFA off, algebraic ISW, batches of 16", flush=True)
    rel, errs = _fixture_chain(deb, golden, batched=True, approximations=False, isw="algebraic")
    for name, e in errs.items():
        if name.startswith("theta_ell"):
            continue
        if name.endswith("_k0") or name.endswith("_k63"):
            assert e < 1e-5, name
        elif name.endswith("_k127"):
            assert e < 0.2, name
    d = errs["theta_ell_k_min_by_l"]
    assert d[10:].max() < 1e-6 and d[5:10].max() < 5e-6 and d[:5].max() < 2e-2
    assert np.median(rel) < 2e-3 and rel.max() < 5e-3


def test_cell_spectrum_driver_composes_the_stages(deb):
    """compute_Cell_spectrum_from_cosmo_params (cmb.py:930-1090) equals the stages called one by one."""
    print("case: C_l driver nmodes=32 ellmax=200 n_k_dense=128 n_fftlog=4096", flush=True)
    kw = dict(ellmax=200, nmodes=32, kmin=1e-4, kmax=0.5, n_k_dense=128, n_fftlog=4096)
    cl, p = deb.compute_Cell_spectrum_from_cosmo_params(disco_notebook(), **kw)
    q = deb.evolve_background(param=disco_notebook(), num_thermo=1024)
    aexp = deb.cmb_aexp_out()
    y, k, q = deb.evolve_perturbations_batched(param=q, aexp_out=aexp, kmin=1e-4, kmax=0.5, num_k=32, rtol=1e-4, atol=1e-4,
                                               return_full=True, batch_size=16, k_sampling_method="camb", **LM)
    yp = deb.compute_time_derivatives(y, q["tau_out"], k, q, *LM.values())
    src = deb.compute_source_function_from_states(y, yp, k, aexp, q["tau_out"], q, **LM)
    th, kk = deb.compute_theta_ell(200, k, q["tau_out"], src["S"], float(q["tau0"]), n_fftlog=4096, n_k_dense=128)
    cl2 = deb.compute_Cell(th, kk, q["n_s"], q["k_p"])
    assert torch.equal(cl, cl2)
    assert p["nout"] == 320 and (cl[2:] > 0).all()


def _reference_dell_from_records(golden):
    """The reference's own D_l^TT (l = 2..2500) recovered from its two recorded per-l relative
    differences to CAMB and to CLASS (tests/test_cmb/test_cmb_vs_{camb,class}_DISCO_Notebook_.csv):
    of the candidates D_X (1 +- r_X), the pair that agrees.  Returns (D_ref, pair disagreement)."""
    ell = np.arange(2, 2501)
    out = []
    for code, col in (("camb", "D_ell"), ("class", "D_ell_TT")):
        g = golden(f"{code}_dell")
        D = np.interp(ell, g["ell"], g[col])
        r = g["ref_relative_error"]
        out.append(np.stack([D * (1 + r), D * (1 - r)]))
    d = np.abs(out[0][:, None, :] - out[1][None, :, :])
    flat = d.reshape(4, -1).argmin(axis=0)
    a, b = flat // 2, flat % 2
    idx = np.arange(ell.size)
    Dref = 0.5 * (out[0][a, idx] + out[1][b, idx])
    return Dref, d.reshape(4, -1)[flat, idx] / Dref


def test_dell_against_camb_class_and_the_reference(deb, golden):
    """The reference's CMB acceptance test (tests/test_cmb.py:1410-1472, 1764-1859):
    compute_Cell_spectrum_from_cosmo_params(DISCO-Notebook, ellmax=2500, nmodes=128, kmin=1e-4,
    kmax=1), radiation-streaming and ultra-relativistic-fluid approximations on, D_l^TT against
    CAMB and CLASS at mean relative error < 1% (the reference records 0.712% / 0.620%).  The same
    D_l is then compared with the reference's own, recovered from those two records."""
    print("case: D_l^TT, l = 2..2500, DISCO-Notebook, 128 camb-sampled k, n_k_dense 8192, n_fftlog 16384", flush=True)
    cl, p = deb.compute_Cell_spectrum_from_cosmo_params(disco_notebook(), ellmax=2500, nmodes=128, kmin=1e-4, kmax=1.0)
    ell, dl = deb.compute_Dell(cl, A_s=p["A_s"], Tcmb=p["Tcmb"])
    ell, dl = _np(ell), _np(dl)
    stats = {}
    for code, col in (("camb", "D_ell"), ("class", "D_ell_TT")):
        g = golden(f"{code}_dell")
        ref = np.interp(ell, g["ell"], g[col])
        rel = np.abs(dl - ref) / ref
        stats[code] = (rel.mean(), rel.max(), g["ref_relative_error"].mean())
        print(f"  vs {code.upper()}: mean rel {rel.mean():.4%}, max {rel.max():.4%}  (reference: mean {stats[code][2]:.4%})")
        assert rel.mean() < 0.01, code
    Dref, spread = _reference_dell_from_records(golden)
    rel = np.abs(dl - Dref) / Dref
    print(f"  vs the reference's own D_l (records agree to median {np.median(spread):.1e}): median rel {np.median(rel):.2e}, "
          f"90% {np.quantile(rel, 0.9):.2e}, max {rel.max():.2e}")
    assert np.median(rel) < 2e-3
