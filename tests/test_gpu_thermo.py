"""The alternative thermal histories of evolve_background on the GPU (thermo_module 'MB95' and
'CLASS', background.py:191-345): MB95 against the oracle's restatement on the GPU's own
background scalars and neutrino spline; the CLASS input by feeding RECFAST's own thermal
history back in as a CLASS table, which must rebuild the same opacity and visibility."""
import numpy as np
import pytest
import torch

from conftest import disco_notebook, load_package
from oracle import mb95_oracle as M
from test_gpu_parity import oracle_on_gpu_tables

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _np(t):
    return t.detach().cpu().numpy()


def test_mb95_matches_oracle(deb, oracle):
    print("case: MB95 thermal history, num_thermo 1024, DISCO-Notebook", flush=True)
    p = deb.evolve_background(param=disco_notebook(), thermo_module="MB95", num_thermo=1024)
    bg = oracle_on_gpu_tables(deb, oracle, p)
    sc = dict(disco_notebook(), **{s: p[s] for s in ["grhom", "grhog", "grhor", "OmegaDE", "adotrad", "taumin", "taumax"]})
    lr = lambda x: float(bg.spline("logrhonu_of_loga", np.array([x]))[0])
    h = M.thermal_history(sc, lr, 1024, p["taumin"], p["taumax"])
    for key, ref in [("aexp", h["a"]), ("tau", h["tau"]), ("xe", h["xe"]), ("Tm", h["tb"]), ("cs2", h["cs2"]), ("xeHI", h["xHII"]),
                     ("xeHeI", h["xHeII"]), ("xeHeII", h["xHeIII"])]:
        err = np.abs(_np(p[key]) - ref).max() / np.abs(ref).max()
        print(f"  {key}: max scaled err {err:.1e}")
        assert err < 1e-10, key
    assert p["amin"] == 1e-9 and p["amax"] == 1.01
    assert float(p["gvis"].max()) > 0 and np.isfinite(_np(p["gvis"])).all()


def test_mb95_power_spectrum_close_to_recfast(deb):
    """P(k, z = 0) with the MB95 thermal history differs from RECFAST's by the baryon-photon
    decoupling only: below 1% at 16 k in 1e-4..1 (lmax 12, rtol 1e-5)."""
    print("case: P(k) with MB95 vs RECFAST, 16 k", flush=True)
    out = {}
    for mod in ("RECFAST", "MB95"):
        p = deb.evolve_background(param=disco_notebook(), thermo_module=mod, num_thermo=1024)
        y, k, p = deb.evolve_perturbations(param=p, aexp_out=[1.0], kmin=1e-4, kmax=1.0, num_k=16, lmaxg=12, lmaxgp=12, lmaxr=12,
                                           rtol=1e-5, atol=1e-10, k_sampling_method="log")
        out[mod] = _np(deb.get_power(k=k, y=y[:, 0, :], idx=4, param=p))
    rel = np.abs(out["MB95"] / out["RECFAST"] - 1)
    print(f"  max rel difference {rel.max():.2e}")
    assert rel.max() < 1e-2 and rel.max() > 0


def test_class_input_rebuilds_the_recfast_background(deb, oracle):
    """RECFAST's own (a, tau, x_e, T_b, c_b^2) handed back in CLASS order as class_thermo: the
    splines, opacity, optical depth and visibility are rebuilt from the same knots (1e-12); the
    a-range becomes the table's (amin = its first a): see below for what that moves."""
    print("case: CLASS-format input from RECFAST's thermal history, num_thermo 1024", flush=True)
    p = deb.evolve_background(param=disco_notebook(), num_thermo=1024)
    ct = {"scale factor a": _np(p["aexp"])[::-1], "conf. time [Mpc]": _np(p["tau"])[::-1], "x_e": _np(p["xe"])[::-1],
          "Tb [K]": _np(p["Tm"])[::-1], "c_b^2": _np(p["cs2"])[::-1]}
    q = deb.evolve_background(param=disco_notebook(), thermo_module="CLASS", class_thermo=ct)
    assert q["amin"] == float(ct["scale factor a"][-1]) and q["amax"] == float(ct["scale factor a"][0])
    for key in ("aexp", "tau", "xe", "opac", "optical_depth", "gvis", "gvisprime", "gvispprime", "xeprime"):
        err = np.abs(_np(q[key]) - _np(p[key])).max() / np.abs(_np(p[key])).max()
        print(f"  {key}: max scaled diff {err:.1e}")
        assert err < 1e-12, key
    assert abs(q["tau0"] / p["tau0"] - 1) < 1e-12
    # the a-range is the table's: taumin = amin / adotrad moves from 4.63e-4 to 5.81e-4, the
    # perturbations' start-time bisection over [taumin, tau(0.1)] (perturbations.py:784-835) lands
    # 2-5% elsewhere, and P(k) moves by the start time's effect on the adiabatic initial state
    # (~2e-4 at rtol 1e-5) -- as in the reference, which reads param['taumin'] there
    k = torch.as_tensor(np.geomspace(1e-3, 0.5, 6), dtype=torch.float64, device="cuda")
    uc = torch.zeros(6, dtype=torch.int32, device="cuda")
    lm = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=8, nqmax=3)
    pk = {}
    for name, par in (("recfast", p), ("class", q)):
        tau = par["tau_of_a_spline"].evaluate(torch.tensor([1.0], dtype=torch.float64, device="cuda")).reshape(1, -1)
        y, st, na, nr, log = deb.evolve_units(par["_deb_tables"], 1024, uc, k, tau, rtol=1e-5, atol=1e-10, step_log=True, **lm)
        pk[name] = _np(deb.get_power(k=k, y=y[:, 0, :], idx=4, param=par))
    rel = np.abs(pk["class"] / pk["recfast"] - 1).max()
    print(f"  P(k) max rel diff {rel:.1e} (start times moved with taumin)")
    assert rel < 1e-3
    # the solve on the CLASS-input tables is the oracle's on those tables (same steps)
    bg = oracle_on_gpu_tables(deb, oracle, q)
    kn, yn, lg = _np(k), _np(y), _np(log)
    for i in range(6):
        ns = int(na[i] + nr[i])
        keep = (lg[i, :ns, 2] < 1.0).astype(np.int32)
        yo, _ = bg.replay(kn[i], [1.0], lg[i, :ns, :2], keep, rtol=1e-5, atol=1e-10, **lm)
        err = np.abs(yn[i, 0, [0, 2, 4, 6, 8, 12]] - yo[0, [0, 2, 4, 6, 8, 12]]).max() / np.abs(yo[0, [4, 6, 8]]).max()
        assert err < 1e-8, (i, err)


def test_thermo_module_errors(deb):
    with pytest.raises(ValueError):
        deb.evolve_background(param=disco_notebook(), thermo_module="HYREC")
    with pytest.raises(ValueError):
        deb.evolve_background(param=disco_notebook(), thermo_module="CLASS")
    with pytest.raises(ValueError):
        deb.evolve_background(param=disco_notebook(), thermo_module="CLASS", class_thermo={"x_e": np.ones(20)})
