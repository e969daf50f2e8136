"""The source epilogue (deb_evolve_sources): perturbations -> time derivatives -> source function
in one launch, with no full states stored.  It must equal the staged chain (deb_evolve with
return_full, deb_time_derivatives, deb_cmb_sources) on the same solver: the same steps are
taken, the same projected state is interpolated, and the right-hand side and the source algebra
are the same code, so the planes are compared bit for bit.  The staged chain itself is checked
against the oracle in test_gpu_cmb.py and test_gpu_cmb_cfg3.py.
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from test_gpu_parity import lcdm_param

pytestmark = pytest.mark.gpu

PLANES = ("S", "S1", "S2", "S3", "S4", "SE", "alpha", "alphaprime")


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _staged(deb, p, k, aexp, lm, solver, isw, batch_size=16, threads=0):
    """The unfused chain over all k at once, keeping the states."""
    tau = p["tau_of_a_spline"].evaluate(torch.as_tensor(aexp, dtype=torch.float64, device="cuda"))
    tables = p["_deb_tables"]
    uc = torch.zeros(k.numel(), dtype=torch.int32, device="cuda")
    kw = dict(rtol=1e-4, atol=1e-4, return_full=True, **lm)
    if solver == "per_mode":
        y, st, na, nr = deb.evolve_units(tables, p["_deb_num_thermo"], uc, k, tau.reshape(1, -1),
                                         schedule={"order": "k_desc", "threads": threads}, **kw)
    else:
        y, st, na, nr = deb.evolve_units_batched(tables, p["_deb_num_thermo"], uc, k, tau.reshape(1, -1), batch_size=batch_size,
                                                 threads=threads, **kw)
    assert int(st.abs().max()) == 0
    yp = deb.compute_time_derivatives(y, tau, k, p, *lm.values())
    src = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, isw=isw, **lm)
    return src, na, nr


def _compare(fused, staged, na, nr, label):
    assert torch.equal(fused["n_accept"], na) and torch.equal(fused["n_reject"], nr), label
    worst = 0.0
    for name in PLANES:
        a, b = fused[name], staged[name]
        assert torch.isfinite(a).all(), (label, name)
        d = float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))
        worst = max(worst, d)
        if not torch.equal(a, b):
            print(f"  {label} {name}: not bitwise, max scaled difference {d:.1e}")
    print(f"  {label}: planes {'bit-identical' if worst == 0.0 else f'max scaled difference {worst:.1e}'}")
    assert worst <= 1e-14, label


def test_fused_equals_staged_at_the_cmb_truncation(deb):
    """l = 24 photon hierarchies (99 state entries), 24 modes of the 3000-mode grid in 1e-5..0.5,
    the 2500 output times of the CMB workload: per-mode steps at 256 and 128 threads, and batched
    steps (batches of 8), each with the spline and the algebraic ISW term."""
    lm = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
    aexp = deb.cmb_aexp_out(250, 500, 1000, 750)
    p = deb.evolve_background(param=lcdm_param(), num_thermo=1024)
    k = torch.as_tensor(np.geomspace(1e-5, 0.5, 3000)[::125].copy(), dtype=torch.float64, device="cuda")
    for solver, threads, isw in (("per_mode", 256, "spline"), ("per_mode", 128, "algebraic"), ("batched", 64, "spline"),
                                 ("batched", 256, "algebraic")):
        label = f"{solver} {threads} threads, isw {isw}"
        print(f"case: fused sources, lmax 24, 24 k x 2500 tau, {label}", flush=True)
        fused = deb.compute_sources_fused(p, k, aexp, solver=solver, batch_size=8, isw=isw, threads=threads, **lm)
        staged, na, nr = _staged(deb, p, k, aexp, lm, solver, isw, batch_size=8, threads=threads)
        _compare(fused, staged, na, nr, label)


def test_fused_with_massive_neutrinos_and_w0wa(deb):
    """One massive neutrino on 8 momentum bins (the generic bin-count kernel), a w0wa fluid, RSA and
    UFA on: 32 modes log-spaced in 1e-4..1, 300 output times of the CMB split."""
    h = 0.674
    P = dict(Omegam=(0.120 + 0.0224) / h ** 2, Omegab=0.0224 / h ** 2, H0=100 * h, Tcmb=2.7255, YHe=0.248, Neff=2.046, Nmnu=1.0,
             mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)
    lm = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8)
    aexp = deb.cmb_aexp_out(30, 60, 120, 90)
    p = deb.evolve_background(param=P, num_thermo=1024)
    k = torch.as_tensor(np.geomspace(1e-4, 1.0, 32), dtype=torch.float64, device="cuda")
    print("case: fused sources, massive nu nq 8 + w0wa, 32 k x 300 tau, per mode", flush=True)
    fused = deb.compute_sources_fused(p, k, aexp, **lm)
    staged, na, nr = _staged(deb, p, k, aexp, lm, "per_mode", "spline")
    _compare(fused, staged, na, nr, "per_mode, nq 8")


def test_driver_fused_equals_streamed(deb):
    """compute_Cell_spectrum_from_cosmo_params with fused sources equals the streamed staged chain."""
    lm = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
    aexp = deb.cmb_aexp_out(250, 500, 1000, 750)
    kw = dict(ellmax=500, nmodes=48, kmin=1e-5, kmax=0.5, n_k_dense=0, n_fftlog=4096, k_sampling_method="log", solver="per_mode",
              aexp_out=aexp, **lm)
    print("case: driver fused vs streamed, 48 log k, ellmax 500", flush=True)
    cl_f, p = deb.compute_Cell_spectrum_from_cosmo_params(lcdm_param(), fused=True, **kw)
    cl_s, _ = deb.compute_Cell_spectrum_from_cosmo_params(lcdm_param(), fused=False, k_chunk=16, **kw)
    assert torch.equal(cl_f, cl_s)
    assert p["nout"] == 2500
