"""The CMB workload at its own truncation and sizes: flat LCDM with massless neutrinos, photon
hierarchies to l = 24 (lmaxg = lmaxgp = 24, lmaxr = 12: 99 state entries), 2500 output times on
the reference's recombination-weighted split (its 32/64/128/96 pieces scaled by 7.8125,
cmb.py:999-1014), k on geomspace(1e-5, 0.5, 3000), l up to 2500 at n_fftlog 16384.

The states at l = 24 are checked by replaying the GPU's step mesh in the oracle; the chain on a
24-mode subset of the 3000-mode grid against the oracle on the GPU's own states at l <= 2500;
streaming the chain over k chunks against the unchunked chain (bit-identical).
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from oracle import cmb_oracle as C
from test_gpu_parity import _gpu_run, lcdm_param, oracle_on_gpu_tables

pytestmark = pytest.mark.gpu

LM = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3)
LMAX = [24, 24, 12, 8, 3]


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _np(t):
    return t.detach().cpu().numpy() if torch.is_tensor(t) else np.asarray(t)


@pytest.fixture(scope="module")
def aexp(deb):
    a = deb.cmb_aexp_out(250, 500, 1000, 750)
    assert a.size == 2500
    return a


def test_states_at_l24_replay_in_the_oracle(deb, oracle, aexp):
    """Full states (99 entries) at three k of the grid over the 2500 output times, rtol = atol =
    1e-4, replayed by the oracle on the GPU's accepted-step mesh: every entry agrees to 1e-10 of
    the state's largest perturbation (measured 3e-14..6e-12), and every field that reaches 1e-4
    of that scale agrees to 1e-8 of its own largest value (measured 2e-13..4e-9; theta_c,
    identically zero in this gauge, aside).  The smaller fields at k = 1e-5 (theta_b ~ 3e-14,
    photon polarisation, neutrino multipoles ~ (k tau)^l) differ by 1e-8..3e-6 of their own
    scale: rounding of the structured tail solve against the oracle's dense LU, 1e-16..1e-14 in
    absolute terms, where a change of k by 1e-13 already moves the polarisation moments by 2e-8."""
    k = np.geomspace(1e-5, 0.5, 3000)[[0, 1500, 2999]]
    print("case: cfg3 states, lmax 24, 3 k x 2500 tau, replay", flush=True)
    p, yg, na, nr, log = _gpu_run(deb, lcdm_param(), k, aexp, LMAX, 1e-4, 1e-4, full=True, num_thermo=1024)
    bg = oracle_on_gpu_tables(deb, oracle, p)
    print(f"  accepted steps {na.tolist()}")
    kw = dict(lmaxg=24, lmaxgp=24, lmaxr=12, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4)
    for i in range(len(k)):
        ns = int(na[i] + nr[i])
        keep = (log[i, :ns, 2] < 1.0).astype(np.int32)
        yo, _ = bg.replay(k[i], aexp, log[i, :ns, :2], keep, return_full=True, **kw)
        d = np.abs(yg[i] - yo).max(axis=0)
        g = np.abs(yo[:, 1:]).max()
        own = np.abs(yo).max(axis=0)
        big = own >= 1e-4 * g
        big[4] = False
        e_own = (d[big] / own[big]).max()
        print(f"  k {k[i]:.2e}: max err / state scale {d.max() / g:.1e}; {big.sum()} fields above 1e-4 of it: max err on own scale {e_own:.1e}")
        assert d.max() < 1e-10 * g
        assert e_own < 1e-8


def test_chain_at_l2500_matches_oracle_and_streams_exactly(deb, oracle, aexp):
    """24 modes (every 125th of the 3000-mode grid): sources from the GPU states vs the oracle's
    on the same states (1e-9), Theta_l for l <= 2500 at n_fftlog 16384 (1e-10) and C_l (1e-12);
    the chain streamed over k in chunks of 8 equals the unchunked chain bit for bit."""
    print("case: cfg3 chain, 24 k x 2500 tau, ellmax 2500, n_fftlog 16384", flush=True)
    p = deb.evolve_background(param=lcdm_param(), num_thermo=1024)
    k = torch.as_tensor(np.geomspace(1e-5, 0.5, 3000)[::125].copy(), dtype=torch.float64, device="cuda")
    whole = deb.compute_sources_streamed(p, k, aexp, k_chunk=None, **LM)
    chunked = deb.compute_sources_streamed(p, k, aexp, k_chunk=8, **LM)
    for name in ("S", "S1", "S2", "S3", "S4"):
        assert torch.equal(whole[name], chunked[name]), name
    assert torch.equal(whole["n_accept"], chunked["n_accept"])
    # the same chain stage by stage, with the states kept, against the oracle
    tau = whole["tau"]
    tables = p["_deb_tables"]
    tau_all = tau.reshape(1, -1)
    uc = torch.zeros(k.numel(), dtype=torch.int32, device="cuda")
    y, st, na, nr = deb.evolve_units(tables, 1024, uc, k, tau_all, rtol=1e-4, atol=1e-4, return_full=True, schedule={"order": "k_desc"},
                                     **LM)
    yp = deb.compute_time_derivatives(y, tau, k, p, *LM.values())
    src = deb.compute_source_function_from_states(y, yp, k, aexp, tau, p, **LM)
    assert torch.equal(src["S"], whole["S"])
    from test_gpu_cmb import _bgd
    d = _bgd(deb, oracle, p, aexp)
    o = C.sources_from_states(_np(y), _np(yp), _np(k), aexp, _np(tau), d, 24, 24, 12, 3)
    err = np.abs(_np(src["S"]) - o["S"]).max() / np.abs(o["S"]).max()
    print(f"  S vs oracle max scaled err {err:.2e}")
    assert err < 1e-9
    tau0 = float(p["tau0"])
    th, kk = deb.compute_theta_ell(2500, k, tau, src["S"], tau0, n_fftlog=16384)
    tho, kko = C.compute_theta_ell(2500, _np(k), _np(tau), _np(src["S"]), tau0, n_fftlog=16384)
    err = np.abs(_np(th) - tho).max() / np.abs(tho).max()
    print(f"  Theta_l (l <= 2500) max scaled err {err:.2e}")
    assert err < 1e-10
    cl = _np(deb.compute_Cell(th, kk, p["n_s"], p["k_p"]))
    clo = C.compute_Cell(tho, kko, p["n_s"], p["k_p"])
    err = np.abs(cl - clo).max() / np.abs(clo).max()
    print(f"  C_l max scaled err {err:.2e}")
    assert err < 1e-12


def test_driver_streamed_per_mode_matches_stages(deb, aexp):
    """compute_Cell_spectrum_from_cosmo_params with solver='per_mode', k_chunk: the streamed
    chain through the driver equals compute_sources_streamed -> compute_theta_ell -> compute_Cell."""
    print("case: driver, per-mode solver streamed in chunks of 16, 48 log k, ellmax 500", flush=True)
    kw = dict(ellmax=500, nmodes=48, kmin=1e-5, kmax=0.5, n_k_dense=0, n_fftlog=4096, k_sampling_method="log", solver="per_mode",
              k_chunk=16, aexp_out=aexp, **LM)
    cl, p = deb.compute_Cell_spectrum_from_cosmo_params(lcdm_param(), **kw)
    q = deb.evolve_background(param=lcdm_param(), num_thermo=1024)
    k = torch.as_tensor(deb.k_grid(q, 1e-5, 0.5, 48, "log"), dtype=torch.float64, device="cuda")
    src = deb.compute_sources_streamed(q, k, aexp, k_chunk=48, **LM)
    th, kk = deb.compute_theta_ell(500, k, src["tau"], src["S"], float(q["tau0"]), n_fftlog=4096)
    assert torch.equal(cl, deb.compute_Cell(th, kk, q["n_s"], q["k_p"]))
    assert p["nout"] == 2500 and p["lmaxg"] == 24
