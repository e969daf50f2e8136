"""Batched shared-step evolution (evolve_perturbations_batched) on the HIP path vs the oracle.

The batch's step mesh is logged on the GPU and replayed mode by mode in the oracle: the states
must then agree to rounding, and the GPU's shared error norm must be the RMS of the oracle's
per-mode norms on that mesh (sqrt(mean_j en_j^2) over the batch).  The free-running oracle
batched solver must reach the same endpoint to the 1e-5 transfer-function tolerance.
"""
import numpy as np
import pytest
import torch

from conftest import disco_notebook, load_package
from test_gpu_parity import DOMINANT, lcdm_param, nu_w0wa_param, oracle_on_gpu_tables

pytestmark = pytest.mark.gpu

TOL_TRANSFER = 1e-5  # relative tolerance for transfer functions and P(k)
TOL_REPLAY = 1e-8  # scaled state difference on an identical step mesh (rounding only)
TOL_NORM = 1e-3  # relative, shared error norm vs the RMS of the oracle's per-mode norms on the same mesh
THETA_C_OUT, THETA_C_FULL = 9, 4  # theta_c / aH among the 20 outputs, theta_c in the full state


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _lm(lmax):
    return dict(lmaxg=lmax[0], lmaxgp=lmax[1], lmaxr=lmax[2], lmaxnu=lmax[3], nqmax=lmax[4])


def _run_batched(deb, P, k, aout, lmax, B, rtol, atol, full=False):
    p = deb.evolve_background(param=dict(P), num_thermo=256)
    tables = p["_deb_tables"]
    tau = p["tau_of_a_spline"].evaluate(np.asarray(aout, dtype=np.float64)).reshape(1, -1)
    kt = torch.as_tensor(k, dtype=torch.float64, device=tables.device)
    uc = torch.zeros(len(k), dtype=torch.int32, device=tables.device)
    y, st, na, nr, log = deb.evolve_units_batched(tables, 256, uc, kt, tau, batch_size=B, rtol=rtol, atol=atol, return_full=full,
                                                  step_log=True, **_lm(lmax))
    torch.cuda.synchronize()
    assert (st.cpu().numpy() == 0).all()
    return p, y.cpu().numpy(), na.cpu().numpy(), nr.cpu().numpy(), log.cpu().numpy()


def _check_batch_replay(bg, k, aout, lmax, B, rtol, atol, yg, na, nr, log, full=False, label=""):
    kw = dict(_lm(lmax), rtol=rtol, atol=atol)
    worst = worst_en = 0.0
    for g in range(len(k) // B):
        sl = slice(g * B, (g + 1) * B)
        assert np.all(na[sl] == na[g * B]) and np.all(nr[sl] == nr[g * B])
        ns = int(na[g * B] + nr[g * B])
        mesh = log[g, :ns, :2]
        en_gpu = log[g, :ns, 2]
        keep = (en_gpu < 1.0).astype(np.int32)
        assert keep.sum() == na[g * B]
        # common start: the earliest of the modes' own start times
        t0 = min(bg.start_tau(kk, aout) for kk in k[sl])
        assert abs(mesh[0, 0] - t0) <= 1e-13 * t0
        en2 = np.zeros(ns)
        for j in range(g * B, (g + 1) * B):
            yo, en = bg.replay(k[j], aout, mesh, keep, return_full=full, **kw)
            fin = np.isfinite(yo)
            assert np.array_equal(fin, np.isfinite(yg[j]))
            scale = np.abs(np.where(fin, yo, 0)).max(axis=0) + 1e-300
            e = np.abs(np.where(fin, yg[j] - yo, 0)) / scale
            e[:, THETA_C_FULL if full else THETA_C_OUT] = 0.0  # zero in synchronous gauge: rounding noise only
            worst = max(worst, float(e.max()))
            en2 += en ** 2
        en_rms = np.sqrt(en2 / B)
        # The error estimate k8 is a small difference of stage values whose rounding is amplified by
        # the conditioning of I / (gamma h) - J (relative scatter ~1e-5 between two FP64 solvers
        # on the same mesh); a wrong normalisation of the shared norm would be O(1).  Norms far
        # below 1 are rounding in both solvers and decide nothing.
        big = en_gpu > 1e-3
        worst_en = max(worst_en, float(np.max(np.abs(en_rms[big] - en_gpu[big]) / en_gpu[big])))
        flips = int(np.sum((en_rms < 1.0) != keep.astype(bool)))
        assert flips <= max(1, ns // 200), (g, flips)
    print(f"  batched replay {label}: max scaled state err {worst:.2e}, shared norm vs RMS of per-mode norms {worst_en:.2e}")
    assert worst < TOL_REPLAY
    assert worst_en < TOL_NORM


def test_batched_lcdm_matches_oracle(deb, oracle):
    """Plain LCDM, 8 modes in 2 batches of 4, rtol 1e-5 / atol 1e-10, outputs from a = 1e-6."""
    k = oracle.geomspace(1e-4, 1.0, 8)
    aout = oracle.geomspace(1e-6, 1.0, 12)
    lmax = [12, 12, 12, 8, 3]
    print("case: batched LCDM nk=8 B=4 nout=12", flush=True)
    p, yg, na, nr, log = _run_batched(deb, lcdm_param(), k, aout, lmax, 4, 1e-5, 1e-10)
    bg = oracle_on_gpu_tables(deb, oracle, p)
    _check_batch_replay(bg, k, aout, lmax, 4, 1e-5, 1e-10, yg, na, nr, log, label="lcdm")
    yo, _, sto, nao, nro = bg.evolve(k, aout, rtol=1e-5, atol=1e-10, batch_size=4, **_lm(lmax))
    print("  accepted steps per batch gpu", na[::4].tolist(), "oracle", nao[::4].tolist())
    assert np.all(sto == 0)
    end = np.abs(yg[:, -1, DOMINANT] - yo[:, -1, DOMINANT]) / np.abs(yo[:, -1, DOMINANT])
    print(f"  free-running endpoint (a=1) dominant fields max rel err {end.max():.2e}")
    assert end.max() < TOL_TRANSFER


def test_batched_massive_nu_w0wa_matches_oracle(deb, oracle):
    """Massive neutrino on 8 momentum bins (lmax_nu 17) + w0wa, 6 modes in 3 batches of 2; P(k) at z = 0."""
    k = oracle.geomspace(1e-4, 2.0, 6)
    aout = [0.5, 1.0]
    lmax = [12, 12, 12, 17, 8]
    P = nu_w0wa_param()
    print("case: batched massive nu + w0wa nk=6 B=2 nout=2", flush=True)
    p, yg, na, nr, log = _run_batched(deb, P, k, aout, lmax, 2, 1e-5, 1e-10)
    bg = oracle_on_gpu_tables(deb, oracle, p)
    _check_batch_replay(bg, k, aout, lmax, 2, 1e-5, 1e-10, yg, na, nr, log, label="nu_w0wa")
    yo, _, sto, nao, nro = bg.evolve(k, aout, rtol=1e-5, atol=1e-10, batch_size=2, **_lm(lmax))
    pk_g = deb.get_power(k=torch.as_tensor(k, device="cuda"), y=torch.as_tensor(yg[:, -1, :], device="cuda"), idx=4,
                         param=P).cpu().numpy()
    pk_o = oracle.get_power(k, yo[:, -1, :], 4, P)
    rel = np.abs(pk_g / pk_o - 1).max()
    print(f"  free-running P(k, z=0) max rel err {rel:.2e}")
    assert rel < TOL_TRANSFER


def test_batched_full_state_cmb_settings(deb, oracle):
    """The CMB driver's settings (rtol = atol = 1e-4, full state with the RSA projection), one batch of 16."""
    k = oracle.geomspace(1e-3, 0.3, 16)
    aout = [1e-3, 2e-3, 0.1, 1.0]
    lmax = [11, 11, 11, 8, 3]
    print("case: batched full state nk=16 B=16 rtol 1e-4", flush=True)
    p, yg, na, nr, log = _run_batched(deb, disco_notebook(), k, aout, lmax, 16, 1e-4, 1e-4, full=True)
    bg = oracle_on_gpu_tables(deb, oracle, p)
    _check_batch_replay(bg, k, aout, lmax, 16, 1e-4, 1e-4, yg, na, nr, log, full=True, label="full B=16")


def test_batch_of_one_is_the_per_mode_solve(deb):
    """With batch_size 1 the shared-step solver reduces to the per-mode one: same step counts and
    the same states to rounding (the two kernels inline the same stage code).  The per-mode solve
    retries a rejected step with the Jacobian pieces formed at that point, the shared-step one forms
    them again, so with rejections present this also checks that the kept pieces are the ones a
    full pass gives."""
    P = lcdm_param()
    tab = deb.background_tables([P], 256)
    tau = deb.tau_out_all(tab, 256, [0.01, 1.0])
    k = torch.as_tensor(np.geomspace(1e-4, 1.0, 5), device="cuda")
    uc = torch.zeros(5, dtype=torch.int32, device="cuda")
    kw = dict(lmaxg=10, lmaxgp=10, lmaxr=10, lmaxnu=8, nqmax=3, rtol=1e-5, atol=1e-10)
    y1, st1, na1, nr1 = deb.evolve_units(tab, 256, uc, k, tau, **kw)
    yb, stb, nab, nrb = deb.evolve_units_batched(tab, 256, uc, k, tau, batch_size=1, **kw)
    assert torch.equal(stb, st1) and torch.equal(nab, na1) and torch.equal(nrb, nr1)
    assert int(nr1.sum()) > 0, "no rejected steps: the retry path is not exercised"
    diff = float(((yb - y1).abs() / (y1.abs().amax(dim=(0, 1)) + 1e-300)).max())
    print(f"  batch of one vs per-mode: max scaled diff {diff:.1e} (bit-identical: {torch.equal(yb, y1)})")
    assert diff < 1e-12


def test_batched_workgroup_width_gives_identical_results(deb):
    """The shared-step kernel at 64, 128 and 256 threads per batch: the extra wavefronts only
    evaluate rows and the next stage's background beside wavefront 0, so the step mesh, the step
    counts and every output bit agree (each instantiation rounds every expression alike)."""
    P = nu_w0wa_param()
    p = deb.evolve_background(param=dict(P), num_thermo=256)
    tau = p["tau_of_a_spline"].evaluate(np.asarray([1 / 3, 0.5, 1.0])).reshape(1, -1)
    k = torch.as_tensor(np.geomspace(1e-5, 5.0, 6), device="cuda")
    uc = torch.zeros(6, dtype=torch.int32, device="cuda")
    kw = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, rtol=1e-5, atol=1e-10)
    runs = {}
    for nt in (64, 128, 256):
        y, st, na, nr, log = deb.evolve_units_batched(p["_deb_tables"], 256, uc, k, tau, batch_size=3, step_log=True,
                                                      threads=nt, **kw)
        torch.cuda.synchronize()
        assert (st.cpu().numpy() == 0).all(), nt
        runs[nt] = [t.cpu().numpy() for t in (y, na, nr, log)]
    for nt in (128, 256):
        for a, b in zip(runs[nt], runs[64]):
            assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f"), nt


def test_batched_reference_surface_and_errors(deb):
    P = disco_notebook()
    p = deb.evolve_background(param=dict(P))
    y, kmodes, p2 = deb.evolve_perturbations_batched(param=p, aexp_out=[0.5, 1.0], kmin=1e-3, kmax=0.2, num_k=32, batch_size=16,
                                                     k_sampling_method="log")
    assert y.shape == (32, 2, 20) and kmodes.shape == (32,)
    assert p2["nout"] == 2 and p2["lmaxg"] == 11 and p2["tau_out"].shape == (2,)
    assert torch.isfinite(y).all()
    with pytest.raises(ValueError, match="multiple of batch_size"):
        deb.evolve_perturbations_batched(param=p, aexp_out=[1.0], kmin=1e-3, kmax=0.2, num_k=10, batch_size=4, k_sampling_method="log")
    with pytest.raises(RuntimeError, match="max_steps"):
        deb.evolve_perturbations_batched(param=p, aexp_out=[1.0], kmin=0.1, kmax=0.2, num_k=2, batch_size=2, max_steps=5,
                                         k_sampling_method="log")
    # units of two cosmologies in one batch are refused (status 5) without being integrated
    tab = deb.background_tables([P, lcdm_param()], 256)
    tau = deb.tau_out_all(tab, 256, [1.0])
    uc = torch.tensor([0, 1, 1, 1], dtype=torch.int32, device="cuda")
    k = torch.full((4,), 0.01, dtype=torch.float64, device="cuda")
    yb, st, na, nr = deb.evolve_units_batched(tab, 256, uc, k, tau, batch_size=2)
    assert st.cpu().tolist() == [5, 5, 0, 0]
    assert torch.isnan(yb[:2]).all() and torch.isfinite(yb[2:]).all()
