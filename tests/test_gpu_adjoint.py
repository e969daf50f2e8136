"""Discrete adjoint kernel (deb_evolve_adjoint) against the oracle.

The oracle's adjoint (oracle/deb_oracle.cpp adjoint_replay: dense LU, complex-step Jacobians,
differences at the level of one right-hand side) runs on the GPU's own background tables and
accepted-step mesh, so the two compute the same frozen-mesh derivative by independent routes:
the parameter gradient, the cotangent of the initial state and of the integration range must
agree.  The oracle's adjoint is itself pinned against central differences of the frozen-mesh map
(tests/test_oracle_adjoint.py).
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from test_gpu_parity import TABLE_OFFSETS

pytestmark = pytest.mark.gpu
LM = dict(lmaxg=8, lmaxgp=8, lmaxr=8, lmaxnu=8, nqmax=3)
SOLVER = dict(rtol=1e-5, atol=1e-10)


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def base_param():
    h = 0.674
    return dict(Omegam=(0.120 + 0.0224) / h ** 2, Omegab=0.0224 / h ** 2, H0=100 * h, Tcmb=2.7255, YHe=0.248, Neff=2.046, Nmnu=1.0,
                mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)


def oracle_bg(deb, oracle, block, N=256):
    """Oracle background read verbatim from one HIP table block (a device tensor)."""
    offs = [deb._lib.table_offset(name, N) for name in TABLE_OFFSETS]
    return oracle.Background.from_tables(block.cpu().numpy(), N, offs)


def _problem(deb, k, aexp, lm, rel_step=3e-5):
    P = base_param()
    th = deb.param_to_theta(P)
    S = deb.grad._setup(th[None, :], P, k, np.asarray(aexp, dtype=np.float64), lm, SOLVER, 256, rel_step)
    return th, S


def test_adjoint_matches_oracle_adjoint(deb, oracle):
    lmax = [LM[k] for k in ("lmaxg", "lmaxgp", "lmaxr", "lmaxnu", "nqmax")]
    k = np.array([1e-3, 0.05, 0.3])
    aexp = [0.5, 1.0]
    print(f"case: adjoint vs oracle adjoint, k {k.tolist()}, a_out {aexp}, lmax {lmax}", flush=True)
    th, S = _problem(deb, k, aexp, LM)
    nk = k.size
    rng = np.random.default_rng(0)
    gbar = np.zeros((nk, len(aexp), 20))
    for f in (0, 4, 9, 12):  # eta, delta_m, theta_c / aH, delta_gamma
        gbar[:, :, f] = rng.standard_normal((nk, len(aexp)))
    dth = torch.as_tensor(S["steps"], dtype=torch.float64, device="cuda")
    g, st, lam0, tb = deb.evolve_units_adjoint(S["t0"], 256, S["tab"], dth, S["uc0"], S["uk0"], S["tau0"], S["tau"], S["mesh"], S["mlen"],
                                               torch.as_tensor(gbar, device="cuda"), **LM, return_parts=True)
    torch.cuda.synchronize()
    assert int(st.abs().max()) == 0
    g, lam0, tb = g.cpu().numpy(), lam0.cpu().numpy(), tb.cpu().numpy()
    bg = oracle_bg(deb, oracle, S["t0"][0])
    sh = [oracle_bg(deb, oracle, S["tab"][i]) for i in range(12)]
    mesh, mlen = S["mesh"].cpu().numpy(), S["mlen"].cpu().numpy()
    go, parts, lo, tbo = bg.adjoint(sh, S["steps"][0], k, S["tau0"][0].cpu().numpy(), S["tau"].cpu().numpy(), mesh, mlen, gbar, lmax)
    for u in range(nk):
        eg = np.abs(g[u] - go[u]).max() / np.abs(go[u]).max()
        el = np.abs(lam0[u] - lo[u]).max() / np.abs(lo[u]).max()
        et = np.abs(tb[u] - tbo[u]).max() / np.abs(tbo[u]).max()
        print(f"  k {k[u]:.3g}: gradient {eg:.2e}, initial-state cotangent {el:.2e}, range cotangent {et:.2e}")
        print(f"    gpu    {g[u]}\n    oracle {go[u]}")
        # the oracle takes the time and scale-factor derivatives of J by central differences
        # (relative step 1e-6), which lose digits in the stiff late-time regime of the larger k;
        # the kernel uses dual numbers.  The gradient itself is pinned to 1e-5.
        assert el < 1e-6 and et < 1e-4
        assert eg < 1e-5


def test_adjoint_jacobian_matches_differences(deb):
    """Full Jacobian by the adjoint vs central differences of the frozen-mesh map (both GPU)."""
    k = deb.k_grid({}, 1e-4, 0.5, 5, "log")
    print("case: adjoint vs difference Jacobian, 5 k, a_out (0.5, 1)", flush=True)
    P = base_param()
    th = deb.param_to_theta(P)
    pa, Ja = deb.pk_jacobian(th, P, k, aexp_out=(0.5, 1.0), method="adjoint", **LM, **SOLVER)
    pf, Jf = deb.pk_jacobian(th, P, k, aexp_out=(0.5, 1.0), method="fd", **LM, **SOLVER)
    assert torch.equal(pa, pf)
    Ja, Jf = Ja.cpu().numpy(), Jf.cpu().numpy()
    for j in deb.grad.ODE_PARAMS:
        rel = np.abs(Ja[:, :, j] - Jf[:, :, j]).max() / np.abs(Jf[:, :, j]).max()
        print(f"  d P / d {deb.PARAM9[j]:8s}: max rel diff {rel:.2e}")
        assert rel < 1e-4, deb.PARAM9[j]
    np.testing.assert_array_equal(Ja[:, :, [3, 4, 8]], Jf[:, :, [3, 4, 8]])
