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


# ---------------------------------------------------------------------------------------------
# the differentiation configuration: massive neutrino (nq 8, lmax_nu 17) + w0wa, lmax 12 (192
# state entries), 256 k log-spaced in 1e-4..1/Mpc, P(k, a = 1), cotangent default_rng(0)
# ---------------------------------------------------------------------------------------------
LM2 = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8)


@pytest.fixture(scope="module")
def cfg2(deb):
    P = base_param()
    th = deb.param_to_theta(P)
    k = deb.k_grid({}, 1e-4, 1.0, 256, "log")
    cot = np.random.default_rng(0).standard_normal(256)
    print("case: 256 k, n = 192, P(k, a = 1) and its 9-parameter VJP by the adjoint", flush=True)
    pk, v = deb.pk_and_vjp(th, P, k, cot, **LM2, **SOLVER)
    _, J = deb.pk_jacobian(th, P, k, **LM2, **SOLVER)
    S = deb.grad._setup(th[None, :], P, k, np.array([1.0]), LM2, SOLVER, 256, deb.grad.REL_STEP)
    return dict(P=P, th=th, k=k, cot=cot, pk=pk.cpu().numpy()[:, 0], v=v.cpu().numpy(), J=J.cpu().numpy()[:, 0, :], S=S)


def _trusted(d1, d2, tol=1e-5):
    """Entries where central differences at two step sizes agree (no approximation-switch or
    thermal-history step decision moved inside either stencil), scaled by the column's largest."""
    return np.abs(d1 - d2) <= tol * np.abs(d2).max(axis=0, keepdims=True)


def test_vjp_matches_oracle_frozen_mesh_differences(deb, oracle, cfg2):
    """The oracle replays the GPU's accepted-step mesh (fractions of each mode's range) on its own
    backgrounds of shifted cosmologies, at every fourth k, for relative steps 3e-5 and 1e-5.  A
    difference quotient that jumps (a switch decision moving inside the stencil) changes with the
    step; where the two agree the adjoint's Jacobian entry must match to 1e-4 of the column's
    largest, and so must the VJP over those entries.  At least 90% of the entries are trusted."""
    c = cfg2
    P, th = c["P"], c["th"]
    sel = np.arange(0, 256, 4)
    k = c["k"][sel]
    lmax = [LM2[x] for x in ("lmaxg", "lmaxgp", "lmaxr", "lmaxnu", "nqmax")]
    mesh, mlen = c["S"]["mesh"].cpu().numpy()[sel], c["S"]["mlen"].cpu().numpy()[sel]
    fds = []
    for rel in (3e-5, 1e-5):
        fd = np.zeros((k.size, 9))
        for j in deb.grad.ODE_PARAMS:
            d = rel * max(abs(th[j]), 1e-2)
            pks = []
            for sgn in (1.0, -1.0):
                t = th.copy()
                t[j] += sgn * d
                p = deb.theta_to_param(t, P)
                bg = oracle.Background(p, 256)
                y = bg.replay_norm(k, bg.spline("tau_of_a", np.array([1.0])), mesh, mlen, lmax, nthreads=16)
                pks.append(oracle.get_power(k, y[:, 0, :], 4, p))
            fd[:, j] = (pks[0] - pks[1]) / (2 * d)
        fds.append(fd)
    cols = list(deb.grad.ODE_PARAMS)
    ok = _trusted(fds[0][:, cols], fds[1][:, cols])
    Ja = c["J"][sel][:, cols]
    err = np.abs(Ja - fds[1][:, cols]) / np.abs(fds[1][:, cols]).max(axis=0, keepdims=True)
    cot = c["cot"][sel]
    for i, j in enumerate(cols):
        m = ok[:, i]
        va, vf = cot[m] @ Ja[m, i], cot[m] @ fds[1][m, j]
        print(f"  {deb.PARAM9[j]:8s} trusted {m.sum():2d}/{m.size}  max entry diff {err[m, i].max():.1e} (untrusted {err[~m, i].max() if (~m).any() else 0:.1e})"
              f"  VJP adjoint {va: .8e} differences {vf: .8e}")
        assert err[m, i].max() <= 1e-4
        assert abs(va - vf) <= 1e-4 * np.abs(cot[m] * fds[1][m, j]).sum()
    assert ok.mean() >= 0.9
    # the VJP launch (cotangent-weighted adjoint) and the Jacobian rows (unit cotangents) agree
    np.testing.assert_allclose(c["v"][cols], c["cot"] @ c["J"][:, cols], rtol=1e-10)
    np.testing.assert_allclose(c["v"][[3, 4]], [c["cot"] @ (c["pk"] * np.log(c["k"] / 0.05)), c["cot"] @ c["pk"]], rtol=1e-12)


def test_dot_product_jvp_vjp(deb, cfg2):
    """<cot, J u> = <u, J^T cot>: J u by central differences of the frozen-mesh map along a random
    direction u of the six ODE parameters (GPU replays at two steps; modes whose quotient changes
    with the step by more than 1e-4 of the largest are masked out of cot), J^T cot by one adjoint
    launch with the masked cotangent."""
    c = cfg2
    P, th, k = c["P"], c["th"], c["k"]
    S = c["S"]
    u = np.zeros(9)
    u[list(deb.grad.ODE_PARAMS)] = np.random.default_rng(5).standard_normal(6) * np.abs(th[list(deb.grad.ODE_PARAMS)])
    jvps = []
    for eps in (1e-5, 3e-6):
        params = [deb.theta_to_param(th + eps * u, P), deb.theta_to_param(th - eps * u, P)]
        tab = deb.background_tables(params, 256)
        tau = deb.tau_out_all(tab, 256, [1.0])
        uc, uk, _ = deb.batch.units_for([k, k], "cuda")
        y, st, _, _ = deb.evolve_units(tab, 256, uc, uk, tau, **LM2, **SOLVER,
                                       replay=(torch.cat([S["mesh"], S["mesh"]]), torch.cat([S["mlen"], S["mlen"]])))
        pk = deb.power_units(tab, 256, uc, uk, y, 4).cpu().numpy()[:, 0].reshape(2, -1)
        jvps.append((pk[0] - pk[1]) / (2 * eps))
    ok = _trusted(jvps[0][:, None], jvps[1][:, None], tol=1e-4)[:, 0]
    cot = np.where(ok, c["cot"], 0.0)
    _, v = deb.pk_and_vjp(th, P, k, cot, **LM2, **SOLVER)
    lhs, rhs = cot @ jvps[1], u @ v.cpu().numpy()
    print(f"  trusted modes {ok.sum()}/{ok.size}   <cot, J u> {lhs:.10e}   <u, J^T cot> {rhs:.10e}   rel {abs(lhs - rhs) / abs(rhs):.2e}")
    assert ok.mean() >= 0.8
    assert abs(lhs - rhs) <= 1e-4 * abs(rhs)


def test_lambda_base_gradient(deb):
    """At w0 = -1, wa = 0 (cosmological constant) the gradient is defined: the w0 and wa columns by
    differences around the fluid's vanishing perturbations, the rest by the adjoint."""
    P = dict(base_param(), w_DE_0=-1.0, w_DE_a=0.0)
    th = deb.param_to_theta(P)
    k = deb.k_grid({}, 1e-3, 0.3, 4, "log")
    print("case: gradient at a cosmological constant, 4 k", flush=True)
    pk, J = deb.pk_jacobian(th, P, k, **LM, **SOLVER)
    J = J.cpu().numpy()
    assert np.all(np.isfinite(J))
    _, Jf = deb.pk_jacobian(th, P, k, method="fd", **LM, **SOLVER)
    Jf = Jf.cpu().numpy()
    for j in deb.grad.ODE_PARAMS:
        rel = np.abs(J[:, 0, j] - Jf[:, 0, j]).max() / np.abs(Jf[:, 0, j]).max()
        print(f"  d P / d {deb.PARAM9[j]:8s}: adjoint vs differences {rel:.2e}")
        assert rel < 1e-4


def test_phantom_cosmology_inside_a_batch(deb):
    """One cosmology whose w(a) crosses -1 gets NaN and status 4; the others are unaffected."""
    P = base_param()
    th0 = deb.param_to_theta(P)
    bad = th0.copy()
    bad[6], bad[7] = -1.05, 0.2  # w(1) = -1.05, w(0) = -0.85: crosses -1
    k = deb.k_grid({}, 1e-3, 0.3, 3, "log")
    print("case: phantom-divide cosmology in a batch of three", flush=True)
    pk, J, st = deb.pk_jacobian_many(np.stack([th0, bad, th0]), P, k, return_status=True, **LM, **SOLVER)
    assert list(st) == [0, 4, 0]
    assert torch.isnan(J[1]).all() and torch.isnan(pk[1]).all()
    assert torch.equal(J[0], J[2]) and torch.isfinite(J[0]).all()
    with pytest.raises(ValueError):
        deb.pk_jacobian(bad, P, k, **LM, **SOLVER)


def test_many_units_batch_matches_single_calls(deb):
    """Above 512 units the base solves run at 128 threads per unit: 3 cosmologies x 200 k give,
    cosmology by cosmology, bitwise the single-cosmology P(k) and Jacobian."""
    P = base_param()
    th0 = deb.param_to_theta(P)
    thetas = np.stack([th0, th0 * (1 + 0.02 * np.array([1, -1, 1, 0, 0, 1, 1, -1, 0])), th0 * (1 - 0.03 * np.array([1, 1, -1, 0, 0, -1, 1, 1, 0]))])
    k = deb.k_grid({}, 1e-4, 0.5, 200, "log")
    print("case: 3 cosmologies x 200 k (600 units) vs one call per cosmology", flush=True)
    pkm, Jm = deb.pk_jacobian_many(thetas, P, k, **LM, **SOLVER)
    for c in range(3):
        pk, J = deb.pk_jacobian(thetas[c], P, k, **LM, **SOLVER)
        assert torch.equal(pkm[c], pk), c
        # the base solves ran at different widths (results independent of it), the adjoint alike
        assert torch.equal(Jm[c], J), c
