"""P(k) gradients with respect to the nine sampling parameters (HIP path).

The gradient is that of the frozen-mesh map (the base solve's accepted step ends held as
fractions of each mode's range), by the discrete adjoint (the default method).  Checked against
the oracle's replay of the same normalized mesh on shifted cosmologies (central differences,
same steps), against the closed forms for A_s, n_s and tau_reion, for the step of the tables'
differences, and through torch.autograd.
"""
import numpy as np
import pytest
import torch

from conftest import load_package

pytestmark = pytest.mark.gpu
LM = dict(lmaxg=8, lmaxgp=8, lmaxr=8, lmaxnu=8, nqmax=3)
SOLVER = dict(rtol=1e-5, atol=1e-10)


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def base_param():
    """Massive neutrino (0.06 eV) + w0wa (-0.95, 0.1) around omega_c h^2 = 0.120, omega_b h^2 = 0.0224, h = 0.674."""
    h = 0.674
    return dict(Omegam=(0.120 + 0.0224) / h ** 2, Omegab=0.0224 / h ** 2, H0=100 * h, Tcmb=2.7255, YHe=0.248, Neff=2.046, Nmnu=1.0,
                mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)


@pytest.fixture(scope="module")
def jac(deb):
    th = deb.param_to_theta(base_param())
    k = deb.k_grid({}, 1e-4, 0.5, 5, "log")
    pk, J = deb.pk_jacobian(th, base_param(), k, **LM, **SOLVER)
    return th, k, pk.cpu().numpy(), J.cpu().numpy()


def test_replay_of_own_mesh_reproduces_the_adaptive_solve(deb):
    print("case: replay of the base mesh, 5 k", flush=True)
    P = base_param()
    k = torch.as_tensor(deb.k_grid({}, 1e-4, 0.5, 5, "log"), device="cuda")
    tab = deb.background_tables([P], 256)
    tau = deb.tau_out_all(tab, 256, [0.5, 1.0])
    uc = torch.zeros(5, dtype=torch.int32, device="cuda")
    y, st, na, nr, log = deb.evolve_units(tab, 256, uc, k, tau, **LM, **SOLVER, step_log=True)
    mesh, mlen = deb.grad._accepted_mesh(log, na, np.full(5, float(tau.max().cpu())))
    yr, st2, na2, nr2 = deb.evolve_units(tab, 256, uc, k, tau, **LM, **SOLVER, replay=(mesh, mlen))
    assert int(st2.abs().max()) == 0 and torch.equal(na2.cpu(), na.cpu()) and int(nr2.abs().max()) == 0
    err = (yr - y).abs().max() / y.abs().max()
    print(f"  replay vs adaptive max scaled diff {float(err):.2e}")
    assert float(err) < 1e-10


def test_closed_form_columns(deb, jac):
    th, k, pk, J = jac
    np.testing.assert_array_equal(J[:, :, 4], pk)
    np.testing.assert_allclose(J[:, :, 3], pk * np.log(k / 0.05)[:, None], rtol=1e-15)
    assert np.all(J[:, :, 8] == 0.0)


def test_gradient_matches_oracle_frozen_mesh(deb, oracle, jac):
    """The oracle replays the same normalized accepted-step mesh on the twelve shifted
    cosmologies; the central differences agree to 1e-4 relative (the gradient tolerance)."""
    print("case: frozen-mesh gradient vs oracle, 5 k, 6 ODE parameters", flush=True)
    th, k, pk, J = jac
    P = base_param()
    tab = deb.background_tables([P], 256)
    tau = deb.tau_out_all(tab, 256, [1.0])
    kt = torch.as_tensor(k, device="cuda")
    uc = torch.zeros(k.size, dtype=torch.int32, device="cuda")
    y, st, na, nr, log = deb.evolve_units(tab, 256, uc, kt, tau, **LM, **SOLVER, step_log=True)
    mesh, mlen = deb.grad._accepted_mesh(log, na, np.full(k.size, float(tau.max().cpu())))
    mesh, mlen = mesh.cpu().numpy(), mlen.cpu().numpy()
    kw = dict(LM, **SOLVER)
    for j in deb.grad.ODE_PARAMS:
        d = deb.grad.REL_STEP * max(abs(th[j]), 1e-2)
        pks = []
        for sgn in (1.0, -1.0):
            t = th.copy()
            t[j] += sgn * d
            p = deb.theta_to_param(t, P)
            bg = oracle.Background(p, 256)
            t1 = float(bg.spline("tau_of_a", np.array([1.0]))[0])
            row = []
            for i, kk in enumerate(k):
                t0 = bg.start_tau(kk, [1.0])
                s = mesh[i, :mlen[i]]
                ends = np.where(s >= 1.0, t1, t0 + s * (t1 - t0))
                starts = np.concatenate([[t0], ends[:-1]])
                yo, _ = bg.replay(kk, [1.0], np.stack([starts, ends], axis=1), np.ones(ends.size, dtype=np.int32), **kw)
                row.append(oracle.get_power(np.array([kk]), yo[None, -1, :], 4, p)[0])
            pks.append(np.array(row))
        go = (pks[0] - pks[1]) / (2 * d)
        rel = np.abs(J[:, 0, j] - go) / np.abs(go).max()
        print(f"  d P / d {deb.PARAM9[j]:8s}: max rel diff {rel.max():.2e}")
        assert rel.max() < 1e-4, deb.PARAM9[j]


def test_step_size_convergence(deb, jac):
    print("case: central differences at 3e-5 and 1e-5 relative steps", flush=True)
    th, k, pk, J = jac
    _, J3 = deb.pk_jacobian(th, base_param(), k, rel_step=1e-5, **LM, **SOLVER)
    J3 = J3.cpu().numpy()
    for j in deb.grad.ODE_PARAMS:
        rel = np.abs(J3[:, 0, j] - J[:, 0, j]) / np.abs(J[:, 0, j]).max()
        print(f"  {deb.PARAM9[j]:8s}: max rel change {rel.max():.2e}")
        assert rel.max() < 1e-4, deb.PARAM9[j]


def test_autograd_function_returns_the_vjp(deb, jac):
    th, k, pk, J = jac
    g = np.random.default_rng(0).standard_normal(k.size)
    theta = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    out = deb.PkFunction.apply(theta, k, base_param(), dict(LM, **SOLVER))
    (out[:, 0].cpu() * torch.tensor(g)).sum().backward()
    # backward runs the adjoint with this cotangent; J's rows came from unit cotangents (the
    # kernel is linear in the cotangent, so the two agree to rounding)
    np.testing.assert_allclose(theta.grad.numpy(), J[:, 0, :].T @ g, rtol=1e-10, atol=0)
    np.testing.assert_array_equal(out.detach().cpu().numpy(), pk)


def test_many_cosmologies_match_one_at_a_time(deb):
    """pk_jacobian_many (two perturbation launches for all cosmologies) gives, cosmology by
    cosmology, bitwise the P(k) and Jacobian of separate pk_jacobian calls."""
    print("case: 3 cosmologies x 4 k, batched gradient vs one call per cosmology", flush=True)
    P = base_param()
    th0 = deb.param_to_theta(P)
    thetas = np.stack([th0, th0 * (1 + 0.02 * np.array([1, -1, 1, 0, 0, 1, 1, -1, 0])),
                       th0 * (1 - 0.03 * np.array([1, 1, -1, 0, 0, -1, 1, 1, 0]))])
    k = deb.k_grid({}, 1e-4, 0.5, 4, "log")
    pkm, Jm = deb.pk_jacobian_many(thetas, P, k, aexp_out=(0.5, 1.0), **LM, **SOLVER)
    assert pkm.shape == (3, 4, 2) and Jm.shape == (3, 4, 2, 9)
    for c in range(3):
        pk, J = deb.pk_jacobian(thetas[c], P, k, aexp_out=(0.5, 1.0), **LM, **SOLVER)
        assert torch.equal(pkm[c], pk) and torch.equal(Jm[c], J), c
    assert not torch.equal(Jm[0], Jm[1])


def test_many_units_match_one_at_a_time(deb):
    """3 cosmologies x 200 k = 600 units: the batched base solve takes its 128-thread schedule (more
    than 512 units) and the adjoint's forward replay runs two 128-thread units per compute unit
    (more units than compute units), while one cosmology's 200 units take 256 threads throughout.
    The P(k) and Jacobian of every cosmology are bitwise those of a separate pk_jacobian call."""
    print("case: 3 cosmologies x 200 k (600 units), batched adjoint vs one call per cosmology", flush=True)
    P = base_param()
    th0 = deb.param_to_theta(P)
    thetas = np.stack([th0, th0 * (1 + 0.02 * np.array([1, -1, 1, 0, 0, 1, 1, -1, 0])),
                       th0 * (1 - 0.03 * np.array([1, 1, -1, 0, 0, -1, 1, 1, 0]))])
    k = deb.k_grid({}, 1e-4, 0.5, 200, "log")
    pkm, Jm = deb.pk_jacobian_many(thetas, P, k, aexp_out=(1.0,), **LM, **SOLVER)
    assert pkm.shape == (3, 200, 1) and Jm.shape == (3, 200, 1, 9)
    assert torch.isfinite(Jm).all()
    for c in range(3):
        pk, J = deb.pk_jacobian(thetas[c], P, k, aexp_out=(1.0,), **LM, **SOLVER)
        assert torch.equal(pkm[c], pk) and torch.equal(Jm[c], J), c


@pytest.mark.parametrize("method", ["adjoint", "fd"])
def test_ragged_grids_match_one_at_a_time(deb, method):
    """Each cosmology on its own k-grid (lengths 5, 3, 6, 4), one of them across the phantom divide
    and one an exact cosmological constant: the batched Jacobian of every valid cosmology equals a
    separate pk_jacobian call bit for bit; the phantom one is NaN with status 4.  pk_and_vjp_many
    on the same grids equals per-cosmology pk_and_vjp to rounding (sums in another order)."""
    print(f"case: ragged grids 5/3/6/4 k, 4 cosmologies (one phantom, one Lambda), method {method}", flush=True)
    P = base_param()
    th0 = deb.param_to_theta(P)
    th_ph = th0.copy()
    th_ph[6], th_ph[7] = -1.02, 0.1  # w crosses -1
    th_l = th0.copy()
    th_l[6], th_l[7] = -1.0, 0.0
    thetas = np.stack([th0, th_ph, th0 * (1 + 0.02 * np.array([1, -1, 1, 0, 0, 1, 1, -1, 0])), th_l])
    grids = [deb.k_grid({}, 1e-4, 0.5, n, "log") for n in (5, 3, 6)] + [deb.k_grid({}, 2e-4, 0.3, 4, "log")]
    pkm, Jm, st = deb.pk_jacobian_many(thetas, P, grids, aexp_out=(0.5, 1.0), return_status=True, method=method, **LM, **SOLVER)
    assert isinstance(pkm, list) and [x.shape[0] for x in pkm] == [5, 3, 6, 4]
    assert st.tolist()[1] == 4 and torch.isnan(Jm[1]).all() and torch.isnan(pkm[1]).all()
    for c in (0, 2, 3):
        pk, J = deb.pk_jacobian(thetas[c], P, grids[c], aexp_out=(0.5, 1.0), method=method, **LM, **SOLVER)
        assert st[c] == 0 and torch.equal(pkm[c], pk) and torch.equal(Jm[c], J), c
    if method == "adjoint":
        rng = np.random.default_rng(5)
        cots = [rng.standard_normal((g.size, 1)) for g in grids]
        pkv, v, sv = deb.pk_and_vjp_many(thetas, P, grids, cots, **LM, **SOLVER)
        assert sv[1] == 4 and torch.isnan(v[1]).all()
        for c in (0, 2, 3):
            _, vc = deb.pk_and_vjp(thetas[c], P, grids[c], cots[c][:, 0], **LM, **SOLVER)
            np.testing.assert_allclose(v[c].cpu().numpy(), vc.cpu().numpy(), rtol=1e-12, atol=0)
