"""Work distribution, many-cosmology batches and graph capture (HIP path).

Every unit's arithmetic is independent of how units are spread over workgroups, of which
other cosmologies share the launch and of graph capture, so all comparisons here are bitwise.
Agreement of the single-cosmology path with the oracle is covered in test_gpu_parity.py.
"""
import numpy as np
import pytest
import torch

from conftest import load_package

pytestmark = pytest.mark.gpu

LMAX = dict(lmaxg=8, lmaxgp=8, lmaxr=8, lmaxnu=8, nqmax=3)
SOLVER = dict(rtol=1e-4, atol=1e-8)
AOUT = [0.1, 0.5, 1.0]


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def lcdm():
    h = 0.674
    return dict(Omegam=(0.120 + 0.0224) / h ** 2, Omegab=0.0224 / h ** 2, H0=100 * h, Tcmb=2.7255, YHe=0.248, Neff=3.046,
                Nmnu=0.0, mnu=0.0, w_DE_0=-1.0, w_DE_a=0.0, cs2_DE=1.0, Omegak=0.0, A_s=2.1e-9, n_s=0.965, k_p=0.05)


def nu_w0wa():
    p = lcdm()
    p.update(Neff=2.046, Nmnu=1.0, mnu=0.06, w_DE_0=-0.95, w_DE_a=0.1)
    return p


def phantom_crossing():
    p = nu_w0wa()
    p.update(w_DE_0=-1.05, w_DE_a=0.1)  # w(a) runs from -0.95 (a -> 0) to -1.05 (a = 1)
    return p


def _same(a, b):
    a, b = a.cpu(), b.cpu()
    return a.shape == b.shape and bool(torch.equal(torch.nan_to_num(a, nan=7.0), torch.nan_to_num(b, nan=7.0)))


def test_schedules_give_identical_results(deb):
    print("case: launch orders index / largest k first / random permutation", flush=True)
    params = [lcdm(), nu_w0wa()]
    kgrids = [np.geomspace(1e-4, 0.5, 5), np.geomspace(3e-4, 2.0, 7)]
    tables = deb.background_tables(params, 256)
    tau = deb.tau_out_all(tables, 256, AOUT)
    uc, uk, _ = deb.batch.units_for(kgrids, tables.device)
    ref = deb.evolve_units(tables, 256, uc, uk, tau, **LMAX, **SOLVER)
    assert int(ref[1].abs().max()) == 0
    perm = torch.randperm(uk.numel(), generator=torch.Generator().manual_seed(5)).to(torch.int32).cuda()
    for sch in [{"order": "k_desc"}, {"order": None}, {"order": perm}]:
        got = deb.evolve_units(tables, 256, uc, uk, tau, **LMAX, **SOLVER, schedule=sch)
        for a, b in zip(got, ref):
            assert _same(a, b), sch
    with pytest.raises(ValueError):
        deb.evolve_units(tables, 256, uc, uk, tau, **LMAX, **SOLVER,
                         schedule={"order": torch.zeros(uk.numel(), dtype=torch.int32, device="cuda")})


def test_evolve_many_matches_single_cosmology_path(deb):
    print("case: evolve_many 3 cosmologies (one crossing w = -1) vs evolve_perturbations", flush=True)
    params = [lcdm(), nu_w0wa(), phantom_crossing()]
    # the same k values as evolve_perturbations' log sampling (a last-bit change in k moves the
    # adaptive step sequence, so bitwise comparisons need identical inputs)
    kgrids = [deb.k_grid({}, 1e-4, 0.5, 4, "log"), deb.k_grid({}, 1e-4, 1.0, 6, "log"), deb.k_grid({}, 1e-3, 1.0, 3, "log")]
    ys, pks, st = deb.evolve_many(params, kgrids, AOUT, **LMAX, **SOLVER)
    st = st.cpu().numpy()
    assert (st[:10] == 0).all() and (st[10:] == 4).all()
    assert torch.isnan(ys[2]).all() and torch.isnan(pks[2]).all()
    for c in range(2):
        p = deb.evolve_background(param=dict(params[c]), num_thermo=256)
        y, k, p = deb.evolve_perturbations(param=p, aexp_out=AOUT, kmin=kgrids[c][0], kmax=kgrids[c][-1], num_k=len(kgrids[c]),
                                           k_sampling_method="log", **LMAX, **SOLVER)
        assert np.array_equal(k.cpu().numpy(), kgrids[c])
        assert _same(ys[c], y)
        pk = deb.get_power(k=k, y=y, idx=4, param=params[c])
        assert _same(pks[c], pk)
    p = deb.evolve_background(param=dict(params[2]), num_thermo=256)
    with pytest.raises(ValueError, match="crosses -1"):
        deb.evolve_perturbations(param=p, aexp_out=AOUT, kmin=1e-3, kmax=1.0, num_k=3, k_sampling_method="log", **LMAX, **SOLVER)


def test_graph_replays_match_eager(deb):
    print("case: PkGraph capture + 3 replays with parameter changes", flush=True)
    params = [lcdm(), nu_w0wa()]
    kgrids = [np.geomspace(1e-4, 0.5, 4), np.geomspace(1e-4, 1.0, 5)]
    g = deb.PkGraph(params, kgrids, AOUT, **LMAX, **SOLVER, schedule={"order": "k_desc"})
    variants = [params, [dict(params[0], H0=68.5, A_s=2.3e-9), dict(params[1], mnu=0.08, n_s=0.95)], params]
    for it, ps in enumerate(variants):
        g.set_params(ps)
        pk, st = g.replay()
        torch.cuda.synchronize()
        pk = pk.clone()
        assert int(st.abs().max()) == 0
        _, pks_eager, _ = deb.evolve_many(ps, kgrids, AOUT, **LMAX, **SOLVER, schedule=None)  # eager work between replays
        torch.cuda.synchronize()
        for c, part in enumerate(g.split(pk)):
            assert _same(part, pks_eager[c]), (it, c)
    # the parameter change moved P(k): replays are not stale
    g.set_params(variants[1])
    pk1 = g.replay()[0].clone()
    g.set_params(variants[0])
    pk0 = g.replay()[0].clone()
    assert not torch.equal(pk0, pk1)


@pytest.mark.skip(reason="the persistent work-queue kernel is removed: its 8-momentum-bin instantiation took a GPU memory "
                         "aperture violation (256 units) and gave wrong results on replay, with no cause found in the source; "
                         "units are launched one workgroup each in cost order instead")
def test_queue_matches_static_for_eight_momentum_bins(deb):
    """Work-queue kernel versus one workgroup per unit, 8 momentum bins (kept as a record)."""


def test_cost_order_matches_index_order_for_eight_momentum_bins(deb):
    """The largest dense block (41 unknowns, 8 momentum bins, lmax_nu 17), k up to 5/Mpc
    (radiation-streaming and fluid approximations active), launched in index order and largest-k
    first: bitwise the same results."""
    print("case: nq=8 lmax_nu=17 index order vs largest k first, 6 k 1e-5..5", flush=True)
    p = deb.evolve_background(param=dict(nu_w0wa()), num_thermo=256)
    tau = p["tau_of_a_spline"].evaluate(np.asarray([1 / 3, 0.5, 1.0])).reshape(1, -1)
    kt = torch.as_tensor(deb.k_grid({}, 1e-5, 5.0, 6, "log"), device="cuda")
    uc = torch.zeros(6, dtype=torch.int32, device="cuda")
    kw = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8, rtol=1e-5, atol=1e-10)
    ref = deb.evolve_units(p["_deb_tables"], 256, uc, kt, tau, **kw)
    assert int(ref[1].abs().max()) == 0
    got = deb.evolve_units(p["_deb_tables"], 256, uc, kt, tau, **kw, schedule={"order": "k_desc"})
    for a, b in zip(got, ref):
        assert _same(a, b)


@pytest.mark.parametrize("nq,lmaxnu", [(8, 17), (5, 10), (3, 8), (4, 6)])
def test_workgroup_width_gives_identical_results(deb, nq, lmaxnu):
    """One, two or four wavefronts per unit perform the same operations on every entry (rows,
    matrix elements and Jacobian columns are only distributed differently): bitwise the same
    states, step counts and step meshes, per-mode and batched; nq = 3, 5, 8 run the bordered
    solve, nq = 4 the generic (run-time block size) factorisation."""
    print(f"case: threads per unit 64/128/256, nq={nq} lmax_nu={lmaxnu}, 6 k 1e-5..5", flush=True)
    p = deb.evolve_background(param=dict(nu_w0wa()), num_thermo=256)
    tau = p["tau_of_a_spline"].evaluate(np.asarray([1 / 3, 0.5, 1.0])).reshape(1, -1)
    kt = torch.as_tensor(deb.k_grid({}, 1e-5, 5.0, 6, "log"), device="cuda")
    uc = torch.zeros(6, dtype=torch.int32, device="cuda")
    kw = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=lmaxnu, nqmax=nq, rtol=1e-5, atol=1e-10)
    ref = deb.evolve_units(p["_deb_tables"], 256, uc, kt, tau, **kw, step_log=True, schedule={"threads": 64})
    assert int(ref[1].abs().max()) == 0
    refb = deb.evolve_units_batched(p["_deb_tables"], 256, uc, kt, tau, batch_size=3, **kw, step_log=True, threads=64)
    assert int(refb[1].abs().max()) == 0
    for nt in (128, 256):
        got = deb.evolve_units(p["_deb_tables"], 256, uc, kt, tau, **kw, step_log=True, schedule={"order": "k_desc", "threads": nt})
        for a, b in zip(got, ref):
            assert _same(a, b), nt
        gotb = deb.evolve_units_batched(p["_deb_tables"], 256, uc, kt, tau, batch_size=3, **kw, step_log=True, threads=nt)
        for a, b in zip(gotb, refb):
            assert _same(a, b), nt
