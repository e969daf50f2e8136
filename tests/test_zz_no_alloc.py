"""deb_evolve_batched_noalloc (the batched solve on the library's static per-device scratch)
against deb_evolve_batched with a caller workspace: bit-identical outputs, status, step counts
and step meshes, on two calls in a row (the second smaller, so it runs on what the first left in
the scratch), and a clear error for a problem larger than the scratch.
"""
import numpy as np
import pytest
import torch

from conftest import load_package
from test_gpu_parity import lcdm_param, nu_w0wa_param

pytestmark = pytest.mark.gpu

LM = dict(lmaxg=12, lmaxgp=12, lmaxr=12, lmaxnu=17, nqmax=8)  # nvar 192


@pytest.fixture(scope="module")
def deb():
    if not torch.cuda.is_available():
        pytest.skip("no ROCm device")
    return load_package()


def _units(deb, tab, n_per_cosmo, kmin, kmax):
    k1 = torch.as_tensor(np.geomspace(kmin, kmax, n_per_cosmo), dtype=torch.float64, device=tab.device)
    uk = torch.cat([k1, k1 * 1.1])
    uc = torch.cat([torch.zeros(n_per_cosmo, dtype=torch.int32), torch.ones(n_per_cosmo, dtype=torch.int32)]).to(tab.device)
    return uc, uk


def _same(a, b):
    for x, y in zip(a, b):
        assert torch.equal(torch.nan_to_num(x, nan=7.0), torch.nan_to_num(y, nan=7.0))


def test_static_scratch_matches_workspace_on_two_calls(deb):
    tab = deb.background_tables([lcdm_param(), nu_w0wa_param()], 256)
    tau = deb.tau_out_all(tab, 256, [0.1, 0.5, 1.0])
    kw = dict(rtol=1e-5, atol=1e-10, step_log=True, **LM)
    L = deb._lib.lib()
    print(f"case: static scratch {L.deb_noalloc_scratch_doubles()} doubles; call 1: 2 cosmologies x 48 k, batches of 8", flush=True)
    uc, uk = _units(deb, tab, 48, 1e-4, 1.0)
    ref = deb.evolve_units_batched(tab, 256, uc, uk, tau, batch_size=8, **kw)
    got = deb.evolve_units_batched(tab, 256, uc, uk, tau, batch_size=8, static_scratch=True, **kw)
    torch.cuda.synchronize()
    assert (ref[1] == 0).all()
    _same(got, ref)
    print("case: call 2: 2 cosmologies x 8 k (other k), batches of 4, full state", flush=True)
    uc, uk = _units(deb, tab, 8, 3e-4, 0.5)
    kw["return_full"] = True
    ref = deb.evolve_units_batched(tab, 256, uc, uk, tau, batch_size=4, **kw)
    got = deb.evolve_units_batched(tab, 256, uc, uk, tau, batch_size=4, static_scratch=True, **kw)
    torch.cuda.synchronize()
    assert (ref[1] == 0).all()
    _same(got, ref)


def test_problem_larger_than_the_scratch_is_rejected(deb):
    """n_units x 2 nvar + 32 > deb_noalloc_scratch_doubles(): refused before anything is launched."""
    L = deb._lib.lib()
    tab = deb.background_tables([lcdm_param()], 256)
    tau = deb.tau_out_all(tab, 256, [1.0])
    n = (L.deb_noalloc_scratch_doubles() - 32) // (2 * 192) + 1
    n += (-n) % 16
    uk = torch.full((n,), 0.01, dtype=torch.float64, device=tab.device)
    uc = torch.zeros(n, dtype=torch.int32, device=tab.device)
    with pytest.raises(ValueError, match="static scratch"):
        deb.evolve_units_batched(tab, 256, uc, uk, tau, batch_size=16, static_scratch=True, **LM)
