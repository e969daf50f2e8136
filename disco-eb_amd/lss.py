"""Large-scale-structure observables on the solver's output, on the GPU: the FFTLog correlation
function, Savitzky-Golay smoothing, the smoothed power spectrum, the Kaiser anisotropic spectrum
and its multipoles (the reference's perturbations.py:1409-1569 and util.py:291-328), with the
reference's keyword signatures.  y is one output time of evolve_perturbations: (num_k, 20) or a
full state whose fields 4 and 5 are delta_m and theta_m / aH.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _dev_vec(x, dev=None):
    t = torch.as_tensor(x, dtype=torch.float64)
    return t.to(dev if dev is not None else (t.device if t.is_cuda else _lib.device())).reshape(-1).contiguous()


def _dev_fields(y, n):
    yt = torch.as_tensor(y, dtype=torch.float64)
    yt = (yt if yt.is_cuda else yt.to(_lib.device())).contiguous()
    if yt.ndim != 2 or yt.shape[0] != n:
        raise ValueError(f"y must be (num_k, n_fields) with num_k = {n}")
    return yt


def savgol_filter(*, y, window_length, polyorder):
    """Savitzky-Golay filter (util.py:291-328): the local least-squares polynomial of degree
    polyorder over window_length points evaluated at each point, by 'same'-length convolution
    (zero beyond the ends).  Returns a float64 device tensor."""
    yt = _dev_vec(y)
    n = yt.numel()
    w = int(window_length)
    L = _lib.lib()
    c = torch.empty(max(w, 1), dtype=torch.float64, device=yt.device)
    _lib.check(L.deb_savgol_coefficients(w, int(polyorder), _lib.ptr(c), _lib.stream()))
    out = torch.empty_like(yt)
    _lib.check(L.deb_savgol_filter(_lib.ptr(yt), n, _lib.ptr(c), w, _lib.ptr(out), _lib.stream()))
    return out


def _window(k, dlogk):
    """window_length of get_power_smoothed (perturbations.py:1496-1497): dlogk over the first log-k
    spacing, rounded half to even, made odd."""
    k01 = torch.as_tensor(k, dtype=torch.float64).reshape(-1)[:2].cpu().numpy()
    w = int(np.round(dlogk / (np.log(k01[1]) - np.log(k01[0]))))
    return w + (w + 1) % 2


def get_power_smoothed(*, k, y, dlogk, idx, param):
    """Savitzky-Golay smoothed power spectrum of field idx (perturbations.py:1473-1505): cubic
    filter of log P over a window of dlogk in log k; the points within half a window of the ends
    keep the unsmoothed P.  A window below 5 points (dlogk under about 4 k-spacings) or above num_k
    raises ValueError."""
    kt = _dev_vec(k)
    n = kt.numel()
    yt = _dev_fields(y, n)
    w = _window(kt, dlogk)
    L = _lib.lib()
    out = torch.empty(n, dtype=torch.float64, device=kt.device)
    work = torch.empty(L.deb_power_smoothed_workspace_doubles(n, w), dtype=torch.float64, device=kt.device)
    _lib.check(L.deb_power_smoothed(_lib.ptr(kt), _lib.ptr(yt), n, yt.shape[1], int(idx), float(param["A_s"]), float(param["n_s"]),
                                    float(param.get("k_p", 0.05)), w, _lib.ptr(out), _lib.ptr(work), _lib.stream()))
    return out


def power_Kaiser(*, y, kmodes, bias, mu_sampling=True, smooth_dlogk=None, nmu, param):
    """Anisotropic spectrum (b delta_m - mu^2 theta_m)^2 (perturbations.py:1507-1543).  Returns
    (P(k, mu) (num_k, nmu), mu (nmu,)) device tensors."""
    kt = _dev_vec(kmodes)
    n = kt.numel()
    yt = _dev_fields(y, n)
    L = _lib.lib()
    pd = pt = None
    if smooth_dlogk is not None:
        pd = get_power_smoothed(k=kt, y=yt, dlogk=smooth_dlogk, idx=4, param=param)
        pt = get_power_smoothed(k=kt, y=yt, dlogk=smooth_dlogk, idx=5, param=param)
    pkmu = torch.empty((n, int(nmu)), dtype=torch.float64, device=kt.device)
    mu = torch.empty(int(nmu), dtype=torch.float64, device=kt.device)
    _lib.check(L.deb_power_kaiser(_lib.ptr(kt), _lib.ptr(yt), n, yt.shape[1], _lib.ptr(pd) if pd is not None else None,
                                  _lib.ptr(pt) if pt is not None else None, float(bias), float(param["A_s"]), float(param["n_s"]),
                                  float(param.get("k_p", 0.05)), int(nmu), 1 if mu_sampling else 0, _lib.ptr(pkmu), _lib.ptr(mu),
                                  _lib.stream()))
    return pkmu, mu


def power_multipoles(*, y, kmodes, b, param):
    """Monopole, quadrupole and hexadecapole of the Kaiser spectrum (perturbations.py:1547-1569)."""
    kt = _dev_vec(kmodes)
    n = kt.numel()
    yt = _dev_fields(y, n)
    out = [torch.empty(n, dtype=torch.float64, device=kt.device) for _ in range(3)]
    _lib.check(_lib.lib().deb_power_multipoles(_lib.ptr(kt), _lib.ptr(yt), n, yt.shape[1], float(b), float(param["A_s"]),
                                               float(param["n_s"]), float(param.get("k_p", 0.05)), *[_lib.ptr(o) for o in out],
                                               _lib.stream()))
    return tuple(out)


def get_xi_from_P(*, k, Pk, N=None, ell=0):
    """Correlation-function multipole xi_l(r) from P(k) on a log-uniform k-grid by FFTLog
    (perturbations.py:1409-1443; N is taken from len(k), as there).  len(k) must be even.
    Returns (xi, r) device tensors in increasing r = 2 pi / k."""
    kt = _dev_vec(k)
    n = kt.numel()
    pt = _dev_vec(Pk, kt.device)
    if pt.numel() != n:
        raise ValueError("k and Pk must have the same length")
    L = _lib.lib()
    xi = torch.empty(n, dtype=torch.float64, device=kt.device)
    r = torch.empty(n, dtype=torch.float64, device=kt.device)
    work = torch.empty(max(L.deb_xi_workspace_doubles(n), 1), dtype=torch.float64, device=kt.device)
    _lib.check(L.deb_xi_from_power(_lib.ptr(kt), _lib.ptr(pt), n, int(ell), _lib.ptr(xi), _lib.ptr(r), _lib.ptr(work), _lib.stream()))
    return xi, r


__all__ = ["savgol_filter", "get_power_smoothed", "power_Kaiser", "power_multipoles", "get_xi_from_P"]
