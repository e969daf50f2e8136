"""evolve_background on the GPU (reference: src/discoeb/background.py:191-345).

The returned ``param`` dict carries the same keys as the reference's: scalars as Python floats,
arrays as float64 device tensors (views into one HBM table block) and splines as
``DeviceSpline`` objects with the reference's evaluate/derivative/derivative2/integral methods.
"""
import numpy as np
import torch

from . import _lib


class DeviceSpline:
    """Natural cubic spline on the background tables (spline_interpolation.py:6-276)."""

    def __init__(self, tables, num_thermo, cosmo, name):
        self._t = tables
        self._n = int(num_thermo)
        self._c = int(cosmo)
        self._sid = _lib.SPLINES.index(name)
        self.name = name

    def _call(self, x, kind):
        scalar = np.ndim(x) == 0 and not (torch.is_tensor(x) and x.ndim > 0)
        xt = torch.as_tensor(x, dtype=torch.float64, device=self._t.device).reshape(-1).contiguous()
        out = torch.empty_like(xt)
        _lib.check(_lib.lib().deb_spline_eval(_lib.ptr(self._t), self._n, self._c, self._sid, kind, _lib.ptr(xt), xt.numel(),
                                              _lib.ptr(out), _lib.stream()))
        return out[0] if scalar else out.reshape(torch.as_tensor(x).shape)

    def evaluate(self, x):
        return self._call(x, 0)

    def derivative(self, x):
        return self._call(x, 1)

    def derivative2(self, x):
        return self._call(x, 2)

    def derivative12(self, x):
        return self._call(x, 1), self._call(x, 2)

    def integral(self, x):
        return self._call(x, 3)


def recfast_failures(tables):
    """Per-cosmology count of RECFAST interval solves that exceeded their step limit (header slot)."""
    return tables[:, _lib.HDR["recfast_failures"]]


def background_tables(params, num_thermo=256, check=True):
    """Background tables of several cosmologies at once: one workgroup per cosmology.

    params: list of param dicts.  Returns a (n_cosmo, deb_tables_size) float64 device tensor.
    With check (and outside graph capture) a RECFAST interval solve that exceeded its 128
    steps in any cosmology raises RuntimeError, as attach_background does for one."""
    L = _lib.lib()
    dev = _lib.device()
    pv = torch.tensor([_lib.param_vector(p) for p in params], dtype=torch.float64, device=dev)
    size = L.deb_tables_size(int(num_thermo))
    tables = torch.empty((len(params), size), dtype=torch.float64, device=dev)
    _lib.check(L.deb_background(_lib.ptr(pv), len(params), int(num_thermo), _lib.ptr(tables), _lib.stream()))
    if check and not torch.cuda.is_current_stream_capturing():
        bad = torch.nonzero(recfast_failures(tables) != 0).flatten().cpu().tolist()
        if bad:
            raise RuntimeError(f"RECFAST interval solve exceeded 128 steps (cosmologies {bad})")
    return tables


def attach_background(param, tables, num_thermo, cosmo=0):
    """Fill ``param`` with the keys evolve_background sets (background.py:145-345)."""
    N = int(num_thermo)
    t = tables[cosmo] if tables.ndim == 2 else tables
    hdr = t[:64].detach().cpu().numpy()
    for k, slot in _lib.HDR.items():
        param[k] = float(hdr[slot])
    param["amin"], param["amax"] = 1e-9, 1.01

    def arr(name, n=N):
        off = _lib.table_offset(name, N)
        return t[off:off + n]

    for name in ["aexp", "tau", "xe", "xeHI", "xeHeI", "xeHeII", "cs2", "Tm", "mu", "xeprime", "opac", "optical_depth",
                 "gvis", "gvisprime", "gvispprime"]:
        param[name] = arr(name)
    param["xeprime_recf"] = arr("xeprime_recf")
    param["a"] = torch.exp(arr("nu_loga", _lib.NNU))
    for s in _lib.SPLINES:
        key = s + "_spline"
        param[key] = DeviceSpline(tables, N, cosmo, s)
    param["_deb_tables"] = tables
    param["_deb_num_thermo"] = N
    param["_deb_cosmo"] = cosmo
    if param["recfast_failures"] != 0:
        raise RuntimeError("RECFAST interval solve exceeded 128 steps")
    return param


def evolve_background(*, param, thermo_module="RECFAST", num_thermo: int = 256, rtol: float = 1e-5, atol: float = 1e-7,
                      order: int = 5, class_thermo=None):
    """Background evolution and RECFAST thermal history on the GPU.

    Same signature and returned keys as the reference (background.py:191-345).  Only the
    'RECFAST' thermal history is provided; rtol/atol/order are accepted and, as in the
    reference, not used by the RECFAST path (thermodynamics_recfast.py:461)."""
    if thermo_module != "RECFAST":
        raise NotImplementedError(f"thermo_module={thermo_module!r}: only 'RECFAST' is provided")
    tables = background_tables([param], num_thermo)
    return attach_background(param, tables, num_thermo, 0)


def compute_background_quantities(aexp, param):
    """rho_nu/rho_nu0, rho_Q/rho_Q(1) and w_Q at scale factors aexp (background.py:348-406)."""
    a = torch.as_tensor(aexp, dtype=torch.float64, device=param["_deb_tables"].device).reshape(-1)
    rhonu = torch.exp(param["logrhonu_of_loga_spline"].evaluate(torch.log(a)))
    w0, wa = param["w_DE_0"], param["w_DE_a"]
    rho_Q = a ** (-3 * (1 + w0 + wa)) * torch.exp(3 * (a - 1) * wa)
    return {"rhonu": rhonu, "rho_Q": rho_Q, "w_Q": w0 + wa * (1.0 - a)}
