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


THERMO_MODULES = {"RECFAST": 0, "MB95": 1, "CLASS": 2}
CLASS_COLUMNS = ("scale factor a", "conf. time [Mpc]", "x_e", "Tb [K]", "c_b^2")


def _class_table(class_thermo, dev):
    """The CLASS thermodynamics columns (CLASS order, decreasing a) as the 5 x n block the kernel
    reads: a ascending, conformal time, x_e, T_b, c_b^2 (background.py:271-285 reverses them)."""
    cols = []
    for name in CLASS_COLUMNS:
        if name not in class_thermo:
            raise ValueError(f"class_thermo lacks the column {name!r}")
        cols.append(torch.flip(torch.as_tensor(np.ascontiguousarray(class_thermo[name], dtype=np.float64)), dims=[0]))
    n = cols[0].numel()
    if any(c.numel() != n for c in cols) or n < 16:
        raise ValueError("class_thermo columns must have one common length >= 16")
    if not bool((cols[0][1:] > cols[0][:-1]).all()) or not bool((cols[1][1:] > cols[1][:-1]).all()):
        raise ValueError("class_thermo's scale factor and conformal time must decrease along the table (CLASS order)")
    return torch.stack(cols).to(dev).contiguous(), n


def background_tables(params, num_thermo=256, check=True, thermo_module="RECFAST", class_thermo=None):
    """Background tables of several cosmologies at once: one workgroup per cosmology.

    params: list of param dicts.  thermo_module: "RECFAST", "MB95" or "CLASS" (then class_thermo:
    one dict of CLASS thermodynamics columns, or a list with one per cosmology, each of the same
    length, which replaces num_thermo).  Returns a (n_cosmo, deb_tables_size) float64 device
    tensor.  With check (and outside graph capture) a RECFAST interval solve that exceeded its
    128 steps in any cosmology raises RuntimeError, as attach_background does for one."""
    if thermo_module not in THERMO_MODULES:
        raise ValueError(f"thermo_module must be one of {sorted(THERMO_MODULES)}")
    L = _lib.lib()
    dev = _lib.device()
    pv = torch.tensor([_lib.param_vector(p) for p in params], dtype=torch.float64, device=dev)
    tin = None
    if thermo_module == "CLASS":
        if class_thermo is None:
            raise ValueError("thermo_module='CLASS' needs class_thermo")
        blocks = [_class_table(c, dev) for c in (class_thermo if isinstance(class_thermo, (list, tuple)) else [class_thermo] * len(params))]
        if len(blocks) != len(params) or len({n for _, n in blocks}) != 1:
            raise ValueError("one class_thermo table per cosmology, all of one length")
        num_thermo = blocks[0][1]
        tin = torch.stack([b for b, _ in blocks]).contiguous()
    size = L.deb_tables_size(int(num_thermo))
    tables = torch.empty((len(params), size), dtype=torch.float64, device=dev)
    _lib.check(L.deb_background_thermo(_lib.ptr(pv), len(params), int(num_thermo), THERMO_MODULES[thermo_module],
                                       _lib.ptr(tin) if tin is not None else None, _lib.ptr(tables), _lib.stream()))
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
    """Background evolution and thermal history on the GPU.

    Same signature and returned keys as the reference (background.py:191-345).
    thermo_module: 'RECFAST' (default), 'MB95' (num_thermo points log-uniform in conformal time,
    thermodynamics_mb95.py) or 'CLASS' (the thermal history of class_thermo, a dict of CLASS
    thermodynamics columns 'scale factor a', 'conf. time [Mpc]', 'x_e', 'Tb [K]', 'c_b^2' in CLASS
    order; its length replaces num_thermo and its a-range sets amin, amax).  rtol/atol/order are
    accepted and, as in the reference, not used (thermodynamics_recfast.py:461).  Keys the
    reference sets only for RECFAST (mu, xeprime_recf) hold zeros for the other modules.
    An extension: param["tau_reion"] > 0 adds reionization to the RECFAST history (the tanh
    model of Lewis 2008 as CAMB uses it, hydrogen and both helium ionisations; the reference has
    none), its redshift solved so that the optical depth it adds is tau_reion; param then also
    holds z_reion and tau_reion_achieved.  The other thermal-history modules ignore it."""
    tables = background_tables([param], num_thermo, thermo_module=thermo_module, class_thermo=class_thermo)
    n = len(class_thermo[CLASS_COLUMNS[0]]) if thermo_module == "CLASS" else num_thermo
    return attach_background(param, tables, n, 0)


def compute_background_quantities(aexp, param):
    """rho_nu/rho_nu0, rho_Q/rho_Q(1) and w_Q at scale factors aexp (background.py:348-406)."""
    a = torch.as_tensor(aexp, dtype=torch.float64, device=param["_deb_tables"].device).reshape(-1)
    rhonu = torch.exp(param["logrhonu_of_loga_spline"].evaluate(torch.log(a)))
    w0, wa = param["w_DE_0"], param["w_DE_a"]
    rho_Q = a ** (-3 * (1 + w0 + wa)) * torch.exp(3 * (a - 1) * wa)
    return {"rhonu": rhonu, "rho_Q": rho_Q, "w_Q": w0 + wa * (1.0 - a)}
