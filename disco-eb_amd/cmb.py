"""CMB source functions, line of sight and C_ell on the GPU (reference: src/discoeb/cmb.py).

Same names, arguments and return layouts as the reference where the reference's function
takes arrays (compute_visibility_functions, compute_theta_ell, compute_Cell, compute_Dell,
compute_Cell_spectrum_from_cosmo_params).  The reference's chain of dict-passing helpers
(extract_perturbations .. compute_source_function, cmb.py:17-633) runs as one fused kernel
here, compute_source_function_from_states, which returns the same S1..S4 / S planes.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .background import evolve_background
from .perturbations import _geomspace, compute_time_derivatives, evolve_perturbations_batched


def _dev(x, dev):
    return torch.as_tensor(x, dtype=torch.float64, device=dev).contiguous()


def extract_perturbations(yout, youtprime, lmaxg, lmaxgp, lmaxr):
    """Named views of the state and its time derivative (cmb.py:17-89); pure indexing."""
    idxb, idxg = 5, 7
    idxgp = 7 + (lmaxg + 1)
    idxr = 9 + lmaxg + lmaxgp
    y, yp = yout, youtprime
    return {"idxb": idxb, "idxg": idxg, "idxgp": idxgp, "idxr": idxr, "iq0": 10 + lmaxg + lmaxgp + lmaxr,
            "eta": y[..., 2], "deltac": y[..., 3], "thetac": y[..., 4], "deltab": y[..., idxb], "thetab": y[..., idxb + 1],
            "deltabprime": yp[..., idxb], "thetabprime": yp[..., idxb + 1],
            "deltag": y[..., idxg], "thetag": y[..., idxg + 1], "shearg": y[..., idxg + 2] / 2.0,
            "deltagprime": yp[..., idxg], "thetagprime": yp[..., idxg + 1], "sheargprime": yp[..., idxg + 2] / 2.0,
            "deltar": y[..., idxr], "thetar": y[..., idxr + 1], "shearr": y[..., idxr + 2] / 2.0,
            "deltarprime": yp[..., idxr], "thetarprime": yp[..., idxr + 1], "shearrprime": yp[..., idxr + 2] / 2.0,
            "deltaq": y[..., -2], "thetaq": y[..., -1]}


def _tables(param):
    t = param["_deb_tables"]
    return t, int(param["_deb_num_thermo"]), int(param.get("_deb_cosmo", 0))


def compute_visibility_functions(tau, param, aexp=None):
    """Opacity, its derivative, g, g', g'' and the optical depth at conformal times tau, from
    splines in log tau of the thermal tables (cmb.py:92-166); also returns aprimeoa, rhonu,
    rho_Q and w_Q at the matching scale factors (a from the a_of_tau spline unless given)."""
    t, N, c = _tables(param)
    dev = t.device
    tt = _dev(tau, dev).reshape(-1)
    a = param["a_of_tau_spline"].evaluate(tt) if aexp is None else _dev(aexp, dev).reshape(-1)
    out = torch.empty((10, tt.numel()), dtype=torch.float64, device=dev)
    work = torch.empty(12 * N, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_visibility(_lib.ptr(t), N, c, _lib.ptr(tt), _lib.ptr(a), tt.numel(), _lib.ptr(out), _lib.ptr(work),
                                         _lib.stream()))
    names = ["opac", "opacprime", "gvis", "gvisprime", "gvispprime", "optical_depth", "aprimeoa", "rhonu", "rho_Q", "w_Q"]
    return {nm: out[i] for i, nm in enumerate(names)}


def compute_source_function_from_states(yout, yprime, kmodes, aexp_out, tau, param, lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, isw="spline"):
    """S1 (ISW), S2 (Sachs-Wolfe), S3 (Doppler), S4 (polarisation), S and the metric alpha,
    alpha' for states yout and their time derivatives yprime (n_k, n_tau, nvar), and the E-mode
    source SE (no counterpart in the reference; see compute_source_term_E) -- the chain of
    compute_Cell_spectrum_from_cosmo_params steps 6-10 (cmb.py:1050-1073).  isw: "spline" (the
    reference's d/dtau(eta + alpha') by a spline in tau, cmb.py:437-449) or "algebraic"
    (eta' + alpha'', the form of its docstring, cmb.py:406-408)."""
    if isw not in ("spline", "algebraic"):
        raise ValueError("isw must be 'spline' or 'algebraic'")
    t, N, c = _tables(param)
    dev = t.device
    y, yp = _dev(yout, dev), _dev(yprime, dev)
    kk, aa, tt = _dev(kmodes, dev).reshape(-1), _dev(aexp_out, dev).reshape(-1), _dev(tau, dev).reshape(-1)
    nk, nt = kk.numel(), tt.numel()
    nv = 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nqmax * (lmaxnu + 1) + 2
    if tuple(y.shape) != (nk, nt, nv) or y.shape != yp.shape:
        raise ValueError(f"yout and yprime must have shape (n_k, n_tau, nvar) = ({nk}, {nt}, {nv})")
    L = _lib.lib()
    out = torch.empty((8, nk, nt), dtype=torch.float64, device=dev)
    work = torch.empty(L.deb_cmb_sources_workspace_doubles(N, nk, nt), dtype=torch.float64, device=dev)
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), 1)
    _lib.check(L.deb_cmb_sources(_lib.ptr(t), N, c, _lib.ptr(y), _lib.ptr(yp), _lib.ptr(kk), nk, _lib.ptr(aa), _lib.ptr(tt), nt,
                                 ctypes.byref(prob), 1 if isw == "algebraic" else 0, _lib.ptr(out), _lib.ptr(work), _lib.stream()))
    return {"S1": out[0], "S2": out[1], "S3": out[2], "S4": out[3], "S": out[4], "alpha": out[5], "alphaprime": out[6], "SE": out[7]}


# ---------------------------------------------------------------------------------------------
# the stages of the source function one at a time (cmb.py:170-633)
# ---------------------------------------------------------------------------------------------
def _plane(x, dev):
    return torch.as_tensor(x, dtype=torch.float64).to(dev).contiguous()


def _kt(kmodes, dev):
    return torch.as_tensor(kmodes, dtype=torch.float64).to(dev).reshape(-1).contiguous()


def compute_neutrino_perturbations(yout, yprime, aexp_out, param, nqmax, iq0):
    """Massive-neutrino density, pressure, flux and shear perturbations and the derivatives of the
    density and shear (cmb.py:170-240): momentum integrals of the states yout and yprime
    (n_k, n_tau, nvar) whose bins start at iq0.  Returns the reference's dict (planes (n_k, n_tau),
    aprimeoa (n_tau,))."""
    t, N, c = _tables(param)
    dev = t.device
    y, yp = _dev(yout, dev), _dev(yprime, dev)
    if y.ndim != 3 or y.shape != yp.shape:
        raise ValueError("yout and yprime must be (n_k, n_tau, nvar) of one shape")
    nk, nt, nv = y.shape
    aa = _dev(aexp_out, dev).reshape(-1)
    if aa.numel() != nt:
        raise ValueError("aexp_out must have n_tau entries")
    out = torch.empty((6, nk, nt), dtype=torch.float64, device=dev)
    aH = torch.empty(nt, dtype=torch.float64, device=dev)
    work = torch.empty(32, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_cmb_neutrinos(_lib.ptr(t), N, c, _lib.ptr(y), _lib.ptr(yp), nk, _lib.ptr(aa), nt, nv, int(iq0), int(nqmax),
                                            _lib.ptr(out), _lib.ptr(aH), _lib.ptr(work), _lib.stream()))
    names = ["drhonu", "dpnu", "fnu", "shearnu", "rho_nu_prime", "shear_nu_prime"]
    res = {nm: out[i] for i, nm in enumerate(names)}
    res["aprimeoa"] = aH
    return res


def compute_metric_perturbations(perturbations, neutrinos, background_quantities, param, kmodes, aexp_out):
    """hprime, etaprime, alpha, alpha', alpha'' and the derivative of the conformal Hubble rate
    from the Einstein equations (cmb.py:243-345), on the dicts of extract_perturbations,
    compute_neutrino_perturbations and compute_background_quantities."""
    t, N, c = _tables(param)
    dev = t.device
    kk = _kt(kmodes, dev)
    aa = _dev(aexp_out, dev).reshape(-1).contiguous()
    pt, nu, bq = perturbations, neutrinos, background_quantities
    src = {"eta": pt["eta"], "deltac": pt["deltac"], "thetac": pt["thetac"], "deltab": pt["deltab"], "thetab": pt["thetab"],
           "deltag": pt["deltag"], "thetag": pt["thetag"], "shearg": pt["shearg"], "sheargp": pt["sheargprime"], "deltar": pt["deltar"],
           "thetar": pt["thetar"], "shearr": pt["shearr"], "shearrp": pt["shearrprime"], "deltaq": pt["deltaq"], "thetaq": pt["thetaq"],
           "drhonu": nu["drhonu"], "fnu": nu["fnu"], "shearnu": nu["shearnu"], "rho_nu_prime": nu["rho_nu_prime"],
           "shear_nu_prime": nu["shear_nu_prime"], "aH": nu["aprimeoa"], "rhonu": bq["rhonu"], "rho_Q": bq["rho_Q"], "w_Q": bq["w_Q"]}
    nk, nt = kk.numel(), aa.numel()
    planes = {}
    for name in _lib.METRIC_INPUTS:
        v = _plane(src[name], dev)
        want = (nt,) if name in ("aH", "rhonu", "rho_Q", "w_Q") else (nk, nt)
        if tuple(v.reshape(want).shape) != want:
            raise ValueError(f"{name} must have shape {want}")
        planes[name] = v.reshape(want)
    ins = _lib.DebMetricInputs(*[_lib.ptr(planes[name]) for name in _lib.METRIC_INPUTS])
    out = torch.empty((6, nk, nt), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_cmb_metric(_lib.ptr(t), N, c, ctypes.byref(ins), _lib.ptr(kk), nk, _lib.ptr(aa), nt, _lib.ptr(out),
                                         _lib.stream()))
    return {"hprime": out[0], "etaprime": out[1], "alpha": out[2], "alphaprime": out[3], "alphapprime": out[4], "aprimeoa": planes["aH"],
            "aprimeoaprime": out[5]}


def compute_polarization_terms(perturbations, metric, visibility_functions, yout, yprime, kmodes, lmaxg, lmaxgp):
    """F_2 + G_0 + G_2 and its first and second derivatives (cmb.py:349-395)."""
    dev = _lib.device()
    ys, yps = _dev(yout, dev), _dev(yprime, dev)
    nk, nt, nv = ys.shape
    kk = _kt(kmodes, dev)
    tg = _plane(perturbations["thetagprime"], dev).reshape(nk, nt)
    ap = _plane(metric["alphaprime"], dev).reshape(nk, nt)
    op = _plane(visibility_functions["opac"], dev).reshape(nt)
    opp = _plane(visibility_functions["opacprime"], dev).reshape(nt)
    out = torch.empty((3, nk, nt), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_cmb_polarization(_lib.ptr(ys), _lib.ptr(yps), nk, nt, nv, int(lmaxg), int(lmaxgp), _lib.ptr(kk), _lib.ptr(tg),
                                               _lib.ptr(ap), _lib.ptr(op), _lib.ptr(opp), _lib.ptr(out), _lib.stream()))
    return {"polarization_term": out[0], "polarization_termprime": out[1], "polarization_termpprime": out[2]}


def _source_term(term, kmodes, vis, planes):
    dev = _lib.device()
    kk = _kt(kmodes, dev)
    pol = _plane(planes["pol"], dev)
    nk, nt = pol.shape
    v = _plane(vis, dev).reshape(nt)
    args = [_plane(planes[n], dev).reshape(nk, nt) if planes.get(n) is not None else None
            for n in ("deltag", "alpha", "alphaprime", "thetab", "thetabprime", "pol", "polp", "polpp")]
    out = torch.empty((nk, nt), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_cmb_source_term(int(term), _lib.ptr(kk), nk, nt, _lib.ptr(v), *[_lib.ptr(a) if a is not None else None for a in args],
                                              None, 0.0, _lib.ptr(out), _lib.stream()))
    return out


def compute_source_term_isw(metric, visibility_functions, perturbations, tau):
    """S1 = e^-tau_c d(eta + alpha')/dtau by a natural spline in tau (cmb.py:401-452)."""
    dev = _lib.device()
    tt = _dev(tau, dev).reshape(-1).contiguous()
    eta = _plane(perturbations["eta"], dev)
    nk, nt = eta.shape
    ap = _plane(metric["alphaprime"], dev).reshape(nk, nt)
    od = _plane(visibility_functions["optical_depth"], dev).reshape(nt)
    L = _lib.lib()
    out = torch.empty((nk, nt), dtype=torch.float64, device=dev)
    work = torch.empty(L.deb_cmb_isw_workspace_doubles(nk, nt), dtype=torch.float64, device=dev)
    _lib.check(L.deb_cmb_source_isw(_lib.ptr(tt), nt, nk, _lib.ptr(eta), _lib.ptr(ap), _lib.ptr(od), _lib.ptr(out), _lib.ptr(work),
                                    _lib.stream()))
    return out


def compute_source_term_sachs_wolfe(perturbations, metric, visibility_functions, polarization_terms, kmodes):
    """S2 = g (delta_g / 4 + 2 alpha' + Pi / 16 + (theta_b' + 3 Pi'' / 16) / k^2) (cmb.py:455-496)."""
    return _source_term(2, kmodes, visibility_functions["gvis"],
                        {"deltag": perturbations["deltag"], "alphaprime": metric["alphaprime"], "thetabprime": perturbations["thetabprime"],
                         "pol": polarization_terms["polarization_term"], "polpp": polarization_terms["polarization_termpprime"]})


def compute_source_term_doppler(perturbations, metric, visibility_functions, polarization_terms, kmodes):
    """S3 = g' (alpha + (theta_b + 3 Pi' / 8) / k^2) (cmb.py:499-534)."""
    return _source_term(3, kmodes, visibility_functions["gvisprime"],
                        {"alpha": metric["alpha"], "thetab": perturbations["thetab"], "pol": polarization_terms["polarization_term"],
                         "polp": polarization_terms["polarization_termprime"]})


def compute_source_term_polarization(visibility_functions, polarization_terms, kmodes):
    """S4 = g'' 3 Pi / (16 k^2) (cmb.py:537-563)."""
    return _source_term(4, kmodes, visibility_functions["gvispprime"], {"pol": polarization_terms["polarization_term"]})


def compute_source_term_E(visibility_functions, polarization_terms, kmodes, tau, tau0):
    """E-mode polarisation source S_E = 3 g Pi / (16 (k chi)^2), chi = tau0 - tau, zero within 1 Mpc
    of today.  No counterpart in the reference (which computes TT only): the line-of-sight E-mode
    source of Zaldarriaga & Seljak (1997) in the reference's normalisation of Pi (= F_2 + G_0 + G_2,
    four times the brightness multipoles), to be projected with compute_theta_ell(field="E")."""
    dev = _lib.device()
    kk = _kt(kmodes, dev)
    pol = _plane(polarization_terms["polarization_term"], dev)
    nk, nt = pol.shape
    g = _plane(visibility_functions["gvis"], dev).reshape(nt)
    tt = _dev(tau, dev).reshape(-1).contiguous()
    out = torch.empty((nk, nt), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_cmb_source_term(5, _lib.ptr(kk), nk, nt, _lib.ptr(g), None, None, None, None, None, _lib.ptr(pol), None, None,
                                              _lib.ptr(tt), float(tau0), _lib.ptr(out), _lib.stream()))
    return out


def compute_source_function(perturbations, metric, visibility_functions, yout, yprime, kmodes, lmaxg, lmaxgp, tau):
    """S1..S4, their sum and the polarisation terms (cmb.py:566-633), stage by stage."""
    pol = compute_polarization_terms(perturbations, metric, visibility_functions, yout, yprime, kmodes, lmaxg, lmaxgp)
    S1 = compute_source_term_isw(metric, visibility_functions, perturbations, tau)
    S2 = compute_source_term_sachs_wolfe(perturbations, metric, visibility_functions, pol, kmodes)
    S3 = compute_source_term_doppler(perturbations, metric, visibility_functions, pol, kmodes)
    S4 = compute_source_term_polarization(visibility_functions, pol, kmodes)
    return {"S1": S1, "S2": S2, "S3": S3, "S4": S4, "S": S1 + S2 + S3 + S4, "polarization_terms": pol}


def compute_theta_ell(ellmax, kmodes, tau, S, tau0, n_fftlog=16384, n_k_dense=0, field="T"):
    """Theta_l(k) by FFTLog line-of-sight integration (cmb.py:680-835).  Returns (theta_ell
    (n_k, ellmax + 1), kmodes used) as device tensors.  field="E" (an extension: the reference
    projects temperature only) projects an E-mode source with the spin-2 factor
    sqrt((l + 2)! / (l - 2)!)."""
    if field not in ("T", "E"):
        raise ValueError("field must be 'T' or 'E'")
    S = torch.as_tensor(S, dtype=torch.float64)
    dev = S.device if S.is_cuda else _lib.device()
    S = S.to(dev).contiguous()
    kk, tt = _dev(kmodes, dev).reshape(-1), _dev(tau, dev).reshape(-1)
    nk, nt = kk.numel(), tt.numel()
    if tuple(S.shape) != (nk, nt):
        raise ValueError("S must have shape (n_kmodes, n_times)")
    nku = int(n_k_dense) if n_k_dense > 0 else nk
    L = _lib.lib()
    theta = torch.empty((nku, int(ellmax) + 1), dtype=torch.float64, device=dev)
    kout = torch.empty(nku, dtype=torch.float64, device=dev)
    work = torch.empty(L.deb_los_workspace_doubles(nk, nt, int(ellmax), int(n_fftlog), int(n_k_dense)), dtype=torch.float64, device=dev)
    _lib.check(L.deb_los_theta_ell_field(_lib.ptr(kk), nk, _lib.ptr(tt), nt, _lib.ptr(S), float(tau0), int(ellmax), int(n_fftlog),
                                         int(n_k_dense), 1 if field == "E" else 0, _lib.ptr(theta), _lib.ptr(kout), _lib.ptr(work),
                                         _lib.stream()))
    return theta, kout


def compute_Cell(theta_ell, kmodes, n_s, k_p, theta_ell2=None):
    """C_l = trapezoid over ln k of (k/k_p)^(n_s-1) Theta_l^2 (cmb.py:838-873); with theta_ell2 (an
    extension) the cross spectrum of Theta_l Theta2_l, e.g. C_l^TE."""
    th = torch.as_tensor(theta_ell, dtype=torch.float64)
    dev = th.device if th.is_cuda else _lib.device()
    th = th.to(dev).contiguous()
    kk = _dev(kmodes, dev).reshape(-1)
    cl = torch.empty(th.shape[1], dtype=torch.float64, device=dev)
    if theta_ell2 is None:
        _lib.check(_lib.lib().deb_cell(_lib.ptr(th), _lib.ptr(kk), kk.numel(), th.shape[1], float(n_s), float(k_p), _lib.ptr(cl),
                                       _lib.stream()))
    else:
        th2 = torch.as_tensor(theta_ell2, dtype=torch.float64).to(dev).contiguous()
        if th2.shape != th.shape:
            raise ValueError("theta_ell and theta_ell2 must have one shape")
        _lib.check(_lib.lib().deb_cell_cross(_lib.ptr(th), _lib.ptr(th2), _lib.ptr(kk), kk.numel(), th.shape[1], float(n_s), float(k_p),
                                             _lib.ptr(cl), _lib.stream()))
    return cl


def compute_Dell(Cell, A_s, Tcmb, ellmax=None):
    """(ell, D_l = l (l + 1) C_l A_s 2 T_cmb^2 1e12) for l >= 2 (cmb.py:876-927)."""
    cl = torch.as_tensor(Cell, dtype=torch.float64)
    dev = cl.device if cl.is_cuda else _lib.device()
    cl = cl.to(dev).contiguous()
    if ellmax is None:
        ellmax = cl.numel() - 1
    ell = torch.empty(ellmax - 1, dtype=torch.float64, device=dev)
    dl = torch.empty(ellmax - 1, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib().deb_dell(_lib.ptr(cl), int(ellmax), float(A_s), float(Tcmb), _lib.ptr(ell), _lib.ptr(dl), _lib.stream()))
    return ell.to(torch.int64), dl


def cmb_aexp_out(n_early=32, n_pre_recomb=64, n_recomb=128, n_post=96, n_reion=0):
    """Output scale factors of the CMB driver: log-spaced pieces split at z = 3000, 1400, 600
    (cmb.py:999-1014).  n_reion > 0 (an extension) merges n_reion more points, log-spaced in a
    over z in [0, 30], so that the source function resolves the reionization visibility: the
    reference's 96 points below z = 600 put about one sample in its width, and the line of sight
    then depends on where the samples fall (derivatives with respect to tau_reion wander by tens
    of percent)."""
    a1, a2, a3 = 1.0 / 3001.0, 1.0 / 1401.0, 1.0 / 601.0
    a = np.concatenate([_geomspace(1e-4, a1, n_early, endpoint=False), _geomspace(a1, a2, n_pre_recomb, endpoint=False),
                        _geomspace(a2, a3, n_recomb, endpoint=False), _geomspace(a3, 1.0, n_post)])
    if n_reion > 0:
        a = np.unique(np.concatenate([a, _geomspace(1.0 / 31.0, 1.0, int(n_reion))]))
    return a


# late-time output points the drivers add when reionization is on (see cmb_aexp_out)
REION_OUTPUTS = 200


def _driver_aexp(param_dict, aexp_out):
    if aexp_out is not None:
        return np.asarray(aexp_out, dtype=np.float64)
    return cmb_aexp_out(n_reion=REION_OUTPUTS if float(param_dict.get("tau_reion", 0.0)) > 0.0 else 0)


def compute_sources_fused(param, kmodes, aexp_out, *, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4,
                          max_steps=2048, solver="per_mode", batch_size=16, isw="spline", threads=0, step_log=False, replay=None):
    """The source function S(k, tau) of the chain perturbations -> time derivatives -> sources
    (cmb.py:1030-1073) in one launch (deb_evolve_sources): each mode's solve forms the source
    terms at its output times from the projected state and its right-hand side, so only the
    (n_k, n_tau) planes are stored, never the states (n_k x n_tau x nvar) or their derivatives.
    Equal to compute_sources_streamed's unfused chain on the same solver.  solver: "per_mode"
    (largest k first) or "batched" (batch_size modes on shared steps; num_k a multiple of it);
    threads: threads per mode or batch (0: 64 per mode, 128 per batch -- with thousands of light
    modes, more modes in flight beat wider workgroups: 0.044 s at 64 threads per mode against
    0.054 s at 128 and 0.092 s at 256 for the 3000-mode, 2500-output workload).  param: from evolve_background.  step_log: also return the attempted steps
    ("step_log": (n_k or n_k / batch_size, max_steps, 3), as evolve_units logs them); replay:
    (mesh, mesh_len) to take each mode's frozen step mesh (per-mode solver only; see evolve_units).
    Returns {"S", "S1".."S4", "SE", "alpha", "alphaprime": (n_k, n_tau), "tau": tau_out,
    "n_accept", "n_reject": (n_k,)}."""
    from .perturbations import DEFAULT_SCHEDULE, _raise_on_status, _schedule
    if solver not in ("per_mode", "batched"):
        raise ValueError("solver must be 'per_mode' or 'batched'")
    if isw not in ("spline", "algebraic"):
        raise ValueError("isw must be 'spline' or 'algebraic'")
    t, N, c = _tables(param)
    dev = t.device
    kk = torch.as_tensor(kmodes, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    aexp = torch.as_tensor(np.asarray(aexp_out, dtype=np.float64) if not torch.is_tensor(aexp_out) else aexp_out, dtype=torch.float64,
                           device=dev).reshape(-1).contiguous()
    n_k, n_tau = kk.numel(), aexp.numel()
    bs = 0 if solver == "per_mode" else int(batch_size)
    if bs and n_k % bs:
        raise ValueError("with solver='batched', num_k must be a multiple of batch_size")
    tau_out = param["tau_of_a_spline"].evaluate(aexp).contiguous()
    L = _lib.lib()
    prob = _lib.DebProblem(int(lmaxg), int(lmaxgp), int(lmaxr), int(lmaxnu), int(nqmax), 1)
    cfg = _lib.DebSolverCfg(float(rtol), float(atol), 0.25, 0.80, 0.0, 20.0, 0.3, int(max_steps))
    out = torch.empty((8, n_k, n_tau), dtype=torch.float64, device=dev)
    status = torch.empty(n_k, dtype=torch.int32, device=dev)
    nacc = torch.empty(n_k, dtype=torch.int32, device=dev)
    nrej = torch.empty(n_k, dtype=torch.int32, device=dev)
    work = torch.empty(L.deb_evolve_sources_workspace_doubles(N, n_k, n_tau, bs, ctypes.byref(prob)), dtype=torch.float64, device=dev)
    threads = int(threads) or (64 if bs == 0 else 128)
    sched_dict = dict(DEFAULT_SCHEDULE, threads=threads) if bs == 0 else {"threads": threads}
    sched, order = _schedule(sched_dict, kk)
    log, mesh, mlen, mstride = None, None, None, 0
    if step_log and replay is not None:
        raise ValueError("step_log and replay are exclusive")
    if step_log:
        log = torch.full((n_k if bs == 0 else n_k // bs, int(max_steps), 3), float("nan"), dtype=torch.float64, device=dev)
    if replay is not None:
        from .perturbations import _check_mesh
        if bs:
            raise ValueError("a replay mesh takes the per-mode solver")
        mesh = replay[0].to(device=dev, dtype=torch.float64).contiguous()
        mlen = replay[1].to(device=dev, dtype=torch.int32).contiguous()
        if mesh.ndim != 2 or mesh.shape[0] != n_k or mlen.numel() != n_k:
            raise ValueError("replay mesh must be (num_k, L) with one length per mode")
        _check_mesh(mesh, mlen)
        mstride = mesh.shape[1]
    _lib.check(L.deb_evolve_sources(_lib.ptr(t), N, c, _lib.ptr(kk), n_k, bs, _lib.ptr(aexp), _lib.ptr(tau_out), n_tau, ctypes.byref(prob),
                                    ctypes.byref(cfg), 1 if isw == "algebraic" else 0, _lib.ptr(out), _lib.ptr(status), _lib.ptr(nacc),
                                    _lib.ptr(nrej), _lib.ptr(work), _lib.ptr(log) if log is not None else None,
                                    _lib.ptr(mesh) if mesh is not None else None, mstride, _lib.ptr(mlen) if mlen is not None else None,
                                    ctypes.byref(sched) if sched is not None else None, _lib.stream()))
    del order, work  # stream-ordered: reused by the allocator only after the launches have run
    _raise_on_status(status)
    res = {"S1": out[0], "S2": out[1], "S3": out[2], "S4": out[3], "S": out[4], "alpha": out[5], "alphaprime": out[6], "SE": out[7],
           "tau": tau_out, "n_accept": nacc, "n_reject": nrej}
    if log is not None:
        res["step_log"] = log
    return res


def compute_sources_streamed(param, kmodes, aexp_out, *, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, rtol=1e-4, atol=1e-4,
                             max_steps=2048, solver="per_mode", batch_size=16, k_chunk=512, isw="spline"):
    """The source function S(k, tau) of the chain perturbations -> time derivatives -> sources
    (cmb.py:1030-1073) computed k_chunk modes at a time, so that the full states and their
    derivatives (n_k x n_tau x nvar each) never exist at once: only one chunk's, plus the
    (n_k, n_tau) source planes.  Every stage is independent across k (the ISW spline runs along
    tau), so the result equals the unchunked chain's.  solver: "per_mode" (adaptive steps per
    mode, largest k first) or "batched" (batch_size modes on shared steps, as the reference's
    driver; k_chunk then a multiple of batch_size).  param: from evolve_background.  Returns
    {"S": (n_k, n_tau), "S1".."S4", "SE" (the E-mode source), "tau": tau_out, "n_accept": (n_k,),
    "n_reject": (n_k,)}."""
    from .perturbations import evolve_units, evolve_units_batched, DEFAULT_SCHEDULE, _raise_on_status
    if solver not in ("per_mode", "batched"):
        raise ValueError("solver must be 'per_mode' or 'batched'")
    tables = param["_deb_tables"]
    dev = tables.device
    kk = torch.as_tensor(kmodes, dtype=torch.float64, device=dev).reshape(-1)
    aexp = torch.as_tensor(np.asarray(aexp_out, dtype=np.float64) if not torch.is_tensor(aexp_out) else aexp_out, dtype=torch.float64,
                           device=dev).reshape(-1)
    n_k, n_tau = kk.numel(), aexp.numel()
    k_chunk = int(k_chunk) if k_chunk else n_k
    if solver == "batched" and (k_chunk % batch_size or n_k % batch_size):
        raise ValueError("with solver='batched', num_k and k_chunk must be multiples of batch_size")
    tau_out = param["tau_of_a_spline"].evaluate(aexp)
    cosmo = int(param.get("_deb_cosmo", 0))
    n_cosmo = tables.shape[0] if tables.ndim == 2 else 1
    tau_all = torch.zeros((n_cosmo, n_tau), dtype=torch.float64, device=dev)
    tau_all[cosmo] = tau_out
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    out = {name: torch.empty((n_k, n_tau), dtype=torch.float64, device=dev) for name in ("S", "S1", "S2", "S3", "S4", "SE")}
    nacc = torch.empty(n_k, dtype=torch.int32, device=dev)
    nrej = torch.empty(n_k, dtype=torch.int32, device=dev)
    N = param["_deb_num_thermo"]
    for k0 in range(0, n_k, k_chunk):
        k1 = min(n_k, k0 + k_chunk)
        kc = kk[k0:k1].contiguous()
        uc = torch.full((k1 - k0,), cosmo, dtype=torch.int32, device=dev)
        kw = dict(rtol=rtol, atol=atol, max_steps=max_steps, return_full=True, **lm)
        if solver == "per_mode":
            y, st, na, nr = evolve_units(tables, N, uc, kc, tau_all, schedule=DEFAULT_SCHEDULE, **kw)
        else:
            y, st, na, nr = evolve_units_batched(tables, N, uc, kc, tau_all, batch_size=batch_size, **kw)
        _raise_on_status(st)
        yp = compute_time_derivatives(y, tau_out, kc, param, lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)
        src = compute_source_function_from_states(y, yp, kc, aexp, tau_out, param, isw=isw, **lm)
        for name in out:
            out[name][k0:k1] = src[name]
        nacc[k0:k1], nrej[k0:k1] = na, nr
        del y, yp, src
    out.update(tau=tau_out, n_accept=nacc, n_reject=nrej)
    return out


def compute_Cell_spectrum_from_cosmo_params(param_dict, ellmax=2500, nmodes=128, kmin=1e-4, kmax=1.0, n_k_dense=8192, n_fftlog=16384,
                                            lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, *, k_sampling_method="camb",
                                            solver="batched", k_chunk=None, aexp_out=None, rtol=1e-4, atol=1e-4, fused=True):
    """TT C_l from cosmological parameters (cmb.py:930-1090): background (num_thermo 1024),
    perturbations on the camb k-sampling at rtol = atol = 1e-4 with full states, integrated in
    batches of 16 modes on one shared step sequence as the reference does (nmodes a multiple of
    16), time derivatives, source function, line of sight, C_l.  Returns (Cell, param).

    Keyword-only extensions (the defaults are the reference's behaviour): k_sampling_method;
    solver "per_mode" for independent adaptive steps per mode; k_chunk to stream the chain over
    k (compute_sources_streamed) so that full states are held for k_chunk modes at a time;
    fused (default) forms the sources in the solver's output epilogue (compute_sources_fused: no
    states stored, the same planes bit for bit; fused=False runs the staged chain, streamed when
    k_chunk is given); aexp_out (default cmb_aexp_out(), with REION_OUTPUTS late points merged
    when param_dict switches on reionization); rtol, atol."""
    param = evolve_background(param=dict(param_dict), thermo_module="RECFAST", num_thermo=1024)
    aexp_out = _driver_aexp(param_dict, aexp_out)
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    tau0 = float(param["tau_of_a_spline"].evaluate(np.array([1.0])).cpu()[0])
    if fused:
        from .perturbations import k_grid
        kmodes = torch.as_tensor(k_grid(param, kmin, kmax, nmodes, k_sampling_method), dtype=torch.float64,
                                 device=param["_deb_tables"].device)
        src = compute_sources_fused(param, kmodes, aexp_out, rtol=rtol, atol=atol, solver=solver, **lm)
        tau, S = src["tau"], src["S"]
        param.update(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, nout=int(tau.numel()), tau_out=tau,
                     _deb_steps=(src["n_accept"], src["n_reject"]))
    elif solver == "batched" and k_chunk is None:
        yout, kmodes, param = evolve_perturbations_batched(param=param, aexp_out=aexp_out, kmin=kmin, kmax=kmax, num_k=nmodes, rtol=rtol,
                                                           atol=atol, return_full=True, k_sampling_method=k_sampling_method, **lm)
        tau = param["tau_out"]
        yprime = compute_time_derivatives(yout, tau, kmodes, param, lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax)
        S = compute_source_function_from_states(yout, yprime, kmodes, aexp_out, tau, param, **lm)["S"]
    else:
        from .perturbations import k_grid
        kmodes = torch.as_tensor(k_grid(param, kmin, kmax, nmodes, k_sampling_method), dtype=torch.float64,
                                 device=param["_deb_tables"].device)
        src = compute_sources_streamed(param, kmodes, aexp_out, rtol=rtol, atol=atol, solver=solver, k_chunk=k_chunk, **lm)
        tau, S = src["tau"], src["S"]
        param.update(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, nout=int(tau.numel()), tau_out=tau,
                     _deb_steps=(src["n_accept"], src["n_reject"]))
    theta, kk = compute_theta_ell(ellmax, kmodes, tau, S, tau0, n_fftlog=n_fftlog, n_k_dense=n_k_dense)
    return compute_Cell(theta, kk, param["n_s"], param.get("k_p", 0.05)), param


def compute_Cell_spectra_from_cosmo_params(param_dict, spectra=("TT", "TE", "EE"), ellmax=2500, nmodes=128, kmin=1e-4, kmax=1.0,
                                            n_k_dense=8192, n_fftlog=16384, lmaxg=11, lmaxgp=11, lmaxr=11, lmaxnu=8, nqmax=3, *,
                                            k_sampling_method="camb", solver="batched", k_chunk=None, aexp_out=None, rtol=1e-4, atol=1e-4,
                                            fused=True):
    """C_l^TT, C_l^TE and C_l^EE from cosmological parameters: the chain of
    compute_Cell_spectrum_from_cosmo_params (cmb.py:930-1090) with the E-mode source projected
    beside the temperature source (an extension: the reference computes TT only; reionization
    only when param_dict carries tau_reion > 0, see evolve_background).  Returns ({spectrum: C_l},
    param)."""
    for x in spectra:
        if x not in ("TT", "TE", "EE"):
            raise ValueError("spectra are drawn from 'TT', 'TE', 'EE'")
    from .perturbations import k_grid
    param = evolve_background(param=dict(param_dict), thermo_module="RECFAST", num_thermo=1024)
    aexp_out = _driver_aexp(param_dict, aexp_out)
    lm = dict(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax)
    tau0 = float(param["tau_of_a_spline"].evaluate(np.array([1.0])).cpu()[0])
    kmodes = torch.as_tensor(k_grid(param, kmin, kmax, nmodes, k_sampling_method), dtype=torch.float64, device=param["_deb_tables"].device)
    if fused:
        src = compute_sources_fused(param, kmodes, aexp_out, rtol=rtol, atol=atol, solver=solver, **lm)
    else:
        src = compute_sources_streamed(param, kmodes, aexp_out, rtol=rtol, atol=atol, solver=solver, k_chunk=k_chunk, **lm)
    param.update(lmaxg=lmaxg, lmaxgp=lmaxgp, lmaxr=lmaxr, lmaxnu=lmaxnu, nqmax=nqmax, nout=int(src["tau"].numel()), tau_out=src["tau"])
    th = {}
    if "TT" in spectra or "TE" in spectra:
        th["T"], kk = compute_theta_ell(ellmax, kmodes, src["tau"], src["S"], tau0, n_fftlog=n_fftlog, n_k_dense=n_k_dense)
    if "EE" in spectra or "TE" in spectra:
        th["E"], kk = compute_theta_ell(ellmax, kmodes, src["tau"], src["SE"], tau0, n_fftlog=n_fftlog, n_k_dense=n_k_dense, field="E")
    n_s, k_p = param["n_s"], param.get("k_p", 0.05)
    out = {}
    for x in spectra:
        a, b = th[x[0]], th[x[1]]
        out[x] = compute_Cell(a, kk, n_s, k_p) if x[0] == x[1] else compute_Cell(a, kk, n_s, k_p, theta_ell2=b)
    return out, param


__all__ = ["extract_perturbations", "compute_visibility_functions", "compute_neutrino_perturbations", "compute_metric_perturbations",
           "compute_polarization_terms", "compute_source_term_isw", "compute_source_term_sachs_wolfe", "compute_source_term_doppler",
           "compute_source_term_polarization", "compute_source_function", "compute_source_function_from_states", "compute_theta_ell",
           "compute_Cell", "compute_Dell", "cmb_aexp_out", "compute_sources_streamed", "compute_Cell_spectrum_from_cosmo_params",
           "compute_source_term_E", "compute_Cell_spectra_from_cosmo_params", "compute_sources_fused"]
