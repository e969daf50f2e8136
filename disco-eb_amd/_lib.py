"""ctypes binding of libdiscoeb_amd.so (the C ABI declared in include/discoeb_amd.h).

The library is the only compute path: if it is missing the import of any solver function
raises, and every call requires a ROCm device.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiscoeb_amd.so")

# deb_layout.h: parameter vector order
PARAM_KEYS = [
    "Omegam", "Omegab", "H0", "Tcmb", "YHe", "Neff", "Nmnu", "mnu", "w_DE_0", "w_DE_a", "cs2_DE", "Omegak",
    "A_s", "n_s", "k_p",
    "use_tca", "tca_tau_c_over_tau_h_trigger", "tca_tau_c_over_tau_k_trigger",
    "use_rsa", "rsa_tau_c_over_tau_trigger", "rsa_tau_over_tau_k_trigger",
    "use_ur_fluid", "ur_fluid_tau_over_tau_k_trigger",
    "tau_reion",
]
# defaults of the approximation switches (src/discoeb/approximations.py:9-20) and optional keys
PARAM_DEFAULTS = {
    "Omegak": 0.0, "k_p": 0.05, "cs2_DE": 1.0, "A_s": 2.1e-9, "n_s": 0.965,
    "use_tca": False, "tca_tau_c_over_tau_h_trigger": 0.015, "tca_tau_c_over_tau_k_trigger": 0.010,
    "use_rsa": True, "rsa_tau_c_over_tau_trigger": 10.0, "rsa_tau_over_tau_k_trigger": 80.0,
    "use_ur_fluid": True, "ur_fluid_tau_over_tau_k_trigger": 120.0,
    # reionization optical depth (an extension; 0: none, as the reference's thermal history)
    "tau_reion": 0.0,
}
HDR = {  # header slots written by the background kernel
    "grhom": 24, "grhog": 25, "grhor": 26, "adotrad": 27, "amnu": 28, "Omegamnu": 29, "OmegaDE": 30, "taumin": 31,
    "taumax": 32, "fHe": 33, "tau0": 34, "tau_maxvis": 35, "optical_depth_today": 36, "akthom": 37,
    "recfast_failures": 40, "amin": 44, "amax": 45, "z_reion": 47, "tau_reion_achieved": 48,
}
ARRAY_SLOTS = [
    "nu_loga", "nu_logrho", "nu_logrho_S", "nu_logp", "nu_logp_S", "nu_logpp", "nu_logpp_S",
    "aexp", "loga", "tau", "xe", "xeHI", "xeHeI", "xeHeII", "Tm", "mu", "cs2", "cs2a", "tempba", "xeprime_recf", "xe_vis",
    "xeprime", "opac", "optical_depth", "gvis", "gvisprime", "gvispprime",
    "S_tau_of_a", "S_a_of_tau", "S_xe_of_tau", "S_cs2a_of_tau", "S_tempba_of_tau", "S_xe_of_loga", "S_cs2a_of_loga",
    "S_opac", "I_opac",
    "scr0", "scr1", "scr2", "scr3", "scr4", "scr5", "scr6", "scr7", "scr8",
]
NNU = 512
SPLINES = ["tau_of_a", "a_of_tau", "xe_of_tau", "cs2a_of_tau", "tempba_of_tau", "xe_of_loga", "cs2a_of_loga",
           "logrhonu_of_loga", "logpnu_of_loga", "logppseudonu_of_loga", "opac"]

# symbols declared in include/discoeb_amd.h
EXPORTS = ["deb_version", "deb_last_error", "deb_tables_size", "deb_table_offset", "deb_background", "deb_background_thermo",
           "deb_spline_eval",
           "deb_default_threads_per_unit", "deb_evolve_workspace_doubles", "deb_evolve", "deb_evolve_replay", "deb_adjoint_workspace_doubles",
           "deb_evolve_adjoint", "deb_evolve_batched_workspace_doubles",
           "deb_evolve_batched", "deb_noalloc_scratch_doubles", "deb_evolve_batched_noalloc", "deb_tau_out", "deb_nvar", "deb_power", "deb_power_units",
           "deb_time_derivatives", "deb_visibility", "deb_cmb_sources_workspace_doubles", "deb_cmb_sources",
           "deb_evolve_sources_workspace_doubles", "deb_evolve_sources",
           "deb_cmb_neutrinos", "deb_cmb_metric", "deb_cmb_polarization", "deb_cmb_source_term", "deb_cmb_isw_workspace_doubles",
           "deb_cmb_source_isw", "deb_los_workspace_doubles", "deb_los_theta_ell", "deb_los_theta_ell_field", "deb_cell_cross", "deb_contract", "deb_cell", "deb_dell", "deb_savgol_coefficients",
           "deb_savgol_filter", "deb_power_smoothed_workspace_doubles", "deb_power_smoothed", "deb_power_kaiser", "deb_power_multipoles",
           "deb_xi_workspace_doubles", "deb_xi_from_power"]


class DebProblem(ctypes.Structure):
    _fields_ = [("lmaxg", ctypes.c_int), ("lmaxgp", ctypes.c_int), ("lmaxr", ctypes.c_int), ("lmaxnu", ctypes.c_int),
                ("nqmax", ctypes.c_int), ("return_full", ctypes.c_int)]


class DebSolverCfg(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double), ("pcoeff", ctypes.c_double),
                ("icoeff", ctypes.c_double), ("dcoeff", ctypes.c_double), ("factormax", ctypes.c_double),
                ("factormin", ctypes.c_double), ("max_steps", ctypes.c_int)]


METRIC_INPUTS = ("eta", "deltac", "thetac", "deltab", "thetab", "deltag", "thetag", "shearg", "sheargp", "deltar", "thetar", "shearr",
                 "shearrp", "deltaq", "thetaq", "drhonu", "fnu", "shearnu", "rho_nu_prime", "shear_nu_prime", "aH", "rhonu", "rho_Q", "w_Q")


class DebMetricInputs(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in METRIC_INPUTS]


class DebSchedule(ctypes.Structure):
    _fields_ = [("d_order", ctypes.c_void_p), ("threads_per_unit", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python disco-eb_amd/build.py` "
                           "(or __graft_entry__.build()); there is no fallback path")
    L = ctypes.CDLL(LIB_PATH)
    L.deb_version.restype = ctypes.c_char_p
    L.deb_last_error.restype = ctypes.c_char_p
    L.deb_tables_size.restype = ctypes.c_long
    L.deb_tables_size.argtypes = [ctypes.c_int]
    L.deb_table_offset.restype = ctypes.c_long
    L.deb_table_offset.argtypes = [ctypes.c_int, ctypes.c_int]
    L.deb_evolve_workspace_doubles.restype = ctypes.c_long
    L.deb_nvar.argtypes = [ctypes.c_int] * 5
    L.deb_background.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.deb_background_thermo.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    L.deb_spline_eval.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
    L.deb_evolve.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.POINTER(DebProblem),
                             ctypes.POINTER(DebSolverCfg), _vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(DebSchedule), _vp]
    L.deb_evolve_replay.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.POINTER(DebProblem),
                                    ctypes.POINTER(DebSolverCfg), _vp, ctypes.c_long, _vp, _vp, _vp, _vp, _vp, _vp,
                                    ctypes.POINTER(DebSchedule), _vp]
    L.deb_adjoint_workspace_doubles.restype = ctypes.c_long
    L.deb_adjoint_workspace_doubles.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(DebProblem)]
    L.deb_evolve_adjoint.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                     ctypes.POINTER(DebProblem), _vp, ctypes.c_long, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     ctypes.POINTER(DebSchedule), _vp]
    L.deb_evolve_batched_workspace_doubles.restype = ctypes.c_long
    L.deb_evolve_batched_workspace_doubles.argtypes = [ctypes.c_int, ctypes.POINTER(DebProblem)]
    L.deb_evolve_batched.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                     ctypes.POINTER(DebProblem), ctypes.POINTER(DebSolverCfg), _vp, _vp, _vp, _vp, _vp, _vp,
                                     ctypes.POINTER(DebSchedule), _vp]
    L.deb_noalloc_scratch_doubles.restype = ctypes.c_long
    L.deb_noalloc_scratch_doubles.argtypes = []
    L.deb_evolve_batched_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                             ctypes.POINTER(DebProblem), ctypes.POINTER(DebSolverCfg), _vp, _vp, _vp, _vp, _vp,
                                             ctypes.POINTER(DebSchedule), _vp]
    L.deb_tau_out.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
    _d = ctypes.c_double
    _i = ctypes.c_int
    L.deb_cmb_neutrinos.argtypes = [_vp, _i, _i, _vp, _vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]
    L.deb_cmb_metric.argtypes = [_vp, _i, _i, ctypes.POINTER(DebMetricInputs), _vp, _i, _vp, _i, _vp, _vp]
    L.deb_cmb_polarization.argtypes = [_vp, _vp, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    L.deb_cmb_source_term.argtypes = [_i, _vp, _i, _i] + [_vp] * 10 + [_d, _vp, _vp]
    L.deb_cmb_isw_workspace_doubles.restype = ctypes.c_long
    L.deb_cmb_isw_workspace_doubles.argtypes = [_i, _i]
    L.deb_cmb_source_isw.argtypes = [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp]
    L.deb_savgol_coefficients.argtypes = [ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.deb_savgol_filter.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp]
    L.deb_power_smoothed_workspace_doubles.restype = ctypes.c_long
    L.deb_power_smoothed_workspace_doubles.argtypes = [ctypes.c_int, ctypes.c_int]
    L.deb_power_smoothed.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d, _d, _d, ctypes.c_int, _vp, _vp, _vp]
    L.deb_power_kaiser.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _d, _d, _d, _d, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    L.deb_power_multipoles.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _d, _d, _d, _d, _vp, _vp, _vp, _vp]
    L.deb_xi_workspace_doubles.restype = ctypes.c_long
    L.deb_xi_workspace_doubles.argtypes = [ctypes.c_int]
    L.deb_xi_from_power.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]
    L.deb_visibility.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, _vp, _vp]
    L.deb_cmb_sources_workspace_doubles.restype = ctypes.c_long
    L.deb_cmb_sources_workspace_doubles.argtypes = [ctypes.c_int] * 3
    L.deb_cmb_sources.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                  ctypes.POINTER(DebProblem), ctypes.c_int, _vp, _vp, _vp]
    L.deb_evolve_sources_workspace_doubles.restype = ctypes.c_long
    L.deb_evolve_sources_workspace_doubles.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(DebProblem)]
    L.deb_evolve_sources.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                     ctypes.POINTER(DebProblem), ctypes.POINTER(DebSolverCfg), ctypes.c_int, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp, ctypes.c_long, _vp, ctypes.POINTER(DebSchedule), _vp]
    L.deb_los_workspace_doubles.restype = ctypes.c_long
    L.deb_los_workspace_doubles.argtypes = [ctypes.c_int] * 5
    L.deb_los_theta_ell.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _vp, _vp, _vp, _vp]
    L.deb_los_theta_ell_field.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp]
    L.deb_cell_cross.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp]
    L.deb_contract.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_long, _vp, _vp]
    L.deb_cell.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp]
    L.deb_dell.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]
    L.deb_power_units.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
    L.deb_power.argtypes = [_vp, ctypes.c_int, _vp, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.c_double, ctypes.c_double, _vp, _vp]
    L.deb_time_derivatives.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp, ctypes.c_int,
                                       ctypes.POINTER(DebProblem), _vp, _vp, _vp]
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().deb_last_error().decode()
        if rc < 0:
            raise ValueError(msg)
        raise RuntimeError(msg)


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("discoeb_amd needs a ROCm GPU (MI355X / gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def table_offset(name, num_thermo):
    return lib().deb_table_offset(ARRAY_SLOTS.index(name), int(num_thermo))


def param_vector(param):
    vals = []
    for k in PARAM_KEYS:
        v = param.get(k, PARAM_DEFAULTS.get(k))
        if v is None:
            raise KeyError(f"missing cosmological parameter {k!r}")
        vals.append(float(v))
    return vals
