"""MI355X-native linear perturbation solver with the call surface of DISCO-EB's
background -> evolve_perturbations -> P(k) path (lawrenceberry/DISCO-EB).

All compute runs in hand-written gfx950 HIP kernels (libdiscoeb_amd.so, C ABI in
include/discoeb_amd.h); this package is the Python host side.
"""
from . import _lib
from .background import DeviceSpline, attach_background, background_tables, compute_background_quantities, evolve_background
from .perturbations import (compute_time_derivatives, evolve_perturbations, evolve_perturbations_batched, evolve_units, evolve_units_batched,
                            get_power, k_grid, nvar)
from .batch import PkGraph, autotune_schedule, evolve_many, power_units, tau_out_all
from .grad import (PARAM9, PkFunction, cell_vjp, evolve_units_adjoint, param_to_theta, pk_and_vjp, pk_and_vjp_many, pk_jacobian, pk_jacobian_many,
                   theta_to_param)
from .lss import get_power_smoothed, get_xi_from_P, power_Kaiser, power_multipoles, savgol_filter
from .cmb import (cmb_aexp_out, compute_Cell, compute_Cell_spectrum_from_cosmo_params, compute_Dell, compute_metric_perturbations,
                  compute_neutrino_perturbations, compute_polarization_terms, compute_source_function, compute_source_function_from_states,
                  compute_source_term_doppler, compute_source_term_E, compute_source_term_isw, compute_source_term_polarization,
                  compute_source_term_sachs_wolfe, compute_Cell_spectra_from_cosmo_params,
                  compute_sources_fused, compute_sources_streamed, compute_theta_ell, compute_visibility_functions,
                  extract_perturbations)

__all__ = ["evolve_background", "background_tables", "attach_background", "compute_background_quantities", "DeviceSpline",
           "evolve_perturbations", "evolve_perturbations_batched", "evolve_units", "evolve_units_batched", "get_power", "compute_time_derivatives", "k_grid", "nvar", "evolve_many",
           "power_units", "tau_out_all", "autotune_schedule", "PkGraph", "extract_perturbations", "compute_visibility_functions",
           "compute_source_function_from_states", "compute_sources_fused", "compute_sources_streamed", "compute_theta_ell", "compute_neutrino_perturbations",
           "compute_metric_perturbations", "compute_polarization_terms", "compute_source_term_isw", "compute_source_term_sachs_wolfe",
           "compute_source_term_doppler", "compute_source_term_polarization", "compute_source_function", "compute_source_term_E",
           "compute_Cell_spectra_from_cosmo_params", "compute_Cell", "compute_Dell", "cmb_aexp_out",
           "compute_Cell_spectrum_from_cosmo_params", "PARAM9", "theta_to_param", "param_to_theta", "pk_jacobian", "pk_jacobian_many", "PkFunction",
           "pk_and_vjp", "pk_and_vjp_many", "evolve_units_adjoint", "cell_vjp", "get_power_smoothed", "get_xi_from_P", "power_Kaiser",
           "power_multipoles", "savgol_filter"]
