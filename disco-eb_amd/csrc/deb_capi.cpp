// deb_capi.cpp -- C ABI entry points (include/discoeb_amd.h): argument checks, error strings,
// kernel launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/discoeb_amd.h"

extern "C" int deb_launch_background(const double *, int, int, int, const double *, double *, hipStream_t);
extern "C" int deb_launch_nu_bins(int, double *, hipStream_t);
extern "C" long deb_cmb_sources_work_doubles(int, int, int);
extern "C" int deb_launch_cmb_sources(const double *, int, const double *, const double *, const double *, int, const double *, const double *,
                                      int, const DebProblem *, int, double *, double *, hipStream_t);
extern "C" int deb_launch_visibility(const double *, int, const double *, const double *, int, double *, double *, hipStream_t);
extern "C" long deb_los_work_doubles(int, int, int, int, int);
extern "C" int deb_launch_los(const double *, int, const double *, int, const double *, double, int, int, int, int, double *, double *, double *,
                              hipStream_t);
extern "C" int deb_launch_contract(const double *, const double *, int, int, long, double *, hipStream_t);
extern "C" int deb_launch_cell(const double *, const double *, const double *, int, int, double, double, double *, hipStream_t);
extern "C" int deb_launch_dell(const double *, int, double, double, double *, double *, hipStream_t);

extern "C" int deb_launch_tau_out(const double *, int, int, const double *, int, double *, hipStream_t);
extern "C" int deb_launch_power_units(const double *, int, const int *, const double *, int, int, const double *, int, int, double *,
                                      hipStream_t);
extern "C" int deb_launch_perturb_batched(const double *, int, const int *, const double *, int, int, const double *, int, const double *,
                                          const DebProblem *, const DebSolverCfg *, double *, double *, int *, int *, int *, double *, int,
                                          hipStream_t);
extern "C" int deb_launch_perturb(const double *, int, const int *, const double *, int, const double *, int, const double *,
                                  const DebProblem *, const DebSolverCfg *, double *, double *, int *, int *, int *, const int *, int,
                                  const double *, long, const int *, hipStream_t);
extern "C" long deb_perturb_lds_bytes(int, int, int, int, int);
extern "C" long deb_evolve_sources_work_doubles(int, int, int, int, int);
extern "C" int deb_launch_evolve_sources(const double *, int, const double *, int, int, const double *, const double *, int, const DebProblem *,
                                         const DebSolverCfg *, int, double *, int *, int *, int *, double *, const int *, int, double *,
                                         const double *, long, const int *, hipStream_t);
extern "C" long deb_adjoint_lds_bytes(const DebProblem *, int);
extern "C" long deb_adjoint_record_doubles(const DebProblem *);
extern "C" int deb_launch_adjoint(const double *, int, const double *, double *, const double *, int, const int *, const double *, int,
                                  const double *, const double *, int, const double *, const DebProblem *, const double *, long, const int *,
                                  const double *, double *, long, double *, double *, double *, int *, const int *, hipStream_t);
extern "C" int deb_launch_spline_eval(const double *, int, int, int, const double *, int, double *, hipStream_t);
extern "C" int deb_launch_power(const double *, int, const double *, long, int, int, double, double, double, double *, hipStream_t);
extern "C" int deb_launch_cmb_neutrinos(const double *, const double *, const double *, int, const double *, int, int, int, int, double *,
                                        double *, double *, hipStream_t);
extern "C" int deb_launch_cmb_metric(const double *, const DebMetricInputs *, const double *, int, const double *, int, double *, hipStream_t);
extern "C" int deb_launch_cmb_polarization(const double *, const double *, int, int, int, int, int, const double *, const double *,
                                           const double *, const double *, const double *, double *, hipStream_t);
extern "C" int deb_launch_cmb_source_term(int, const double *, int, int, const double *, const double *, const double *, const double *,
                                          const double *, const double *, const double *, const double *, const double *, const double *, double,
                                          double *, hipStream_t);
extern "C" int deb_launch_cmb_isw(const double *, int, int, const double *, const double *, const double *, double *, double *, hipStream_t);
extern "C" int deb_launch_savgol_coefficients(int, int, double *, hipStream_t);
extern "C" int deb_launch_savgol_apply(const double *, int, const double *, int, double *, hipStream_t);
extern "C" int deb_launch_power_smoothed(const double *, const double *, int, int, int, double, double, double, int, double *, double *,
                                         hipStream_t);
extern "C" int deb_launch_kaiser(const double *, const double *, int, int, const double *, const double *, double, double, double, double, int,
                                 int, double *, double *, hipStream_t);
extern "C" int deb_launch_multipoles(const double *, const double *, int, int, double, double, double, double, double *, double *, double *,
                                     hipStream_t);
extern "C" int deb_launch_xi(const double *, const double *, int, int, double *, double *, double *, hipStream_t);
extern "C" int deb_launch_time_deriv(const double *, const double *, const double *, int, const double *, int, const double *,
                                     const DebProblem *, double *, hipStream_t);

namespace {
thread_local std::string g_err;
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
int hip_rc(int rc, const char *what) {
  if (rc != 0) {
    g_err = std::string(what) + ": " + hipGetErrorString((hipError_t)rc);
    return rc;
  }
  g_err.clear();
  return 0;
}
int check_problem(const DebProblem *p) {
  if (!p) return fail(-1, "problem descriptor is NULL");
  if (p->lmaxg < 3 || p->lmaxgp < 3 || p->lmaxr < 3 || p->lmaxnu < 3)
    return fail(-1, "hierarchy truncations lmaxg, lmaxgp, lmaxr, lmaxnu must be >= 3");
  if (p->lmaxg > 250 || p->lmaxgp > 250 || p->lmaxr > 250 || p->lmaxnu > 250) return fail(-1, "hierarchy truncations must be <= 250");
  if (p->nqmax < 1 || p->nqmax > 16) return fail(-1, "nqmax must be in [1, 16]");
  return 0;
}
}  // namespace

namespace deb {
// Error reporting for entry points defined in other source files (deb_noalloc.hip).
int capi_fail(int code, const char *msg) { return fail(code, msg); }
}  // namespace deb

extern "C" {

const char *deb_version(void) { return "discoeb_amd 0.1 (gfx950)"; }
const char *deb_last_error(void) { return g_err.c_str(); }
long deb_tables_size(int num_thermo) { return deb_table_doubles(num_thermo); }
long deb_table_offset(int slot, int num_thermo) { return deb_array_offset(slot, num_thermo); }
int deb_default_threads_per_unit(void) { return DEB_DEFAULT_THREADS_PER_UNIT; }

long deb_evolve_workspace_doubles(void) { return 32; }  // momentum bins: 16 nodes, 16 weights

int deb_nvar(int lmaxg, int lmaxgp, int lmaxr, int lmaxnu, int nqmax) {
  return 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nqmax * (lmaxnu + 1) + 2;
}

int deb_background(const double *d_params, int n_cosmo, int num_thermo, double *d_tables, void *stream) {
  if (n_cosmo < 0) return fail(-1, "n_cosmo must be >= 0");
  if (num_thermo < 16) return fail(-1, "num_thermo must be >= 16");
  if (n_cosmo == 0) return 0;
  if (!d_params || !d_tables) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_background(d_params, n_cosmo, num_thermo, 0, nullptr, d_tables, (hipStream_t)stream), "deb_background");
}

int deb_background_thermo(const double *d_params, int n_cosmo, int num_thermo, int thermo_module, const double *d_thermo_in,
                          double *d_tables, void *stream) {
  if (thermo_module < 0 || thermo_module > 2) return fail(-1, "thermo_module must be 0 (RECFAST), 1 (MB95) or 2 (given table)");
  if (n_cosmo < 0) return fail(-1, "n_cosmo must be >= 0");
  if (num_thermo < 16) return fail(-1, "num_thermo must be >= 16");
  if (n_cosmo == 0) return 0;
  if (!d_params || !d_tables || (thermo_module == 2 && !d_thermo_in)) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_background(d_params, n_cosmo, num_thermo, thermo_module, d_thermo_in, d_tables, (hipStream_t)stream),
                "deb_background_thermo");
}

int deb_spline_eval(const double *d_tables, int num_thermo, int cosmo, int spline_id, int kind, const double *d_x, int nx,
                    double *d_out, void *stream) {
  if (spline_id < 0 || spline_id >= DEB_SPL_COUNT) return fail(-1, "unknown spline id");
  if (kind < 0 || kind > 3) return fail(-1, "kind must be 0 (value), 1, 2 (derivatives) or 3 (integral)");
  if (kind == 3 && spline_id != DEB_SPL_OPAC) return fail(-1, "integrals are tabulated for the opacity spline only");
  if (cosmo < 0) return fail(-1, "cosmology index must be >= 0");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  return hip_rc(deb_launch_spline_eval(t, num_thermo, spline_id, kind, d_x, nx, d_out, (hipStream_t)stream), "deb_spline_eval");
}

// Argument checks shared by the evolution entry points (0 = fine).
static int check_solve(const double *d_tables, const int *d_unit_cosmo, const double *d_unit_k, int n_units, const double *d_tau_out,
                       int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status, int *d_n_accept,
                       int *d_n_reject, double *d_work) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (!cfg) return fail(-1, "solver descriptor is NULL");
  if (!(cfg->rtol > 0) || !(cfg->atol >= 0)) return fail(-1, "rtol must be > 0 and atol >= 0");
  if (cfg->max_steps < 1) return fail(-1, "max_steps must be >= 1");
  if (!(cfg->factormin > 0) || !(cfg->factormax >= 1)) return fail(-1, "factormin must be > 0 and factormax >= 1");
  if (nout < 1) return fail(-1, "at least one output time is required");
  if (n_units < 0) return fail(-1, "n_units must be >= 0");
  if (n_units == 0) return 0;
  if (!d_tables || !d_unit_cosmo || !d_unit_k || !d_tau_out || !d_y_out || !d_status || !d_n_accept || !d_n_reject || !d_work)
    return fail(-1, "NULL device pointer");
  long lds = deb_perturb_lds_bytes(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax);
  if (lds > 160 * 1024) return fail(-1, "state too large for the 160 KiB LDS of one CU (lower the hierarchy truncations)");
  return 0;
}

static int evolve_common(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                         const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status,
                         int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const DebSchedule *sched,
                         const double *d_replay, long replay_stride, const int *d_replay_len, void *stream) {
  int rc = check_solve(d_tables, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout, prob, cfg, d_y_out, d_status, d_n_accept, d_n_reject,
                       d_work);
  if (rc || n_units == 0) return rc;
  rc = deb_launch_nu_bins(prob->nqmax, d_work, (hipStream_t)stream);
  if (rc) return hip_rc(rc, "deb_evolve (momentum bins)");
  const int *order = sched ? sched->d_order : nullptr;
  const int nt = sched ? sched->threads_per_unit : 0;
  rc = deb_launch_perturb(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout, d_work, prob, cfg, d_step_log, d_y_out,
                          d_status, d_n_accept, d_n_reject, order, nt, d_replay, replay_stride, d_replay_len, (hipStream_t)stream);
  if (rc == -2) return fail(-1, "state too large for LDS");
  if (rc == -4) return fail(-1, "threads_per_unit must be 0, 64, 128 or 256");
  return hip_rc(rc, "deb_evolve");
}

int deb_evolve(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units, const double *d_tau_out,
               int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status, int *d_n_accept, int *d_n_reject,
               double *d_work, double *d_step_log, const DebSchedule *sched, void *stream) {
  return evolve_common(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout, prob, cfg, d_y_out, d_status, d_n_accept,
                       d_n_reject, d_work, d_step_log, sched, nullptr, 0, nullptr, stream);
}

int deb_evolve_replay(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                      const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, const double *d_mesh, long mesh_stride,
                      const int *d_mesh_len, double *d_y_out, int *d_status, int *d_n_accept, int *d_n_reject, double *d_work,
                      const DebSchedule *sched, void *stream) {
  if (n_units > 0 && (!d_mesh || !d_mesh_len)) return fail(-1, "NULL device pointer");
  if (mesh_stride < 1) return fail(-1, "mesh_stride must be >= 1");
  return evolve_common(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout, prob, cfg, d_y_out, d_status, d_n_accept,
                       d_n_reject, d_work, nullptr, sched, d_mesh, mesh_stride, d_mesh_len, stream);
}

long deb_adjoint_workspace_doubles(int n_units, int n_cosmo, int num_thermo, int max_mesh_len, const DebProblem *prob) {
  if (!prob || n_units < 0 || n_cosmo < 0 || max_mesh_len < 1) return -1;
  const long n = deb_nvar(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax);
  return 32 + 6L * n_cosmo * deb_table_doubles(num_thermo) + (long)n_units * ((max_mesh_len + 1) * n + (long)max_mesh_len * deb_adjoint_record_doubles(prob));
}

int deb_evolve_adjoint(const double *d_tables, int num_thermo, int n_cosmo, const double *d_shifted_tables, const double *d_dtheta,
                       const int *d_unit_cosmo, const double *d_unit_k, int n_units, const double *d_tau_out, const double *d_tau_out_shifted,
                       int nout, const DebProblem *prob, const double *d_mesh, long mesh_stride, int max_mesh_len, const int *d_mesh_len,
                       const double *d_gbar, double *d_grad, int *d_status, double *d_work, double *d_lam0, double *d_tbar,
                       const DebSchedule *sched, void *stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (n_units < 0 || n_cosmo < 1) return fail(-1, "n_units must be >= 0 and n_cosmo >= 1");
  if (n_units == 0) return 0;
  if (nout < 1 || nout > 64) return fail(-1, "the adjoint takes 1..64 output times");
  if (max_mesh_len < 1 || mesh_stride < max_mesh_len) return fail(-1, "max_mesh_len must lie in 1..mesh_stride");
  if (!d_tables || !d_shifted_tables || !d_dtheta || !d_unit_cosmo || !d_unit_k || !d_tau_out || !d_tau_out_shifted || !d_mesh ||
      !d_mesh_len || !d_gbar || !d_grad || !d_status || !d_work)
    return fail(-1, "NULL device pointer");
  if (sched && sched->threads_per_unit != 0 && sched->threads_per_unit != 256) return fail(-1, "the adjoint runs 256 threads per unit");
  if (deb_adjoint_lds_bytes(prob, 256) > 160 * 1024) return fail(-1, "state too large for LDS");
  rc = deb_launch_nu_bins(prob->nqmax, d_work, (hipStream_t)stream);
  if (rc) return hip_rc(rc, "deb_evolve_adjoint (momentum bins)");
  const long n = deb_nvar(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax);
  double *dtab = d_work + 32;
  double *ckpt = dtab + 6L * n_cosmo * deb_table_doubles(num_thermo);
  rc = deb_launch_adjoint(d_tables, num_thermo, d_shifted_tables, dtab, d_dtheta, n_cosmo, d_unit_cosmo, d_unit_k, n_units, d_tau_out,
                          d_tau_out_shifted, nout, d_work, prob, d_mesh, mesh_stride, d_mesh_len, d_gbar, ckpt,
                          (long)(max_mesh_len + 1) * n + (long)max_mesh_len * deb_adjoint_record_doubles(prob),
                          d_grad, d_lam0, d_tbar, d_status, sched ? sched->d_order : nullptr, (hipStream_t)stream);
  if (rc == -2) return fail(-1, "state too large for LDS");
  if (rc == -5) return fail(-1, "the adjoint takes 1..64 output times");
  return hip_rc(rc, "deb_evolve_adjoint");
}

long deb_evolve_batched_workspace_doubles(int n_units, const DebProblem *prob) {
  if (!prob || n_units < 0) return -1;
  return 32 + 2L * deb_nvar(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax) * n_units;
}

int deb_evolve_batched(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units, int batch_size,
                       const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status,
                       int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const DebSchedule *sched, void *stream) {
  if (batch_size < 1) return fail(-1, "batch_size must be >= 1");
  if (sched && sched->d_order) return fail(-1, "batches are consecutive units: the schedule's order must be NULL");
  if (n_units % batch_size != 0) return fail(-1, "the number of units must be a multiple of batch_size");
  int rc = check_solve(d_tables, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout, prob, cfg, d_y_out, d_status, d_n_accept, d_n_reject,
                       d_work);
  if (rc || n_units == 0) return rc;
  rc = deb_launch_nu_bins(prob->nqmax, d_work, (hipStream_t)stream);
  if (rc) return hip_rc(rc, "deb_evolve_batched (momentum bins)");
  rc = deb_launch_perturb_batched(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, batch_size, d_tau_out, nout, d_work, prob, cfg,
                                  d_step_log, d_y_out, d_status, d_n_accept, d_n_reject, d_work + 32, sched ? sched->threads_per_unit : 0,
                                  (hipStream_t)stream);
  if (rc == -2) return fail(-1, "state too large for LDS");
  if (rc == -4) return fail(-1, "threads_per_unit must be 0, 64, 128 or 256");
  if (rc == -3) return fail(-1, "the number of units must be a multiple of batch_size");
  return hip_rc(rc, "deb_evolve_batched");
}

int deb_tau_out(const double *d_tables, int num_thermo, int n_cosmo, const double *d_aexp_out, int nout, double *d_tau_out, void *stream) {
  if (n_cosmo < 0 || nout < 0) return fail(-1, "n_cosmo and nout must be >= 0");
  if (n_cosmo == 0 || nout == 0) return 0;
  if (!d_tables || !d_aexp_out || !d_tau_out) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_tau_out(d_tables, num_thermo, n_cosmo, d_aexp_out, nout, d_tau_out, (hipStream_t)stream), "deb_tau_out");
}

int deb_power(const double *d_k, int k_repeat, const double *d_y, long n, int ystride, int idx, double A_s, double n_s, double k_p,
              double *d_pk, void *stream) {
  if (k_repeat < 1 || ystride < 1 || idx < 0 || idx >= ystride) return fail(-1, "invalid stride, repeat or field index");
  return hip_rc(deb_launch_power(d_k, k_repeat, d_y, n, ystride, idx, A_s, n_s, k_p, d_pk, (hipStream_t)stream), "deb_power");
}

int deb_power_units(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units, int nout,
                    const double *d_y, int ystride, int idx, double *d_pk, void *stream) {
  if (n_units < 0 || nout < 0) return fail(-1, "n_units and nout must be >= 0");
  if (ystride < 1 || idx < 0 || idx >= ystride) return fail(-1, "invalid stride or field index");
  if ((long)n_units * nout == 0) return 0;
  if (!d_tables || !d_unit_cosmo || !d_unit_k || !d_y || !d_pk) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_power_units(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, nout, d_y, ystride, idx, d_pk,
                                       (hipStream_t)stream),
                "deb_power_units");
}

int deb_time_derivatives(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_tau, int n_tau,
                         const double *d_k, int n_k, const DebProblem *prob, double *d_dy, double *d_work, void *stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (n_k <= 0 || n_tau <= 0) return 0;
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  rc = deb_launch_nu_bins(prob->nqmax, d_work, (hipStream_t)stream);
  if (rc) return hip_rc(rc, "deb_time_derivatives (momentum bins)");
  return hip_rc(deb_launch_time_deriv(t, d_y, d_tau, n_tau, d_k, n_k, d_work, prob, d_dy, (hipStream_t)stream), "deb_time_derivatives");
}

// ------------------------------------------------------------------ CMB chain
long deb_cmb_sources_workspace_doubles(int num_thermo, int n_k, int n_tau) { return deb_cmb_sources_work_doubles(num_thermo, n_k, n_tau); }

int deb_cmb_sources(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_yprime, const double *d_k, int n_k,
                    const double *d_aexp, const double *d_tau, int n_tau, const DebProblem *prob, int isw_mode, double *d_out, double *d_work,
                    void *stream) {
  int rc = check_problem(prob);
  if (rc) return rc;
  if (isw_mode != 0 && isw_mode != 1) return fail(-1, "isw_mode must be 0 (spline derivative) or 1 (algebraic)");
  if (n_k < 1 || n_tau < 3) return fail(-1, "need n_k >= 1 and n_tau >= 3");
  if (cosmo < 0) return fail(-1, "cosmology index must be >= 0");
  if (!d_tables || !d_y || !d_yprime || !d_k || !d_aexp || !d_tau || !d_out || !d_work) return fail(-1, "NULL device pointer");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  return hip_rc(deb_launch_cmb_sources(t, num_thermo, d_y, d_yprime, d_k, n_k, d_aexp, d_tau, n_tau, prob, isw_mode, d_out, d_work,
                                       (hipStream_t)stream),
                "deb_cmb_sources");
}

long deb_evolve_sources_workspace_doubles(int num_thermo, int n_k, int n_tau, int batch_size, const DebProblem *prob) {
  if (!prob || n_k < 0 || n_tau < 0 || batch_size < 0) return -1;
  return deb_evolve_sources_work_doubles(num_thermo, n_k, n_tau, batch_size,
                                         deb_nvar(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax));
}

int deb_evolve_sources(const double *d_tables, int num_thermo, int cosmo, const double *d_k, int n_k, int batch_size, const double *d_aexp,
                       const double *d_tau, int n_tau, const DebProblem *prob, const DebSolverCfg *cfg, int isw_mode, double *d_out, int *d_status,
                       int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const double *d_mesh, long mesh_stride,
                       const int *d_mesh_len, const DebSchedule *sched, void *stream) {
  if (isw_mode != 0 && isw_mode != 1) return fail(-1, "isw_mode must be 0 (spline derivative) or 1 (algebraic)");
  if (d_mesh && batch_size > 0) return fail(-1, "a replay mesh takes per-mode steps (batch_size 0)");
  if (d_mesh && (!d_mesh_len || mesh_stride < 1)) return fail(-1, "a replay mesh needs its lengths and a stride >= 1");
  if (d_mesh && d_step_log) return fail(-1, "step log and replay mesh are exclusive");
  if (n_tau < 3) return fail(-1, "need n_tau >= 3");
  if (cosmo < 0) return fail(-1, "cosmology index must be >= 0");
  if (batch_size < 0) return fail(-1, "batch_size must be >= 0");
  if (batch_size > 0 && n_k % batch_size != 0) return fail(-1, "with batched steps n_k must be a multiple of batch_size");
  if (batch_size > 0 && sched && sched->d_order) return fail(-1, "batched steps take no launch order");
  int rc = check_solve(d_tables, &cosmo, d_k, n_k, d_tau, n_tau, prob, cfg, d_out, d_status, d_n_accept, d_n_reject, d_work);
  if (rc || n_k == 0) return rc;
  if (!d_aexp) return fail(-1, "NULL device pointer");
  const int nt = sched ? sched->threads_per_unit : 0;
  if (nt != 0 && nt != 64 && nt != 128 && nt != 256) return fail(-1, "threads_per_unit must be 0, 64, 128 or 256");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  rc = deb_launch_evolve_sources(t, num_thermo, d_k, n_k, batch_size, d_aexp, d_tau, n_tau, prob, cfg, isw_mode, d_out, d_status, d_n_accept,
                                 d_n_reject, d_work, sched ? sched->d_order : nullptr, nt, d_step_log, d_mesh, mesh_stride, d_mesh_len,
                                 (hipStream_t)stream);
  if (rc == -2) return fail(-1, "state too large for LDS");
  return hip_rc(rc, "deb_evolve_sources");
}

int deb_visibility(const double *d_tables, int num_thermo, int cosmo, const double *d_tau, const double *d_aexp, int n_tau, double *d_out,
                   double *d_work, void *stream) {
  if (n_tau < 1) return 0;
  if (cosmo < 0) return fail(-1, "cosmology index must be >= 0");
  if (!d_tables || !d_tau || !d_aexp || !d_out || !d_work) return fail(-1, "NULL device pointer");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  return hip_rc(deb_launch_visibility(t, num_thermo, d_tau, d_aexp, n_tau, d_out, d_work, (hipStream_t)stream), "deb_visibility");
}

long deb_los_workspace_doubles(int n_k, int n_tau, int ellmax, int n_fftlog, int n_k_dense) {
  return deb_los_work_doubles(n_k, n_tau, ellmax + 1, n_fftlog, n_k_dense);
}

int deb_los_theta_ell(const double *d_k, int n_k, const double *d_tau, int n_tau, const double *d_S, double tau0, int ellmax, int n_fftlog,
                      int n_k_dense, double *d_theta, double *d_k_out, double *d_work, void *stream) {
  return deb_los_theta_ell_field(d_k, n_k, d_tau, n_tau, d_S, tau0, ellmax, n_fftlog, n_k_dense, 0, d_theta, d_k_out, d_work, stream);
}

int deb_los_theta_ell_field(const double *d_k, int n_k, const double *d_tau, int n_tau, const double *d_S, double tau0, int ellmax, int n_fftlog,
                            int n_k_dense, int field, double *d_theta, double *d_k_out, double *d_work, void *stream) {
  if (field != 0 && field != 1) return fail(-1, "field must be 0 (temperature) or 1 (E-mode)");
  if (ellmax < 0) return fail(-1, "ellmax must be >= 0");
  if (n_fftlog < 128 || n_fftlog > 16384 || (n_fftlog & (n_fftlog - 1))) return fail(-1, "n_fftlog must be a power of two in [128, 16384]");
  if (n_k < 2 || n_tau < 3) return fail(-1, "need n_k >= 2 and n_tau >= 3");
  if (n_k_dense < 0 || n_k_dense == 1) return fail(-1, "n_k_dense must be 0 or >= 2");
  if (!d_k || !d_tau || !d_S || !d_theta || !d_k_out || !d_work) return fail(-1, "NULL device pointer");
  int rc = deb_launch_los(d_k, n_k, d_tau, n_tau, d_S, tau0, ellmax, n_fftlog, n_k_dense, field, d_theta, d_k_out, d_work, (hipStream_t)stream);
  if (rc == -2) return fail(-1, "n_fftlog must be a power of two in [128, 16384]");
  return hip_rc(rc, "deb_los_theta_ell");
}

int deb_contract(const double *d_A, const double *d_B, int m, int n, long k, double *d_C, void *stream) {
  if (m < 0 || n < 0 || k < 0 || (k % 16)) return fail(-1, "m, n >= 0 and k a non-negative multiple of 16");
  if (((uintptr_t)d_A | (uintptr_t)d_B) & 15) return fail(-1, "d_A and d_B must be 16-byte aligned");
  if ((long)m * n == 0) return 0;
  if (!d_A || !d_B || !d_C) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_contract(d_A, d_B, m, n, k, d_C, (hipStream_t)stream), "deb_contract");
}

int deb_cell(const double *d_theta, const double *d_k, int n_k, int n_ell, double n_s, double k_p, double *d_cell, void *stream) {
  if (n_k < 2 || n_ell < 1) return fail(-1, "need n_k >= 2 and n_ell >= 1");
  if (!d_theta || !d_k || !d_cell) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_cell(d_theta, nullptr, d_k, n_k, n_ell, n_s, k_p, d_cell, (hipStream_t)stream), "deb_cell");
}

int deb_cell_cross(const double *d_theta1, const double *d_theta2, const double *d_k, int n_k, int n_ell, double n_s, double k_p, double *d_cell,
                   void *stream) {
  if (n_k < 2 || n_ell < 1) return fail(-1, "need n_k >= 2 and n_ell >= 1");
  if (!d_theta1 || !d_theta2 || !d_k || !d_cell) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_cell(d_theta1, d_theta2, d_k, n_k, n_ell, n_s, k_p, d_cell, (hipStream_t)stream), "deb_cell_cross");
}

int deb_dell(const double *d_cell, int ellmax, double A_s, double Tcmb, double *d_ell, double *d_dell, void *stream) {
  if (ellmax < 2) return fail(-1, "ellmax must be >= 2");
  if (!d_cell || !d_ell || !d_dell) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_dell(d_cell, ellmax, A_s, Tcmb, d_ell, d_dell, (hipStream_t)stream), "deb_dell");
}

// ------------------------------------------------------------------ CMB source stages

int deb_cmb_neutrinos(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_yprime, int n_k,
                      const double *d_aexp, int n_tau, int nvar, int iq0, int nqmax, double *d_out, double *d_aprimeoa, double *d_work,
                      void *stream) {
  if (nqmax < 1 || nqmax > 16) return fail(-1, "nqmax must be in [1, 16]");
  if (iq0 < 0 || iq0 + 3 * nqmax > nvar) return fail(-1, "iq0 + 3 nqmax must not exceed nvar");
  if (n_k < 1 || n_tau < 1) return fail(-1, "need n_k >= 1 and n_tau >= 1");
  if (!d_tables || !d_y || !d_yprime || !d_aexp || !d_out || !d_aprimeoa || !d_work) return fail(-1, "NULL device pointer");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  return hip_rc(deb_launch_cmb_neutrinos(t, d_y, d_yprime, n_k, d_aexp, n_tau, nvar, iq0, nqmax, d_out, d_aprimeoa, d_work, (hipStream_t)stream),
                "deb_cmb_neutrinos");
}

int deb_cmb_metric(const double *d_tables, int num_thermo, int cosmo, const DebMetricInputs *in, const double *d_k, int n_k,
                   const double *d_aexp, int n_tau, double *d_out, void *stream) {
  if (!in) return fail(-1, "metric inputs are NULL");
  if (n_k < 1 || n_tau < 1) return fail(-1, "need n_k >= 1 and n_tau >= 1");
  const double *const ptrs[] = {in->eta,    in->deltac,  in->thetac,  in->deltab,  in->thetab,       in->deltag,        in->thetag,
                                in->shearg, in->sheargp, in->deltar,  in->thetar,  in->shearr,       in->shearrp,       in->deltaq,
                                in->thetaq, in->drhonu,  in->fnu,     in->shearnu, in->rho_nu_prime, in->shear_nu_prime, in->aH,
                                in->rhonu,  in->rho_Q,   in->w_Q,     d_tables,    d_k,              d_aexp,            d_out};
  for (const double *q : ptrs)
    if (!q) return fail(-1, "NULL device pointer");
  const double *t = d_tables + (long)cosmo * deb_table_doubles(num_thermo);
  return hip_rc(deb_launch_cmb_metric(t, in, d_k, n_k, d_aexp, n_tau, d_out, (hipStream_t)stream), "deb_cmb_metric");
}

int deb_cmb_polarization(const double *d_y, const double *d_yprime, int n_k, int n_tau, int nvar, int lmaxg, int lmaxgp, const double *d_k,
                         const double *d_thetagprime, const double *d_alphaprime, const double *d_opac, const double *d_opacprime,
                         double *d_out, void *stream) {
  if (lmaxg < 3 || lmaxgp < 3 || 10 + lmaxg + lmaxgp > nvar) return fail(-1, "need lmaxg, lmaxgp >= 3 within nvar");
  if (n_k < 1 || n_tau < 1) return fail(-1, "need n_k >= 1 and n_tau >= 1");
  if (!d_y || !d_yprime || !d_k || !d_thetagprime || !d_alphaprime || !d_opac || !d_opacprime || !d_out)
    return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_cmb_polarization(d_y, d_yprime, n_k, n_tau, nvar, 7, 8 + lmaxg, d_k, d_thetagprime, d_alphaprime, d_opac, d_opacprime,
                                            d_out, (hipStream_t)stream),
                "deb_cmb_polarization");
}

int deb_cmb_source_term(int term, const double *d_k, int n_k, int n_tau, const double *d_vis, const double *d_deltag, const double *d_alpha,
                        const double *d_alphaprime, const double *d_thetab, const double *d_thetabprime, const double *d_pol,
                        const double *d_polprime, const double *d_polpprime, const double *d_tau, double tau0, double *d_out, void *stream) {
  if (term < 2 || term > 5) return fail(-1, "term must be 2 (Sachs-Wolfe), 3 (Doppler), 4 (polarisation) or 5 (E-mode)");
  if (n_k < 1 || n_tau < 1) return fail(-1, "need n_k >= 1 and n_tau >= 1");
  bool ok = d_k && d_vis && d_pol && d_out;
  if (term == 2) ok = ok && d_deltag && d_alphaprime && d_thetabprime && d_polpprime;
  if (term == 3) ok = ok && d_alpha && d_thetab && d_polprime;
  if (term == 5) ok = ok && d_tau;
  if (!ok) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_cmb_source_term(term, d_k, n_k, n_tau, d_vis, d_deltag, d_alpha, d_alphaprime, d_thetab, d_thetabprime, d_pol,
                                           d_polprime, d_polpprime, d_tau, tau0, d_out, (hipStream_t)stream),
                "deb_cmb_source_term");
}

long deb_cmb_isw_workspace_doubles(int n_k, int n_tau) { return 2L * n_k * n_tau + n_tau + 64; }

int deb_cmb_source_isw(const double *d_tau, int n_tau, int n_k, const double *d_eta, const double *d_alphaprime, const double *d_optical_depth,
                       double *d_S1, double *d_work, void *stream) {
  if (n_k < 1 || n_tau < 3) return fail(-1, "need n_k >= 1 and n_tau >= 3");
  if (!d_tau || !d_eta || !d_alphaprime || !d_optical_depth || !d_S1 || !d_work) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_cmb_isw(d_tau, n_tau, n_k, d_eta, d_alphaprime, d_optical_depth, d_S1, d_work, (hipStream_t)stream),
                "deb_cmb_source_isw");
}

// ------------------------------------------------------------------ large-scale structure

int deb_savgol_coefficients(int window_length, int polyorder, double *d_coef, void *stream) {
  if (polyorder < 0 || polyorder > 8) return fail(-1, "polyorder must lie in 0..8");
  if (window_length < polyorder + 1) return fail(-1, "window_length must exceed polyorder");
  if (!d_coef) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_savgol_coefficients(window_length, polyorder, d_coef, (hipStream_t)stream), "deb_savgol_coefficients");
}

int deb_savgol_filter(const double *d_y, int n, const double *d_coef, int window_length, double *d_out, void *stream) {
  if (n < 1 || window_length < 1 || window_length > n) return fail(-1, "need 1 <= window_length <= n");
  if (!d_y || !d_coef || !d_out) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_savgol_apply(d_y, n, d_coef, window_length, d_out, (hipStream_t)stream), "deb_savgol_filter");
}

long deb_power_smoothed_workspace_doubles(int n, int window_length) { return (long)n + window_length; }

int deb_power_smoothed(const double *d_k, const double *d_y, int n, int ystride, int idx, double A_s, double n_s, double k_p,
                       int window_length, double *d_out, double *d_work, void *stream) {
  if (window_length < 5 || window_length % 2 == 0 || window_length > n)
    return fail(-1, "window_length must be odd and lie in 5..n (the smoothing scale spans too few or too many k-modes)");
  if (ystride < 1 || idx < 0 || idx >= ystride) return fail(-1, "invalid stride or field index");
  if (!d_k || !d_y || !d_out || !d_work) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_power_smoothed(d_k, d_y, n, ystride, idx, A_s, n_s, k_p, window_length, d_out, d_work, (hipStream_t)stream),
                "deb_power_smoothed");
}

int deb_power_kaiser(const double *d_k, const double *d_y, int n, int ystride, const double *d_pdelta, const double *d_ptheta, double bias,
                     double A_s, double n_s, double k_p, int nmu, int mu_sampling, double *d_pkmu, double *d_mu, void *stream) {
  if (n < 1 || nmu < 1) return fail(-1, "need n >= 1 and nmu >= 1");
  if ((d_pdelta == nullptr) != (d_ptheta == nullptr)) return fail(-1, "give both smoothed spectra or neither");
  if (!d_pdelta && (!d_y || ystride < 6)) return fail(-1, "the unsmoothed form reads fields 4 and 5 of d_y");
  if (!d_k || !d_pkmu || !d_mu) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_kaiser(d_k, d_y, n, ystride, d_pdelta, d_ptheta, bias, A_s, n_s, k_p, nmu, mu_sampling ? 1 : 0, d_pkmu, d_mu,
                                  (hipStream_t)stream),
                "deb_power_kaiser");
}

int deb_power_multipoles(const double *d_k, const double *d_y, int n, int ystride, double b, double A_s, double n_s, double k_p, double *d_p0,
                         double *d_p2, double *d_p4, void *stream) {
  if (n < 1 || ystride < 6) return fail(-1, "need n >= 1 and fields 4 and 5 in d_y");
  if (!d_k || !d_y || !d_p0 || !d_p2 || !d_p4) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_multipoles(d_k, d_y, n, ystride, b, A_s, n_s, k_p, d_p0, d_p2, d_p4, (hipStream_t)stream), "deb_power_multipoles");
}

long deb_xi_workspace_doubles(int n) { return 2L * (n / 2 + 1); }

int deb_xi_from_power(const double *d_k, const double *d_pk, int n, int ell, double *d_xi, double *d_r, double *d_work, void *stream) {
  if (n < 4 || n % 2) return fail(-1, "n must be even and >= 4");
  if (ell < 0) return fail(-1, "ell must be >= 0");
  if (!d_k || !d_pk || !d_xi || !d_r || !d_work) return fail(-1, "NULL device pointer");
  return hip_rc(deb_launch_xi(d_k, d_pk, n, ell, d_xi, d_r, d_work, (hipStream_t)stream), "deb_xi_from_power");
}

}  // extern "C"
