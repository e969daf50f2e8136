/* discoeb_amd.h -- C ABI of the MI355X (gfx950) linear-perturbation solver.
 *
 * Every pointer argument named d_* is a device pointer (HBM) owned by the caller; the library
 * never allocates, frees or retains caller memory.  Work is enqueued on the HIP stream passed
 * as `stream` (a hipStream_t; NULL = default stream) and is asynchronous unless stated.
 * Return codes: 0 success, < 0 invalid argument (see deb_last_error()), > 0 hipError_t.
 *
 * Each entry point names the function of the reference (lawrenceberry/DISCO-EB, paths relative
 * to its repository root) whose role it takes over.
 */
#ifndef DISCOEB_AMD_H
#define DISCOEB_AMD_H

#include "../disco-eb_amd/csrc/deb_layout.h"
#include "../disco-eb_amd/csrc/deb_perturb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library version string. */
const char *deb_version(void);

/* Message of the last failing call on this host thread ("" if none). */
const char *deb_last_error(void);

/* Doubles per cosmology in the background table block (deb_layout.h) for num_thermo knots,
 * and the offset of one array slot inside a block. */
long deb_tables_size(int num_thermo);
long deb_table_offset(int slot, int num_thermo);

/* Background evolution and RECFAST thermal history for n_cosmo cosmologies.
 * d_params: n_cosmo x DEB_NPARAM doubles (deb_layout.h order).  d_tables: n_cosmo x
 * deb_tables_size(num_thermo) doubles, written.
 * Replaces evolve_background(param=..., thermo_module='RECFAST', num_thermo=...)
 * (src/discoeb/background.py:191-345, thermodynamics_recfast.py:456-492). */
int deb_background(const double *d_params, int n_cosmo, int num_thermo, double *d_tables, void *stream);

/* deb_background with the thermal-history module chosen: thermo_module 0 RECFAST (as
 * deb_background), 1 MB95 (num_thermo points log-uniform in conformal time; the reference's
 * compute_thermo, src/discoeb/thermodynamics_mb95.py:92-212), 2 a given thermal history:
 * d_thermo_in holds per cosmology 5 x num_thermo doubles -- scale factor (ascending), conformal
 * time [Mpc], x_e, T_b [K], c_b^2 -- from which the splines, opacity and visibility are built and
 * whose a-range sets the background's (the CLASS input of background.py:230-236, 271-285).
 * Replaces evolve_background's thermo_module / class_thermo options (src/discoeb/background.py:191-345). */
int deb_background_thermo(const double *d_params, int n_cosmo, int num_thermo, int thermo_module, const double *d_thermo_in,
                          double *d_tables, void *stream);

/* Spline queries on the background tables of one cosmology: kind 0 value, 1 first derivative,
 * 2 second derivative, 3 integral from the first knot (opacity spline only).
 * Replaces spline_interpolation.evaluate/derivative/derivative2/integral
 * (src/discoeb/spline_interpolation.py:147-276) on the param[...] splines. */
enum {
  DEB_SPL_TAU_OF_A, DEB_SPL_A_OF_TAU, DEB_SPL_XE_OF_TAU, DEB_SPL_CS2A_OF_TAU, DEB_SPL_TEMPBA_OF_TAU, DEB_SPL_XE_OF_LOGA,
  DEB_SPL_CS2A_OF_LOGA, DEB_SPL_LOGRHONU_OF_LOGA, DEB_SPL_LOGPNU_OF_LOGA, DEB_SPL_LOGPPNU_OF_LOGA, DEB_SPL_OPAC, DEB_SPL_COUNT
};
int deb_spline_eval(const double *d_tables, int num_thermo, int cosmo, int spline_id, int kind, const double *d_x, int nx,
                    double *d_out, void *stream);

/* Threads per unit deb_evolve and deb_evolve_batched use when no schedule sets them. */
int deb_default_threads_per_unit(void);

/* Doubles of device workspace deb_evolve needs (neutrino momentum bins). */
long deb_evolve_workspace_doubles(void);

/* Adaptive Rodas5 evolution of n_units independent (cosmology, k) modes.
 * d_unit_cosmo[u], d_unit_k[u]: cosmology index and wavenumber [1/Mpc] of unit u.
 * d_tau_out: n_cosmo x nout conformal output times (ascending per cosmology).
 * d_y_out: n_units x nout x (return_full ? nvar : 20) doubles, written.
 * d_status[u]: 0 ok, 1 max_steps exceeded, 2 non-finite error norm (outputs NaN from there),
 * 4 w(a) = w0 + wa (1 - a) crosses -1 in (0, 1] (singular fluid equations; not integrated, NaN outputs).
 * d_n_accept / d_n_reject: accepted / rejected step counts per unit.
 * d_work: deb_evolve_workspace_doubles() doubles.
 * d_step_log: NULL, or n_units x max_steps x 3 doubles receiving (t_prev, t_next, error norm)
 * of every attempted step (the step mesh, for replay checks and frozen-mesh derivatives).
 * sched: NULL (units in index order, library-default width) or a DebSchedule (deb_perturb.h)
 * giving the launch order of the units, one workgroup each (most expensive first shortens the
 * tail), and the workgroup width (threads_per_unit).  Results depend on neither: every thread
 * count performs the same operations on each entry.
 * Replaces evolve_perturbations (src/discoeb/perturbations.py:1124-1226) with
 * evolve_one_mode (:882-949) per unit. */
int deb_evolve(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
               const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status,
               int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const DebSchedule *sched, void *stream);

/* Doubles of device workspace deb_evolve_batched needs for n_units units: the momentum bins
 * plus two state vectors per unit. */
long deb_evolve_batched_workspace_doubles(int n_units, const DebProblem *prob);

/* Batched shared-step Rodas5 evolution: the units g * batch_size .. (g + 1) * batch_size - 1
 * (n_units a multiple of batch_size, each batch within one cosmology) are integrated together
 * from the earliest of their start times with one step size and one accept decision, taken on
 * the RMS of the filtered error over all batch_size x 6 entries.  Outputs, status and step
 * counts as deb_evolve (every unit of a batch has the batch's counts); status 5 marks a batch
 * whose units have different cosmologies (not integrated, NaN outputs).
 * d_work: deb_evolve_batched_workspace_doubles(n_units, prob) doubles.
 * d_step_log: NULL, or (n_units / batch_size) x max_steps x 3 doubles receiving (t_prev,
 * t_next, error norm) of every attempted step of each batch.
 * sched: NULL or a DebSchedule with d_order NULL; its threads_per_unit sets the width of the
 * workgroup that runs each batch.
 * Replaces evolve_perturbations_batched (src/discoeb/perturbations.py:1228-1337) with
 * evolve_modes_batched (:1001-1120), determine_start_times_per_batch (:954-999) and
 * Rodas5Batched (src/discoeb/ode_integrators_stiff.py:904-1059). */
int deb_evolve_batched(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units, int batch_size,
                       const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, double *d_y_out, int *d_status,
                       int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const DebSchedule *sched, void *stream);

/* Doubles of the static device scratch deb_evolve_batched_noalloc uses on each device (8 MiB). */
#define DEB_NOALLOC_SCRATCH_DOUBLES (1L << 20)
long deb_noalloc_scratch_doubles(void);

/* deb_evolve_batched with no workspace argument and no device allocation: its workspace (the
 * momentum bins and two state vectors per unit, deb_evolve_batched_workspace_doubles(n_units,
 * prob) doubles) is a file-scope device array of DEB_NOALLOC_SCRATCH_DOUBLES doubles, one per
 * device, held until the process ends; larger problems are rejected (error -1).  Arguments,
 * outputs and status as deb_evolve_batched, on the device current at the call.  Calls that
 * overlap on one device (two streams or two host threads) share the array and race; calls
 * ordered on one stream do not.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int deb_evolve_batched_noalloc(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                               int batch_size, const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg,
                               double *d_y_out, int *d_status, int *d_n_accept, int *d_n_reject, double *d_step_log,
                               const DebSchedule *sched, void *stream);

/* deb_evolve on a frozen step mesh: unit u takes exactly d_mesh_len[u] steps whose ends are
 * t0 + s (t1 - t0) for the increasing s = d_mesh[u * mesh_stride + i] (the last one 1), with
 * no error control or rejection; t0 and t1 are the unit's own start and final output times.
 * The map from parameters to outputs with the mesh held fixed is what gradients differentiate.
 * Replaces, for differentiation, the forward-mode pass of jax.jacfwd through diffeqsolve
 * (src/discoeb/perturbations.py:904-925 with DirectAdjoint), with step sizes held fixed. */
int deb_evolve_replay(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                      const double *d_tau_out, int nout, const DebProblem *prob, const DebSolverCfg *cfg, const double *d_mesh, long mesh_stride,
                      const int *d_mesh_len, double *d_y_out, int *d_status, int *d_n_accept, int *d_n_reject, double *d_work,
                      const DebSchedule *sched, void *stream);

/* Discrete adjoint of deb_evolve_replay: per unit, the gradient of <gbar, outputs> with respect to
 * the six parameters that enter the background and the perturbations, theta_p for p = 0..5 =
 * (omega_c, omega_b, h, m_nu, w0, wa).  Reverse Rodas5 sweep over checkpointed states (the
 * transposed structured solve of every stage, vector-Jacobian products of the right-hand side);
 * the initial state, start and end times and output times by central differences.
 * d_shifted_tables: n_cosmo x 12 table blocks, cosmology c's block 2 p + 0 at theta_p + dtheta[c, p]
 *   and 2 p + 1 at theta_p - dtheta[c, p] (the other parameters at cosmology c's values);
 * d_dtheta: n_cosmo x 6 steps; d_tau_out_shifted: n_cosmo x 12 x nout output times of those;
 * d_mesh, mesh_stride, d_mesh_len: as deb_evolve_replay (mesh_len <= max_mesh_len);
 * d_gbar: n_units x nout x 20 cotangents of the output fields (nout <= 64);
 * d_grad: n_units x 6, written (NaN and status 4 for a unit whose w(a) crosses -1, status 7 for a
 *   unit whose mesh_len is outside 1..max_mesh_len);
 * d_work: deb_adjoint_workspace_doubles(...) doubles (momentum bins, derivative tables,
 *   checkpoints: per unit the state at every step start and a record of every step's stages and
 *   factorisation, so the reverse sweep does not recompute the steps); d_lam0 (n_units x nvar) and d_tbar (n_units x 2), optional (NULL): the
 *   cotangents of the initial state and of the start and end times.  Two launches: the forward replay
 *   (256 threads per unit, or 128 when the units outnumber the compute units) and the reverse sweep
 *   (256 threads per unit).
 * Replaces the forward-mode derivative jax.jacfwd takes through diffeqsolve (DirectAdjoint,
 * src/discoeb/perturbations.py:904-925) and the Rodas5 step (src/discoeb/ode_integrators_stiff.py:
 * 776-892), with the step mesh held fixed; reverse mode gives all six parameter derivatives of
 * one mode in one sweep. */
long deb_adjoint_workspace_doubles(int n_units, int n_cosmo, int num_thermo, int max_mesh_len, const DebProblem *prob);
int deb_evolve_adjoint(const double *d_tables, int num_thermo, int n_cosmo, const double *d_shifted_tables, const double *d_dtheta,
                       const int *d_unit_cosmo, const double *d_unit_k, int n_units, const double *d_tau_out, const double *d_tau_out_shifted,
                       int nout, const DebProblem *prob, const double *d_mesh, long mesh_stride, int max_mesh_len, const int *d_mesh_len,
                       const double *d_gbar, double *d_grad, int *d_status, double *d_work, double *d_lam0, double *d_tbar,
                       const DebSchedule *sched, void *stream);

/* Conformal output times tau_out[c, j] = tau(a = d_aexp_out[j]) of every cosmology c, from its
 * tau_of_a spline; d_tau_out: n_cosmo x nout doubles.
 * Replaces tau_out = param['tau_of_a_spline'].evaluate(aexp_out) (perturbations.py:1200-1206). */
int deb_tau_out(const double *d_tables, int num_thermo, int n_cosmo, const double *d_aexp_out, int nout, double *d_tau_out, void *stream);

/* Number of state variables for the given truncations (perturbations.py:893). */
int deb_nvar(int lmaxg, int lmaxgp, int lmaxr, int lmaxnu, int nqmax);

/* P(k) = 2 pi^2 A_s (k/k_p)^(n_s-1) k^-3 y[idx]^2 for n wavenumbers; y row stride `ystride`
 * doubles, k of row r is d_k[r / k_repeat].
 * Replaces get_power (src/discoeb/perturbations.py:1446-1468). */
int deb_power(const double *d_k, int k_repeat, const double *d_y, long n, int ystride, int idx, double A_s, double n_s, double k_p,
              double *d_pk, void *stream);

/* P(k) for the output of deb_evolve over many cosmologies: d_pk[u, j] from d_y[u, j, idx]
 * (row stride ystride) with k = d_unit_k[u] and A_s, n_s, k_p of cosmology d_unit_cosmo[u]
 * read from its table block, so a captured graph stays valid when parameters change.
 * Replaces get_power (src/discoeb/perturbations.py:1446-1468) applied per cosmology. */
int deb_power_units(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units, int nout,
                    const double *d_y, int ystride, int idx, double *d_pk, void *stream);

/* dy/dtau re-evaluated at saved states: d_y is n_k x n_tau x nvar, d_tau (n_tau), d_k (n_k);
 * d_dy has the same shape.  Replaces compute_time_derivatives
 * (src/discoeb/perturbations.py:1339-1406). */
int deb_time_derivatives(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_tau, int n_tau,
                         const double *d_k, int n_k, const DebProblem *prob, double *d_dy, double *d_work, void *stream);

/* ---------------------------------------------------------------- CMB chain ---------------- */

/* Visibility functions and background quantities at conformal times d_tau (scale factors
 * d_aexp), n_tau of each: d_out is 10 x n_tau doubles, rows opac, opac', g, g', g'', tau_c,
 * aH, rho_nu/rho_nu0, rho_Q/rho_Q(1), w_Q.  d_work: 12 num_thermo doubles.
 * Replaces compute_visibility_functions (src/discoeb/cmb.py:92-166) and
 * compute_background_quantities (src/discoeb/background.py:348-406). */
int deb_visibility(const double *d_tables, int num_thermo, int cosmo, const double *d_tau, const double *d_aexp, int n_tau, double *d_out,
                   double *d_work, void *stream);

/* Doubles of workspace deb_cmb_sources needs. */
long deb_cmb_sources_workspace_doubles(int num_thermo, int n_k, int n_tau);

/* CMB temperature source function from full states d_y and their time derivatives d_yprime
 * (both n_k x n_tau x nvar, as deb_evolve with return_full and deb_time_derivatives write
 * them) at scale factors d_aexp / conformal times d_tau (n_tau, increasing).  d_out: 8 x n_k x
 * n_tau doubles, planes S1 (ISW), S2 (Sachs-Wolfe), S3 (Doppler), S4 (polarisation), S, alpha,
 * alpha', S_E (the E-mode source 3 g Pi / (16 (k chi)^2), chi = tau0 - tau, zero for chi < 1 Mpc;
 * no counterpart in the reference).  isw_mode 0: S1 = e^-tau_c d/dtau(eta + alpha') from a natural spline in tau (the
 * reference's compute_source_term_isw, src/discoeb/cmb.py:437-449); 1: S1 = e^-tau_c (eta' +
 * alpha''), the algebraic form of its docstring (:406-408).  Replaces extract_perturbations ..
 * compute_source_function (src/discoeb/cmb.py:17-633) as chained in
 * compute_Cell_spectrum_from_cosmo_params (:1050-1073). */
int deb_cmb_sources(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_yprime, const double *d_k, int n_k,
                    const double *d_aexp, const double *d_tau, int n_tau, const DebProblem *prob, int isw_mode, double *d_out, double *d_work,
                    void *stream);

/* Doubles of workspace deb_evolve_sources needs (batch_size 0: per-mode steps). */
long deb_evolve_sources_workspace_doubles(int num_thermo, int n_k, int n_tau, int batch_size, const DebProblem *prob);

/* The chain perturbations -> time derivatives -> source function in one pass, for n_k modes d_k of
 * cosmology `cosmo` at n_tau output times d_tau (its conformal times of the scale factors d_aexp,
 * increasing): each mode's solve, at every output time, projects the interpolated state (RSA),
 * evaluates the right-hand side there and forms the source terms from the two, so the full states
 * and their derivatives (n_k x n_tau x nvar each) are never stored.  d_out: 8 x n_k x n_tau
 * doubles, the planes of deb_cmb_sources, equal to what deb_evolve (return_full) +
 * deb_time_derivatives + deb_cmb_sources write.  batch_size 0: adaptive steps per mode (as
 * deb_evolve; sched gives the launch order and width); > 0: batches of batch_size modes on shared
 * steps (as deb_evolve_batched; n_k a multiple of batch_size, no launch order).  d_status,
 * d_n_accept, d_n_reject per mode as deb_evolve.  prob->return_full is ignored.  d_step_log: NULL or
 * the attempted steps as deb_evolve (per mode) / deb_evolve_batched (per batch) log them.  d_mesh,
 * mesh_stride, d_mesh_len: NULL, or a frozen step mesh per mode as deb_evolve_replay takes it (per-mode
 * steps only; exclusive with the step log): the map that frozen-mesh derivatives differentiate.
 * Replaces the chain of compute_Cell_spectrum_from_cosmo_params (src/discoeb/cmb.py:1030-1073):
 * evolve_perturbations_batched(return_full=True) (perturbations.py:1229-1337),
 * compute_time_derivatives (perturbations.py:1339-1406), compute_source_function (cmb.py:17-633). */
int deb_evolve_sources(const double *d_tables, int num_thermo, int cosmo, const double *d_k, int n_k, int batch_size, const double *d_aexp,
                       const double *d_tau, int n_tau, const DebProblem *prob, const DebSolverCfg *cfg, int isw_mode, double *d_out, int *d_status,
                       int *d_n_accept, int *d_n_reject, double *d_work, double *d_step_log, const double *d_mesh, long mesh_stride,
                       const int *d_mesh_len, const DebSchedule *sched, void *stream);

/* The stages of the source function one at a time (the reference's per-stage functions).  Planes are
 * n_k x n_tau doubles, row-major by k; per-time arrays n_tau doubles.
 *
 * Massive-neutrino moments of the states (nvar entries each; momentum bins from iq0, nqmax of
 * them per multipole): d_out planes 0..5 = drhonu, dpnu, fnu, shearnu, rho_nu_prime,
 * shear_nu_prime; d_aprimeoa the conformal Hubble rate at d_aexp; d_work 32 doubles.
 * Replaces compute_neutrino_perturbations (src/discoeb/cmb.py:170-240). */
int deb_cmb_neutrinos(const double *d_tables, int num_thermo, int cosmo, const double *d_y, const double *d_yprime, int n_k,
                      const double *d_aexp, int n_tau, int nvar, int iq0, int nqmax, double *d_out, double *d_aprimeoa, double *d_work,
                      void *stream);

/* Einstein equations: d_out planes 0..5 = hprime, etaprime, alpha, alphaprime, alphapprime,
 * aprimeoaprime.  Replaces compute_metric_perturbations (src/discoeb/cmb.py:243-345). */
int deb_cmb_metric(const double *d_tables, int num_thermo, int cosmo, const DebMetricInputs *in, const double *d_k, int n_k,
                   const double *d_aexp, int n_tau, double *d_out, void *stream);

/* Polarisation combination F_2 + G_0 + G_2 and its first and second derivatives: d_out planes 0..2.
 * Replaces compute_polarization_terms (src/discoeb/cmb.py:349-395). */
int deb_cmb_polarization(const double *d_y, const double *d_yprime, int n_k, int n_tau, int nvar, int lmaxg, int lmaxgp, const double *d_k,
                         const double *d_thetagprime, const double *d_alphaprime, const double *d_opac, const double *d_opacprime,
                         double *d_out, void *stream);

/* One source term: term 2 Sachs-Wolfe (d_vis = g; reads deltag, alphaprime, thetabprime, pol,
 * polpprime), 3 Doppler (d_vis = g'; alpha, thetab, polprime), 4 polarisation (d_vis = g''; pol),
 * 5 the E-mode source (d_vis = g; pol, d_tau and tau0; no counterpart in the reference); pointers
 * a term does not read may be NULL.  Replaces compute_source_term_sachs_wolfe, _doppler and
 * _polarization (src/discoeb/cmb.py:455-563). */
int deb_cmb_source_term(int term, const double *d_k, int n_k, int n_tau, const double *d_vis, const double *d_deltag, const double *d_alpha,
                        const double *d_alphaprime, const double *d_thetab, const double *d_thetabprime, const double *d_pol,
                        const double *d_polprime, const double *d_polpprime, const double *d_tau, double tau0, double *d_out, void *stream);

/* Doubles of workspace deb_cmb_source_isw needs. */
long deb_cmb_isw_workspace_doubles(int n_k, int n_tau);

/* S1 = e^-tau_c d(eta + alpha')/dtau by a natural spline of the potential in tau (n_tau >= 3).
 * Replaces compute_source_term_isw (src/discoeb/cmb.py:401-452). */
int deb_cmb_source_isw(const double *d_tau, int n_tau, int n_k, const double *d_eta, const double *d_alphaprime, const double *d_optical_depth,
                       double *d_S1, double *d_work, void *stream);

/* Doubles of workspace deb_los_theta_ell needs. */
long deb_los_workspace_doubles(int n_k, int n_tau, int ellmax, int n_fftlog, int n_k_dense);

/* Theta_l(k), l = 0..ellmax, by FFTLog line-of-sight integration of the source d_S (n_k x n_tau)
 * on wavenumbers d_k and increasing conformal times d_tau, tau0 = conformal time today;
 * n_fftlog a power of two in [128, 16384]; n_k_dense > 0 first interpolates S onto that many
 * log-uniform wavenumbers.  d_theta: n_k_used x (ellmax + 1); d_k_out: the n_k_used wavenumbers.
 * Replaces compute_theta_ell (src/discoeb/cmb.py:680-835) with
 * _fftlog_bessel_integral_coefficients (:636-677). */
int deb_los_theta_ell(const double *d_k, int n_k, const double *d_tau, int n_tau, const double *d_S, double tau0, int ellmax, int n_fftlog,
                      int n_k_dense, double *d_theta, double *d_k_out, double *d_work, void *stream);

/* deb_los_theta_ell for field 0 (temperature, as deb_los_theta_ell) or 1 (E-mode polarisation:
 * Theta_E,l = sqrt((l + 2)! / (l - 2)!) int S_E j_l(k chi) dtau, zero for l < 2; no counterpart in
 * the reference, which computes TT only). */
int deb_los_theta_ell_field(const double *d_k, int n_k, const double *d_tau, int n_tau, const double *d_S, double tau0, int ellmax, int n_fftlog,
                            int n_k_dense, int field, double *d_theta, double *d_k_out, double *d_work, void *stream);

/* C = A B^T in FP64 on the matrix cores: A m x k, B n x k (row-major, k a multiple of 16, both 16-byte aligned), C
 * m x n.  The contraction of compute_theta_ell (src/discoeb/cmb.py:805-833, the einsums over
 * the FFTLog frequencies), exposed on its own. */
int deb_contract(const double *d_A, const double *d_B, int m, int n, long k, double *d_C, void *stream);

/* C_l = trapezoid over ln k of (k/k_p)^(n_s-1) Theta_l(k)^2: d_theta n_k x n_ell, d_cell n_ell.
 * Replaces compute_Cell (src/discoeb/cmb.py:838-873). */
int deb_cell(const double *d_theta, const double *d_k, int n_k, int n_ell, double n_s, double k_p, double *d_cell, void *stream);

/* Cross spectrum: trapezoid over ln k of (k/k_p)^(n_s-1) Theta1_l Theta2_l (C_l^TE from the
 * temperature and E-mode Theta; no counterpart in the reference). */
int deb_cell_cross(const double *d_theta1, const double *d_theta2, const double *d_k, int n_k, int n_ell, double n_s, double k_p, double *d_cell,
                   void *stream);

/* D_l = l (l + 1) C_l A_s 2 T_cmb^2 1e12 for l = 2..ellmax: d_ell, d_dell (ellmax - 1 each).
 * Replaces compute_Dell (src/discoeb/cmb.py:876-927). */
int deb_dell(const double *d_cell, int ellmax, double A_s, double Tcmb, double *d_ell, double *d_dell, void *stream);

/* ---------------------------------------------------------------------------------------------
 * Large-scale-structure observables on the solver's output.  d_y: n x ystride doubles (one output
 * time of evolve, the 20 output fields or the full state's first columns: delta_m at 4,
 * theta_m / aH at 5); A_s, n_s, k_p the primordial parameters.
 * ------------------------------------------------------------------------------------------- */

/* Savitzky-Golay taps (window_length >= polyorder + 1 >= 1, polyorder <= 8): d_coef, window_length
 * doubles.  Replaces the coefficient part of savgol_filter (src/discoeb/util.py:291-328). */
int deb_savgol_coefficients(int window_length, int polyorder, double *d_coef, void *stream);

/* d_out[i] = sum_m d_coef[m] d_y[i + (window_length - 1) / 2 - m] (zero beyond the ends), the
 * 'same'-length convolution of savgol_filter (src/discoeb/util.py:327); window_length <= n. */
int deb_savgol_filter(const double *d_y, int n, const double *d_coef, int window_length, double *d_out, void *stream);

/* Doubles of workspace deb_power_smoothed needs. */
long deb_power_smoothed_workspace_doubles(int n, int window_length);

/* P(k) of field idx smoothed in log P by a cubic Savitzky-Golay filter of window_length (odd,
 * 5..n); the first window_length / 2 and the last (window_length + 1) / 2 points keep the
 * unsmoothed P.  Replaces get_power_smoothed (src/discoeb/perturbations.py:1473-1505), whose
 * window length the caller computes from dlogk and the k spacing. */
int deb_power_smoothed(const double *d_k, const double *d_y, int n, int ystride, int idx, double A_s, double n_s, double k_p,
                       int window_length, double *d_out, double *d_work, void *stream);

/* P(k, mu) = (bias delta_m - mu^2 theta_m)^2 on n x nmu points, mu uniform on [-1, 1]
 * (mu_sampling 1) or cos of uniform angles on [0, pi] (0); d_mu (nmu) written.  With d_pdelta
 * and d_ptheta (smoothed spectra of fields 4 and 5) delta_m = sqrt(pdelta), theta_m =
 * -sqrt(ptheta); NULL: the scaled transfer functions.  Replaces power_Kaiser
 * (src/discoeb/perturbations.py:1507-1543). */
int deb_power_kaiser(const double *d_k, const double *d_y, int n, int ystride, const double *d_pdelta, const double *d_ptheta, double bias,
                     double A_s, double n_s, double k_p, int nmu, int mu_sampling, double *d_pkmu, double *d_mu, void *stream);

/* Monopole, quadrupole and hexadecapole of the Kaiser spectrum with linear bias b.  Replaces
 * power_multipoles (src/discoeb/perturbations.py:1547-1569). */
int deb_power_multipoles(const double *d_k, const double *d_y, int n, int ystride, double b, double A_s, double n_s, double k_p, double *d_p0,
                         double *d_p2, double *d_p4, void *stream);

/* Doubles of workspace deb_xi_from_power needs. */
long deb_xi_workspace_doubles(int n);

/* The correlation-function multipole xi_l(r) of P(k) on the log-uniform grid d_k (n even, >= 4) by
 * FFTLog: d_xi and d_r (n each, increasing r = 2 pi / k).  Replaces get_xi_from_P
 * (src/discoeb/perturbations.py:1409-1443). */
int deb_xi_from_power(const double *d_k, const double *d_pk, int n, int ell, double *d_xi, double *d_r, double *d_work, void *stream);

#ifdef __cplusplus
}
#endif
#endif
