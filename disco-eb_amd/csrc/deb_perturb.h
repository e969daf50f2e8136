/* deb_perturb.h -- problem and solver descriptors passed by value to the perturbation kernel */
#ifndef DEB_PERTURB_H
#define DEB_PERTURB_H
#ifdef __cplusplus
extern "C" {
#endif
typedef struct DebProblem {
  int lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax; /* hierarchy truncations, momentum bins */
  int return_full;                         /* 1: projected state (nvar), 0: 20 output fields */
} DebProblem;

typedef struct DebSolverCfg {
  double rtol, atol, pcoeff, icoeff, dcoeff, factormax, factormin;
  int max_steps;
} DebSolverCfg;

/* Threads of the workgroup that solves one unit when a schedule does not say (see below). */
#define DEB_DEFAULT_THREADS_PER_UNIT 256

/* Launch order and width: one workgroup per unit, workgroup g solving unit d_order[g] (device
 * array of n_units distinct unit indices; NULL: 0..n_units-1).  Workgroups start in index order
 * as compute slots free up, so listing the most expensive units first (largest k) shortens the
 * tail of a launch with more units than resident slots.  threads_per_unit: 64, 128 or 256
 * threads (1, 2 or 4 wavefronts) share one unit's rows, matrix entries and Jacobian columns,
 * shortening each step at the cost of fewer units in flight; 0 = the library default. */
typedef struct DebSchedule {
  const int *d_order;
  int threads_per_unit;
} DebSchedule;

/* Inputs of the Einstein-equation stage of the CMB sources (compute_metric_perturbations): planes
 * n_k x n_tau (row-major by k) of the perturbations and neutrino moments, and n_tau values of the
 * background at the output times. */
typedef struct DebMetricInputs {
  const double *eta, *deltac, *thetac, *deltab, *thetab, *deltag, *thetag, *shearg, *sheargp, *deltar, *thetar, *shearr, *shearrp,
      *deltaq, *thetaq;
  const double *drhonu, *fnu, *shearnu, *rho_nu_prime, *shear_nu_prime;
  const double *aH, *rhonu, *rho_Q, *w_Q;
} DebMetricInputs;
#ifdef __cplusplus
}
#endif
#endif
