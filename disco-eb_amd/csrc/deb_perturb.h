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

/* Launch order of the units: one workgroup (one wavefront) per unit, workgroup g solving unit
 * d_order[g] (device array of n_units distinct unit indices; NULL: 0..n_units-1).  Workgroups
 * start in index order as compute slots free up, so listing the most expensive units first
 * (largest k) shortens the tail of a launch with more units than resident slots. */
typedef struct DebSchedule {
  const int *d_order;
} DebSchedule;
#ifdef __cplusplus
}
#endif
#endif
