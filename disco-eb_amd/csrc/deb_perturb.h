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

/* How units are spread over workgroups (one wavefront each).  persistent = 0: one workgroup
 * per unit.  persistent = 1: n_workgroups resident workgroups (0: as many as fit on the
 * device) claim `chunk` units at a time from a device counter, in the order d_order
 * (device array of n_units unit indices, NULL: 0..n_units-1). */
typedef struct DebSchedule {
  int persistent, chunk, n_workgroups, reserved;
  const int *d_order;
} DebSchedule;
#ifdef __cplusplus
}
#endif
#endif
