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
#ifdef __cplusplus
}
#endif
#endif
