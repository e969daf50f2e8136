/* deb_layout.h -- layout of one cosmology's background tables in HBM (doubles).
 *
 * One contiguous block per cosmology, stride deb_table_doubles(N) doubles, N = num_thermo.
 * Scalars first (DEB_HDR_* slots), then the 512-knot neutrino tables (uniform in log a),
 * then the N-knot thermal-history arrays and the second derivatives of every spline the
 * reference builds in evolve_background (background.py:248-256, 320).  Spline y-arrays alias
 * the data arrays (e.g. tau_of_a uses x = aexp, y = tau).
 */
#ifndef DEB_LAYOUT_H
#define DEB_LAYOUT_H

/* offsets are computed identically by host code and kernels */
#if defined(__HIPCC__)
#define DEB_HOST_DEVICE __host__ __device__
#else
#define DEB_HOST_DEVICE
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* cosmology parameter vector (input), same order as the reference's param keys */
enum {
  DEB_P_OMEGAM, DEB_P_OMEGAB, DEB_P_H0, DEB_P_TCMB, DEB_P_YHE, DEB_P_NEFF, DEB_P_NMNU, DEB_P_MNU, DEB_P_W0, DEB_P_WA,
  DEB_P_CS2DE, DEB_P_OMEGAK, DEB_P_AS, DEB_P_NS, DEB_P_KP,
  DEB_P_USE_TCA, DEB_P_TCA_H, DEB_P_TCA_K, DEB_P_USE_RSA, DEB_P_RSA_C, DEB_P_RSA_K, DEB_P_USE_UFA, DEB_P_UFA_K,
  DEB_P_TAU_REION, /* optical depth of reionization; <= 0: none (the reference's thermal history has none) */
  DEB_NPARAM
};

/* header scalars: slots 0..DEB_NPARAM-1 repeat the input parameters */
enum {
  DEB_HDR_GRHOM = 24, DEB_HDR_GRHOG, DEB_HDR_GRHOR, DEB_HDR_ADOTRAD, DEB_HDR_AMNU, DEB_HDR_OMEGAMNU, DEB_HDR_OMEGADE,
  DEB_HDR_TAUMIN, DEB_HDR_TAUMAX, DEB_HDR_FHE, DEB_HDR_TAU0, DEB_HDR_TAU_MAXVIS, DEB_HDR_OPTDEPTH_TODAY, DEB_HDR_AKTHOM,
  DEB_HDR_LAMBDA_DE, DEB_HDR_N, DEB_HDR_RECFAST_FAIL, DEB_HDR_NU_X0, DEB_HDR_NU_DXINV, DEB_HDR_A_START, DEB_HDR_AMIN,
  DEB_HDR_AMAX, DEB_HDR_THERMO, DEB_HDR_Z_REION, DEB_HDR_TAU_REION_OUT,
  DEB_HDR_SIZE = 64
};

#define DEB_NNU 512

/* array slots, in order; each has N entries unless listed in the DEB_NU_ group (512) */
enum {
  DEB_A_NU_LOGA, DEB_A_NU_LOGRHO, DEB_A_NU_LOGRHO_S, DEB_A_NU_LOGP, DEB_A_NU_LOGP_S, DEB_A_NU_LOGPP, DEB_A_NU_LOGPP_S,
  DEB_A_NU_COUNT,
  DEB_A_AEXP = DEB_A_NU_COUNT, DEB_A_LOGA, DEB_A_TAU, DEB_A_XE, DEB_A_XEHI, DEB_A_XEHEI, DEB_A_XEHEII, DEB_A_TM, DEB_A_MU,
  DEB_A_CS2, DEB_A_CS2A, DEB_A_TEMPBA, DEB_A_XEPRIME_RECF, DEB_A_XE_VIS, DEB_A_XEPRIME, DEB_A_OPAC, DEB_A_OPTDEPTH,
  DEB_A_GVIS, DEB_A_GVISP, DEB_A_GVISPP,
  /* spline second derivatives */
  DEB_A_S_TAU_OF_A, DEB_A_S_A_OF_TAU, DEB_A_S_XE_OF_TAU, DEB_A_S_CS2A_OF_TAU, DEB_A_S_TEMPBA_OF_TAU, DEB_A_S_XE_OF_LOGA,
  DEB_A_S_CS2A_OF_LOGA, DEB_A_S_OPAC, DEB_A_I_OPAC,
  /* scratch for the background kernel (recfast derivative rows, Thomas workspace) */
  DEB_A_SCR0, DEB_A_SCR1, DEB_A_SCR2, DEB_A_SCR3, DEB_A_SCR4, DEB_A_SCR5, DEB_A_SCR6, DEB_A_SCR7, DEB_A_SCR8,
  DEB_A_COUNT
};

static inline DEB_HOST_DEVICE long deb_array_offset(int slot, int N) {
  if (slot < DEB_A_NU_COUNT) return DEB_HDR_SIZE + (long)slot * DEB_NNU;
  return DEB_HDR_SIZE + (long)DEB_A_NU_COUNT * DEB_NNU + (long)(slot - DEB_A_NU_COUNT) * (N > DEB_NNU ? N : DEB_NNU);
}
static inline DEB_HOST_DEVICE long deb_table_doubles(int N) { return deb_array_offset(DEB_A_COUNT, N); }

#ifdef __cplusplus
}
#endif
#endif
