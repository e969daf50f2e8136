// deb_cosmo.hpp -- read-only view of one cosmology's background tables and the background
// functions every kernel needs: H(a) (background.py:101-123) and dtau/da (:72-81).
#pragma once
#include "deb_device.hpp"

namespace deb {

struct BgView {
  const double *t;  // table base of this cosmology
  int N;
  double Omegam, Omegab, H0, Tcmb, YHe, Neff, Nmnu, mnu, w0, wa, cs2DE, Omegak, A_s, n_s, k_p;
  bool use_tca, use_rsa, use_ufa, lambda_de;
  double tca_h, tca_k, rsa_c, rsa_k, ufa_k;
  double grhom, grhog, grhor, adotrad, amnu, OmegaDE, taumin, fHe, akthom;
  double nu_x0, nu_dxinv;
  const double *nu_loga, *nu_lr, *nu_lrS, *nu_lp, *nu_lpS;
  const double *aexp, *loga, *tau, *xe, *cs2a;
  const double *S_tau_of_a, *S_a_of_tau, *S_xe_of_tau, *S_xe_of_loga, *S_cs2a_of_loga;

  __device__ void init(const double *base) {
    t = base;
    N = (int)base[DEB_HDR_N];
    Omegam = t[DEB_P_OMEGAM]; Omegab = t[DEB_P_OMEGAB]; H0 = t[DEB_P_H0]; Tcmb = t[DEB_P_TCMB]; YHe = t[DEB_P_YHE];
    Neff = t[DEB_P_NEFF]; Nmnu = t[DEB_P_NMNU]; mnu = t[DEB_P_MNU]; w0 = t[DEB_P_W0]; wa = t[DEB_P_WA];
    cs2DE = t[DEB_P_CS2DE]; Omegak = t[DEB_P_OMEGAK]; A_s = t[DEB_P_AS]; n_s = t[DEB_P_NS]; k_p = t[DEB_P_KP];
    use_tca = t[DEB_P_USE_TCA] != 0.0; tca_h = t[DEB_P_TCA_H]; tca_k = t[DEB_P_TCA_K];
    use_rsa = t[DEB_P_USE_RSA] != 0.0; rsa_c = t[DEB_P_RSA_C]; rsa_k = t[DEB_P_RSA_K];
    use_ufa = t[DEB_P_USE_UFA] != 0.0; ufa_k = t[DEB_P_UFA_K];
    grhom = t[DEB_HDR_GRHOM]; grhog = t[DEB_HDR_GRHOG]; grhor = t[DEB_HDR_GRHOR]; adotrad = t[DEB_HDR_ADOTRAD];
    amnu = t[DEB_HDR_AMNU]; OmegaDE = t[DEB_HDR_OMEGADE]; taumin = t[DEB_HDR_TAUMIN]; fHe = t[DEB_HDR_FHE];
    akthom = t[DEB_HDR_AKTHOM]; lambda_de = t[DEB_HDR_LAMBDA_DE] != 0.0;
    nu_x0 = t[DEB_HDR_NU_X0]; nu_dxinv = t[DEB_HDR_NU_DXINV];
    nu_loga = t + deb_array_offset(DEB_A_NU_LOGA, N);
    nu_lr = t + deb_array_offset(DEB_A_NU_LOGRHO, N);
    nu_lrS = t + deb_array_offset(DEB_A_NU_LOGRHO_S, N);
    nu_lp = t + deb_array_offset(DEB_A_NU_LOGP, N);
    nu_lpS = t + deb_array_offset(DEB_A_NU_LOGP_S, N);
    aexp = t + deb_array_offset(DEB_A_AEXP, N);
    loga = t + deb_array_offset(DEB_A_LOGA, N);
    tau = t + deb_array_offset(DEB_A_TAU, N);
    xe = t + deb_array_offset(DEB_A_XE, N);
    cs2a = t + deb_array_offset(DEB_A_CS2A, N);
    S_tau_of_a = t + deb_array_offset(DEB_A_S_TAU_OF_A, N);
    S_a_of_tau = t + deb_array_offset(DEB_A_S_A_OF_TAU, N);
    S_xe_of_tau = t + deb_array_offset(DEB_A_S_XE_OF_TAU, N);
    S_xe_of_loga = t + deb_array_offset(DEB_A_S_XE_OF_LOGA, N);
    S_cs2a_of_loga = t + deb_array_offset(DEB_A_S_CS2A_OF_LOGA, N);
  }

  // log rho_nu / rho_nu0 of one massive flavour (uniform spline in log a)
  template <class T> __device__ T logrhonu(const T &la) const {
    int i = spl_index_uniform(nu_x0, nu_dxinv, DEB_NNU, val(la));
    return spl_eval_at(nu_loga, nu_lr, nu_lrS, i, la);
  }
  template <class T> __device__ T logpnu(const T &la) const {
    int i = spl_index_uniform(nu_x0, nu_dxinv, DEB_NNU, val(la));
    return spl_eval_at(nu_loga, nu_lp, nu_lpS, i, la);
  }
  // dark-energy density rho_Q(a)/rho_Q(1) for w(a) = w0 + wa (1 - a)
  template <class T> __device__ T rho_Q(const T &a) const {
    return dpow(a, -3.0 * (1.0 + w0 + wa)) * dexp(3.0 * (a - 1.0) * wa);
  }
  // conformal Hubble rate a'/a
  template <class T> __device__ T aprimeoa(const T &a) const {
    T rhonu = dexp(logrhonu(dlog(a)));
    T grho = grhom * Omegam / a + (grhog + grhor * (Neff + Nmnu * rhonu)) / (a * a) + grhom * OmegaDE * rho_Q(a) * a * a +
             grhom * Omegak;
    return dsqrt(grho / 3.0);
  }
  __device__ double dtauda(double a) const {
    double rhonu = exp(logrhonu(log(a)));
    double rho_DE = pow(a, -3.0 * (1.0 + w0 + wa)) * exp(3.0 * (a - 1.0) * wa);
    double grho2 = grhom * Omegam * a + (grhog + grhor * (Neff + Nmnu * rhonu)) + grhom * OmegaDE * rho_DE * a * a * a * a +
                   grhom * Omegak * a * a;
    return sqrt(3.0 / grho2);
  }
  // non-uniform splines used by the perturbation kernel
  __device__ double tau_of_a(double a) const {
    int i = spl_index(aexp, N, a);
    return spl_eval_at(aexp, tau, S_tau_of_a, i, a);
  }
  __device__ double a_of_tau(double tt) const {
    int i = spl_index(tau, N, tt);
    return spl_eval_at(tau, aexp, S_a_of_tau, i, tt);
  }
  __device__ double xe_of_tau(double tt) const {
    int i = spl_index(tau, N, tt);
    return spl_eval_at(tau, xe, S_xe_of_tau, i, tt);
  }
};

}  // namespace deb
