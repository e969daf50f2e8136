// deb_source_pt.hpp -- the CMB source algebra at one (k, tau) point, shared by the source kernels
// (deb_cmb.hip: the fused point kernel and the reference's stages one at a time) and by the
// perturbation kernels' source epilogue (deb_perturb_dev.hpp), so that every path evaluates the
// same expressions in the same order.
//
// Follows the reference's src/discoeb/cmb.py: compute_neutrino_perturbations (:169-240),
// compute_metric_perturbations (:243-345), compute_polarization_terms (:349-395) and the source
// terms S1..S4 (:398-563).
#ifndef DEB_SOURCE_PT_HPP
#define DEB_SOURCE_PT_HPP
#include <hip/hip_runtime.h>

#include "deb_cosmo.hpp"

namespace deb {

// rows of the per-output-time table written by vis_eval_kernel (each n_tau long):
// 0 opac, 1 opac', 2 g, 3 g', 4 g'', 5 tau_c, 6 aH, 7 rho_nu, 8 rho_Q, 9 w_Q
constexpr int NV = 10;

// massive neutrinos (perturbations.py:29-85, cmb.py:170-240): momentum integrals of the state and
// its derivative at one point; Y(i), P(i) row i of the state and of its time derivative, qw the bins
struct NuPt {
  double drhonu, dpnu, fnu, shearnu, rnp, snp;
};
template <class YA, class PA>
__device__ inline NuPt nu_point(const YA &Y, const PA &P, int iq0, int nq, const double *qw, double a, double amnu, double aH) {
  NuPt r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < nq; ++q) {
    const double qq = qw[q], wq = qw[16 + q];
    const double aq = a * amnu / qq;
    const double v = 1.0 / sqrt(1.0 + aq * aq);
    const double vp = -0.5 * aq * aH * v * v * v;
    const double p0 = Y(iq0 + q), p1 = Y(iq0 + nq + q), p2 = Y(iq0 + 2 * nq + q);
    const double p0p = P(iq0 + q), p2p = P(iq0 + 2 * nq + q);
    r.drhonu += wq * p0 / v;
    r.dpnu += wq * p0 * v;
    r.fnu += wq * p1;
    r.shearnu += wq * p2 * v;
    r.rnp += wq * (p0p / v - p0 / (v * v) * vp);
    r.snp += wq * (p2p * v + p2 * vp);
  }
  r.dpnu = r.dpnu / 3.0;
  r.shearnu = r.shearnu * 2.0 / 3.0;
  r.snp = r.snp * 2.0 / 3.0;
  return r;
}

// the perturbations the Einstein equations read at one point (cmb.py:243-345)
struct PertPt {
  double eta, deltac, thetac, deltab, thetab, deltag, thetag, shearg, sheargp, deltar, thetar, shearr, shearrp, deltaq, thetaq;
};
struct MetricPt {
  double hprime, etaprime, alpha, alphap, alphapp, aHp;
};
__device__ inline MetricPt metric_point(const BgView &b, double k, double a, double aH, double rhonu, double rho_Q, double w_Q, const PertPt &x,
                                        double drhonu, double fnu, double shearnu, double rnp, double snp) {
  const double k2 = k * k, a2 = a * a;
  const double Omegac = b.Omegam - b.Omegab;
  const double rpptq = (1.0 + w_Q) * rho_Q * b.grhom * b.OmegaDE * x.thetaq * a2;
  const double grhoprime =
      (-b.grhom * b.Omegam / a - 2.0 * (b.grhog + b.grhor * (b.Neff + b.Nmnu * rhonu)) / a2 +
       b.grhom * b.OmegaDE * (-3.0 * (1.0 + b.w0 + (1.0 - a) * b.wa) * rho_Q * a2 + 2.0 * rho_Q * a2)) *
          aH +
      b.grhor * b.Nmnu * rnp / a2;
  MetricPt r;
  r.aHp = grhoprime / (6.0 * aH);
  const double dgrho = b.grhom * (Omegac * x.deltac + b.Omegab * x.deltab) / a +
                       (b.grhog * x.deltag + b.grhor * (b.Neff * x.deltar + b.Nmnu * drhonu)) / a2 +
                       b.grhom * b.OmegaDE * x.deltaq * rho_Q * a2;
  const double dgtheta = b.grhom * (Omegac * x.thetac + b.Omegab * x.thetab) / a +
                         4.0 / 3.0 * (b.grhog * x.thetag + b.Neff * b.grhor * x.thetar) / a2 + b.Nmnu * b.grhor * k * fnu / a2 + rpptq;
  const double dgshear = 4.0 / 3.0 * (b.grhog * x.shearg + b.Neff * b.grhor * x.shearr) / a2 + b.Nmnu * b.grhor * shearnu / a2;
  const double dgshearp =
      4.0 / 3.0 * (b.grhog * x.sheargp + b.Neff * b.grhor * x.shearrp) / a2 + b.Nmnu * b.grhor * snp / a2 - 2.0 * aH * dgshear;
  r.hprime = (2.0 * k2 * x.eta + dgrho) / aH;
  r.etaprime = 0.5 * dgtheta / k2;
  r.alpha = (r.hprime + 6.0 * r.etaprime) / 2.0 / k2;
  r.alphap = -3.0 * dgshear / (2.0 * k2) + x.eta - 2.0 * aH * r.alpha;
  r.alphapp = -3.0 * dgshearp / (2.0 * k2) + r.etaprime - 2.0 * (r.aHp * r.alpha + aH * r.alphap);
  return r;
}

// polarisation terms (cmb.py:349-395): y2 = (F_2, G_0, G_2), p2 their derivatives, p3 = (F_3', G_1', G_3')
struct PolPt {
  double pol, polp, polpp;
};
__device__ inline PolPt pol_point(double k, double thetagp, double alphap, double opac, double opacp, const double *y2, const double *p2,
                                  const double *p3) {
  const double k2 = k * k;
  const double couplp = 8.0 * (thetagp + k2 * alphap) / 15.0 - k * 0.6 * (p3[0] + p3[1] + p3[2]);
  PolPt r;
  r.pol = y2[0] + y2[1] + y2[2];
  r.polp = p2[0] + p2[1] + p2[2];
  r.polpp = couplp - 0.3 * (opacp * r.pol + opac * r.polp);
  return r;
}

// S2, S3, S4 (cmb.py:455-563)
__device__ inline double s2_point(double k, double g, double deltag, double alphap, double pol, double thetabp, double polpp) {
  return g * (0.25 * deltag + 2.0 * alphap + pol / 16.0 + (thetabp + 3.0 / 16.0 * polpp) / (k * k));
}
__device__ inline double s3_point(double k, double g1, double alpha, double thetab, double polp) {
  return g1 * (alpha + (thetab + 3.0 / 8.0 * polp) / (k * k));
}
__device__ inline double s4_point(double k, double g2, double pol) { return g2 * 3.0 / 16.0 * pol / (k * k); }
// E-mode source (no counterpart in the reference; Zaldarriaga & Seljak 1997 in this normalisation,
// Pi = F_2 + G_0 + G_2 = 4 x the brightness multipoles): S_E = 3 g Pi / (16 (k chi)^2), chi = tau0 - tau,
// zero within 1 Mpc of today (where 1 / chi^2 diverges and g is ~1e-10)
__device__ inline double se_point(double k, double g, double pol, double chi) {
  return chi >= 1.0 ? g * 3.0 / 16.0 * pol / (k * k * chi * chi) : 0.0;
}


// every output of the fused source chain at one point: Y(i) row i of the (RSA-projected) state,
// P(i) row i of its time derivative (the right-hand side at Y), a the output scale factor, V the visibility table at
// this output time (V[r * vstride], rows as NV above), chi = tau0 - tau; qw: the momentum bins
// (nodes at [q], weights at [16 + q])
struct SrcPt {
  double S1, S2, S3, S4, alpha, alphap, SE, pot;  // S1: the algebraic form e^-tau_c (eta' + alpha''); pot: eta + alpha'
};
template <class YA, class PA>
__device__ inline SrcPt source_point(const BgView &b, const YA &Y, const PA &P, int n, int idxg, int idxgp, int idxr, int iq0, int nq,
                                     const double *qw, double k, double a, const double *V, long vstride, double chi) {
  const double opac = V[0 * vstride], opacp = V[1 * vstride], g = V[2 * vstride], g1 = V[3 * vstride], g2 = V[4 * vstride];
  const double aH = V[6 * vstride], rhonu = V[7 * vstride], rho_Q = V[8 * vstride], w_Q = V[9 * vstride];
  const NuPt nu = nu_point(Y, P, iq0, nq, qw, a, b.amnu, aH);
  const PertPt x{Y(2),           Y(3),           Y(4),           Y(5),     Y(6),     Y(idxg),  Y(idxg + 1), Y(idxg + 2) / 2.0,
                 P(idxg + 2) / 2.0, Y(idxr), Y(idxr + 1), Y(idxr + 2) / 2.0, P(idxr + 2) / 2.0, Y(n - 2), Y(n - 1)};
  const MetricPt mt = metric_point(b, k, a, aH, rhonu, rho_Q, w_Q, x, nu.drhonu, nu.fnu, nu.shearnu, nu.rnp, nu.snp);
  const double y2[3] = {Y(idxg + 2), Y(idxgp), Y(idxgp + 2)}, p2[3] = {P(idxg + 2), P(idxgp), P(idxgp + 2)};
  const double p3[3] = {P(idxg + 3), P(idxgp + 1), P(idxgp + 3)};
  const PolPt pl = pol_point(k, P(idxg + 1), mt.alphap, opac, opacp, y2, p2, p3);
  SrcPt r;
  r.S2 = s2_point(k, g, x.deltag, mt.alphap, pl.pol, P(6), pl.polpp);
  r.S3 = s3_point(k, g1, mt.alpha, x.thetab, pl.polp);
  r.S4 = s4_point(k, g2, pl.pol);
  r.alpha = mt.alpha;
  r.alphap = mt.alphap;
  r.SE = se_point(k, g, pl.pol, chi);
  r.S1 = exp(-V[5 * vstride]) * (mt.etaprime + mt.alphapp);
  r.pot = x.eta + mt.alphap;
  return r;
}

}  // namespace deb
#endif
