// deb_cmb.hip -- CMB source functions and the FFTLog line-of-sight projection.
//
// Follows the reference's src/discoeb/cmb.py: compute_visibility_functions (:92-166),
// compute_neutrino_perturbations (:169-240), compute_metric_perturbations (:243-345),
// compute_polarization_terms (:349-395), the source terms S1..S4 (:398-633),
// _fftlog_bessel_integral_coefficients (:636-677) with util.py:13-45, compute_theta_ell
// (:680-835), compute_Cell (:838-873) and compute_Dell (:876-927).
//
// Kernels:
//   * vis_build / vis_eval: natural splines in log tau of the thermal tables (one thread per
//     spline), evaluated with the background quantities at every output time;
//   * source_point: one thread per (k, tau) -- neutrino moments, metric, polarisation terms,
//     S2, S3, S4 and the ISW potential eta + alpha';
//   * source_isw: one thread per k -- natural spline of the potential in tau, its derivative
//     at the knots times exp(-tau_c), and the total S;
//   * densify_build / densify_eval: S interpolated in log k onto the dense log-uniform grid;
//   * fftlog: one workgroup per k-row -- the row's spline in log chi, the biased samples on the
//     FFTLog grid written bit-reversed into LDS, an in-place radix-2 FFT of N/2 complex points
//     (N/2 * 16 bytes of LDS), the real-FFT split and the phase / k-power epilogue, written as
//     the K-contiguous left operand of the contraction;
//   * bessel_coef: M_l(p_n) by the Lanczos complex log-gamma, the right operand;
//   * contract: Theta = A B^T by FP64 MFMA (v_mfma_f64_16x16x4f64), 128 x 128 output tiles;
//   * cell / dell: trapezoid over ln k, and the D_l scaling.
#include <hip/hip_runtime.h>

#include "deb_complex.hpp"
#include "deb_cosmo.hpp"
#include "deb_perturb.h"
#include "deb_source_pt.hpp"

namespace deb {
namespace {
using namespace cplx;

constexpr double kPi = 3.141592653589793;

// ---------------------------------------------------------------------------------------------
// natural cubic splines (spline_interpolation.py:41-82): cp = modified upper diagonal of the
// interior system, a function of the knots only
// ---------------------------------------------------------------------------------------------
__device__ void spline_cp(const double *__restrict__ x, int n, double *__restrict__ cp) {
  const int m = n - 2;
  for (int i = 0; i < m; ++i) {
    double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    double b = 2.0 * (h0 + h1);
    double den = (i == 0) ? b : b - h0 * cp[i - 1];
    cp[i] = (i < m - 1) ? h1 / den : 0.0;
  }
}

// second derivatives of the natural spline through (x[i], y[i * ys]) into S[i * ss]; the
// forward sweep keeps dp in S, the back substitution overwrites it
__device__ void spline_build(const double *__restrict__ x, const double *__restrict__ y, long ys, int n, const double *__restrict__ cp,
                             double *__restrict__ S, long ss) {
  S[0] = 0.0;
  if (n > 1) S[(long)(n - 1) * ss] = 0.0;
  const int m = n - 2;
  if (m <= 0) return;
  double dprev = 0.0;
  for (int i = 0; i < m; ++i) {
    double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
    double y0 = y[(long)i * ys], y1 = y[(long)(i + 1) * ys], y2 = y[(long)(i + 2) * ys];
    double d = 6.0 * ((y2 - y1) / h1 - (y1 - y0) / h0);
    double b = 2.0 * (h0 + h1);
    double dp = (i == 0) ? d / b : (d - h0 * dprev) / (b - h0 * cp[i - 1]);
    S[(long)(i + 1) * ss] = dp;
    dprev = dp;
  }
  double nxt = S[(long)m * ss];
  for (int i = m - 2; i >= 0; --i) {
    double v = S[(long)(i + 1) * ss] - cp[i] * nxt;
    S[(long)(i + 1) * ss] = v;
    nxt = v;
  }
}

__device__ __forceinline__ double spl_value(const double *x, const double *y, long ys, const double *S, long ss, int i, double xq) {
  double h = x[i + 1] - x[i];
  double t = (xq - x[i]) / h;
  double A = 1.0 - t, B = t;
  return A * y[(long)i * ys] + B * y[(long)(i + 1) * ys] +
         ((A * A * A - A) * S[(long)i * ss] + (B * B * B - B) * S[(long)(i + 1) * ss]) * (h * h / 6.0);
}

__device__ __forceinline__ double spl_slope(const double *x, const double *y, long ys, const double *S, long ss, int i, double xq) {
  double h = x[i + 1] - x[i];
  double d = xq - x[i];
  double S0 = S[(long)i * ss], S1 = S[(long)(i + 1) * ss];
  double b = (y[(long)(i + 1) * ys] - y[(long)i * ys]) / h - h * (S1 + 2.0 * S0) / 6.0;
  return b + 2.0 * (S0 / 2.0) * d + 3.0 * ((S1 - S0) / (6.0 * h)) * d * d;
}

// ---------------------------------------------------------------------------------------------
// visibility functions and background quantities at the output times
// ---------------------------------------------------------------------------------------------
// workspace layout (doubles): [0, N) log tau knots, [N, 2N) cp, then for spline s = 0..4
// (opac, gvis, gvis', gvis'', log max(tau_c, 1e-10)) its values at [(2 + 2 s) N, ...) and
// second derivatives at [(3 + 2 s) N, ...)
__global__ void vis_build_kernel(const double *__restrict__ t, double *__restrict__ w) {
  const int N = (int)t[DEB_HDR_N];
  if (threadIdx.x == 0) {
    const double *tau = t + deb_array_offset(DEB_A_TAU, N);
    for (int i = 0; i < N; ++i) w[i] = log(tau[i]);
    spline_cp(w, N, w + N);
  }
  __syncthreads();
  const int s = threadIdx.x;
  if (s >= 5) return;
  const int slot[5] = {DEB_A_OPAC, DEB_A_GVIS, DEB_A_GVISP, DEB_A_GVISPP, DEB_A_OPTDEPTH};
  const double *src = t + deb_array_offset(slot[s], N);
  double *Y = w + (long)(2 + 2 * s) * N, *S = w + (long)(3 + 2 * s) * N;
  double last = src[0];
  for (int i = 0; i < N; ++i) {  // NaN entries take the previous value (spline_interpolation.py:32-40)
    double v = src[i];
    if (s == 4 && !isnan(v)) v = log(v > 1e-10 ? v : 1e-10);
    if (isnan(v)) v = last;
    Y[i] = v;
    last = v;
  }
  spline_build(w, Y, 1, N, w + N, S, 1);
}

// per output time j: [0] opac [1] opac' [2] g [3] g' [4] g'' [5] tau_c [6] aH [7] rho_nu
// [8] rho_Q [9] w_Q, each a row of n_tau (NV rows)
__global__ void vis_eval_kernel(const double *__restrict__ t, const double *__restrict__ w, const double *__restrict__ tau_q,
                                const double *__restrict__ aexp, int n_tau, double *__restrict__ V) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_tau) return;
  BgView b;
  b.init(t);
  const int N = b.N;
  const double lq = log(tau_q[j]);
  const int i = spl_index(w, N, lq);
  double v[6];
  for (int s = 0; s < 5; ++s) v[s < 1 ? 0 : s + 1] = spl_value(w, w + (long)(2 + 2 * s) * N, 1, w + (long)(3 + 2 * s) * N, 1, i, lq);
  // the loop filled v[0] (opac), v[2..5]; reorder the log optical depth into place
  double od = exp(v[5]);
  double g = v[2], g1 = v[3], g2 = v[4];
  double opacp = spl_slope(w, w + 2L * N, 1, w + 3L * N, 1, i, lq) / tau_q[j];
  const double a = aexp[j];
  V[0L * n_tau + j] = v[0];
  V[1L * n_tau + j] = opacp;
  V[2L * n_tau + j] = g;
  V[3L * n_tau + j] = g1;
  V[4L * n_tau + j] = g2;
  V[5L * n_tau + j] = od;
  V[6L * n_tau + j] = b.aprimeoa(a);
  V[7L * n_tau + j] = exp(b.logrhonu(log(a)));
  V[8L * n_tau + j] = b.rho_Q(a);
  V[9L * n_tau + j] = b.w0 + b.wa * (1.0 - a);
}

// ---------------------------------------------------------------------------------------------
// source terms: the point functions (deb_source_pt.hpp) are shared by the fused kernel (every
// stage of one point in one thread), the stage kernels (the reference's compute_* functions one
// at a time) and the perturbation kernels' source epilogue
// ---------------------------------------------------------------------------------------------
// outputs (each n_k x n_tau, row-major by k): 0 S1, 1 S2, 2 S3, 3 S4, 4 S, 5 alpha, 6 alpha', 7 S_E;
// pot: n_tau x n_k (transposed) potential eta + alpha' for the ISW spline
__global__ void source_point_kernel(const double *__restrict__ t, const double *__restrict__ y, const double *__restrict__ yp,
                                    const double *__restrict__ kk, int n_k, int n_tau, const double *__restrict__ aexp,
                                    const double *__restrict__ V, const double *__restrict__ qw, DebProblem prob, int isw,
                                    const double *__restrict__ tau, double *__restrict__ out, double *__restrict__ pot) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau), ik = (int)(idx / n_tau);
  BgView b;
  b.init(t);
  const int lmaxg = prob.lmaxg, lmaxgp = prob.lmaxgp, lmaxr = prob.lmaxr, nq = prob.nqmax;
  const int n = 7 + (lmaxg + 1) + (lmaxgp + 1) + (lmaxr + 1) + nq * (prob.lmaxnu + 1) + 2;
  const int idxg = 7, idxgp = 8 + lmaxg, idxr = 9 + lmaxg + lmaxgp, iq0 = 10 + lmaxg + lmaxgp + lmaxr;
  const double *ys = y + idx * n, *yps = yp + idx * n;
  const SrcPt s = source_point(b, [&](int i) { return ys[i]; }, [&](int i) { return yps[i]; }, n, idxg, idxgp, idxr, iq0, nq, qw, kk[ik],
                               aexp[j], V + j, n_tau, t[DEB_HDR_TAU0] - tau[j]);
  const long plane = (long)n_k * n_tau;
  out[1 * plane + idx] = s.S2;
  out[2 * plane + idx] = s.S3;
  out[3 * plane + idx] = s.S4;
  out[5 * plane + idx] = s.alpha;
  out[6 * plane + idx] = s.alphap;
  out[7 * plane + idx] = s.SE;
  if (isw == 1) {  // S1 = e^-tau_c (eta' + alpha''), the algebraic form (cmb.py:406-408)
    out[0 * plane + idx] = s.S1;
    out[4 * plane + idx] = s.S1 + s.S2 + s.S3 + s.S4;
  } else {
    pot[(long)j * n_k + ik] = s.pot;
  }
}

// compute_neutrino_perturbations (cmb.py:170-240) on its own: planes 0..5 (n_k x n_tau) drhonu,
// dpnu, fnu, shearnu, rho_nu_prime, shear_nu_prime; aH (n_tau) the conformal Hubble rate
__global__ void nu_stage_kernel(const double *__restrict__ t, const double *__restrict__ y, const double *__restrict__ yp, int n_k, int n_tau,
                                const double *__restrict__ aexp, const double *__restrict__ qw, int n, int iq0, int nq, double *__restrict__ out,
                                double *__restrict__ aHo) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau);
  BgView b;
  b.init(t);
  const double a = aexp[j], aH = b.aprimeoa(a);
  const double *ys = y + idx * n, *yps = yp + idx * n;
  const NuPt nu = nu_point([&](int i) { return ys[i]; }, [&](int i) { return yps[i]; }, iq0, nq, qw, a, b.amnu, aH);
  const long plane = (long)n_k * n_tau;
  out[0 * plane + idx] = nu.drhonu;
  out[1 * plane + idx] = nu.dpnu;
  out[2 * plane + idx] = nu.fnu;
  out[3 * plane + idx] = nu.shearnu;
  out[4 * plane + idx] = nu.rnp;
  out[5 * plane + idx] = nu.snp;
  if (idx < n_tau) aHo[idx] = b.aprimeoa(aexp[idx]);
}

// compute_metric_perturbations (cmb.py:243-345) from the dictionaries' planes: planes 0..5 of out
// hprime, etaprime, alpha, alphaprime, alphapprime, aprimeoaprime
__global__ void metric_stage_kernel(const double *__restrict__ t, DebMetricInputs in, const double *__restrict__ kk, int n_k,
                                    const double *__restrict__ aexp, int n_tau, double *__restrict__ out) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau), ik = (int)(idx / n_tau);
  BgView b;
  b.init(t);
  const PertPt x{in.eta[idx],    in.deltac[idx], in.thetac[idx],  in.deltab[idx], in.thetab[idx], in.deltag[idx],
                 in.thetag[idx], in.shearg[idx], in.sheargp[idx], in.deltar[idx], in.thetar[idx], in.shearr[idx],
                 in.shearrp[idx], in.deltaq[idx], in.thetaq[idx]};
  const MetricPt mt = metric_point(b, kk[ik], aexp[j], in.aH[j], in.rhonu[j], in.rho_Q[j], in.w_Q[j], x, in.drhonu[idx], in.fnu[idx],
                                   in.shearnu[idx], in.rho_nu_prime[idx], in.shear_nu_prime[idx]);
  const long plane = (long)n_k * n_tau;
  out[0 * plane + idx] = mt.hprime;
  out[1 * plane + idx] = mt.etaprime;
  out[2 * plane + idx] = mt.alpha;
  out[3 * plane + idx] = mt.alphap;
  out[4 * plane + idx] = mt.alphapp;
  out[5 * plane + idx] = mt.aHp;
}

// compute_polarization_terms (cmb.py:349-395): planes 0..2 polarization_term and its first and
// second derivatives, from the state, its derivative, theta_gamma', alpha' and the opacity
__global__ void pol_stage_kernel(const double *__restrict__ y, const double *__restrict__ yp, int n_k, int n_tau, int n, int idxg, int idxgp,
                                 const double *__restrict__ kk, const double *__restrict__ thetagp, const double *__restrict__ alphap,
                                 const double *__restrict__ opac, const double *__restrict__ opacp, double *__restrict__ out) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau), ik = (int)(idx / n_tau);
  const double *Y = y + idx * n, *P = yp + idx * n;
  const double y2[3] = {Y[idxg + 2], Y[idxgp], Y[idxgp + 2]}, p2[3] = {P[idxg + 2], P[idxgp], P[idxgp + 2]};
  const double p3[3] = {P[idxg + 3], P[idxgp + 1], P[idxgp + 3]};
  const PolPt pl = pol_point(kk[ik], thetagp[idx], alphap[idx], opac[j], opacp[j], y2, p2, p3);
  const long plane = (long)n_k * n_tau;
  out[0 * plane + idx] = pl.pol;
  out[1 * plane + idx] = pl.polp;
  out[2 * plane + idx] = pl.polpp;
}

// compute_source_term_sachs_wolfe / _doppler / _polarization (cmb.py:455-563): term 2, 3 or 4;
// term 5 the E-mode source
__global__ void source_term_kernel(int term, const double *__restrict__ kk, int n_k, int n_tau, const double *__restrict__ gv,
                                   const double *__restrict__ deltag, const double *__restrict__ alpha, const double *__restrict__ alphap,
                                   const double *__restrict__ thetab, const double *__restrict__ thetabp, const double *__restrict__ pol,
                                   const double *__restrict__ polp, const double *__restrict__ polpp, const double *__restrict__ tau, double tau0,
                                   double *__restrict__ out) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau), ik = (int)(idx / n_tau);
  const double k = kk[ik];
  if (term == 2) out[idx] = s2_point(k, gv[j], deltag[idx], alphap[idx], pol[idx], thetabp[idx], polpp[idx]);
  else if (term == 3) out[idx] = s3_point(k, gv[j], alpha[idx], thetab[idx], polp[idx]);
  else if (term == 4) out[idx] = s4_point(k, gv[j], pol[idx]);
  else out[idx] = se_point(k, gv[j], pol[idx], tau0 - tau[j]);
}

// the ISW potential eta + alpha' transposed to n_tau x n_k for source_isw_kernel
__global__ void isw_potential_kernel(const double *__restrict__ eta, const double *__restrict__ alphap, int n_k, int n_tau,
                                     double *__restrict__ pot) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_k * n_tau) return;
  const int j = (int)(idx % n_tau), ik = (int)(idx / n_tau);
  pot[(long)j * n_k + ik] = eta[idx] + alphap[idx];
}

// S1 = exp(-tau_c) d(eta + alpha')/dtau from the natural spline of the potential in tau
// (cmb.py:398-452), then S = S1 + S2 + S3 + S4.  One thread per k; pot and its second
// derivatives are stored tau-major (n_tau x n_k) so neighbouring threads touch neighbouring words.
__global__ void source_isw_kernel(const double *__restrict__ tau, int n_tau, int n_k, const double *__restrict__ cp,
                                  const double *__restrict__ pot, double *__restrict__ Spot, const double *__restrict__ V,
                                  double *__restrict__ out) {
  const int ik = blockIdx.x * blockDim.x + threadIdx.x;
  if (ik >= n_k) return;
  spline_build(tau, pot + ik, n_k, n_tau, cp, Spot + ik, n_k);
  const long plane = (long)n_k * n_tau;
  for (int j = 0; j < n_tau; ++j) {
    const int i = spl_index(tau, n_tau, tau[j]);
    const double d = spl_slope(tau, pot + ik, n_k, Spot + ik, n_k, i, tau[j]);
    const long o = (long)ik * n_tau + j;
    const double S1 = exp(-V[5L * n_tau + j]) * d;
    out[0 * plane + o] = S1;
    out[4 * plane + o] = S1 + out[1 * plane + o] + out[2 * plane + o] + out[3 * plane + o];
  }
}

// compute_source_term_isw (cmb.py:401-452) on its own: S1 = exp(-tau_c) d(eta + alpha')/dtau, the
// potential pot (n_tau x n_k) from isw_potential_kernel, optical depth od (n_tau)
__global__ void isw_stage_kernel(const double *__restrict__ tau, int n_tau, int n_k, const double *__restrict__ cp,
                                 const double *__restrict__ pot, double *__restrict__ Spot, const double *__restrict__ od,
                                 double *__restrict__ S1) {
  const int ik = blockIdx.x * blockDim.x + threadIdx.x;
  if (ik >= n_k) return;
  spline_build(tau, pot + ik, n_k, n_tau, cp, Spot + ik, n_k);
  for (int j = 0; j < n_tau; ++j) {
    const int i = spl_index(tau, n_tau, tau[j]);
    const double d = spl_slope(tau, pot + ik, n_k, Spot + ik, n_k, i, tau[j]);
    S1[(long)ik * n_tau + j] = exp(-od[j]) * d;
  }
}

// spline_cp on one block of 256 threads: cp_i = h1_i / (b_i - h0_i cp_{i-1}) is the Moebius map of
// [[0, h1_i], [-h0_i, b_i]] applied to (cp_{i-1}, 1); each thread composes its chunk's maps, a
// block-wide inclusive scan of the (max-normalised) 2 x 2 products gives every chunk its incoming
// value, and the chunks are replayed.
__global__ __launch_bounds__(256) void spline_cp_kernel(const double *__restrict__ x, int n, double *__restrict__ cp) {
  __shared__ double s00[256], s01[256], s10[256], s11[256];
  const int tid = threadIdx.x, nth = blockDim.x;
  const int m = n - 2;
  if (m <= 0) return;
  const int C = (m + nth - 1) / nth, c0 = tid * C, c1 = min(m, c0 + C);
  double a = 1.0, b = 0.0, c = 0.0, d = 1.0;
  for (int i = c0; i < c1; ++i) {
    const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1], bb = 2.0 * (h0 + h1);
    const double na = h1 * c, nb = h1 * d, nc = bb * c - h0 * a, nd = bb * d - h0 * b;
    const double sc = 1.0 / fmax(fmax(fabs(na), fabs(nb)), fmax(fabs(nc), fabs(nd)));
    a = na * sc; b = nb * sc; c = nc * sc; d = nd * sc;
  }
  s00[tid] = a; s01[tid] = b; s10[tid] = c; s11[tid] = d;
  __syncthreads();
  for (int off = 1; off < nth; off <<= 1) {  // P_t = M_t P_{t - off}
    double pa = a, pb = b, pc = c, pd = d;
    if (tid >= off) {
      const double qa = s00[tid - off], qb = s01[tid - off], qc = s10[tid - off], qd = s11[tid - off];
      pa = a * qa + b * qc; pb = a * qb + b * qd; pc = c * qa + d * qc; pd = c * qb + d * qd;
      const double sc = 1.0 / fmax(fmax(fabs(pa), fabs(pb)), fmax(fabs(pc), fabs(pd)));
      pa *= sc; pb *= sc; pc *= sc; pd *= sc;
    }
    __syncthreads();
    s00[tid] = a = pa; s01[tid] = b = pb; s10[tid] = c = pc; s11[tid] = d = pd;
    __syncthreads();
  }
  double p = 0.0, q = 1.0;  // (cp_{-1}, 1)
  if (tid > 0) { p = s01[tid - 1]; q = s11[tid - 1]; }
  for (int i = c0; i < c1; ++i) {
    const double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1], bb = 2.0 * (h0 + h1);
    const double np_ = h1 * q, nq = bb * q - h0 * p;
    const double sc = 1.0 / fmax(fabs(np_), fabs(nq));
    p = np_ * sc; q = nq * sc;
    cp[i] = (i < m - 1) ? p / q : 0.0;
  }
}

// ---------------------------------------------------------------------------------------------
// line of sight
// ---------------------------------------------------------------------------------------------
// logs of the input wavenumbers and the dense log-uniform grid (jax.numpy.linspace arithmetic:
// start (1 - s) + stop s, s = i / (n - 1), last point exactly stop)
__global__ void log_k_kernel(const double *__restrict__ k, int n_k, int n_kd, double *__restrict__ lk, double *__restrict__ lkd,
                             double *__restrict__ kd) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_k) lk[i] = log(k[i]);
  if (i < n_kd) {
    double l0 = log(k[0]), l1 = log(k[n_k - 1]);
    double s = (double)i / (double)(n_kd - 1);
    double v = (i == n_kd - 1) ? l1 : l0 * (1.0 - s) + l1 * s;
    lkd[i] = v;
    kd[i] = exp(v);
  }
}

// one thread per time slice: spline of S[:, j] in log k (S is n_k x n_tau)
__global__ void densify_build_kernel(const double *__restrict__ lk, int n_k, int n_tau, const double *__restrict__ cp,
                                     const double *__restrict__ S, double *__restrict__ Sk) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_tau) return;
  spline_build(lk, S + j, n_tau, n_k, cp, Sk + j, n_tau);
}

__global__ void densify_eval_kernel(const double *__restrict__ lk, int n_k, int n_tau, const double *__restrict__ S,
                                    const double *__restrict__ Sk, const double *__restrict__ lkd, int n_kd, double *__restrict__ Sd) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n_kd * n_tau) return;
  const int j = (int)(idx % n_tau), id = (int)(idx / n_tau);
  const double xq = lkd[id];
  const int i = spl_index(lk, n_k, xq);
  Sd[idx] = spl_value(lk, S + j, n_tau, Sk + j, n_tau, i, xq);
}

// knots of the log-chi splines: chi_i = tau0 - tau[n - 1 - i], log max(chi, 1e-10)
__global__ void chi_knots_kernel(const double *__restrict__ tau, int n_tau, double tau0, double *__restrict__ lco) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_tau) return;
  double chi = tau0 - tau[n_tau - 1 - i];
  lco[i] = log(chi > 1e-10 ? chi : 1e-10);
}

// FFTLog grid of cmb.py:753-775: chi = tau0 - tau (increasing after reversal), log chi_min =
// log max(chi_0, 1) - 2, log chi_max = log chi_last + 2, N points of width dl
__device__ __forceinline__ void fft_grid(const double *__restrict__ tau, int n_tau, double tau0, int N, double &lmin, double &dl) {
  const double c0 = tau0 - tau[n_tau - 1], c1 = tau0 - tau[0];
  lmin = log(c0 > 1.0 ? c0 : 1.0) - 2.0;
  const double lmax = log(c1) + 2.0;
  dl = (lmax - lmin) / (double)N;
}

// e^{-2 pi i t / N} for t < N / 2
__global__ void twiddle_kernel(int N, double2 *__restrict__ tw) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N / 2) return;
  double s, c;
  sincospi(-2.0 * (double)t / (double)N, &s, &c);
  tw[t] = make_double2(c, s);
}


// Row-independent tables of the FFTLog stage, built once per launch (one block):
//   knot tables of the log-chi spline solve: rden_i = 1 / den_i (den_i the forward-elimination
//   pivot of spline_build), bco_i = -h0_i / den_i (cp comes from spline_cp);
//   per sample m of the log-chi grid: its knot interval I[m] (the index rule of spl_index) and
//   the four spline-basis weights of spl_value times chi_m^1.1 (W[4 m .. 4 m + 3]);
//   per output frequency n <= N/2: e^{-i eta_n lmin} / N (ph[n]).
__global__ __launch_bounds__(256) void fftlog_setup_kernel(const double *__restrict__ lco, int n_tau, const double *__restrict__ cp,
                                                           const double *__restrict__ tau, double tau0, int N, double *__restrict__ rden,
                                                           double *__restrict__ bco, int *__restrict__ I, double *__restrict__ W,
                                                           double2 *__restrict__ ph) {
  const int tid = threadIdx.x;
  const int m = n_tau - 2;
  for (int i = tid; i < m; i += blockDim.x) {  // cp is known: the pivots are independent
    const double h0 = lco[i + 1] - lco[i], h1 = lco[i + 2] - lco[i + 1];
    const double b = 2.0 * (h0 + h1);
    const double den = (i == 0) ? b : b - h0 * cp[i - 1];
    rden[i] = 1.0 / den;
    bco[i] = (i == 0) ? 0.0 : -h0 / den;
  }
  double lmin, dl;
  fft_grid(tau, n_tau, tau0, N, lmin, dl);
  for (int q = tid; q < N; q += blockDim.x) {
    const double lx = lmin + (double)q * dl;
    const int i = spl_index(lco, n_tau, lx);
    const double h = lco[i + 1] - lco[i];
    const double t = (lx - lco[i]) / h;
    const double A = 1.0 - t, B = t, e = exp(1.1 * lx), c = h * h / 6.0;
    I[q] = i;
    W[4 * q] = A * e;
    W[4 * q + 1] = B * e;
    W[4 * q + 2] = (A * A * A - A) * c * e;
    W[4 * q + 3] = (B * B * B - B) * c * e;
  }
  for (int n = tid; n <= N / 2; n += blockDim.x) {
    const double eta = 2.0 * kPi * (double)n / ((double)N * dl);
    double s1, c1;
    sincos(-eta * lmin, &s1, &c1);
    ph[n] = make_double2(c1 / (double)N, s1 / (double)N);
  }
}

// inclusive scan over the block of affine maps x -> a + b x, composed in thread order
// (entry t after entry t - 1); sa, sb: blockDim.x doubles each in LDS
__device__ void affine_scan(double *sa, double *sb, int tid, int nth) {
  for (int off = 1; off < nth; off <<= 1) {
    double a = sa[tid], b = sb[tid];
    if (tid >= off) {
      const double pa = sa[tid - off], pb = sb[tid - off];
      a = a + b * pa;
      b = b * pb;
    }
    __syncthreads();
    sa[tid] = a;
    sb[tid] = b;
    __syncthreads();
  }
}

// FFTLog coefficients of one k-row per workgroup (cmb.py:753-803).  The row's natural spline in
// log chi (knots lco, the row reversed so chi increases) is solved by two block-wide scans of the
// affine recurrences of spline_build (forward elimination dp_i = d_i rden_i + bco_i dp_{i-1}, back
// substitution v_i = dp_i - cp_i v_{i+1}) into the second derivatives Sd (LDS when they fit, else
// the row's slot of Schi); samples f_m = S(chi_m) chi_m^1.1 on log chi_m = lmin + m dl are four
// multiply-adds with the setup's weights, packed as z_j = f_2j + i f_2j+1 at bit-reversed LDS
// slots; a radix-2 FFT of M = N/2 points gives Z, the real-FFT split X_n = rfft(f)_n for n <= N/2,
// and ck_n = X_n e^{-i eta_n lmin} / N * k^{-(p_n + 1)} is written as A[row, 2n] = w_n Re ck,
// A[row, 2n + 1] = -w_n Im ck (w = 1 for n = 0, N/2, else 2), zero beyond 2 (N/2 + 1).
template <int LOGM>
__global__ __launch_bounds__(256) void fftlog_kernel(const double *__restrict__ S, int n_tau, const double *__restrict__ lco,
                                                     const double *__restrict__ cp, const double *__restrict__ rden,
                                                     const double *__restrict__ bco, const int *__restrict__ Iq,
                                                     const double *__restrict__ Wq, const double2 *__restrict__ ph, int sd_in_lds,
                                                     double *__restrict__ Schi, const double *__restrict__ kk, const double *__restrict__ tau,
                                                     double tau0, const double2 *__restrict__ tw, double *__restrict__ A, long Kp) {
  constexpr int M = 1 << LOGM, N = 2 * M;
  extern __shared__ double2 z[];
  double *sa = reinterpret_cast<double *>(z + M), *sb = sa + 256;
  const int row = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
  double *Sd = sd_in_lds ? sb + 256 : Schi + (long)row * n_tau;
  double lmin, dl;
  fft_grid(tau, n_tau, tau0, N, lmin, dl);
  const double *Yr = S + (long)row * n_tau + (n_tau - 1);  // Yr[-i] = value at knot i
  const int m = n_tau - 2;
  const int C = (m + nth - 1) / nth;
  const int c0 = tid * C, c1 = min(m, c0 + C);
  // forward elimination: a_i = d_i rden_i kept in Sd[i + 1], then the chunk's composite map
  {
    double ca = 0.0, cb = 1.0;
    for (int i = c0; i < c1; ++i) {
      const double h0 = lco[i + 1] - lco[i], h1 = lco[i + 2] - lco[i + 1];
      const double y0 = Yr[-i], y1 = Yr[-(i + 1)], y2 = Yr[-(i + 2)];
      const double d = 6.0 * ((y2 - y1) / h1 - (y1 - y0) / h0);
      const double a = d * rden[i], b = bco[i];
      Sd[i + 1] = a;
      ca = a + b * ca;
      cb = b * cb;
    }
    sa[tid] = ca;
    sb[tid] = cb;
  }
  __syncthreads();
  affine_scan(sa, sb, tid, nth);
  {
    double x = (tid == 0) ? 0.0 : sa[tid - 1];  // dp before the chunk (dp_{-1} = 0)
    for (int i = c0; i < c1; ++i) {
      x = Sd[i + 1] + bco[i] * x;
      Sd[i + 1] = x;
    }
  }
  __syncthreads();
  // back substitution in reverse order: chunk of thread t' = nth - 1 - t taken from the end
  {
    const int tr = nth - 1 - tid;
    const int r0 = tr * C, r1 = min(m, r0 + C);  // indices m-1-r for r in [r0, r1)
    double ca = 0.0, cb = 1.0;
    for (int r = r0; r < r1; ++r) {
      const int i = m - 1 - r;
      const double a = Sd[i + 1], b = -cp[i];
      ca = a + b * ca;
      cb = b * cb;
    }
    __syncthreads();
    sa[tr] = ca;
    sb[tr] = cb;
    __syncthreads();
    affine_scan(sa, sb, tid, nth);
    double x = (tr == 0) ? 0.0 : sa[tr - 1];  // v_m = 0
    for (int r = r0; r < r1; ++r) {
      const int i = m - 1 - r;
      x = Sd[i + 1] - cp[i] * x;
      Sd[i + 1] = x;
    }
    if (tid == 0) { Sd[0] = 0.0; Sd[n_tau - 1] = 0.0; }
  }
  __syncthreads();
  for (int j = tid; j < M; j += nth) {
    double f[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q = 2 * j + h;
      const int i = Iq[q];
      const double *w = Wq + 4 * (long)q;
      f[h] = w[0] * Yr[-i] + w[1] * Yr[-(i + 1)] + w[2] * Sd[i] + w[3] * Sd[i + 1];
    }
    const int r = (int)(__brev((unsigned)j) >> (32 - LOGM));
    z[r] = make_double2(f[0], f[1]);
  }
  __syncthreads();
  for (int s = 1; s <= LOGM; ++s) {  // in-place decimation in time
    const int half = 1 << (s - 1);
    const int stride = N >> s;  // twiddle index step: e^{-2 pi i j / 2^s} = tw[j N / 2^s]
    for (int bfly = tid; bfly < M / 2; bfly += nth) {
      const int j = bfly & (half - 1);
      const int i0 = ((bfly >> (s - 1)) << s) | j, i1 = i0 + half;
      const double2 w = tw[j * stride];
      const double2 u = z[i0], v = cmul(z[i1], w);
      z[i0] = make_double2(u.x + v.x, u.y + v.y);
      z[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
  const double lk = log(kk[row]);
  const double kpow = exp(-(-1.1 + 1.0) * lk);
  double *Ar = A + (long)row * Kp;
  for (int nn = tid; nn <= M; nn += nth) {
    const double2 Zn = z[nn & (M - 1)];
    const double2 Zm = z[(M - nn) & (M - 1)];
    const double2 Zc = make_double2(Zm.x, -Zm.y);
    const double2 E = make_double2(0.5 * (Zn.x + Zc.x), 0.5 * (Zn.y + Zc.y));
    const double2 D = make_double2(0.5 * (Zn.x - Zc.x), 0.5 * (Zn.y - Zc.y));
    const double2 Wt = (nn < M) ? tw[nn] : make_double2(-1.0, 0.0);
    const double2 O = cmul(Wt, make_double2(D.y, -D.x));  // e^{-2 pi i n / N} (-i) D
    double2 X = make_double2(E.x + O.x, E.y + O.y);
    const double eta = 2.0 * kPi * (double)nn / ((double)N * dl);
    double s2, c2;
    sincos(-eta * lk, &s2, &c2);
    X = cmul(X, ph[nn]);
    X = cmul(X, make_double2(kpow * c2, kpow * s2));
    const double wn = (nn == 0 || nn == M) ? 1.0 : 2.0;
    Ar[2 * nn] = wn * X.x;
    Ar[2 * nn + 1] = -wn * X.y;
  }
  for (long c = (long)N + 2 + tid; c < Kp; c += nth) Ar[c] = 0.0;
}

// B[l, 2n] = Re M_l(p_n), B[l, 2n + 1] = Im M_l(p_n), p_n = -1.1 + i eta_n (cmb.py:636-677)
__global__ void bessel_coef_kernel(int n_ell, int N, const double *__restrict__ tau, int n_tau, double tau0, double *__restrict__ B,
                                   long Kp, int field) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double lmin, dl;
  fft_grid(tau, n_tau, tau0, N, lmin, dl);
  const int nf = N / 2 + 1;
  if (idx >= (long)n_ell * Kp / 2) return;
  const int l = (int)(idx / (Kp / 2)), nn = (int)(idx % (Kp / 2));
  double2 M = make_double2(0.0, 0.0);
  if (nn < nf) {
    const double eta = 2.0 * kPi * (double)nn / ((double)N * dl);
    const double2 p = make_double2(-1.1, eta);
    const double2 a1 = make_double2((1.0 + l + p.x) / 2.0, p.y / 2.0), a2 = make_double2((2.0 + l - p.x) / 2.0, -p.y / 2.0);
    double2 lg = make_double2((p.x - 1.0) * log(2.0), p.y * log(2.0));
    lg.x += 0.5 * log(kPi);
    lg = cadd(lg, lngamma_complex(a1));
    lg = csub(lg, lngamma_complex(a2));
    const double e = exp(lg.x);
    double s, c;
    sincos(lg.y, &s, &c);
    M = make_double2(e * c, e * s);
    if (e == 0.0) M = make_double2(0.0, 0.0);  // the reflection branch overflowed: Gamma -> 0
    if (field == 1) {  // E-mode: Theta_E,l = sqrt((l + 2)! / (l - 2)!) int S_E j_l (no counterpart in the reference)
      const double fl = l >= 2 ? sqrt((double)(l + 2) * (l + 1) * l * (l - 1)) : 0.0;
      M = make_double2(M.x * fl, M.y * fl);
    }
  }
  B[(long)l * Kp + 2 * nn] = M.x;
  B[(long)l * Kp + 2 * nn + 1] = M.y;
}

// ---------------------------------------------------------------------------------------------
// Theta = A B^T with FP64 MFMA.  A: M x K, B: N x K (both K-contiguous, K a multiple of 16),
// C: M x N row-major.  Workgroup: 512 lanes (8 waves in 4 x 2), tile 128 x 128, each wave a
// 32 x 64 block of 2 x 4 v_mfma_f64_16x16x4f64 accumulators (small enough for two waves per SIMD).
// Operand layout of the 16x16x4 f64 MFMA: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// the accumulator's element v of lane l is C[4 v + l / 16][l % 16] (checked on the device).
// K is staged through LDS 32 at a time; the next 32 are loaded from HBM into registers while the
// current ones are multiplied.  Inside a K-tile, MFMA step s of lane (l % 16, l / 16) takes
// k = 8 (l / 16) + s for both operands (a permutation of the tile's k, the same on both sides, so
// the sum is unchanged), which makes each lane's two steps s, s + 1 one 16-byte LDS read.
// Block b runs on XCD b % 8 (the dispatcher's round robin): the tiles are numbered so that the
// blocks of one XCD cover one compact rectangle of tiles (gm tile rows by about half the tile
// columns), whose operand panels its L2 then shares.
// ---------------------------------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int TM = 128, TN = 128, TK = 32, TKP = TK + 2;  // +2: 16-byte rows offset against LDS bank conflicts
constexpr int NXCD = 8;

constexpr int CT = 512;  // threads per contraction workgroup
__global__ __launch_bounds__(CT) void contract_kernel(const double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C,
                                                       int Mr, int Nc, long K, int tiles_m, int tiles_n, int per_xcd, int gm) {
  __shared__ __align__(16) double As[TM * TKP], Bs[TN * TKP];
  // virtual tile number: XCD-major, then grouped column-major in bands of gm tile rows
  const int b = blockIdx.x;
  const long v = (long)(b % NXCD) * per_xcd + b / NXCD;
  if (v >= (long)tiles_m * tiles_n) return;
  const long band = v / ((long)gm * tiles_n), within = v % ((long)gm * tiles_n);
  const int rows_in_band = (int)min((long)gm, (long)tiles_m - band * gm);
  const int tm = (int)(band * gm + within % rows_in_band), tn = (int)(within / rows_in_band);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
  const long m0 = (long)tm * TM, n0 = (long)tn * TN;
  v4d acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
  const int lr = lane & 15, lk = lane >> 4;
  // global -> register staging: thread covers double pairs e = tid + CT t, row e / 16, columns 2 (e % 16), + 1
  constexpr int NL = TM * TK / 2 / CT;
  double2 ga[NL], gb[NL];
  auto load = [&](long k0) {
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int e = tid + t * CT;
      const int r = e >> 4, c = 2 * (e & 15);
      const long ra = m0 + r, rb = n0 + r;
      const bool kin = k0 + c < K;  // K is even: a pair is wholly inside or outside
      ga[t] = (ra < Mr && kin) ? *reinterpret_cast<const double2 *>(A + ra * K + k0 + c) : make_double2(0.0, 0.0);
      gb[t] = (rb < Nc && kin) ? *reinterpret_cast<const double2 *>(B + rb * K + k0 + c) : make_double2(0.0, 0.0);
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int e = tid + t * CT;
      const int r = e >> 4, c = 2 * (e & 15);
      *reinterpret_cast<double2 *>(As + r * TKP + c) = ga[t];
      *reinterpret_cast<double2 *>(Bs + r * TKP + c) = gb[t];
    }
  };
  load(0);
  store();
  __syncthreads();
  for (long k0 = 0; k0 < K; k0 += TK) {
    const bool more = k0 + TK < K;
    if (more) load(k0 + TK);
#pragma unroll
    for (int s = 0; s < TK / 4; s += 2) {
      double2 a[2], bb[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const double2 *>(As + (wm + i * 16 + lr) * TKP + lk * 8 + s);
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = *reinterpret_cast<const double2 *>(Bs + (wn + j * 16 + lr) * TKP + lk * 8 + s);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, bb[j].x, acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, bb[j].y, acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (more) {
      store();
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v4 = 0; v4 < 4; ++v4) {
        const long r = m0 + wm + i * 16 + 4 * v4 + lk, c = n0 + wn + j * 16 + lr;
        if (r < Mr && c < Nc) C[r * Nc + c] = acc[i][j][v4];
      }
}

// grid of contract_kernel: NXCD x per_xcd blocks (the surplus exits), bands of gm tile rows
struct ContractGrid {
  int tiles_m, tiles_n, per_xcd, gm, blocks;
};
static ContractGrid contract_grid(int m, int n) {
  ContractGrid g;
  g.tiles_m = (m + TM - 1) / TM;
  g.tiles_n = (n + TN - 1) / TN;
  const int tiles = g.tiles_m * g.tiles_n;
  g.per_xcd = (tiles + NXCD - 1) / NXCD;
  const int half_cols = (g.tiles_n + 1) / 2;
  g.gm = (g.per_xcd + half_cols - 1) / half_cols;
  if (g.gm < 1) g.gm = 1;
  if (g.gm > g.tiles_m) g.gm = g.tiles_m;
  g.blocks = NXCD * g.per_xcd;
  return g;
}

// C_l = sum_i (ln k_{i+1} - ln k_i) (f_i + f_{i+1}) / 2, f = (k / k_p)^(n_s - 1) Theta_l(k)^2
__global__ void cell_kernel(const double *__restrict__ th, const double *__restrict__ th2, const double *__restrict__ kk, int n_k, int n_ell,
                            double n_s, double k_p, double *__restrict__ cl) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_ell) return;
  double acc = 0.0, fprev = 0.0, lkprev = 0.0;
  for (int i = 0; i < n_k; ++i) {
    const double k = kk[i], t = th[(long)i * n_ell + l], t2 = th2[(long)i * n_ell + l];
    const double f = pow(k / k_p, n_s - 1.0) * t * t2, lk = log(k);
    if (i > 0) acc += (lk - lkprev) * (f + fprev) / 2.0;
    fprev = f;
    lkprev = lk;
  }
  cl[l] = acc;
}

__global__ void dell_kernel(const double *__restrict__ cl, int ellmax, double A_s, double Tcmb, double *__restrict__ ell,
                            double *__restrict__ dl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > ellmax - 2) return;
  const double l = (double)(i + 2);
  ell[i] = l;
  dl[i] = l * (l + 1.0) * cl[i + 2] * A_s * 2.0 * Tcmb * Tcmb * 1e12;
}

}  // namespace
}  // namespace deb

// ---------------------------------------------------------------------------------------------
// launchers (stream-ordered, no allocation)
// ---------------------------------------------------------------------------------------------
extern "C" int deb_launch_nu_bins(int nqmax, double *d_qw, hipStream_t stream);

extern "C" long deb_cmb_sources_work_doubles(int num_thermo, int n_k, int n_tau) {
  return 12L * num_thermo + (long)deb::NV * n_tau + 2L * n_k * n_tau + n_tau + 64;
}

extern "C" int deb_launch_cmb_sources(const double *d_t, int num_thermo, const double *d_y, const double *d_yp, const double *d_k, int n_k,
                                      const double *d_aexp, const double *d_tau, int n_tau, const DebProblem *prob, int isw, double *d_out,
                                      double *d_work, hipStream_t stream) {
  double *vw = d_work, *V = vw + 12L * num_thermo, *pot = V + (long)deb::NV * n_tau, *Spot = pot + (long)n_k * n_tau;
  double *cp = Spot + (long)n_k * n_tau, *qw = cp + n_tau;
  int rc = deb_launch_nu_bins(prob->nqmax, qw, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(deb::vis_build_kernel, dim3(1), dim3(64), 0, stream, d_t, vw);
  hipLaunchKernelGGL(deb::vis_eval_kernel, dim3((n_tau + 127) / 128), dim3(128), 0, stream, d_t, vw, d_tau, d_aexp, n_tau, V);
  long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::source_point_kernel, dim3((unsigned)((pts + 127) / 128)), dim3(128), 0, stream, d_t, d_y, d_yp, d_k, n_k, n_tau,
                     d_aexp, V, qw, *prob, isw, d_tau, d_out, pot);
  if (isw == 0) {
    hipLaunchKernelGGL(deb::spline_cp_kernel, dim3(1), dim3(256), 0, stream, d_tau, n_tau, cp);
    hipLaunchKernelGGL(deb::source_isw_kernel, dim3((n_k + 63) / 64), dim3(64), 0, stream, d_tau, n_tau, n_k, cp, pot, Spot, V, d_out);
  }
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cmb_neutrinos(const double *d_t, const double *d_y, const double *d_yp, int n_k, const double *d_aexp, int n_tau,
                                        int nvar, int iq0, int nqmax, double *d_out, double *d_aH, double *d_qw, hipStream_t stream) {
  int rc = deb_launch_nu_bins(nqmax, d_qw, stream);
  if (rc) return rc;
  const long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::nu_stage_kernel, dim3((unsigned)((pts + 127) / 128)), dim3(128), 0, stream, d_t, d_y, d_yp, n_k, n_tau, d_aexp, d_qw,
                     nvar, iq0, nqmax, d_out, d_aH);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cmb_metric(const double *d_t, const DebMetricInputs *in, const double *d_k, int n_k, const double *d_aexp, int n_tau,
                                     double *d_out, hipStream_t stream) {
  const long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::metric_stage_kernel, dim3((unsigned)((pts + 127) / 128)), dim3(128), 0, stream, d_t, *in, d_k, n_k, d_aexp, n_tau,
                     d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cmb_polarization(const double *d_y, const double *d_yp, int n_k, int n_tau, int nvar, int idxg, int idxgp,
                                           const double *d_k, const double *d_thetagp, const double *d_alphap, const double *d_opac,
                                           const double *d_opacp, double *d_out, hipStream_t stream) {
  const long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::pol_stage_kernel, dim3((unsigned)((pts + 127) / 128)), dim3(128), 0, stream, d_y, d_yp, n_k, n_tau, nvar, idxg, idxgp,
                     d_k,
                     d_thetagp, d_alphap, d_opac, d_opacp, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cmb_source_term(int term, const double *d_k, int n_k, int n_tau, const double *d_g, const double *d_deltag,
                                          const double *d_alpha, const double *d_alphap, const double *d_thetab, const double *d_thetabp,
                                          const double *d_pol, const double *d_polp, const double *d_polpp, const double *d_tau, double tau0,
                                          double *d_out, hipStream_t stream) {
  const long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::source_term_kernel, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, stream, term, d_k, n_k, n_tau, d_g, d_deltag,
                     d_alpha, d_alphap, d_thetab, d_thetabp, d_pol, d_polp, d_polpp, d_tau, tau0, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cmb_isw(const double *d_tau, int n_tau, int n_k, const double *d_eta, const double *d_alphap, const double *d_od,
                                  double *d_S1, double *d_work, hipStream_t stream) {
  double *pot = d_work, *Spot = pot + (long)n_k * n_tau, *cp = Spot + (long)n_k * n_tau;
  const long pts = (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::isw_potential_kernel, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, stream, d_eta, d_alphap, n_k, n_tau, pot);
  hipLaunchKernelGGL(deb::spline_cp_kernel, dim3(1), dim3(256), 0, stream, d_tau, n_tau, cp);
  hipLaunchKernelGGL(deb::isw_stage_kernel, dim3((n_k + 63) / 64), dim3(64), 0, stream, d_tau, n_tau, n_k, cp, pot, Spot, d_od, d_S1);
  return (int)hipGetLastError();
}

// The ISW term from a potential plane already written (n_tau x n_k, as source_point_kernel writes
// it with isw 0): S1 and S of d_out (8 planes of n_k x n_tau) from the spline in tau.  d_work:
// n_tau x n_k + n_tau doubles.
extern "C" int deb_launch_cmb_isw_finish(const double *d_tau, int n_tau, int n_k, const double *d_pot, const double *d_V, double *d_out,
                                         double *d_work, hipStream_t stream) {
  double *Spot = d_work, *cp = Spot + (long)n_k * n_tau;
  hipLaunchKernelGGL(deb::spline_cp_kernel, dim3(1), dim3(256), 0, stream, d_tau, n_tau, cp);
  hipLaunchKernelGGL(deb::source_isw_kernel, dim3((n_k + 63) / 64), dim3(64), 0, stream, d_tau, n_tau, n_k, cp, d_pot, Spot, d_V, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_visibility(const double *d_t, int num_thermo, const double *d_tau, const double *d_aexp, int n_tau, double *d_V,
                                     double *d_work, hipStream_t stream) {
  hipLaunchKernelGGL(deb::vis_build_kernel, dim3(1), dim3(64), 0, stream, d_t, d_work);
  hipLaunchKernelGGL(deb::vis_eval_kernel, dim3((n_tau + 127) / 128), dim3(128), 0, stream, d_t, d_work, d_tau, d_aexp, n_tau, d_V);
  return (int)hipGetLastError();
}

static long los_kp(int n_fftlog) { return (((long)n_fftlog + 2 + 31) / 32) * 32; }

extern "C" long deb_los_work_doubles(int n_k, int n_tau, int n_ell, int n_fftlog, int n_k_dense) {
  const long nk = n_k_dense > 0 ? n_k_dense : n_k, Kp = los_kp(n_fftlog);
  long w = (long)n_fftlog + 3L * n_k + 2L * n_tau;  // twiddles, log k / cp_k, lco / cp_chi
  w += 2L * n_tau + 6L * n_fftlog + 64;               // the FFTLog setup tables
  if (n_k_dense > 0) w += (long)n_k * n_tau + 2L * n_k_dense + 2L * n_k_dense * n_tau;
  w += nk * n_tau + nk * Kp + (long)n_ell * Kp;
  return w + 64;
}

extern "C" int deb_launch_los(const double *d_k, int n_k, const double *d_tau, int n_tau, const double *d_S, double tau0, int ellmax,
                              int n_fftlog, int n_k_dense, int field, double *d_theta, double *d_kout, double *d_work, hipStream_t stream) {
  const int n_ell = ellmax + 1;
  const long Kp = los_kp(n_fftlog);
  int logm = 0;
  while ((2 << logm) < n_fftlog) ++logm;
  if ((2 << logm) != n_fftlog || logm < 6 || logm > 13) return -2;  // power of two, 128 .. 16384
  double *w = d_work;
  double2 *tw = (double2 *)w;
  w += n_fftlog;
  double *lk = w; w += n_k;
  double *cpk = w; w += 2L * n_k;
  double *lco = w; w += n_tau;
  double *cpchi = w; w += n_tau;
  double *rden = w; w += n_tau;
  double *bco = w; w += n_tau;
  double *Wq = w; w += 4L * n_fftlog;
  w += (w - d_work) & 1;  // 16-byte alignment of the phase table
  double2 *ph = (double2 *)w; w += 2L * (n_fftlog / 2 + 1);
  int *Iq = (int *)w; w += (n_fftlog + 1) / 2 + 1;
  const double *S = d_S;
  const double *kuse = d_k;
  int nk = n_k;
  double *lkd = nullptr;
  hipLaunchKernelGGL(deb::twiddle_kernel, dim3((n_fftlog / 2 + 255) / 256), dim3(256), 0, stream, n_fftlog, tw);
  if (n_k_dense > 0) {
    double *Sk = w; w += (long)n_k * n_tau;
    lkd = w; w += n_k_dense;
    double *kd = w; w += n_k_dense;
    double *Sd = w; w += 2L * n_k_dense * n_tau;  // densified S (the second half is slack)
    int mx = n_k > n_k_dense ? n_k : n_k_dense;
    hipLaunchKernelGGL(deb::log_k_kernel, dim3((mx + 255) / 256), dim3(256), 0, stream, d_k, n_k, n_k_dense, lk, lkd, kd);
    hipLaunchKernelGGL(deb::spline_cp_kernel, dim3(1), dim3(256), 0, stream, lk, n_k, cpk);
    hipLaunchKernelGGL(deb::densify_build_kernel, dim3((n_tau + 127) / 128), dim3(128), 0, stream, lk, n_k, n_tau, cpk, d_S, Sk);
    long pts = (long)n_k_dense * n_tau;
    hipLaunchKernelGGL(deb::densify_eval_kernel, dim3((unsigned)((pts + 255) / 256)), dim3(256), 0, stream, lk, n_k, n_tau, d_S, Sk, lkd,
                       n_k_dense, Sd);
    S = Sd;
    kuse = kd;
    nk = n_k_dense;
    hipError_t e = hipMemcpyAsync(d_kout, kd, sizeof(double) * n_k_dense, hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return (int)e;
  } else {
    hipError_t e = hipMemcpyAsync(d_kout, d_k, sizeof(double) * n_k, hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return (int)e;
  }
  double *Schi = w; w += (long)nk * n_tau;
  w += (w - d_work) & 1;  // 16-byte alignment of the contraction operands (Kp is even)
  double *A = w; w += (long)nk * Kp;
  double *B = w; w += (long)n_ell * Kp;
  hipLaunchKernelGGL(deb::chi_knots_kernel, dim3((n_tau + 255) / 256), dim3(256), 0, stream, d_tau, n_tau, tau0, lco);
  hipLaunchKernelGGL(deb::spline_cp_kernel, dim3(1), dim3(256), 0, stream, lco, n_tau, cpchi);
  hipLaunchKernelGGL(deb::fftlog_setup_kernel, dim3(1), dim3(256), 0, stream, lco, n_tau, cpchi, d_tau, tau0, n_fftlog, rden, bco, Iq, Wq,
                     ph);
  // LDS: the FFT buffer, the scan's 2 x 256 doubles, and the row's spline second derivatives if
  // they fit in the 160 KiB of a compute unit (else they stay in Schi)
  const size_t lds_base = sizeof(double2) * (size_t)(n_fftlog / 2) + 2 * 256 * sizeof(double);
  const int sd_lds = (lds_base + sizeof(double) * (size_t)n_tau <= 160 * 1024) ? 1 : 0;
  const size_t lds = lds_base + (sd_lds ? sizeof(double) * (size_t)n_tau : 0);
  auto launch_fft = [&](auto kern) -> int {
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(kern, dim3(nk), dim3(256), lds, stream, S, n_tau, lco, cpchi, rden, bco, Iq, Wq, ph, sd_lds, Schi, kuse, d_tau, tau0,
                       tw, A, Kp);
    return (int)hipGetLastError();
  };
  int rc = 0;
  switch (logm) {
    case 6: rc = launch_fft(deb::fftlog_kernel<6>); break;
    case 7: rc = launch_fft(deb::fftlog_kernel<7>); break;
    case 8: rc = launch_fft(deb::fftlog_kernel<8>); break;
    case 9: rc = launch_fft(deb::fftlog_kernel<9>); break;
    case 10: rc = launch_fft(deb::fftlog_kernel<10>); break;
    case 11: rc = launch_fft(deb::fftlog_kernel<11>); break;
    case 12: rc = launch_fft(deb::fftlog_kernel<12>); break;
    default: rc = launch_fft(deb::fftlog_kernel<13>); break;
  }
  if (rc) return rc;
  long cells = (long)n_ell * (Kp / 2);
  hipLaunchKernelGGL(deb::bessel_coef_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, stream, n_ell, n_fftlog, d_tau, n_tau,
                     tau0, B, Kp, field);
  const deb::ContractGrid g = deb::contract_grid(nk, n_ell);
  hipLaunchKernelGGL(deb::contract_kernel, dim3(g.blocks), dim3(deb::CT), 0, stream, A, B, d_theta, nk, n_ell, Kp, g.tiles_m, g.tiles_n, g.per_xcd,
                     g.gm);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_contract(const double *d_A, const double *d_B, int m, int n, long k, double *d_C, hipStream_t stream) {
  if (m == 0 || n == 0) return 0;
  const deb::ContractGrid g = deb::contract_grid(m, n);
  hipLaunchKernelGGL(deb::contract_kernel, dim3(g.blocks), dim3(deb::CT), 0, stream, d_A, d_B, d_C, m, n, k, g.tiles_m, g.tiles_n, g.per_xcd,
                     g.gm);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_cell(const double *d_theta, const double *d_theta2, const double *d_k, int n_k, int n_ell, double n_s, double k_p,
                               double *d_cl, hipStream_t stream) {
  hipLaunchKernelGGL(deb::cell_kernel, dim3((n_ell + 127) / 128), dim3(128), 0, stream, d_theta, d_theta2 ? d_theta2 : d_theta, d_k, n_k,
                     n_ell, n_s, k_p, d_cl);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_dell(const double *d_cl, int ellmax, double A_s, double Tcmb, double *d_ell, double *d_dl, hipStream_t stream) {
  if (ellmax < 2) return 0;
  hipLaunchKernelGGL(deb::dell_kernel, dim3((ellmax + 127) / 128), dim3(128), 0, stream, d_cl, ellmax, A_s, Tcmb, d_ell, d_dl);
  return (int)hipGetLastError();
}
