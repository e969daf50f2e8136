// deb_perturb_dev.hpp -- device building blocks of the per-mode Rodas5 solver, shared by the
// forward kernels (deb_perturb.hip) and the discrete adjoint (deb_adjoint.hip).
//
// deb_perturb.hip -- synchronous-gauge perturbations of one k-mode per workgroup (one, two or
// four wavefronts), integrated with the linearly implicit Rodas5 method under a PI step-size
// controller.
//
// Follows the reference's evolve_one_mode / evolve_modes_batched (src/discoeb/perturbations.py:
// 882-1120), model_synchronous (:88-525), approximations.py, Rodas5Transformed
// (ode_integrators_stiff.py:678-892) and diffrax's PIDController + SaveAt(ts) with linear
// interpolation between accepted steps.
//
// Solve of W = I/(gamma h) - J, exact up to rounding, exploiting the structure of J:
//   * row 0 (the scale factor) depends on a only, so x_0 = b_0 / W_00 is solved first and the
//     column of a is moved to the right-hand side;
//   * every Boltzmann hierarchy tail (l >= 3: photon temperature, photon polarisation,
//     massless neutrinos, each massive-neutrino momentum bin) is a tridiagonal chain that
//     touches the rest of the system only through its l = 2 variable; a bottom-up sweep
//     eliminates each chain into the dense block;
//   * the dense block D (metric, CDM, baryons, l <= 2 moments, dark energy: 17 + 3 nq
//     unknowns) receives one diagonal correction per chain (Schur complement); with a
//     compile-time bin count its massive-neutrino rows are eliminated through h' and eta'
//     (bordered reduction, reduce_dense) and the remaining 19 x 19 system is inverted by one
//     wavefront (Gauss-Jordan with partial pivoting, rows in registers).  Each of the 8 stage
//     solves is then two chain sweeps and small mat-vecs on wavefront 0, while wavefront 1
//     evaluates the next stage's background quantities.
// The Jacobian column of a and df/dtau come from dual-number evaluations of the RHS; the
// dense block's columns from lane-parallel evaluation of the (linear) RHS rows on unit
// vectors; chain entries are the hierarchy coefficients themselves.
#pragma once
#include <hip/hip_runtime.h>

#include "deb_cosmo.hpp"
#include "deb_perturb.h"
#include "deb_source_pt.hpp"

namespace deb {
namespace {

constexpr int WAVE = 64;
constexpr int NQMAX = 16;
constexpr int VQ_LANE0 = 16;  // lanes 16..16+nq-1 hold the momentum bins' velocity factors
constexpr int NCORE = 17;     // dense-block unknowns outside the massive-neutrino l = 0, 1, 2 rows
constexpr int NRED = NCORE + 2;  // reduced system: those plus h' and eta'

// Rodas5 tableau (transformed form), ode_integrators_stiff.py:680-745
namespace r5 {
constexpr double gamma = 0.19;
constexpr double a21 = 2.0, a31 = 3.040894194418781, a32 = 1.041747909077569, a41 = 2.576417536461461,
                 a42 = 1.622083060776640, a43 = -0.9089668560264532, a51 = 2.760842080225597, a52 = 1.446624659844071,
                 a53 = -0.3036980084553738, a54 = 0.2877498600325443, a61 = -14.09640773051259, a62 = 6.925207756232704,
                 a63 = -41.47510893210728, a64 = 2.343771018586405, a65 = 24.13215229196062;
constexpr double C21 = -10.31323885133993, C31 = -21.04823117650003, C32 = -7.234992135176716, C41 = 32.22751541853323,
                 C42 = -4.943732386540191, C43 = 19.44922031041879, C51 = -20.69865579590063, C52 = -8.816374604402768,
                 C53 = 1.260436877740897, C54 = -0.7495647613787146, C61 = -46.22004352711257, C62 = -17.49534862857472,
                 C63 = -289.6389582892057, C64 = 93.60855400400906, C65 = 318.3822534212147, C71 = 34.20013733472935,
                 C72 = -14.15535402717690, C73 = 57.82335640988400, C74 = 25.83362985412365, C75 = 1.408950972071624,
                 C76 = -6.551835421242162, C81 = 42.57076742291101, C82 = -13.80770672017997, C83 = 93.98938432427124,
                 C84 = 18.77919633714503, C85 = -31.58359187223370, C86 = -6.685968952921985, C87 = -5.810979938412932;
constexpr double c2 = 0.38, c3 = 0.3878509998321533, c4 = 0.4839718937873840, c5 = 0.4570477008819580;
constexpr double d1 = gamma, d2 = -0.1823079225333714636, d3 = -0.319231832186874912, d4 = 0.3449828624725343,
                 d5 = -0.377417564392089818;
}  // namespace r5

// ---------------------------------------------------------------------------------------
// state layout (perturbations.py:134-186, 893)
// ---------------------------------------------------------------------------------------
struct PLay {
  int lmaxg, lmaxgp, lmaxr, lmaxnu, nq, n, idxg, idxgp, idxr, iq0, m, nch;
  const double *q, *w, *dlfdlq;  // momentum bins (LDS)
};

// ---------------------------------------------------------------------------------------
// background and thermodynamics at (tau, a) for one mode
// ---------------------------------------------------------------------------------------
// value of lane src (uniform src) of every wave lane, whatever the current exec mask
__device__ __forceinline__ double lane_bcast(double x, int src) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src), hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int N> __device__ __forceinline__ Dn<N> lane_bcast(const Dn<N> &x, int src) {
  Dn<N> r;
  r.v = lane_bcast(x.v, src);
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = lane_bcast(x.d[i], src);
  return r;
}

template <class T> struct PBg {
  T a, tau, aH, aHH, cs2, xe, xeprime, opac, tauc, taucprime, pb43, rho_Q, w_Q, ca2_Q;
  T amnu_a;  // a m_nu / T_nu (in units of the momentum grid)
  T ia, ia2, iaH, ipb, ipbm;  // 1 / a, 1 / a^2, 1 / max(aH, 1e-30), 1 / (1 + pb43), 1 / max(pb43, 1e-30)
  T vl;      // lane VQ_LANE0 + q: velocity factor of bin q (q < nq), computed once per evaluation
  T ivl;     // the same lanes: its reciprocal, eps / q = sqrt(1 + (a m / q)^2)
  bool rsa, ufa, tca;
  // velocity factor q / eps of momentum bin q
  __device__ __forceinline__ T vq(const PLay &L, int q) const {
    T aq = amnu_a / L.q[q];
    return 1.0 / dsqrt(1.0 + aq * aq);
  }
  // the same, read from lane VQ_LANE0 + q: q uniform across the wave
  __device__ __forceinline__ T vq_u(int q) const { return lane_bcast(vl, VQ_LANE0 + q); }
  __device__ __forceinline__ T ivq_u(int q) const { return lane_bcast(ivl, VQ_LANE0 + q); }
  // the same for a bin that may differ between lanes (selected among the broadcasts; an array
  // indexed by a run-time bin would live in scratch memory)
  __device__ __forceinline__ T vq_l(const PLay &L, int q) const {
    T r = lane_bcast(vl, VQ_LANE0);
    for (int j = 1; j < L.nq; ++j) {
      const T t = lane_bcast(vl, VQ_LANE0 + j);
      if (q == j) r = t;
    }
    return r;
  }
};

// Per-wave cache of the spline intervals the RHS reads: the thermal-history splines in log a
// (x_e and c_s^2 a share knots) and the neutrino tables.  Successive evaluations move a little
// in a, so the interval and its coefficients are reused; a miss falls back to the full search,
// and the interval chosen is always the one spl_index would return.
struct SplCache {
  int i, j;
  double x[2], ye[2], Se[2], yc[2], Sc[2];      // log-a knots, x_e, c_s^2 a
  double nx[2], nr[2], nSr[2], np[2], nSp[2];   // neutrino tables
};

// coefficients of the current intervals, in registers of every thread
struct SplCoef {
  double x[2], ye[2], Se[2], yc[2], Sc[2], nx[2], nr[2], nSr[2], np[2], nSp[2];
};

// Orders the LDS accesses of one wavefront: what a lane wrote before it is visible to every
// lane of the same wavefront after it (LDS operations of one wavefront complete in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Interval coefficients for log a = la.  A hit reads the LDS entry; a miss loads from the
// tables on every thread, thread 0 republishes the entry and a barrier orders that write before
// any later read.  The hit/miss decision is uniform across the workgroup (same inputs, same
// entry), so the barriers are reached by all threads.  WL: called by one wavefront alone while
// no other wavefront uses the cache; its lane 0 writes and wave_sync orders the accesses.
template <bool WL = false>
__device__ __forceinline__ void cache_lookup(const BgView &b, SplCache *c, double la, int tid, SplCoef &k) {
  const int N = b.N;
  bool refreshed = false;
  int i = c->i;
  bool hit = (i >= 0) && (i == 0 || c->x[0] < la) && (i == N - 2 || la <= c->x[1]);
  if (hit) {
    k.x[0] = c->x[0]; k.x[1] = c->x[1]; k.ye[0] = c->ye[0]; k.ye[1] = c->ye[1]; k.Se[0] = c->Se[0]; k.Se[1] = c->Se[1];
    k.yc[0] = c->yc[0]; k.yc[1] = c->yc[1]; k.Sc[0] = c->Sc[0]; k.Sc[1] = c->Sc[1];
  } else {
    i = spl_index(b.loga, N, la);
    k.x[0] = b.loga[i]; k.x[1] = b.loga[i + 1];
    k.ye[0] = b.xe[i]; k.ye[1] = b.xe[i + 1]; k.Se[0] = b.S_xe_of_loga[i]; k.Se[1] = b.S_xe_of_loga[i + 1];
    k.yc[0] = b.cs2a[i]; k.yc[1] = b.cs2a[i + 1]; k.Sc[0] = b.S_cs2a_of_loga[i]; k.Sc[1] = b.S_cs2a_of_loga[i + 1];
    refreshed = true;
  }
  int j = spl_index_uniform(b.nu_x0, b.nu_dxinv, DEB_NNU, la);
  bool nhit = (j == c->j);
  if (nhit) {
    k.nx[0] = c->nx[0]; k.nx[1] = c->nx[1]; k.nr[0] = c->nr[0]; k.nr[1] = c->nr[1]; k.nSr[0] = c->nSr[0]; k.nSr[1] = c->nSr[1];
    k.np[0] = c->np[0]; k.np[1] = c->np[1]; k.nSp[0] = c->nSp[0]; k.nSp[1] = c->nSp[1];
  } else {
    k.nx[0] = b.nu_loga[j]; k.nx[1] = b.nu_loga[j + 1];
    k.nr[0] = b.nu_lr[j]; k.nr[1] = b.nu_lr[j + 1]; k.nSr[0] = b.nu_lrS[j]; k.nSr[1] = b.nu_lrS[j + 1];
    k.np[0] = b.nu_lp[j]; k.np[1] = b.nu_lp[j + 1]; k.nSp[0] = b.nu_lpS[j]; k.nSp[1] = b.nu_lpS[j + 1];
    refreshed = true;
  }
  if (refreshed) {
    if constexpr (WL) wave_sync(); else __syncthreads();  // every thread has read the old entry before it is overwritten
    if ((WL ? (tid & (WAVE - 1)) : tid) == 0) {
      if (!hit) {
        c->i = i; c->x[0] = k.x[0]; c->x[1] = k.x[1]; c->ye[0] = k.ye[0]; c->ye[1] = k.ye[1]; c->Se[0] = k.Se[0];
        c->Se[1] = k.Se[1]; c->yc[0] = k.yc[0]; c->yc[1] = k.yc[1]; c->Sc[0] = k.Sc[0]; c->Sc[1] = k.Sc[1];
      }
      if (!nhit) {
        c->j = j; c->nx[0] = k.nx[0]; c->nx[1] = k.nx[1]; c->nr[0] = k.nr[0]; c->nr[1] = k.nr[1]; c->nSr[0] = k.nSr[0];
        c->nSr[1] = k.nSr[1]; c->np[0] = k.np[0]; c->np[1] = k.np[1]; c->nSp[0] = k.nSp[0]; c->nSp[1] = k.nSp[1];
      }
    }
    if constexpr (WL) wave_sync(); else __syncthreads();  // the new entry is visible before any later lookup
  }
}

template <class T>
__device__ __forceinline__ T cubic_at(double x0, double x1, double y0, double y1, double S0, double S1, const T &xq) {
  double hl = x1 - x0;
  T t = (xq - x0) / hl;
  T A = 1.0 - t;
  T B = t;
  return A * y0 + B * y1 + ((A * A * A - A) * S0 + (B * B * B - B) * S1) * (hl * hl / 6.0);
}
template <class T>
__device__ __forceinline__ T cubic_deriv_at(double x0, double x1, double y0, double y1, double S0, double S1, const T &xq) {
  double hl = x1 - x0;
  T d = xq - x0;
  double bb = (y1 - y0) / hl - hl * (S1 + 2.0 * S0) / 6.0;
  return bb + 2.0 * (S0 / 2.0) * d + 3.0 * ((S1 - S0) / (6.0 * hl)) * d * d;
}

template <class T>
__device__ __forceinline__ void pbg_eval(const BgView &b, const PLay &L, double k, const T &tau, const T &a, PBg<T> &g, SplCache *sc, int tid) {
  g.a = a;
  g.tau = tau;
  T loga = dlog(a);
  SplCoef c;
  cache_lookup(b, sc, val(loga), tid, c);
  g.cs2 = cubic_at(c.x[0], c.x[1], c.yc[0], c.yc[1], c.Sc[0], c.Sc[1], loga) / a;
  g.xe = cubic_at(c.x[0], c.x[1], c.ye[0], c.ye[1], c.Se[0], c.Se[1], loga);
  T dxe = cubic_deriv_at(c.x[0], c.x[1], c.ye[0], c.ye[1], c.Se[0], c.Se[1], loga);
  g.pb43 = 4.0 / 3.0 * (b.grhog / (b.grhom * b.Omegab * a));
  T rhonu = dexp(cubic_at(c.nx[0], c.nx[1], c.nr[0], c.nr[1], c.nSr[0], c.nSr[1], loga));
  T pnu = dexp(cubic_at(c.nx[0], c.nx[1], c.np[0], c.np[1], c.nSp[0], c.nSp[1], loga));
  g.w_Q = b.w0 + b.wa * (1.0 - a);
  g.rho_Q = b.rho_Q(a);
  T a2 = a * a;
  T grho = b.grhom * b.Omegam / a + (b.grhog + b.grhor * (b.Neff + b.Nmnu * rhonu)) / a2 + b.grhom * b.OmegaDE * g.rho_Q * a2 +
           b.grhom * b.Omegak;
  g.aH = dsqrt(grho / 3.0);
  g.xeprime = dxe * g.aH;
  T gpres = ((b.grhog + b.grhor * b.Neff) / 3.0 + b.grhor * b.Nmnu * pnu) / a2 + g.w_Q * b.grhom * b.OmegaDE * g.rho_Q * a2;
  g.aHH = 0.5 * (g.aH * g.aH - gpres);
  T wprime = -b.wa * g.aH * a;
  g.ca2_Q = g.w_Q - wprime / 3.0 / ((1.0 + g.w_Q) + 1e-6) / g.aH;
  g.opac = g.xe * b.akthom / a2;
  g.tauc = 1.0 / dmax(g.opac, 1e-30);
  g.taucprime = g.tauc * (2.0 * g.aH - g.xeprime / dmax(g.xe, 1e-30));
  // approximation switches (approximations.py:50-75)
  double tv = val(tau), tc = val(g.tauc), th = 1.0 / fmax(val(g.aH), 1e-30), tk = 1.0 / fmax(k, 1e-30);
  g.rsa = b.use_rsa && (tc / fmax(tv, 1e-30) > b.rsa_c) && (tv * k > b.rsa_k);
  g.ufa = !g.rsa && b.use_ufa && (tv * k > b.ufa_k);
  bool uncoupled = (tc / fmax(tk, 1e-30) > b.tca_k) || ((tc / fmax(th, 1e-30) > b.tca_h) && (tc / fmax(tk, 1e-30) > 0.1 * b.tca_k));
  g.tca = !g.rsa && b.use_tca && !uncoupled;
  g.amnu_a = a * b.amnu;
  g.ia = 1.0 / a;
  g.ia2 = 1.0 / a2;
  g.iaH = 1.0 / dmax(g.aH, 1e-30);
  g.ipb = 1.0 / (1.0 + g.pb43);
  g.ipbm = 1.0 / dmax(g.pb43, 1e-30);
  const int lane = tid & (WAVE - 1);  // all lanes take part here: the bins are read from lanes VQ_LANE0..
  const int qb = lane - VQ_LANE0;
  const T aq = g.amnu_a / L.q[(qb >= 0 && qb < L.nq) ? qb : 0];
  g.ivl = dsqrt(1.0 + aq * aq);
  g.vl = 1.0 / g.ivl;
}

// pbg_eval for plain doubles with the independent transcendentals and divisions spread over
// lanes: one exp serves rho_nu, p_nu (lanes 0, 1) and the dark-energy density (lane 2, as
// exp(-3 (1 + w0 + wa) ln a + 3 (a - 1) wa)); then five rounds of one division per lane and
// one square root, each result broadcast with readlane.  Every other quantity is the same
// expression as in pbg_eval.  All 64 lanes of every wave must be active.  WL: one wavefront
// evaluates alone (see cache_lookup).
template <bool WL = false>
__device__ __forceinline__ void pbg_eval_lanes(const BgView &b, const PLay &L, double k, double tau, double a, PBg<double> &g,
                                               SplCache *sc, int tid) {
  const int lane = tid & (WAVE - 1);
  const int qb = lane - VQ_LANE0;
  const bool binlane = qb >= 0 && qb < L.nq;
  g.a = a;
  g.tau = tau;
  const double loga = log(a);
  SplCoef c;
  cache_lookup<WL>(b, sc, loga, tid, c);
  const double cs2c = cubic_at(c.x[0], c.x[1], c.yc[0], c.yc[1], c.Sc[0], c.Sc[1], loga);
  g.xe = cubic_at(c.x[0], c.x[1], c.ye[0], c.ye[1], c.Se[0], c.Se[1], loga);
  const double dxe = cubic_deriv_at(c.x[0], c.x[1], c.ye[0], c.ye[1], c.Se[0], c.Se[1], loga);
  const double a2 = a * a;
  g.w_Q = b.w0 + b.wa * (1.0 - a);
  g.amnu_a = a * b.amnu;
  // one exp: lanes 0, 1 the neutrino density and pressure splines, lane 2 the dark energy
  const bool lp = lane == 1;
  double x = cubic_at(c.nx[0], c.nx[1], lp ? c.np[0] : c.nr[0], lp ? c.np[1] : c.nr[1], lp ? c.nSp[0] : c.nSr[0],
                      lp ? c.nSp[1] : c.nSr[1], loga);
  if (lane == 2) x = -3.0 * (1.0 + b.w0 + b.wa) * loga + 3.0 * (a - 1.0) * b.wa;
  const double ex = exp(x);
  const double rhonu = lane_bcast(ex, 0), pnu = lane_bcast(ex, 1);
  g.rho_Q = lane_bcast(ex, 2);
  // divisions, lane-parallel pass 1
  double num = 1.0, den = 1.0;
  switch (lane) {
    case 0: num = cs2c; den = a; break;
    case 1: num = b.grhom * b.Omegam; den = a; break;
    case 2: num = b.grhog + b.grhor * (b.Neff + b.Nmnu * rhonu); den = a2; break;
    case 3: num = b.grhog; den = b.grhom * b.Omegab * a; break;
    case 4: num = b.grhog + b.grhor * b.Neff; den = 3.0; break;
    case 5: num = g.xe * b.akthom; den = a2; break;
    case 6: num = 1.0; den = fmax(k, 1e-30); break;
    case 7: num = 1.0; den = a; break;
    case 8: num = 1.0; den = a2; break;
    default: if (binlane) { num = g.amnu_a; den = L.q[qb]; } break;
  }
  double r = num / den;
  g.cs2 = lane_bcast(r, 0);
  const double t1 = lane_bcast(r, 1), t2 = lane_bcast(r, 2);
  g.pb43 = 4.0 / 3.0 * lane_bcast(r, 3);
  const double gin = lane_bcast(r, 4);
  g.opac = lane_bcast(r, 5);
  const double tk = lane_bcast(r, 6);
  g.ia = lane_bcast(r, 7);
  g.ia2 = lane_bcast(r, 8);
  const double aq = r;  // bin lanes
  const double grho = t1 + t2 + b.grhom * b.OmegaDE * g.rho_Q * a2 + b.grhom * b.Omegak;
  // division pass 2
  num = 1.0; den = 1.0;
  switch (lane) {
    case 0: num = grho; den = 3.0; break;
    case 1: num = gin + b.grhor * b.Nmnu * pnu; den = a2; break;
    case 2: num = 1.0; den = fmax(g.opac, 1e-30); break;
    case 3: num = 1.0; den = 1.0 + g.pb43; break;
    case 4: num = 1.0; den = fmax(g.pb43, 1e-30); break;
    default: if (binlane) num = 1.0 + aq * aq; break;
  }
  r = num / den;
  const double gp1 = lane_bcast(r, 1);
  g.tauc = lane_bcast(r, 2);
  g.ipb = lane_bcast(r, 3);
  g.ipbm = lane_bcast(r, 4);
  const double sq = sqrt(r);  // lane 0: aH; bin lanes: sqrt(1 + aq^2)
  g.ivl = sq;
  g.aH = lane_bcast(sq, 0);
  g.xeprime = dxe * g.aH;
  const double gpres = gp1 + g.w_Q * b.grhom * b.OmegaDE * g.rho_Q * a2;
  g.aHH = 0.5 * (g.aH * g.aH - gpres);
  const double wprime = -b.wa * g.aH * a;
  const double tv = tau, tc = g.tauc;
  // division pass 3
  num = 1.0; den = 1.0;
  switch (lane) {
    case 0: num = wprime; den = 3.0; break;
    case 1: num = g.xeprime; den = fmax(g.xe, 1e-30); break;
    case 2: num = 1.0; den = fmax(g.aH, 1e-30); break;
    case 3: num = tc; den = fmax(tv, 1e-30); break;
    case 4: num = tc; den = fmax(tk, 1e-30); break;
    default: if (binlane) den = sq; break;
  }
  r = num / den;
  const double w3 = lane_bcast(r, 0);
  g.taucprime = g.tauc * (2.0 * g.aH - lane_bcast(r, 1));
  const double th = lane_bcast(r, 2), r_tv = lane_bcast(r, 3), r_tk = lane_bcast(r, 4);
  g.iaH = th;
  g.vl = r;  // bin lanes: 1 / sqrt(1 + aq^2)
  // division pass 4
  num = 1.0; den = 1.0;
  switch (lane) {
    case 0: num = w3; den = (1.0 + g.w_Q) + 1e-6; break;
    case 1: num = tc; den = fmax(th, 1e-30); break;
    default: break;
  }
  r = num / den;
  const double w4 = lane_bcast(r, 0), r_th = lane_bcast(r, 1);
  g.ca2_Q = g.w_Q - w4 / g.aH;  // division pass 5: one division, every lane
  // approximation switches (approximations.py:50-75)
  g.rsa = b.use_rsa && (r_tv > b.rsa_c) && (tv * k > b.rsa_k);
  g.ufa = !g.rsa && b.use_ufa && (tv * k > b.ufa_k);
  bool uncoupled = (r_tk > b.tca_k) || ((r_th > b.tca_h) && (r_tk > 0.1 * b.tca_k));
  g.tca = !g.rsa && b.use_tca && !uncoupled;
}

// A background that one lane evaluates for its own (tau, a), independently of the other lanes
// (the output-parallel source epilogue): the bins' velocity factors are the lane's own.
template <int NQ> struct PBgOwn : PBg<double> {
  static constexpr int NB = NQ > 0 ? NQ : NQMAX;
  double vo[NB], ivo[NB];
  __device__ __forceinline__ double vq_u(int q) const { return vo[q]; }
  __device__ __forceinline__ double ivq_u(int q) const { return ivo[q]; }
};

// pbg_eval_lanes for one lane alone: the same expressions, each evaluated by this lane (no lane
// broadcasts, no shared spline-interval cache: the intervals come from spl_index and
// spl_index_uniform, the indices a cache hit would hold), so the fields agree with
// pbg_eval_lanes to the bit.  Lanes may differ in (tau, a) and may be inactive.
template <int NQ>
__device__ __forceinline__ void pbg_eval_own(const BgView &b, const PLay &L, double k, double tau, double a, PBgOwn<NQ> &g) {
  g.a = a;
  g.tau = tau;
  const double loga = log(a);
  const int i = spl_index(b.loga, b.N, loga);
  const double x0 = b.loga[i], x1 = b.loga[i + 1];
  const double cs2c = cubic_at(x0, x1, b.cs2a[i], b.cs2a[i + 1], b.S_cs2a_of_loga[i], b.S_cs2a_of_loga[i + 1], loga);
  g.xe = cubic_at(x0, x1, b.xe[i], b.xe[i + 1], b.S_xe_of_loga[i], b.S_xe_of_loga[i + 1], loga);
  const double dxe = cubic_deriv_at(x0, x1, b.xe[i], b.xe[i + 1], b.S_xe_of_loga[i], b.S_xe_of_loga[i + 1], loga);
  const int j = spl_index_uniform(b.nu_x0, b.nu_dxinv, DEB_NNU, loga);
  const double a2 = a * a;
  g.w_Q = b.w0 + b.wa * (1.0 - a);
  g.amnu_a = a * b.amnu;
  const double rhonu = exp(cubic_at(b.nu_loga[j], b.nu_loga[j + 1], b.nu_lr[j], b.nu_lr[j + 1], b.nu_lrS[j], b.nu_lrS[j + 1], loga));
  const double pnu = exp(cubic_at(b.nu_loga[j], b.nu_loga[j + 1], b.nu_lp[j], b.nu_lp[j + 1], b.nu_lpS[j], b.nu_lpS[j + 1], loga));
  const double xq = -3.0 * (1.0 + b.w0 + b.wa) * loga + 3.0 * (a - 1.0) * b.wa;
  g.rho_Q = exp(xq);
  // division pass 1
  double num = cs2c, den = a;
  g.cs2 = num / den;
  num = b.grhom * b.Omegam;
  const double t1 = num / den;
  num = b.grhog + b.grhor * (b.Neff + b.Nmnu * rhonu);
  const double t2 = num / a2;
  num = b.grhog;
  den = b.grhom * b.Omegab * a;
  g.pb43 = 4.0 / 3.0 * (num / den);
  num = b.grhog + b.grhor * b.Neff;
  const double gin = num / 3.0;
  num = g.xe * b.akthom;
  g.opac = num / a2;
  const double tk = 1.0 / fmax(k, 1e-30);
  g.ia = 1.0 / a;
  g.ia2 = 1.0 / a2;
  const double grho = t1 + t2 + b.grhom * b.OmegaDE * g.rho_Q * a2 + b.grhom * b.Omegak;
  // division pass 2
  num = gin + b.grhor * b.Nmnu * pnu;
  const double gp1 = num / a2;
  g.tauc = 1.0 / fmax(g.opac, 1e-30);
  g.ipb = 1.0 / (1.0 + g.pb43);
  g.ipbm = 1.0 / fmax(g.pb43, 1e-30);
  g.aH = sqrt(grho / 3.0);
#pragma unroll
  for (int q = 0; q < PBgOwn<NQ>::NB; ++q) {
    if (q >= L.nq) break;
    const double aq = g.amnu_a / L.q[q];
    num = 1.0 + aq * aq;
    g.ivo[q] = sqrt(num / 1.0);
    g.vo[q] = 1.0 / g.ivo[q];
  }
  g.vl = 0.0;
  g.ivl = 0.0;
  g.xeprime = dxe * g.aH;
  const double gpres = gp1 + g.w_Q * b.grhom * b.OmegaDE * g.rho_Q * a2;
  g.aHH = 0.5 * (g.aH * g.aH - gpres);
  const double wprime = -b.wa * g.aH * a;
  const double tv = tau, tc = g.tauc;
  // division pass 3
  const double w3 = wprime / 3.0;
  g.taucprime = g.tauc * (2.0 * g.aH - g.xeprime / fmax(g.xe, 1e-30));
  const double th = 1.0 / fmax(g.aH, 1e-30);
  const double r_tv = tc / fmax(tv, 1e-30), r_tk = tc / fmax(tk, 1e-30);
  g.iaH = th;
  // division pass 4
  const double w4 = w3 / ((1.0 + g.w_Q) + 1e-6), r_th = tc / fmax(th, 1e-30);
  g.ca2_Q = g.w_Q - w4 / g.aH;
  // approximation switches (approximations.py:50-75)
  g.rsa = b.use_rsa && (r_tv > b.rsa_c) && (tv * k > b.rsa_k);
  g.ufa = !g.rsa && b.use_ufa && (tv * k > b.ufa_k);
  bool uncoupled = (r_tk > b.tca_k) || ((r_th > b.tca_h) && (r_tk > 0.1 * b.tca_k));
  g.tca = !g.rsa && b.use_tca && !uncoupled;
}

// A PBg<double> in LDS, handed from the wavefront that evaluated it to the whole workgroup:
// the uniform fields, the per-bin velocity factors (lanes VQ_LANE0 + q) and the switches.
// x0 carries the row-0 unknown of the current stage solve.
struct PBgShared {
  double f[20];
  double vl[NQMAX], ivl[NQMAX];
  double x0;
  int flags;
};

__device__ __forceinline__ void pbg_store(const PBg<double> &g, PBgShared *s, const PLay &L, int lane) {
  if (lane == 0) {
    s->f[0] = g.a; s->f[1] = g.tau; s->f[2] = g.aH; s->f[3] = g.aHH; s->f[4] = g.cs2; s->f[5] = g.xe; s->f[6] = g.xeprime;
    s->f[7] = g.opac; s->f[8] = g.tauc; s->f[9] = g.taucprime; s->f[10] = g.pb43; s->f[11] = g.rho_Q; s->f[12] = g.w_Q;
    s->f[13] = g.ca2_Q; s->f[14] = g.amnu_a;
    s->f[15] = g.ia; s->f[16] = g.ia2; s->f[17] = g.iaH; s->f[18] = g.ipb; s->f[19] = g.ipbm;
    s->flags = (g.rsa ? 1 : 0) | (g.ufa ? 2 : 0) | (g.tca ? 4 : 0);
  }
  const int qb = lane - VQ_LANE0;
  if (qb >= 0 && qb < L.nq) { s->vl[qb] = g.vl; s->ivl[qb] = g.ivl; }
}

__device__ __forceinline__ void pbg_load(PBg<double> &g, const PBgShared *s, const PLay &L, int lane) {
  g.a = s->f[0]; g.tau = s->f[1]; g.aH = s->f[2]; g.aHH = s->f[3]; g.cs2 = s->f[4]; g.xe = s->f[5]; g.xeprime = s->f[6];
  g.opac = s->f[7]; g.tauc = s->f[8]; g.taucprime = s->f[9]; g.pb43 = s->f[10]; g.rho_Q = s->f[11]; g.w_Q = s->f[12];
  g.ca2_Q = s->f[13]; g.amnu_a = s->f[14];
  g.ia = s->f[15]; g.ia2 = s->f[16]; g.iaH = s->f[17]; g.ipb = s->f[18]; g.ipbm = s->f[19];
  const int fl = s->flags;
  g.rsa = fl & 1; g.ufa = (fl & 2) != 0; g.tca = (fl & 4) != 0;
  const int qb = lane - VQ_LANE0;
  const bool bin = qb >= 0 && qb < L.nq;  // only the bin lanes are ever read
  g.vl = bin ? s->vl[qb] : 0.0;
  g.ivl = bin ? s->ivl[qb] : 0.0;
}

// metric perturbations and effective (RSA/TCA) fields, perturbations.py:292-351,407-427
template <class T> struct Met {
  T hprime, etaprime, alpha, dgrho, dgpres, dg, tg, sg, dr, tr, sr, dbp_tca, tbp_tca, dgp_tca, tgp_tca;
};

// BV: BgView, or a view whose density parameters carry derivatives (the adjoint's); G: PBg<T>
// or a type derived from it (its vq_u / ivq_u give the bins' velocity factors)
template <class T, class Y, class BV, class G>
__device__ __forceinline__ void metric(const BV &b, const PLay &L, double k, const G &g, const Y &y, Met<T> &m) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0;
  const auto Omegac = b.Omegam - b.Omegab;
  const double eta = y(2), deltac = y(3), thetac = y(4), deltab = y(5), thetab = y(6);
  const double deltag = y(7), thetag = y(8), shearg = y(9) / 2.0;
  const double deltar = y(L.idxr), thetar = y(L.idxr + 1), shearr = y(L.idxr + 2) / 2.0;
  const double deltaq = y(n - 2), thetaq = y(n - 1);
  const double ik2 = 1.0 / (k * k);
  T drhonu(0.0), dpnu(0.0), fnu(0.0), shearnu(0.0);
  for (int q = 0; q < nq; ++q) {
    double p0 = y(iq0 + q), p1 = y(iq0 + nq + q), p2 = y(iq0 + 2 * nq + q);
    const T v = g.vq_u(q), iv = g.ivq_u(q);
    drhonu += L.w[q] * p0 * iv;
    dpnu += L.w[q] * p0 * v;
    fnu += T(L.w[q] * p1);
    shearnu += L.w[q] * p2 * v;
  }
  dpnu = dpnu * (1.0 / 3.0);
  shearnu = shearnu * (2.0 / 3.0);
  const T a = g.a, a2 = g.a * g.a;
  const double cs2_Q = b.cs2DE;
  T rpptq = (1.0 + g.w_Q) * g.rho_Q * b.grhom * b.OmegaDE * thetaq * a2;
  T de_rho = b.grhom * b.OmegaDE * deltaq * g.rho_Q * a2;
  T dgrho = b.grhom * (Omegac * deltac + b.Omegab * deltab) * g.ia + (b.grhog * deltag + b.grhor * (b.Neff * deltar + b.Nmnu * drhonu)) * g.ia2 +
            de_rho;
  T dgrho_wo = b.grhom * (Omegac * deltac + b.Omegab * deltab) * g.ia + b.grhor * b.Nmnu * drhonu * g.ia2 + de_rho;
  T hp_rsa = (2.0 * k * k * eta + dgrho_wo) * g.iaH;
  if (g.rsa) {  // compute_fields_rsa (approximations.py:28-47)
    const double sk2 = fmax(k * k, 1e-30);
    T stc = dmax(g.tauc, 1e-30);
    m.dg = 4.0 / sk2 * (g.aH * hp_rsa - sk2 * eta) + 4.0 / (sk2 * stc) * (thetab + 0.5 * hp_rsa);
    m.tg = -0.5 * hp_rsa + 3.0 / (sk2 * stc) * (-g.taucprime / stc * (thetab + 0.5 * hp_rsa) + (-g.aH * thetab + g.cs2 * sk2 * deltab + sk2 * eta));
    m.sg = T(0.0);
    m.dr = 4.0 / sk2 * (g.aH * hp_rsa - sk2 * eta);
    m.tr = -0.5 * hp_rsa;
    m.sr = T(0.0);
  } else {
    m.dg = T(deltag); m.tg = T(thetag); m.sg = T(shearg);
    m.dr = T(deltar); m.tr = T(thetar); m.sr = T(shearr);
  }
  m.dgpres = (b.grhog * m.dg + b.grhor * b.Neff * m.dr) * g.ia2 * (1.0 / 3.0) + b.grhor * b.Nmnu * dpnu * g.ia2 +
             (cs2_Q * de_rho + (cs2_Q - g.ca2_Q) * (3.0 * g.aH * rpptq * ik2));
  T dgtheta = b.grhom * (Omegac * thetac + b.Omegab * thetab) * g.ia + 4.0 / 3.0 * (b.grhog * m.tg + b.Neff * b.grhor * m.tr) * g.ia2 +
              b.Nmnu * b.grhor * k * fnu * g.ia2 + rpptq;
  m.dgrho = dgrho;
  m.hprime = g.rsa ? hp_rsa : (2.0 * k * k * eta + dgrho) * g.iaH;
  m.etaprime = 0.5 * dgtheta * ik2;
  m.alpha = (m.hprime + 6.0 * m.etaprime) * 0.5 * ik2;
  if (g.tca) {  // _photon_tca (perturbations.py:407-435)
    const double k2 = k * k;
    m.dbp_tca = -thetab - 0.5 * m.hprime;
    T tbp = (-g.aH * thetab + k2 * g.cs2 * deltab + k2 * g.pb43 * (0.25 * deltag - shearg)) * g.ipb;
    m.dgp_tca = 4.0 / 3.0 * (-thetag - 0.5 * m.hprime);
    T slip = 2.0 * g.pb43 * g.ipb * g.aH * (thetab - thetag) +
             g.tauc * (-g.aHH * thetab - 0.5 * g.aH * k2 * deltag + k2 * (g.cs2 * m.dbp_tca - 0.25 * m.dgp_tca)) * g.ipb;
    tbp = tbp + g.pb43 * g.ipb * slip;
    m.tbp_tca = tbp;
    m.tgp_tca = (-tbp - g.aH * thetab + k2 * g.cs2 * deltab) * g.ipbm + k2 * (0.25 * deltag - shearg);
  }
}

// RHS row i for i outside the l >= 3 tails (row 0 included)
template <class T, class Y, class V, class BV>
__device__ __forceinline__ T core_row(int i, const BV &b, const PLay &L, double k, const PBg<T> &g, const Met<T> &m, const Y &y,
                                      const V &vrow) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0, idxgp = L.idxgp, idxr = L.idxr;
  const double k2 = k * k;
  if (i == 0) return g.aH * g.a;
  if (i == 1) return -(m.dgrho + 3.0 * m.dgpres) * g.a;
  if (i == 2) return m.etaprime;
  if (i == 3) return -y(4) - 0.5 * m.hprime;
  if (i == 4) return -g.aH * y(4);
  if (i == 5) return -y(6) - 0.5 * m.hprime;
  if (i == 6) {
    if (g.tca) return m.tbp_tca;
    return -g.aH * y(6) + k2 * g.cs2 * y(5) + g.pb43 * g.opac * (m.tg - y(6));
  }
  if (i <= 9) {  // photon l = 0, 1, 2
    if (g.rsa) return T(0.0);
    if (g.tca) return (i == 7) ? m.dgp_tca : (i == 8 ? m.tgp_tca : T(0.0));
    if (i == 7) return 4.0 / 3.0 * (-m.tg - 0.5 * m.hprime);
    if (i == 8) return k2 * (0.25 * m.dg - m.sg) - g.opac * (m.tg - y(6));
    double polter = y(9) + y(idxgp) + y(idxgp + 2);
    return 8.0 / 15.0 * (m.tg + k2 * m.alpha) - 3.0 / 5.0 * k * y(10) - g.opac * (y(9) - 0.1 * polter);
  }
  if (i >= idxgp && i <= idxgp + 2) {  // polarisation l = 0, 1, 2
    if (g.rsa || g.tca) return T(0.0);
    double polter = y(9) + y(idxgp) + y(idxgp + 2);
    int l = i - idxgp;
    if (l == 0) return -k * y(idxgp + 1) - g.opac * y(idxgp) + g.opac * polter / 2.0;
    if (l == 1) return k * (1.0 / 3.0) * (y(idxgp) - 2.0 * y(idxgp + 2)) - g.opac * y(idxgp + 1);
    return k * 0.2 * (2.0 * y(idxgp + 1) - 3.0 * y(idxgp + 3)) - g.opac * y(idxgp + 2) + g.opac * polter * 0.1;
  }
  if (i >= idxr && i <= idxr + 2) {  // massless neutrinos l = 0, 1, 2
    if (g.rsa) return T(0.0);
    int l = i - idxr;
    if (l == 0) return 4.0 / 3.0 * (-m.tr - 0.5 * m.hprime);
    if (l == 1) return k2 * (0.25 * m.dr - m.sr);
    if (g.ufa) return -3.0 / dmax(g.tau, 1e-30) * m.sr + 4.0 / 15.0 * (m.tr + 0.5 * m.hprime);
    return 8.0 / 15.0 * (m.tr + k2 * m.alpha) - 0.6 * k * y(idxr + 3);
  }
  if (i >= iq0 && i < iq0 + 3 * nq) {  // massive neutrinos l = 0, 1, 2
    int l = (i - iq0) / nq, q = (i - iq0) - l * nq;
    T kv = k * vrow(q);
    double dl = L.dlfdlq[q];
    if (l == 0) return -kv * y(iq0 + nq + q) + m.hprime * dl * (1.0 / 6.0);
    if (l == 1) return kv * (y(iq0 + q) - 2.0 * y(iq0 + 2 * nq + q)) * (1.0 / 3.0);
    return kv * (2.0 * y(iq0 + nq + q) - 3.0 * y(iq0 + 3 * nq + q)) * 0.2 - (m.hprime * (1.0 / 15.0) + 0.4 * m.etaprime) * dl;
  }
  if (b.lambda_de) return T(0.0);
  const double cs2_Q = b.cs2DE;
  if (i == n - 2)  // BL10 (3.5)
    return -(1.0 + g.w_Q) * (y(n - 1) + 0.5 * m.hprime) - 3.0 * (cs2_Q - g.w_Q) * g.aH * y(n - 2) -
           9.0 * (1.0 + g.w_Q) * (cs2_Q - g.ca2_Q) * g.aH * g.aH / k2 * y(n - 1);
  return -(1.0 - 3.0 * cs2_Q) * g.aH * y(n - 1) + cs2_Q / (1.0 + g.w_Q) * k2 * y(n - 2);  // BL10 (3.6)
}

// tail chain coefficients: f_l = lo y_{l-1} + di y_l + up y_{l+1}, l = 3..lmax
// ca, cb: l / (2l + 1) and -(l + 1) / (2l + 1); vrow(q): velocity factor of momentum bin q
template <class T, class V>
__device__ __forceinline__ void tail_coef(int ch, int l, const PLay &L, double k, const PBg<T> &g, double ca, double cb, const V &vrow,
                                          T &lo, T &di, T &up) {
  int lmax;
  bool off;
  T kv(k), damp(0.0);
  if (ch == 0) { lmax = L.lmaxg; off = g.rsa || g.tca; damp = g.opac; }
  else if (ch == 1) { lmax = L.lmaxgp; off = g.rsa || g.tca; damp = g.opac; }
  else if (ch == 2) { lmax = L.lmaxr; off = g.rsa || g.ufa; }
  else { lmax = L.lmaxnu; off = false; kv = k * vrow(ch - 3); }
  if (off) { lo = T(0.0); di = T(0.0); up = T(0.0); return; }
  if (l < lmax) {
    lo = kv * ca;
    up = kv * cb;
    di = -damp;
  } else {
    lo = kv;
    up = T(0.0);
    di = -(double)(lmax + 1) / g.tau - damp;
  }
}

// ---------------------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------------------
// per-unit LDS workspace (shared by the workgroup's threads)
struct Ws {
  double *y0, *y1, *u, *f, *J0, *dT, *K, *S, *Jh, *tlo, *tdi, *tup, *tmul, *td, *xd;
  double *Sinv;         // the buffer holding the dense (or reduced) inverse
  double *Jd;           // diagonal of the dense block's Jacobian at the current Jacobian point
  int *dmap, *rk, *chead, *cstride, *clen, *cpar, *piv, *pi, *rmap;
  int *rs, *rbin;       // per row: chain stride (tail rows), momentum bin (massive-neutrino rows)
  int *perm;            // RHS row of each thread slot (slot tid + it NT), -1 for none
  int *perm1;           // the same for the NT - 64 threads of waves 1..
  int *pch;             // per dense row: the chain whose head couples to it (its l = 2 parent), -1 for none
  double *tca, *tcb;    // per tail row l < lmax: l / (2l + 1) and -(l + 1) / (2l + 1)
  // bordered solve of the dense block (fixed block sizes): metric derivatives G (2 per dense
  // column), per-bin 3x3 inverses, Q = M^-1 U, the neutrino part of a right-hand side, the
  // reduced right-hand side and solution, the reduced matrix (inverted in place)
  double *G, *Minv, *Q, *rn, *xr, *R;
  double *P;            // [W_cn; G_n^T] M^-1 (NRED x 3 nq): the reduced right-hand side straight from b
  double *rtd;          // 1 / td of the chain pivots, for the stage solves
  PBgShared *pg;        // background of the next stage, evaluated during the current stage solve
  PBgShared *pgc, *pgn; // this step's background (Jacobian point) and the next step's, if accepted
  SplCache *sc;
};


// Row order of the RHS evaluation for nthreads threads (slot t + it nthreads of thread t): a wave
// pays for every row formula any of its lanes takes, so each class of core rows (row 0 and rows
// 1..9; polarisation, massless-neutrino and dark-energy rows l <= 2; massive-neutrino rows
// l <= 2) gets a wave-sized block of slots of its own where there are enough, and the tail
// rows fill the rest.  perm holds n + 256 entries; run by one thread.
__device__ __forceinline__ void build_perm(const PLay &L, const Ws &W, int *perm, int nthreads) {
  const int n = L.n, nq = L.nq;
  const int nit = (n + nthreads - 1) / nthreads, slots = nit * nthreads, nb = slots / WAVE;
  for (int t = 0; t < slots; ++t) perm[t] = -1;
  // classes: 0 rows 0..9; 1 and 4 the other core rows (the first three polarisation rows and
  // the first massless-neutrino row, then the rest), split over two blocks; 2 massive-neutrino
  // rows l <= 2; 3 tail rows
  auto cls = [&](int i) {
    if (W.rk[i] >= 0) return 3;
    if (i <= 9) return 0;
    if (i >= L.iq0 && i < L.iq0 + 3 * nq) return 2;
    return (i <= L.idxr) ? 1 : 4;
  };
  const int home[5] = {nb - 1, 0, nb > 2 ? 1 : 0, 0, nb > 3 ? 2 : 0};
  int used[64];  // slots taken per block (nb <= 64 for n <= 4096 rows)
  for (int bk = 0; bk < nb; ++bk) used[bk] = 0;
  const int order[5] = {0, 1, 4, 2, 3};
  for (int oc = 0; oc < 5; ++oc)
    for (int i = 0; i < n; ++i) {
      const int c = order[oc];
      if (cls(i) != c) continue;
      int bk = home[c];
      while (used[bk] == WAVE) bk = (bk + 1) % nb;
      perm[bk * WAVE + used[bk]++] = i;
    }
}

// rows of the tail chains: rk[i] = (chain << 8) | l, -1 for rows in the dense block or row 0
template <int NT>
__device__ __forceinline__ void setup_maps(const PLay &L, Ws &W, int tid) {
  const int n = L.n, nq = L.nq;
  if (tid == 0) {
    int r = 0;
    for (int i = 1; i <= 9; ++i) W.dmap[r++] = i;
    for (int i = 0; i < 3; ++i) W.dmap[r++] = L.idxgp + i;
    for (int i = 0; i < 3; ++i) W.dmap[r++] = L.idxr + i;
    for (int i = 0; i < 3 * nq; ++i) W.dmap[r++] = L.iq0 + i;
    W.dmap[r++] = n - 2;
    W.dmap[r++] = n - 1;
    // chains: head index, stride, length, dense-block row of the l = 2 parent
    W.chead[0] = L.idxg + 3; W.cstride[0] = 1; W.clen[0] = L.lmaxg - 2; W.cpar[0] = 8;
    W.chead[1] = L.idxgp + 3; W.cstride[1] = 1; W.clen[1] = L.lmaxgp - 2; W.cpar[1] = 11;
    W.chead[2] = L.idxr + 3; W.cstride[2] = 1; W.clen[2] = L.lmaxr - 2; W.cpar[2] = 14;
    for (int q = 0; q < nq; ++q) {
      W.chead[3 + q] = L.iq0 + 3 * nq + q; W.cstride[3 + q] = nq; W.clen[3 + q] = L.lmaxnu - 2; W.cpar[3 + q] = 15 + 2 * nq + q;
    }
  }
  for (int i = tid; i < n; i += NT) {
    W.rk[i] = -1; W.rs[i] = 0; W.tca[i] = 0.0; W.tcb[i] = 0.0;
    W.rbin[i] = (i >= L.iq0 && i < L.iq0 + 3 * nq) ? (i - L.iq0) % nq : 0;
  }
  __syncthreads();
  for (int c = 0; c < L.nch; ++c)
    for (int j = tid; j < W.clen[c]; j += NT) {
      const int i = W.chead[c] + j * W.cstride[c], l = 3 + j;
      const double inv = 1.0 / (double)(2 * l + 1);
      W.rk[i] = (c << 8) | l;
      W.rs[i] = W.cstride[c];
      W.rbin[i] = c >= 3 ? c - 3 : 0;
      W.tca[i] = (double)l * inv;
      W.tcb[i] = -(double)(l + 1) * inv;
    }
  __syncthreads();
  // row orders for the whole workgroup and, with more than one wavefront, for waves 1.. alone
  // (they evaluate the first stage's rows while wave 0 inverts the dense block)
  if (tid == 0) {
    build_perm(L, W, W.perm, NT);
    if (NT > WAVE) build_perm(L, W, W.perm1, NT - WAVE);
  }
  __syncthreads();
}


// full RHS f(tau, u) (lanes over rows) with the background g at (tau, u[0]); u in LDS.  NT
// threads t = 0 .. NT-1 (whole wavefronts) take the rows in the order of perm (built for NT
// threads by build_perm); emit(i, f_i) receives each row.
template <int NT, class E>
__device__ __forceinline__ void rhs_rows(const BgView &b, const PLay &L, double k, const PBg<double> &g, const double *u, const int *perm,
                                         Ws &W, int tid, const E &emit) {
  auto yv = [&](int i) { return u[i]; };
  Met<double> m;
  metric<double>(b, L, k, g, yv, m);
  const int nit = (L.n + NT - 1) / NT;
  for (int it = 0; it < nit; ++it) {  // a uniform trip count: every lane takes part in the shuffle
    const int i = perm[tid + it * NT];
    const int ic = i >= 0 ? i : 0;
    const double vsh = __shfl(g.vl, VQ_LANE0 + W.rbin[ic]);  // this row's momentum-bin factor
    auto vrow = [&](int) { return vsh; };
    if (i >= 0) {
      const int rk = W.rk[i];
      double r;
      if (rk < 0) {
        r = core_row<double>(i, b, L, k, g, m, yv, vrow);
      } else {
        const int ch = rk >> 8, l = rk & 255, s = W.rs[i];
        double lo, di, up;
        tail_coef<double>(ch, l, L, k, g, W.tca[i], W.tcb[i], vrow, lo, di, up);
        const double yn = (up != 0.0) ? u[i + s] : 0.0;
        r = lo * u[i - s] + di * u[i] + up * yn;
      }
      emit(i, r);
    }
  }
}

template <int NT>
__device__ __forceinline__ void rhs_eval(const BgView &b, const PLay &L, double k, double tau, const double *u, double *out, Ws &W, int tid) {
  PBg<double> g;
  pbg_eval_lanes(b, L, k, tau, u[0], g, W.sc, tid);
  rhs_rows<NT>(b, L, k, g, u, W.perm, W, tid, [&](int i, double r) { out[i] = r; });
}

// d f / da (into outa) and d f / dtau (into outt) at (tau, y): one pass with two dual directions
template <int NT>
__device__ __forceinline__ void rhs_dual(const BgView &b, const PLay &L, double k, double tau, const double *y, double *outa, double *outt, Ws &W,
                         int tid) {
  typedef Dn<2> D2;
  D2 td(tau), ad(y[0]);
  ad.d[0] = 1.0;
  td.d[1] = 1.0;
  PBg<D2> g;
  pbg_eval<D2>(b, L, k, td, ad, g, W.sc, tid);
  auto yv = [&](int i) { return y[i]; };
  Met<D2> m;
  metric<D2>(b, L, k, g, yv, m);
  const int nit = (L.n + NT - 1) / NT;
  for (int it = 0; it < nit; ++it) {  // a uniform trip count: every lane takes part in the shuffle
    const int i = W.perm[tid + it * NT];
    const int ic = i >= 0 ? i : 0;
    const int src = VQ_LANE0 + W.rbin[ic];
    D2 vsh;
    vsh.v = __shfl(g.vl.v, src);
    vsh.d[0] = __shfl(g.vl.d[0], src);
    vsh.d[1] = __shfl(g.vl.d[1], src);
    auto vrow = [&](int) { return vsh; };
    if (i >= 0) {
      const int rk = W.rk[i];
      D2 r;
      if (rk < 0) {
        r = core_row<D2>(i, b, L, k, g, m, yv, vrow);
      } else {
        const int ch = rk >> 8, l = rk & 255, s = W.rs[i];
        D2 lo, di, up;
        tail_coef<D2>(ch, l, L, k, g, W.tca[i], W.tcb[i], vrow, lo, di, up);
        const double yn = (up.v != 0.0 || up.d[0] != 0.0 || up.d[1] != 0.0) ? y[i + s] : 0.0;
        r = lo * y[i - s] + di * y[i] + up * yn;
      }
      outa[i] = r.d[0];
      outt[i] = r.d[1];
    }
  }
}

// Wave-wide maximum of a 64-bit unsigned value, returned to every lane: DPP row shifts by 1, 2,
// 4 and 8 leave each row's maximum in its lane 15, row broadcasts of lanes 15 and 31 carry it
// to lane 63, which readlane returns.  Lanes a DPP move does not write keep their own value.
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long umax_dpp(unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(unsigned)(v >> 32);
  const unsigned tlo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
  const unsigned thi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
  const unsigned long long t = ((unsigned long long)thi << 32) | tlo;
  return t > v ? t : v;
}
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
  v = umax_dpp<0x111, 0xf>(v);  // row_shr:1
  v = umax_dpp<0x112, 0xf>(v);  // row_shr:2
  v = umax_dpp<0x114, 0xf>(v);  // row_shr:4
  v = umax_dpp<0x118, 0xf>(v);  // row_shr:8
  v = umax_dpp<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v = umax_dpp<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, WAVE - 1), hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), WAVE - 1);
  return ((unsigned long long)hi << 32) | lo;
}

// Gauss-Jordan inversion of the N x N row-major matrix A (LDS, N <= 64 known at compile time)
// by one wavefront, in registers: no barriers and no LDS traffic between pivots.  With G =
// min(3, 64 / N) column groups, lane g N + r holds columns g CW .. g CW + CW - 1 (CW = ceil(N / G))
// of row r, so that each pivot step updates CW entries per lane instead of N; the pivot row's
// entries and each row's pivot-column entry reach the other groups by cross-lane reads.  Pivot p
// is the not-yet-used row with the largest |entry| in column p: one wave maximum of a 64-bit key
// over the lanes of p's group, the magnitude's bit pattern with its six lowest bits replaced by
// 63 - lane, so that ties and near-ties go to the first row; rows are never moved.  Each step
// replaces column p by the multipliers, in place: after the last step row rmap[p] (the row that
// pivoted on column p) gives unknown p,
//   x[p] = sum_c Ainv[rmap[p]][c] b[rmap[c]],
// so a right-hand side is scattered through pi = rmap^-1.  Every entry is formed by the same
// operations as with one row per lane.  Ainv may alias A (A is read in full before any write).
template <int N>
__device__ __forceinline__ void gauss_jordan_wave(const double *A, double *Ainv, int *rmap, int *pi, int lane) {
  static_assert(N <= WAVE, "one row per lane");
  constexpr int G = (WAVE / N) < 3 ? (WAVE / N) : 3;
  constexpr int CW = (N + G - 1) / G;
  const int g = lane / N, r = lane - g * N;
  const bool live = g < G;  // lanes G N .. 63 hold nothing
  double a[CW];
#pragma unroll
  for (int c = 0; c < CW; ++c) {
    const int j = g * CW + c;
    a[c] = (live && j < N) ? A[r * N + j] : 0.0;
  }
  wave_sync();  // every lane has read A before any lane writes Ainv (they may alias)
  bool used = !live;
#pragma unroll
  for (int p = 0; p < N; ++p) {
    const int gp = p / CW, cp = p - gp * CW;  // the group and slot holding column p
    const unsigned long long mag = (unsigned long long)__double_as_longlong(a[cp]) & 0x7fffffffffffffc0ull;
    const unsigned long long key = (used || g != gp) ? 0ull : (mag | (unsigned long long)(WAVE - 1 - lane));
    const int bl = WAVE - 1 - (int)(wave_max_u64(key) & (WAVE - 1));  // the pivot's lane, in group gp
    const int bi = bl - gp * N;                                        // its row
    const double ip = 1.0 / lane_bcast(a[cp], bl);
    const bool me = live && r == bi;
    const double f = __shfl(a[cp], gp * N + r);  // this row's entry in column p
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const double pj = __shfl(a[c], g * N + bi) * ip;  // the pivot row's entry in this column, scaled
      double v = me ? pj : a[c] - f * pj;
      if (c == cp && g == gp) v = me ? ip : -f * ip;  // column p becomes the multipliers
      a[c] = v;
    }
    used = used || me;
    if (lane == 0) { rmap[p] = bi; pi[bi] = p; }
  }
  if (live)
#pragma unroll
    for (int c = 0; c < CW; ++c) {
      const int j = g * CW + c;
      if (j < N) Ainv[r * N + j] = a[c];
    }
}

// Inverse of a 3x3 matrix by Gauss-Jordan with partial pivoting, in registers.
__device__ __forceinline__ void inv3(double (&a)[3][6]) {  // a = [M | I] in, [I | M^-1] out
#pragma unroll
  for (int p = 0; p < 3; ++p) {
#pragma unroll
    for (int i = p + 1; i < 3; ++i) {
      if (fabs(a[i][p]) > fabs(a[p][p])) {
#pragma unroll
        for (int j = 0; j < 6; ++j) { const double t = a[i][j]; a[i][j] = a[p][j]; a[p][j] = t; }
      }
    }
    const double ip = 1.0 / a[p][p];
#pragma unroll
    for (int j = 0; j < 6; ++j) a[p][j] *= ip;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (i == p) continue;
      const double f = a[i][p];
#pragma unroll
      for (int j = 0; j < 6; ++j) a[i][j] -= f * a[p][j];
    }
  }
}

// dense-block index of core unknown i (i < NCORE): rows 1..9, polarisation and massless
// neutrino l = 0..2, then the two dark-energy rows after the 3 nq massive-neutrino rows
__device__ __forceinline__ int core_index(int i, int nq) { return i < 15 ? i : 3 * nq + i; }

// Bordered reduction of the dense block W (m = 17 + 3 nq unknowns).  The massive-neutrino rows
// l = 0, 1, 2 of bin q depend on other unknowns only through h' and eta' (linear functionals
// of the state, G^T x): with U the rows' coefficients of (h', eta'),
//   W_nn = M - U G_n^T (M block-diagonal, one 3x3 block per bin),  W_nc = -U G_c^T.
// Writing z = G^T x, the neutrino unknowns are x_n = M^-1 b_n + Q z (Q = M^-1 U), and
//   [ W_cc          W_cn Q      ] [x_c]   [ b_c - W_cn M^-1 b_n ]
//   [ -G_c^T   I - G_n^T Q      ] [ z ] = [ G_n^T M^-1 b_n      ]
// is a 19 x 19 system for every nq.  W (as built), G (from the metric on the unit vectors
// that give W's columns) are numeric; U is (dl/6, 0, -dl/15) for h' and (0, 0, -2 dl/5) for
// eta' in the l = 0, 1, 2 rows (the reference's massive-neutrino equations).  P = [W_cn;
// G_n^T] M^-1 is kept as well, so that a stage solve forms the reduced right-hand side from b
// in one pass.
template <int MC, int NT>
__device__ __forceinline__ void reduce_dense(const PLay &L, Ws &W, int tid) {
  constexpr int NQ = (MC - NCORE) / 3;
  constexpr int NN = 3 * NQ;
  const int m = MC;
  for (int q = tid; q < NQ; q += NT) {  // M_q = W_qq + U_q G_q^T, inverted
    const double dl = L.dlfdlq[q];
    const double uh[3] = {dl / 6.0, 0.0, -dl / 15.0}, ue[3] = {0.0, 0.0, -0.4 * dl};
    double a[3][6];
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const int row = 15 + l * NQ + q;
#pragma unroll
      for (int lp = 0; lp < 3; ++lp) {
        const int col = 15 + lp * NQ + q;
        a[l][lp] = W.S[row * m + col] + uh[l] * W.G[2 * col] + ue[l] * W.G[2 * col + 1];
        a[l][3 + lp] = (l == lp) ? 1.0 : 0.0;
      }
    }
    inv3(a);
#pragma unroll
    for (int l = 0; l < 3; ++l)
#pragma unroll
      for (int lp = 0; lp < 3; ++lp) W.Minv[9 * q + 3 * l + lp] = a[l][3 + lp];
  }
  __syncthreads();
  for (int j = tid; j < NN; j += NT) {  // Q = M^-1 U
    const int l = j / NQ, q = j - l * NQ;
    const double dl = L.dlfdlq[q];
    const double *mi = W.Minv + 9 * q + 3 * l;
    W.Q[2 * j] = mi[0] * (dl / 6.0) + mi[2] * (-dl / 15.0);
    W.Q[2 * j + 1] = mi[2] * (-0.4 * dl);
  }
  __syncthreads();
  for (int e = tid; e < NRED * NRED; e += NT) {
    const int i = e / NRED, j = e - i * NRED;
    double v;
    if (i < NCORE && j < NCORE) {
      v = W.S[core_index(i, NQ) * m + core_index(j, NQ)];
    } else if (i < NCORE) {  // W_cn Q
      const double *row = W.S + core_index(i, NQ) * m + 15;
      v = 0.0;
      for (int jn = 0; jn < NN; ++jn) v += row[jn] * W.Q[2 * jn + (j - NCORE)];
    } else if (j < NCORE) {
      v = -W.G[2 * core_index(j, NQ) + (i - NCORE)];
    } else {  // I - G_n^T Q
      v = (i == j) ? 1.0 : 0.0;
      for (int jn = 0; jn < NN; ++jn) v -= W.G[2 * (15 + jn) + (i - NCORE)] * W.Q[2 * jn + (j - NCORE)];
    }
    W.R[e] = v;
  }
  for (int e = tid; e < NRED * NN; e += NT) {  // P = [W_cn; G_n^T] M^-1, bin by bin
    const int i = e / NN, jn = e - i * NN;
    const int l = jn / NQ, q = jn - l * NQ;
    const double *mi = W.Minv + 9 * q + l;  // column l of M_q^-1
    double a0, a1, a2;
    if (i < NCORE) {
      const double *row = W.S + core_index(i, NQ) * m + 15 + q;
      a0 = row[0]; a1 = row[NQ]; a2 = row[2 * NQ];
    } else {
      const int t = i - NCORE;
      a0 = W.G[2 * (15 + q) + t]; a1 = W.G[2 * (15 + NQ + q) + t]; a2 = W.G[2 * (15 + 2 * NQ + q) + t];
    }
    W.P[e] = a0 * mi[0] + a1 * mi[3] + a2 * mi[6];
  }
  __syncthreads();
}

// Jacobian pieces at (t0, y0) and the structured factorisation of W = I/(gamma dt) - J
// g receives the background at (t0, y0[0]), which is also the first stage's: evaluated here
// (gsrc 0) or loaded from W.pgn (1) or W.pgc (2), where earlier attempts left it; with more than
// one wavefront it is also kept in W.pgc.  With the bordered solve and more than one wavefront,
// waves 1.. run first(tid - 64) while wave 0 inverts the reduced matrix (g and W00 are set by
// then).
// same_point: this attempt retries a rejected one from the same (t0, y0), so the Jacobian
// pieces the rejected attempt formed are still in LDS (J0, dT, the dense block's entries and its
// diagonal Jd, Jh, G, the tail coefficients); with the bordered solve only the diagonal, which
// carries 1/(gamma dt), is formed again, from Jd, and the entries are those a full pass would give.
template <int MC, int NT, class F>
__device__ __forceinline__ void factorise(const BgView &b, const PLay &L, double k, double t0, double dt, Ws &W, int tid, double &W00,
                                          PBg<double> &g, int gsrc, const F &first, bool same_point = false) {
  const int n = L.n, m = L.m, nch = L.nch;
  const double gi = 1.0 / (dt * r5::gamma);
  const bool reuse = MC > 0 && same_point;
  if (!reuse) rhs_dual<NT>(b, L, k, t0, W.y0, W.J0, W.dT, W, tid);
  if (gsrc == 0) {
    pbg_eval_lanes(b, L, k, t0, W.y0[0], g, W.sc, tid);
  } else {
    pbg_load(g, gsrc == 1 ? W.pgn : W.pgc, L, tid & (WAVE - 1));
  }
  if (NT > WAVE && gsrc != 2 && tid < WAVE) pbg_store(g, W.pgc, L, tid);
  if (reuse) {
    for (int r = tid; r < m; r += NT) W.S[r * m + r] = gi - W.Jd[r];
  } else {
    // dense-block columns 0..m-1 and chain-head columns m..m+nch-1, one per lane; wave w takes
    // rows w, w + NR, ... of every column, so a wave evaluates one row formula at a time
    constexpr int NR = NT / WAVE;
    const int rg = __builtin_amdgcn_readfirstlane(tid / WAVE);  // uniform: the row formulas branch on scalars
    for (int col = tid % WAVE; col < m + nch; col += WAVE) {
      int s = (col < m) ? W.dmap[col] : W.chead[col - m];
      auto ev = [&](int i) { return i == s ? 1.0 : 0.0; };
      Met<double> mt;
      metric<double>(b, L, k, g, ev, mt);
      auto vsel = [&](int q) { return g.vq_l(L, q); };
      // the bordered solve reads, of the massive-neutrino rows (dense rows 15 .. 15 + 3 nq - 1),
      // only each bin's own 3 x 3 block: their other entries are not formed
      const int nq = L.nq, cq = col - 15;
      const bool ncol = MC > 0 && col < m && cq >= 0 && cq < 3 * nq;
      for (int r = rg; r < m; r += NR) {
        const int rq = r - 15;
        if (MC > 0 && col < m && rq >= 0 && rq < 3 * nq && !(ncol && rq % nq == cq % nq)) continue;
        // a hierarchy's l = 3 entry (a chain head) appears only in that hierarchy's l = 2 row
        if (col >= m && r != W.cpar[col - m]) continue;
        double v = core_row<double>(__builtin_amdgcn_readfirstlane(W.dmap[r]), b, L, k, g, mt, ev, vsel);
        if (col < m) {
          W.S[r * m + col] = ((r == col) ? gi : 0.0) - v;  // W_DD = I/(gamma h) - J_DD
          if (r == col) W.Jd[r] = v;
        } else {
          W.Jh[(col - m) * m + r] = v;
        }
      }
      if (MC > 0 && col < m && rg == 0) {  // h' and eta' per unit of this unknown
        W.G[2 * col] = mt.hprime;
        W.G[2 * col + 1] = mt.etaprime;
      }
    }
    // tail coefficients
    for (int i = tid; i < n; i += NT) {
      int rk = W.rk[i];
      if (rk >= 0) {
        double lo, di, up;
        auto vsel = [&](int q) { return g.vq_l(L, q); };
        tail_coef<double>(rk >> 8, rk & 255, L, k, g, W.tca[i], W.tcb[i], vsel, lo, di, up);
        W.tlo[i] = lo; W.tdi[i] = di; W.tup[i] = up;
      }
    }
  }
  __syncthreads();
  // bottom-up (UL) elimination of each chain: d_l and multipliers
  for (int c = tid; c < nch; c += NT) {
    int h = W.chead[c], s = W.cstride[c], len = W.clen[c];
    int i = h + (len - 1) * s;
    W.td[i] = gi - W.tdi[i];
    for (int j = len - 2; j >= 0; --j) {
      int ii = h + j * s;
      double mul = (-W.tup[ii]) / W.td[ii + s];
      W.tmul[ii] = mul;
      W.td[ii] = (gi - W.tdi[ii]) - mul * (-W.tlo[ii + s]);
    }
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT)
    if (W.rk[i] >= 0) W.rtd[i] = 1.0 / W.td[i];
  // Schur correction of the parent diagonal column, S[:, p_c] -= Jh_c * g_c * lo_1c, whose one
  // nonzero entry is the parent's diagonal (only that entry of Jh_c is formed)
  for (int c = tid; c < nch; c += NT) {
    const int h = W.chead[c], pr = W.cpar[c];
    W.S[pr * m + pr] -= W.Jh[c * m + pr] * (W.tlo[h] / W.td[h]);
  }
  W00 = gi - W.J0[0];
  __syncthreads();
  // Dense block: with a compile-time bin count, the bordered reduction and a one-wavefront
  // Gauss-Jordan of the 19 x 19 reduced matrix.  Otherwise in-place Gauss-Jordan of the m x m
  // block with partial pivoting (first maximum): per pivot, wave 0 moves row p to the pivot
  // row's slot and snapshots the scaled pivot row and column p, then one element-parallel pass
  // updates all m^2 entries.  Row interchanges leave the inverse with permuted columns; the
  // composite permutation pi is applied to the right-hand side in wsolve.
  if constexpr (MC > 0) {
    reduce_dense<MC, NT>(L, W, tid);
    if (tid < WAVE) gauss_jordan_wave<NRED>(W.R, W.R, W.rmap, W.pi, tid);
    else if constexpr (NT > WAVE) first(tid - WAVE);
    W.Sinv = W.R;
    __syncthreads();
  } else {
    W.Sinv = W.S;
    double *prow = W.u, *colp = W.f;  // free during the factorisation
    const int lane = tid & (WAVE - 1);
    for (int p = 0; p < m; ++p) {
      double ip = 0.0;
      if (NT == WAVE || tid < WAVE) {  // wave 0: pivot search, interchange, snapshots
        double best = -1.0;
        int bi = p;
        for (int r = p + lane; r < m; r += WAVE) {
          double v = fabs(W.S[r * m + p]);
          if (v > best) { best = v; bi = r; }
        }
  #pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          double ob = __shfl_xor(best, o);
          int oi = __shfl_xor(bi, o);
          if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) W.piv[p] = bi;
        ip = 1.0 / W.S[bi * m + p];
        for (int j = lane; j < m; j += WAVE) {
          double vp = W.S[p * m + j], vb = W.S[bi * m + j];
          prow[j] = (j == p) ? ip : vb * ip;
          W.S[bi * m + j] = vp;                                   // old row p moves to the pivot slot
          colp[j] = (j == bi) ? W.S[p * m + p] : W.S[j * m + p];  // column p, after the interchange
        }
      }
      __syncthreads();
      if (NT > WAVE) ip = prow[p];
      for (int e = tid; e < m * m; e += NT) {
        int r = e / m, c = e - r * m;
        double v;
        if (r == p) v = prow[c];
        else if (c == p) v = -colp[r] * ip;
        else v = W.S[e] - colp[r] * prow[c];
        W.S[e] = v;
      }
      __syncthreads();
    }
    if (tid == 0) {  // inverse[:, j] = S[:, pi[j]]
      for (int j = 0; j < m; ++j) { W.pi[j] = j; W.rmap[j] = j; }
      for (int p = m - 1; p >= 0; --p) {
        int q = W.piv[p];
        int t = W.pi[p]; W.pi[p] = W.pi[q]; W.pi[q] = t;
      }
    }
    __syncthreads();
  }
}

// solve W x = v in place (v in LDS, length n), x0 = v[0] / W_00 given and the other rows of v
// already holding b_i + J_i0 x0 (row 0 eliminated): wavefront 0 alone (tid
// is its lane), its phases ordered by wave_sync; the caller's workgroup barrier publishes x.
// Every phase has at most 64 items (chains, dense rows, reduced rows) but the dense rows of a
// 16-bin grid, so the other wavefronts are free for other work meanwhile.
template <int MC>
__device__ __forceinline__ void wsolve(const PLay &L, Ws &W, double x0, double *v, int tid) {
  const int m = L.m, nch = L.nch;
  // Phase 1: bottom-up sweep of each chain, one lane per chain; the next element's operands are
  // loaded before the current store.
  // A chain head couples to its parent row alone, so the chain's lane also forms that dense
  // row's right-hand side b (row 0 and the chain heads eliminated); other lanes form the rest.
  for (int c = tid; c < nch; c += WAVE) {
    const int h = W.chead[c], s = W.cstride[c], len = W.clen[c];
    int ii = h + (len - 1) * s;
    double x = v[ii];
    int j = len - 2;
    for (; j >= 3; j -= 4) {  // four elements: all operands loaded, then the serial recurrence
      const int i0 = ii - s, i1 = ii - 2 * s, i2 = ii - 3 * s, i3 = ii - 4 * s;
      const double a0 = v[i0], a1 = v[i1], a2 = v[i2], a3 = v[i3];
      const double m0 = W.tmul[i0], m1 = W.tmul[i1], m2 = W.tmul[i2], m3 = W.tmul[i3];
      v[ii] = x;
      const double x1 = a0 - m0 * x, x2 = a1 - m1 * x1, x3 = a2 - m2 * x2;
      v[i0] = x1;
      v[i1] = x2;
      v[i2] = x3;
      x = a3 - m3 * x3;
      ii = i3;
    }
    for (; j >= 0; --j) {
      const int in = ii - s;
      const double vn = v[in], mn = W.tmul[in];
      v[ii] = x;
      x = vn - mn * x;
      ii = in;
    }
    v[ii] = x;
    const int pr = W.cpar[c], ip = W.dmap[pr];
    W.xd[MC > 0 ? pr : W.pi[pr]] = v[ip] + W.Jh[c * m + pr] * (x * W.rtd[h]);
  }
  for (int r = tid; r < m; r += WAVE) {
    if (W.pch[r] >= 0) continue;
    W.xd[MC > 0 ? r : W.pi[r]] = v[W.dmap[r]];
  }
  wave_sync();
  if constexpr (MC > 0) {
    // bordered solve (see reduce_dense)
    constexpr int NQ = (MC - NCORE) / 3;
    constexpr int NN = 3 * NQ;
    // Phase 2: M^-1 b_n (items 0 .. NN-1) and the reduced right-hand side (the next NRED
    // items, stored through the pivot permutation): b_c - P_c b_n for the core rows and
    // P_h b_n for h' and eta', one partial sum per l block of b_n
    for (int e = tid; e < NN + NRED; e += WAVE) {
      if (e < NN) {
        const int l = e / NQ, q = e - l * NQ;
        const double *mi = W.Minv + 9 * q + 3 * l;
        W.rn[e] = mi[0] * W.xd[15 + q] + mi[1] * W.xd[15 + NQ + q] + mi[2] * W.xd[15 + 2 * NQ + q];
      } else {
        const int i = e - NN;
        const double *row = W.P + i * NN;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          a0 += row[q] * W.xd[15 + q];
          a1 += row[NQ + q] * W.xd[15 + NQ + q];
          a2 += row[2 * NQ + q] * W.xd[15 + 2 * NQ + q];
        }
        const double pb = (a0 + a1) + a2;
        W.xr[W.pi[i]] = (i < NCORE) ? W.xd[core_index(i, NQ)] - pb : pb;
      }
    }
    wave_sync();
    // Phase 3: the reduced solve, core unknowns and z = (h', eta')
    for (int i = tid; i < NRED; i += WAVE) {
      const double *row = W.Sinv + W.rmap[i] * NRED;
      double acc = 0.0;
      for (int j = 0; j < NRED; ++j) acc += row[j] * W.xr[j];
      if (i < NCORE) v[W.dmap[core_index(i, NQ)]] = acc;
      else W.xr[NRED + i - NCORE] = acc;
    }
  } else {
    // Phase 2: the dense solve, x_D = W_DD^-1 b through the pivot permutation
    for (int r = tid; r < m; r += WAVE) {
      double acc = 0.0;
      const double *row = W.Sinv + W.rmap[r] * m;
      for (int j = 0; j < m; ++j) acc += row[j] * W.xd[j];
      v[W.dmap[r]] = acc;
    }
  }
  if (tid == 0) v[0] = x0;  // read by no one before the workgroup barrier that follows the solve
  wave_sync();
  // Last phase: the neutrino unknowns x_n = M^-1 b_n + Q z, and the top-down sweeps of the
  // chains, four elements at a time as above; a chain whose parent is a neutrino row forms that
  // unknown itself with the same expression
  auto xn = [&](int j) { return fma(W.Q[2 * j + 1], W.xr[NRED + 1], fma(W.Q[2 * j], W.xr[NRED], W.rn[j])); };
  if constexpr (MC > 0)
    for (int j = tid; j < MC - NCORE; j += WAVE) v[W.dmap[15 + j]] = xn(j);
  for (int c = tid; c < nch; c += WAVE) {
    const int h = W.chead[c], s = W.cstride[c], len = W.clen[c];
    const int pr = W.cpar[c];
    const double xp = (MC > 0 && pr >= 15 && pr < MC - 2) ? xn(pr - 15) : v[W.dmap[pr]];
    double x = (v[h] + W.tlo[h] * xp) * W.rtd[h];
    int ii = h;
    int j = 1;
    for (; j + 3 < len; j += 4) {
      const int i0 = ii + s, i1 = ii + 2 * s, i2 = ii + 3 * s, i3 = ii + 4 * s;
      const double v0 = v[i0], v1 = v[i1], v2 = v[i2], v3 = v[i3];
      const double l0 = W.tlo[i0], l1 = W.tlo[i1], l2 = W.tlo[i2], l3 = W.tlo[i3];
      const double r0 = W.rtd[i0], r1 = W.rtd[i1], r2 = W.rtd[i2], r3 = W.rtd[i3];
      v[ii] = x;
      const double x1 = (v0 + l0 * x) * r0, x2 = (v1 + l1 * x1) * r1, x3 = (v2 + l2 * x2) * r2;
      v[i0] = x1;
      v[i1] = x2;
      v[i2] = x3;
      x = (v3 + l3 * x3) * r3;
      ii = i3;
    }
    for (; j < len; ++j) {
      const int in = ii + s;
      const double vn = v[in], ln = W.tlo[in], rn = W.rtd[in];
      v[ii] = x;
      x = (vn + ln * x) * rn;
      ii = in;
    }
    v[ii] = x;
  }
}

// adiabatic initial conditions (perturbations.py:680-781), written by one thread
__device__ __forceinline__ void initial_conditions(const BgView &b, const PLay &L, double tau, double k, double *y) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0;
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  double a = b.a_of_tau(tau);
  double rnu = exp(b.logrhonu(log(a)));
  double rhom = b.grhom * b.Omegam / (a * a * a);
  double rhor = (b.grhog + b.grhor * (b.Neff + b.Nmnu * rnu)) / (a * a * a * a);
  double rhonu = b.grhor * (b.Neff + b.Nmnu * rnu) / (a * a * a * a);
  double fracb = b.Omegab / b.Omegam, fracnu = rhonu / rhor;
  double om = a * rhom / sqrt(rhor);
  const double ci = -1.0, s2 = 1.0;
  double kt = k * tau;
  double deltag = -kt * kt / 3 * (1 - om * tau / 5) * ci * s2;
  double thetag = -kt * kt * kt / tau / 36 * (1 - 3 * (1 + 5 * fracb - fracnu) / 20 / (1 - fracnu) * om * tau) * ci * s2;
  double deltar = deltag;
  double thetar = -kt * kt * kt * kt / tau / 36 / (4 * fracnu + 15) *
                  (4 * fracnu + 11 + 12 - 3 * (8 * fracnu * fracnu + 50 * fracnu + 275) / 20 / (2 * fracnu + 15) * tau * om) * ci;
  double shearr = kt * kt / (45 + 12 * fracnu) * (3 * s2 - 1) * (1 + (4 * fracnu - 5) / 4 / (2 * fracnu + 15) * tau * om) * ci;
  double cs2_Q = b.cs2DE, w_Q = b.w0 + b.wa * (1.0 - a);
  double deltaq = kt * kt / 4 * (1 + w_Q) * (4 - 3 * cs2_Q) / (4 - 6 * w_Q + 3 * cs2_Q) * ci * s2;
  double thetaq = kt * kt * kt * kt / tau / 4 * cs2_Q / (4 - 6 * w_Q + 3 * cs2_Q) * ci * s2;
  if (b.lambda_de) { deltaq = 0.0; thetaq = 0.0; }
  double eta = ci * (1 - kt * kt / 12 / (15 + 4 * fracnu) *
                             (5 + 4 * s2 * fracnu - (16 * fracnu * fracnu + 280 * fracnu + 325) / 10 / (2 * fracnu + 15) * tau * om));
  y[0] = a; y[1] = 0.0; y[2] = eta;
  y[3] = 0.75 * deltag; y[4] = 0.0; y[5] = 0.75 * deltag; y[6] = thetag; y[7] = deltag; y[8] = thetag;
  y[L.idxr] = deltar; y[L.idxr + 1] = thetar; y[L.idxr + 2] = shearr * 2.0;
  for (int q = 0; q < nq; ++q) {
    double aq = a * b.amnu / L.q[q];
    double v = 1.0 / sqrt(1.0 + aq * aq);
    double dl = L.dlfdlq[q];
    y[iq0 + q] = -0.25 * dl * deltar;
    y[iq0 + nq + q] = -dl * thetar / v / k / 3.0;
    y[iq0 + 2 * nq + q] = -0.5 * dl * shearr;
  }
  y[n - 2] = deltaq;
  y[n - 1] = thetaq;
}

// determine_starting_time (perturbations.py:784-835): two 7-step bisections in log tau
__device__ __forceinline__ double starting_time(const BgView &b, double k) {
  const double small_k = 0.0004, large_k = 0.07;
  double tau0 = b.taumin, tau1 = b.tau_of_a(0.1);
  double tau_k = 1.0 / k;
  double akthom2 = 2.3048e-9 * (1.0 - b.YHe) * b.Omegab * b.H0 * b.H0;
  double lt[2];
  for (int which = 0; which < 2; ++which) {
    auto fn = [&](double x) {
      double tt = exp(x);
      double a = b.a_of_tau(tt);
      double aH = b.aprimeoa(a);
      if (which == 0) return (1.0 / aH) / tau_k / large_k - 1.0;
      double xe = b.xe_of_tau(tt);
      double opac = xe * akthom2 / (a * a);
      return (1.0 / opac) / (1.0 / aH) / small_k - 1.0;
    };
    double xl = log(tau0), xr = log(tau1);
    for (int it = 0; it < 7; ++it) {
      double xm = 0.5 * (xl + xr);
      if (fn(xm) * fn(xl) > 0) xl = xm; else xr = xm;
    }
    lt[which] = 0.5 * (xl + xr);
  }
  return exp(fmin(lt[1], lt[0]));
}

// The RSA projection's switch and fields at (tau, y) (approximations.py:84-176): when `on`,
// the photon and massless-neutrino l = 0, 1 moments become (dg, tg) and (dr, tr) and the rest of
// both hierarchies and the polarisation hierarchy zero (rsa_value).  y(i): the state's row i.
struct RsaProj {
  bool on;
  double dg, tg, dr, tr;
};
template <class Y>
__device__ __forceinline__ RsaProj rsa_project(const BgView &b, const PLay &L, double k, double tau, const Y &y) {
  RsaProj r{false, 0.0, 0.0, 0.0, 0.0};
  if (!b.use_rsa) return r;
  const int n = L.n, nq = L.nq, iq0 = L.iq0;
  const double a = y(0), loga = log(a);
  double aH = b.aprimeoa(a);
  int i = spl_index(b.loga, b.N, loga);
  double cs2 = spl_eval_at(b.loga, b.cs2a, b.S_cs2a_of_loga, i, loga) / a;
  double xe = spl_eval_at(b.loga, b.xe, b.S_xe_of_loga, i, loga);
  double xep = spl_deriv_at(b.loga, b.xe, b.S_xe_of_loga, i, loga) * aH;
  double opac = xe * b.akthom / (a * a);
  double tc = 1.0 / fmax(opac, 1e-30);
  double tcp = tc * (2.0 * aH - xep / fmax(xe, 1e-30));
  r.on = (tc / fmax(tau, 1e-30) > b.rsa_c) && (tau * k > b.rsa_k);
  if (r.on) {
    double drhonu = 0.0;
    for (int q = 0; q < nq; ++q) {
      double aq = a * b.amnu / L.q[q];
      drhonu += L.w[q] * y(iq0 + q) * sqrt(1.0 + aq * aq);
    }
    double rho_Q = pow(a, -3.0 * (1.0 + b.w0 + b.wa)) * exp(3.0 * (a - 1.0) * b.wa);
    double eta = y(2), deltab = y(5), thetab = y(6);
    double dgw = b.grhom * ((b.Omegam - b.Omegab) * y(3) + b.Omegab * deltab) / a + b.grhor * b.Nmnu * drhonu / (a * a) +
                 b.grhom * b.OmegaDE * y(n - 2) * rho_Q * a * a;
    double hp = (2.0 * k * k * eta + dgw) / fmax(aH, 1e-30);
    double sk2 = fmax(k * k, 1e-30), stc = fmax(tc, 1e-30);
    r.dg = 4.0 / sk2 * (aH * hp - sk2 * eta) + 4.0 / (sk2 * stc) * (thetab + 0.5 * hp);
    r.tg = -0.5 * hp + 3.0 / (sk2 * stc) * (-tcp / stc * (thetab + 0.5 * hp) + (-aH * thetab + cs2 * sk2 * deltab + sk2 * eta));
    r.dr = 4.0 / sk2 * (aH * hp - sk2 * eta);
    r.tr = -0.5 * hp;
  }
  return r;
}
// row j of the projected state, v its value before the projection (p.on)
__device__ __forceinline__ double rsa_value(const PLay &L, const RsaProj &p, int j, double v) {
  if (j >= L.idxg && j <= L.idxg + L.lmaxg) return (j == L.idxg) ? p.dg : (j == L.idxg + 1 ? p.tg : 0.0);
  if (j >= L.idxgp && j <= L.idxgp + L.lmaxgp) return 0.0;
  if (j >= L.idxr && j <= L.idxr + L.lmaxr) return (j == L.idxr) ? p.dr : (j == L.idxr + 1 ? p.tr : 0.0);
  return v;
}

// RSA output projection (approximations.py:84-176) and output conversion
// (perturbations.py:528-677) of the interpolated state y (LDS), written to dst
template <int NT>
__device__ __forceinline__ void emit_output(const BgView &b, const PLay &L, double k, double tau, double *y, double *dst, int full, int tid) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0;
  const double a = y[0], loga = log(a);
  if (b.use_rsa) {
    const RsaProj pr = rsa_project(b, L, k, tau, [&](int i) { return y[i]; });
    if (pr.on) {  // uniform: every thread has the same inputs
      __syncthreads();
      for (int j = tid; j < n; j += NT) y[j] = rsa_value(L, pr, j, y[j]);
      __syncthreads();
    }
  }
  if (full) {  // dst NULL: the projection in place only
    if (dst != nullptr)
      for (int j = tid; j < n; j += NT) dst[j] = y[j];
    return;
  }
  // 20 output fields
  const double Omegac = b.Omegam - b.Omegab;
  double eta = y[2], deltac = y[3], thetac = y[4], deltab = y[5], thetab = y[6], deltag = y[7], thetag = y[8];
  double deltar = y[L.idxr], thetar = y[L.idxr + 1];
  double rhonu = exp(b.logrhonu(loga)), pnu = exp(b.logpnu(loga));
  double drhonu = 0.0, fnu = 0.0;
  for (int q = 0; q < nq; ++q) {
    double aq = a * b.amnu / L.q[q];
    double v = 1.0 / sqrt(1.0 + aq * aq);
    drhonu += L.w[q] * y[iq0 + q] / v;
    fnu += L.w[q] * y[iq0 + nq + q];
  }
  double deltaq = y[n - 2], thetaq = y[n - 1];
  double w_Q = b.w0 + b.wa * (1.0 - a);
  double rho_Q = pow(a, -3.0 * (1.0 + b.w0 + b.wa)) * exp(3.0 * (a - 1.0) * b.wa);
  double rpptq = (1 + w_Q) * rho_Q * b.grhom * b.OmegaDE * thetaq * a * a;
  double grho = b.grhom * b.Omegam / a + (b.grhog + b.grhor * (b.Neff + b.Nmnu * rhonu)) / (a * a) + b.grhom * b.OmegaDE * rho_Q * a * a +
                b.grhom * b.Omegak;
  double aH = sqrt(grho / 3.0);
  double dgrho = b.grhom * (Omegac * deltac + b.Omegab * deltab) / a +
                 (b.grhog * deltag + b.grhor * (b.Neff * deltar + b.Nmnu * drhonu)) / (a * a) + b.grhom * b.OmegaDE * deltaq * rho_Q * a * a;
  double dgtheta = b.grhom * (Omegac * thetac + b.Omegab * thetab) / a + 4.0 / 3.0 * (b.grhog * thetag + b.Neff * b.grhor * thetar) / (a * a) +
                   b.Nmnu * b.grhor * k * fnu / (a * a) + rpptq;
  double hprime = (2.0 * k * k * eta + dgrho) / aH;
  double etaprime = 0.5 * dgtheta / (k * k);
  double alpha = (hprime + 6. * etaprime) / 2. / (k * k);
  double mden = b.grhom * b.Omegam / a + (b.grhor * b.Nmnu * rhonu) / (a * a);
  double deltam = (b.grhom * (Omegac * deltac + b.Omegab * deltab) / a + (b.grhor * b.Nmnu * drhonu) / (a * a)) / mden;
  double thetam = (b.grhom * (Omegac * thetac + b.Omegab * thetab) / a + b.Nmnu * b.grhor * k * fnu / (a * a)) /
                  (3.0 * (b.grhom * b.Omegam / a + b.grhor * b.Nmnu * rhonu / (a * a)));
  double deltabc = (b.grhom * (Omegac * deltac + b.Omegab * deltab) / a) / (b.grhom * b.Omegam / a);
  double thetabc = (b.grhom * (Omegac * thetac + b.Omegab * thetab) / a) / (3.0 * (b.grhom * b.Omegam / a) / (a * a));
  thetam += alpha * k * k;
  thetabc += alpha * k * k;
  double deltanu = drhonu / rhonu, thetanu = k * fnu / (rhonu + pnu);
  if (tid < 20) {
    double v;
    switch (tid) {
      case 0: v = eta; break;
      case 1: v = etaprime; break;
      case 2: v = hprime; break;
      case 3: v = alpha; break;
      case 4: v = deltam; break;
      case 5: v = thetam / aH; break;
      case 6: v = deltabc; break;
      case 7: v = thetabc / aH; break;
      case 8: v = deltac; break;
      case 9: v = thetac / aH; break;
      case 10: v = deltab; break;
      case 11: v = thetab / aH; break;
      case 12: v = deltag; break;
      case 13: v = thetag / aH; break;
      case 14: v = deltar; break;
      case 15: v = thetar / aH; break;
      case 16: v = deltanu; break;
      case 17: v = thetanu / aH; break;
      case 18: v = deltaq; break;
      default: v = thetaq / aH; break;
    }
    dst[tid] = v;
  }
}

// Source epilogue: the chain perturbations -> time derivatives -> source function
// (compute_time_derivatives, perturbations.py:1339-1406, and cmb.py:17-633 as chained in
// cmb.py:1030-1073) applied to each output as it is emitted, so full states are never stored.
// V: the visibility table of the unit's cosmology at the output times (NV rows of nout);
// planes: 8 planes of n_units x nout, as deb_cmb_sources writes them (S1, S2, S3, S4, S, alpha,
// alpha', S_E); pot: nout x n_units, the ISW potential eta + alpha' (isw 0: S1 and S come later
// from its spline in tau); isw 1 writes S1 = e^-tau_c (eta' + alpha'') and S.
struct SrcOut {
  const double *V, *aexp, *qw;
  double *planes, *pot;
  int isw, n_units;
};

// The outputs j0 .. j1 - 1 of one accepted step (tprev, tnext] of a unit, one output per thread
// (NT at a time): each thread interpolates the step's end states y0, y1 (LDS) to its output time,
// applies the RSA projection, evaluates the background there on its own (pbg_eval_own) and forms
// the source point, reading the right-hand side rows that it needs (rhs_rows' row formulas, on
// the projected state) as it needs them; nothing is shared between the threads, so the outputs
// of a step cost one source point's latency instead of one each.  The arithmetic of every
// quantity is that of the staged chain (emit_output's projection, rhs_eval, source_point).
template <int NQ, int NT>
__device__ __forceinline__ void emit_sources(const BgView &b, const PLay &L, double k, double tprev, double tnext, const double *y0,
                                             const double *y1, const double *__restrict__ tout, int j0, int j1, int unit, int nout,
                                             const SrcOut &so, const Ws &W, int tid) {
  for (int j = j0 + tid; j < j1; j += NT) {
    const double tau = tout[j];
    const double th = (tau - tprev) / (tnext - tprev);
    auto yi = [&](int i) { return y0[i] + th * (y1[i] - y0[i]); };
    const RsaProj pr = rsa_project(b, L, k, tau, yi);
    auto yv = [&](int i) { return pr.on ? rsa_value(L, pr, i, yi(i)) : yi(i); };
    PBgOwn<NQ> g;
    pbg_eval_own<NQ>(b, L, k, tau, yv(0), g);
    Met<double> m;
    metric<double>(b, L, k, g, yv, m);
    auto vrow = [&](int q) { return g.vq_u(q); };
    auto row = [&](int i) -> double {  // right-hand side row i (rhs_rows)
      const int rk = W.rk[i];
      if (rk < 0) return core_row<double>(i, b, L, k, g, m, yv, vrow);
      const int ch = rk >> 8, l = rk & 255, s = W.rs[i];
      double lo, di, up;
      tail_coef<double>(ch, l, L, k, g, W.tca[i], W.tcb[i], vrow, lo, di, up);
      const double yn = (up != 0.0) ? yv(i + s) : 0.0;
      return lo * yv(i - s) + di * yv(i) + up * yn;
    };
    const SrcPt s = source_point(b, yv, row, L.n, L.idxg, L.idxgp, L.idxr, L.iq0, L.nq, so.qw, k, so.aexp[j], so.V + j, nout,
                                 b.t[DEB_HDR_TAU0] - tau);
    const long plane = (long)so.n_units * nout, o = (long)unit * nout + j;
    so.planes[1 * plane + o] = s.S2;
    so.planes[2 * plane + o] = s.S3;
    so.planes[3 * plane + o] = s.S4;
    so.planes[5 * plane + o] = s.alpha;
    so.planes[6 * plane + o] = s.alphap;
    so.planes[7 * plane + o] = s.SE;
    if (so.isw == 1) {
      so.planes[0 * plane + o] = s.S1;
      so.planes[4 * plane + o] = s.S1 + s.S2 + s.S3 + s.S4;
    } else {
      so.pot[(long)j * so.n_units + unit] = s.pot;
    }
  }
}

// NaN source outputs of unit for the output times j0 .. nout - 1 (a unit not integrated, or
// stopped early)
template <int NT>
__device__ __forceinline__ void nan_sources(int unit, int j0, int nout, const SrcOut &so, int tid) {
  const long plane = (long)so.n_units * nout;
  for (int j = j0 + tid; j < nout; j += NT) {
    for (int p = 0; p < 8; ++p) so.planes[p * plane + (long)unit * nout + j] = __builtin_nan("");
    if (so.isw != 1) so.pot[(long)j * so.n_units + unit] = __builtin_nan("");
  }
}

// diffrax PIDController factor (error order 5, safety 0.9)
struct Pid {
  double prev_inv, prev_prev_inv;
  __device__ double adapt(double en, bool keep, const DebSolverCfg &s) {
    const double order = 5.0;
    double c1 = (s.icoeff + s.pcoeff + s.dcoeff) / order;
    double c2 = -(s.pcoeff + 2.0 * s.dcoeff) / order;
    double c3 = s.dcoeff / order;
    double inv = 1.0 / en;
    double f1 = (c1 == 0.0) ? 1.0 : pow(inv, c1);
    double f2 = (c2 == 0.0) ? 1.0 : pow(prev_inv, c2);
    double f3 = (c3 == 0.0) ? 1.0 : pow(prev_prev_inv, c3);
    double factor = fmin(fmax(0.9 * f1 * f2 * f3, keep ? 1.0 : s.factormin), s.factormax);
    if (inv == 0.0 || isinf(inv)) inv = 1.0;
    if (keep) { prev_prev_inv = prev_inv; prev_inv = inv; }
    return factor;
  }
};

// One Rodas5 step attempt of the mode with wavenumber k from W.y0 at tprev over dt: factorises
// W = 1 / (gamma dt) - J, runs the eight stages and writes the candidate y1 = u + k8 to W.y1
// (K + 7 n holds the error estimate k8).  Returns 1 on every thread if y1 has a NaN.
// Rodas5 stage time (stages 6-8 at t0 + dt) and stage input y0 + sum_j a_sj k_j (u + k_5,
// u + k_6 for the last two); kk(j) returns k_j of the row.  Both the stage-input pass and the
// look-ahead of the next stage's row 0 use these, so the two agree to the bit.
__device__ __forceinline__ double stage_time(int s, double tprev, double dt) {
  using namespace r5;
  const double cs = (s == 0) ? 0.0 : (s == 1) ? c2 : (s == 2) ? c3 : (s == 3) ? c4 : c5;
  return (s < 5) ? tprev + cs * dt : tprev + dt;
}
template <class KK>
__device__ __forceinline__ double stage_in(int s, double y0i, double ui, const KK &kk) {
  using namespace r5;
  switch (s) {
    case 0: return y0i;
    case 1: return fma(a21, kk(0), y0i);
    case 2: return fma(a32, kk(1), fma(a31, kk(0), y0i));
    case 3: return fma(a43, kk(2), fma(a42, kk(1), fma(a41, kk(0), y0i)));
    case 4: return fma(a54, kk(3), fma(a53, kk(2), fma(a52, kk(1), fma(a51, kk(0), y0i))));
    case 5: return fma(a65, kk(4), fma(a64, kk(3), fma(a63, kk(2), fma(a62, kk(1), fma(a61, kk(0), y0i)))));
    case 6: return ui + kk(5);
    default: return ui + kk(6);
  }
}

// One Rodas5 attempt from (tprev, y0) with step dt.  The first stage reuses the background the
// Jacobian pass evaluated at (tprev, y0[0]).  With more than one wavefront, the stage solve runs
// on wavefront 0 while wavefront 1 evaluates the next stage's background: it needs only row 0
// of the stage input, and row 0 of the solve (x0 = v[0] / W_00) is known before the solve starts.
// gsrc: where the Jacobian point's background comes from (see factorise).  tspec >= 0: with
// more than one wavefront, wavefront 1 also evaluates the background at (tspec, y1[0]) into
// W.pgn during the last stage solve: the next step's Jacobian point if this step is accepted.
// same_point: the previous attempt was rejected, so its Jacobian pieces are reused (factorise).
template <int NQ, int NT>
__device__ __forceinline__ int rodas5_attempt(const BgView &b, const PLay &L, double k, double tprev, double dt, Ws &W, int tid,
                                              int gsrc = 0, double tspec = -1.0, bool same_point = false) {
  constexpr int MC = NQ > 0 ? 17 + 3 * NQ : 0;
  // the first stage's right-hand side is formed beside the inversion of the reduced matrix
  constexpr bool FIRST_BESIDE_GJ = MC > 0 && NT > WAVE;
  const int n = L.n;
  double *K = W.K;
  double *u = W.u, *y0 = W.y0, *dT = W.dT;
  double W00;
  PBg<double> g;
  using namespace r5;
  // right-hand side of stage s, row i, from f_i: dts df/dtau (stages 1-5) and sum_j (C_sj / dt) k_j
  auto assemble = [&](int s, int i, double r, double dts) {
    switch (s) {
      case 0: r = r + dts * dT[i]; break;
      case 1: r = r + dts * dT[i] + (C21 / dt) * K[i]; break;
      case 2: r = r + dts * dT[i] + ((C31 / dt) * K[i] + (C32 / dt) * K[n + i]); break;
      case 3: r = r + dts * dT[i] + ((C41 / dt) * K[i] + (C42 / dt) * K[n + i] + (C43 / dt) * K[2 * n + i]); break;
      case 4:
        r = r + dts * dT[i] + ((C51 / dt) * K[i] + (C52 / dt) * K[n + i] + (C53 / dt) * K[2 * n + i] + (C54 / dt) * K[3 * n + i]);
        break;
      case 5:
        r = r + ((C61 / dt) * K[i] + (C62 / dt) * K[n + i] + (C63 / dt) * K[2 * n + i] + (C64 / dt) * K[3 * n + i] +
                 (C65 / dt) * K[4 * n + i]);
        break;
      case 6:
        r = r + ((C71 / dt) * K[i] + (C72 / dt) * K[n + i] + (C73 / dt) * K[2 * n + i] + (C74 / dt) * K[3 * n + i] +
                 (C75 / dt) * K[4 * n + i] + (C76 / dt) * K[5 * n + i]);
        break;
      default:
        r = r + ((C81 / dt) * K[i] + (C82 / dt) * K[n + i] + (C83 / dt) * K[2 * n + i] + (C84 / dt) * K[3 * n + i] +
                 (C85 / dt) * K[4 * n + i] + (C86 / dt) * K[5 * n + i] + (C87 / dt) * K[6 * n + i]);
        break;
    }
    return r;
  };
  // Each row's stage right-hand side goes to k_s as the row is evaluated, with row 0 already
  // eliminated: row 0 of W is W_00 alone, so x0 = b_0 / W_00 and b_i + J_i0 x0 is the right-hand
  // side of the rest.  Row 0's f is g.aH g.a (core_row), so every thread forms x0 itself.
  auto emit_stage = [&](int s, double dts) {
    double *ks = K + s * n;
    const double x0 = assemble(s, 0, g.aH * g.a, dts) / W00;
    return [&, s, dts, ks, x0](int i, double r) {
      const double v = assemble(s, i, r, dts);
      if (i == 0) {
        ks[0] = v;
        W.pg->x0 = x0;
      } else {
        ks[i] = v + W.J0[i] * x0;
      }
    };
  };
  auto first = [&](int t) {
    rhs_rows<(NT > WAVE ? NT - WAVE : WAVE)>(b, L, k, g, y0, W.perm1, W, t, emit_stage(0, dt * d1));
  };
  factorise<MC, NT>(b, L, k, tprev, dt, W, tid, W00, g, gsrc, first, same_point);
  for (int s = 0; s < 8; ++s) {
    double *ks = K + s * n;
    // df/dtau weight, selected without a run-time-indexed array (scratch memory)
    const double dts = dt * ((s == 0) ? d1 : (s == 1) ? d2 : (s == 2) ? d3 : (s == 3) ? d4 : d5);
    if (!(FIRST_BESIDE_GJ && s == 0)) {
      for (int i = tid; i < n; i += NT) u[i] = stage_in(s, y0[i], u[i], [&](int j) { return K[j * n + i]; });
      __syncthreads();
      if (s > 0) {
        if constexpr (NT == WAVE) pbg_eval_lanes(b, L, k, stage_time(s, tprev, dt), u[0], g, W.sc, tid);
        else pbg_load(g, W.pg, L, tid & (WAVE - 1));
      }
      rhs_rows<NT>(b, L, k, g, u, W.perm, W, tid, emit_stage(s, dts));
      __syncthreads();
    }
    if (tid < WAVE) {
      wsolve<MC>(L, W, W.pg->x0, ks, tid);
    } else if constexpr (NT > WAVE) {
      if (tid < 2 * WAVE && s < 7) {
        const double x0 = W.pg->x0;
        const double un = stage_in(s + 1, y0[0], u[0], [&](int j) { return j == s ? x0 : K[j * n]; });
        PBg<double> gn;
        pbg_eval_lanes<true>(b, L, k, stage_time(s + 1, tprev, dt), un, gn, W.sc, tid);
        pbg_store(gn, W.pg, L, tid - WAVE);
      } else if (tid < 2 * WAVE && tspec >= 0.0) {  // y1[0] = u[0] + k_8[0] (as below)
        PBg<double> gn;
        pbg_eval_lanes<true>(b, L, k, tspec, u[0] + W.pg->x0, gn, W.sc, tid);
        pbg_store(gn, W.pgn, L, tid - WAVE);
      }
    }
    __syncthreads();
  }
  // candidate y1 = u + k8 (perturbations.py:855-859)
  int nanflag = 0;
  for (int i = tid; i < n; i += NT) {
    double v = u[i] + K[7 * n + i];
    W.y1[i] = v;
    nanflag |= (v != v);
  }
  nanflag = __syncthreads_or(nanflag);
  __syncthreads();
  return nanflag;
}

// Sum of squares of the scaled error over the six filtered entries (0, 2, 3, 5, 6, 7) with
// weights (1, k^2, 1, 1, 1 / k^2, 1) (rms_norm_filtered, perturbations.py:855-859); the scale
// uses |y0| alone when the candidate has a NaN.
__device__ __forceinline__ double filtered_sq(double k, const double *y0, const double *y1, const double *err, int nanflag,
                                              const DebSolverCfg &cfg) {
  const int idx[6] = {0, 2, 3, 5, 6, 7};
  const double wts[6] = {1.0, k * k, 1.0, 1.0, 1.0 / (k * k), 1.0};
  double acc = 0.0;
  for (int j = 0; j < 6; ++j) {
    int i = idx[j];
    double yy = fmax(fabs(y0[i]), fabs(nanflag ? y0[i] : y1[i]));
    double e = err[i] / (cfg.atol + yy * cfg.rtol) * wts[j];
    acc += e * e;
  }
  return acc;
}

// One (cosmology, k) unit from its initial conditions to the last output time.  The LDS
// workspace W and the layout L (with its momentum-bin tables) are set up by the caller.
// SRC: the source epilogue (so) replaces the output conversion; y_out is not written.
template <int NQ, int NT, bool SRC = false>
__device__ __forceinline__ void solve_unit(int unit, const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                           const double *__restrict__ unit_k, const double *__restrict__ tau_out, int nout,
                                           const DebProblem &prob, const DebSolverCfg &cfg, double *__restrict__ step_log,
                                           double *__restrict__ y_out, int *__restrict__ status, int *__restrict__ n_acc,
                                           int *__restrict__ n_rej, const double *__restrict__ replay, long replay_stride,
                                           const int *__restrict__ replay_len, const PLay &L, Ws &W, int tid,
                                           const SrcOut *so = nullptr) {
  const int cosmo = unit_cosmo[unit];
  const double k = unit_k[unit];
  const int n = L.n;
  // the cosmology's background view lives in LDS: its ~100 scalar registers' worth of fields
  // would not fit beside the solver's other uniform values, so each use reads it back
  __shared__ BgView s_bg;
  __syncthreads();  // the previous unit's last LDS reads are done
  if (tid == 0) { W.sc->i = -1; W.sc->j = -1; s_bg.init(tables + (long)cosmo * tstride); }
  __syncthreads();
  const BgView &b = s_bg;
  const int wout = prob.return_full ? n : 20;
  double *yo = SRC ? nullptr : y_out + (long)unit * nout * wout;
  // w(a) = w0 + wa (1 - a) reaching -1 for some a in (0, 1] (the phantom divide) makes the
  // fluid equations singular (their 1 / (1 + w) terms): such units are not integrated
  // (status 4, NaN outputs).  w0 = -1, wa = 0 is the cosmological constant, handled exactly.
  if (!b.lambda_de && (1.0 + b.w0) * (1.0 + b.w0 + b.wa) <= 0.0) {
    if constexpr (SRC) nan_sources<NT>(unit, 0, nout, *so, tid);
    else
      for (int i = tid; i < nout * wout; i += NT) yo[i] = __builtin_nan("");
    if (tid == 0) { status[unit] = 4; n_acc[unit] = 0; n_rej[unit] = 0; }
    return;
  }

  // output times of this cosmology
  const double *tout = tau_out + (long)cosmo * nout;
  double tmin = tout[0], t1 = tout[0];
  for (int j = 1; j < nout; ++j) { tmin = fmin(tmin, tout[j]); t1 = fmax(t1, tout[j]); }
  double t0 = 0.99 * fmin(tmin, starting_time(b, k));
  if (tid == 0) initial_conditions(b, L, t0, k, W.y0);
  __syncthreads();

  double tprev = t0;
  double tnext = fmin(t0 + fmin(t0 / 4.0, 0.5 * (t1 - t0)), t1);
  // replay: the accepted step ends are given as s = (t - t0) / (t1 - t0), increasing to 1; the
  // steps are taken as they are, with no error control (the frozen-mesh map)
  const double *rs = (replay != nullptr) ? replay + (long)unit * replay_stride : nullptr;
  const int rlen = (replay != nullptr) ? replay_len[unit] : 0;
  int ridx = 0;
  auto mesh_t = [&](int i) { return (rs[i] >= 1.0) ? t1 : t0 + rs[i] * (t1 - t0); };
  if (rs != nullptr) tnext = (rlen > 0) ? mesh_t(0) : t1;
  Pid pid;
  pid.prev_inv = 1.0;
  pid.prev_prev_inv = 1.0;
  int isave = 0, steps = 0, na = 0, nr = 0, st = 0;
  int gsrc = 0;  // the next Jacobian point's background: evaluate, W.pgn (accepted), W.pgc (rejected)
  bool same = false;  // the last attempt was rejected: the next one starts from the same point
  double *K = W.K;
  double *u = W.u, *y0 = W.y0;
  while (tprev < t1) {
    if (steps >= cfg.max_steps || (rs != nullptr && ridx >= rlen)) { st = 1; break; }
    const double dt = tnext - tprev;
    const int nanflag = rodas5_attempt<NQ, NT>(b, L, k, tprev, dt, W, tid, gsrc, fmin(tnext, t1), same);
    const double en = sqrt(filtered_sq(k, W.y0, W.y1, K + 7 * n, nanflag, cfg) / 6.0);
    if (step_log != nullptr && tid == 0) {  // (t_prev, t_next, error norm) of every attempted step
      double *lg = step_log + ((long)unit * cfg.max_steps + steps) * 3;
      lg[0] = tprev;
      lg[1] = tnext;
      lg[2] = en;
    }
    if (!isfinite(en)) { st = 2; break; }
    bool keep = (rs != nullptr) || en < 1.0;
    double factor = (rs != nullptr) ? 1.0 : pid.adapt(en, keep, cfg);
    if (keep) {
      if constexpr (SRC) {  // every output of the step at once
        int je = isave;
        while (je < nout && tout[je] <= tnext) ++je;
        emit_sources<NQ, NT>(b, L, k, tprev, tnext, y0, W.y1, tout, isave, je, unit, nout, *so, W, tid);
        isave = je;
      } else {
        while (isave < nout && tout[isave] <= tnext) {
          double th = (tout[isave] - tprev) / (tnext - tprev);
          for (int i = tid; i < n; i += NT) u[i] = y0[i] + th * (W.y1[i] - y0[i]);
          __syncthreads();
          emit_output<NT>(b, L, k, tout[isave], u, yo + (long)isave * wout, prob.return_full, tid);
          __syncthreads();
          ++isave;
        }
      }
      __syncthreads();  // every wave has read y0 for the error norm before it is replaced
      for (int i = tid; i < n; i += NT) y0[i] = W.y1[i];
      ++na;
    } else {
      ++nr;
    }
    __syncthreads();
    double nt0 = keep ? tnext : tprev;
    tprev = fmin(nt0, t1);
    gsrc = (NT > WAVE) ? (keep ? 1 : 2) : 0;
    same = !keep;
    double tn = nt0 + dt * factor;
    tnext = (tn > t1 - 1e-10) ? t1 : tn;
    if (rs != nullptr) {
      ++ridx;
      tnext = (ridx < rlen) ? mesh_t(ridx) : t1;
    }
    ++steps;
  }
  if (st != 0) {
    if constexpr (SRC) nan_sources<NT>(unit, isave, nout, *so, tid);
    else
      for (int j = isave; j < nout; ++j)
        for (int i = tid; i < wout; i += NT) yo[(long)j * wout + i] = __builtin_nan("");
  }
  if (tid == 0) {
    status[unit] = st;
    n_acc[unit] = na;
    n_rej[unit] = nr;
  }
}

// LDS bytes of the forward solver's carve-up (setup_solver): 22 n + m^2 + nch m + 4 m + 3 NRED nq
// + 21 nq + 24 + NRED^2 doubles and three PBgShared, then m + 5 n + 512 + 4 nch + 4 m + 1 ints and
// the spline cache.
__host__ __device__ inline long deb_perturb_lds_bytes_dev(const DebProblem &p) {
  const long n = 7 + (p.lmaxg + 1) + (p.lmaxgp + 1) + (p.lmaxr + 1) + (long)p.nqmax * (p.lmaxnu + 1) + 2;
  const long m = 17 + 3 * (long)p.nqmax, nch = 3 + p.nqmax;
  const long nq = p.nqmax;
  const long dbl = n * 22 + m * m + nch * m + 4 * m + 3 * NRED * nq + 21 * nq + 24 + NRED * NRED + 3 * (long)(sizeof(PBgShared) / 8);
  const long ints = m + 5 * n + 512 + 4 * nch + 4 * m + 1;
  return dbl * 8 + ints * 4 + (long)sizeof(SplCache);
}

// Layout of the state vector for the truncations in prob (NQ > 0: compile-time bin count) and
// the carve-up of the dynamic LDS into the solver workspace, which must match
// deb_perturb_lds_bytes(): 22 n + m^2 + nch m + 4 m + 3 NRED nq + 21 nq + 24 + NRED^2 doubles and three PBgShared, then
// m + 5 n + 512 + 4 nch + 4 m ints and the spline cache.  Returns false (nothing set up) when that exceeds lds_bytes.
template <int NQ, int NT>
__device__ __forceinline__ bool setup_solver(const DebProblem &prob, const double *__restrict__ nu_qw, long lds_bytes, PLay &L, Ws &W,
                                             int tid) {
  L.lmaxg = prob.lmaxg; L.lmaxgp = prob.lmaxgp; L.lmaxr = prob.lmaxr; L.lmaxnu = prob.lmaxnu;
  L.nq = (NQ > 0) ? NQ : prob.nqmax;
  L.n = 7 + (L.lmaxg + 1) + (L.lmaxgp + 1) + (L.lmaxr + 1) + L.nq * (L.lmaxnu + 1) + 2;
  L.idxg = 7; L.idxgp = 7 + (L.lmaxg + 1); L.idxr = 9 + L.lmaxg + L.lmaxgp; L.iq0 = 10 + L.lmaxg + L.lmaxgp + L.lmaxr;
  L.m = 17 + 3 * L.nq;
  L.nch = 3 + L.nq;
  const int n = L.n, m = L.m, nch = L.nch;
  extern __shared__ double smem[];
  double *p = smem;
  W.y0 = p; p += n; W.y1 = p; p += n; W.u = p; p += n; W.f = p; p += n; W.J0 = p; p += n; W.dT = p; p += n;
  W.K = p; p += 8 * n; W.tlo = p; p += n; W.tdi = p; p += n; W.tup = p; p += n; W.tmul = p; p += n; W.td = p; p += n;
  W.S = p; p += (size_t)m * m; W.Jh = p; p += (size_t)nch * m; W.xd = p; p += m; W.Jd = p; p += m; W.P = p; p += NRED * 3 * L.nq;
  W.tca = p; p += n; W.tcb = p; p += n;
  W.G = p; p += 2 * (size_t)m; W.Minv = p; p += 9 * L.nq; W.Q = p; p += 6 * L.nq; W.rn = p; p += 3 * L.nq; W.xr = p; p += 24;
  W.R = p; p += NRED * NRED;
  W.rtd = p; p += n;
  W.pg = (PBgShared *)p; p += sizeof(PBgShared) / 8;
  W.pgc = (PBgShared *)p; p += sizeof(PBgShared) / 8;
  W.pgn = (PBgShared *)p; p += sizeof(PBgShared) / 8;
  double *sq = p; p += L.nq; double *sw = p; p += L.nq; double *sdl = p; p += L.nq;
  int *ip = (int *)p;
  W.dmap = ip; ip += m; W.rk = ip; ip += n; W.chead = ip; ip += nch; W.cstride = ip; ip += nch; W.clen = ip; ip += nch;
  W.cpar = ip; ip += nch; W.piv = ip; ip += m; W.pi = ip; ip += m; W.rmap = ip; ip += m; W.rs = ip; ip += n; W.rbin = ip; ip += n; W.perm = ip; ip += n + 256; W.perm1 = ip; ip += n + 256; W.pch = ip; ip += m;
  ip += (2 - ((long)((char *)ip - (char *)smem) / 4) % 2) % 2;  // 8-byte alignment for the cache
  W.sc = (SplCache *)ip; ip += sizeof(SplCache) / 4;
  if ((long)((char *)ip - (char *)smem) > lds_bytes) return false;
  for (int q = tid; q < L.nq; q += NT) {
    sq[q] = nu_qw[q];
    sw[q] = nu_qw[NQMAX + q];
    sdl[q] = -nu_qw[q] / (1.0 + exp(-nu_qw[q]));
  }
  L.q = sq; L.w = sw; L.dlfdlq = sdl;
  setup_maps<NT>(L, W, tid);
  for (int r = tid; r < m; r += NT) {  // the chain table was published by a barrier in setup_maps
    int c = -1;
    for (int j = 0; j < nch; ++j)
      if (W.cpar[j] == r) c = j;
    W.pch[r] = c;
  }
  return true;
}
// Batched shared-step evolution (evolve_modes_batched, perturbations.py:1002-1120, with
// Rodas5Batched): workgroup g integrates the `batch` consecutive units g * batch .. (g + 1) *
// batch - 1, which must belong to one cosmology, on ONE step sequence: a common start time
// (the smallest of the units' own), one step size, and one accept decision from the RMS of the
// filtered errors over all batch x 6 entries.  Each attempt walks the units in turn through the
// per-mode Rodas5 stages in LDS; their states live in `state` (2 x nvar doubles per unit, the
// current state and the candidate, swapped on accept), which stays in L2 between visits.
// step_log, when given, receives (t_prev, t_next, error norm) of every attempt of batch g at
// step_log[(g * max_steps + step) * 3].  Status 5: the batch's units have different cosmologies.
// SRC: the source epilogue (so) replaces the output conversion; y_out is not written.
template <int NQ, int NT, bool SRC = false>
__device__ __forceinline__ void solve_batch(const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                            const double *__restrict__ unit_k, int n_units, int batch, const double *__restrict__ tau_out,
                                            int nout, const double *__restrict__ nu_qw, const DebProblem &prob, const DebSolverCfg &cfg,
                                            long lds_bytes, double *__restrict__ step_log, double *__restrict__ y_out,
                                            int *__restrict__ status, int *__restrict__ n_acc, int *__restrict__ n_rej,
                                            double *__restrict__ state, const SrcOut *so = nullptr) {
  const int g = blockIdx.x, tid = threadIdx.x;
  const int u0 = g * batch;
  if (u0 + batch > n_units) return;
  PLay L;
  Ws W;
  if (!setup_solver<NQ, NT>(prob, nu_qw, lds_bytes, L, W, tid)) {
    for (int j = tid; j < batch; j += NT) { status[u0 + j] = 3; n_acc[u0 + j] = 0; n_rej[u0 + j] = 0; }
    return;
  }
  const int n = L.n;
  const int cosmo = unit_cosmo[u0];
  __shared__ BgView s_bg;  // in LDS, as in solve_unit
  if (tid == 0) { W.sc->i = -1; W.sc->j = -1; s_bg.init(tables + (long)cosmo * tstride); }
  __syncthreads();
  const BgView &b = s_bg;
  const int wout = prob.return_full ? n : 20;
  double *yo = SRC ? nullptr : y_out + (long)u0 * nout * wout;  // unit j's outputs at yo + j nout wout
  int mixed = 0;
  for (int j = tid; j < batch; j += NT) mixed |= (unit_cosmo[u0 + j] != cosmo);
  mixed = __syncthreads_or(mixed);
  const bool phantom = !b.lambda_de && (1.0 + b.w0) * (1.0 + b.w0 + b.wa) <= 0.0;
  if (mixed || phantom) {  // not integrated: NaN outputs (w crossing -1 as in the per-mode kernel)
    if constexpr (SRC) {
      for (int j = 0; j < batch; ++j) nan_sources<NT>(u0 + j, 0, nout, *so, tid);
    } else {
      for (long i = tid; i < (long)batch * nout * wout; i += NT) yo[i] = __builtin_nan("");
    }
    for (int j = tid; j < batch; j += NT) { status[u0 + j] = mixed ? 5 : 4; n_acc[u0 + j] = 0; n_rej[u0 + j] = 0; }
    return;
  }

  const double *tout = tau_out + (long)cosmo * nout;
  double tmin = tout[0], t1 = tout[0];
  for (int j = 1; j < nout; ++j) { tmin = fmin(tmin, tout[j]); t1 = fmax(t1, tout[j]); }
  // determine_start_times_per_batch (perturbations.py:954-999): the earliest start of the batch
  double t0 = 1e300;
  for (int j = 0; j < batch; ++j) t0 = fmin(t0, 0.99 * fmin(tmin, starting_time(b, unit_k[u0 + j])));
  double *st0 = state + (long)u0 * 2 * n;  // unit j: st0 + 2 n j + {cur, 1 - cur} n
  for (int j = 0; j < batch; ++j) {
    if (tid == 0) initial_conditions(b, L, t0, unit_k[u0 + j], W.y0);
    __syncthreads();
    for (int i = tid; i < n; i += NT) st0[2L * n * j + i] = W.y0[i];
    __syncthreads();
  }

  double tprev = t0;
  double tnext = fmin(t0 + fmin(t0 / 4.0, 0.5 * (t1 - t0)), t1);
  Pid pid;
  pid.prev_inv = 1.0;
  pid.prev_prev_inv = 1.0;
  int isave = 0, steps = 0, na = 0, nr = 0, st = 0, cur = 0;
  while (tprev < t1) {
    if (steps >= cfg.max_steps) { st = 1; break; }
    const double dt = tnext - tprev;
    double acc = 0.0, acc_nan = 0.0;  // scaled with max(|y0|, |y1|), and with |y0| alone
    int nanany = 0;
    for (int j = 0; j < batch; ++j) {
      const double k = unit_k[u0 + j];
      const double *ys = st0 + 2L * n * j + (long)cur * n;
      double *yd = st0 + 2L * n * j + (long)(1 - cur) * n;
      for (int i = tid; i < n; i += NT) W.y0[i] = ys[i];
      __syncthreads();
      const int nanflag = rodas5_attempt<NQ, NT>(b, L, k, tprev, dt, W, tid);
      for (int i = tid; i < n; i += NT) yd[i] = W.y1[i];
      acc += filtered_sq(k, W.y0, W.y1, W.K + 7 * n, 0, cfg);
      acc_nan += filtered_sq(k, W.y0, W.y1, W.K + 7 * n, 1, cfg);
      nanany |= nanflag;
      __syncthreads();
    }
    const double en = sqrt((nanany ? acc_nan : acc) / (6.0 * batch));
    if (step_log != nullptr && tid == 0) {
      double *lg = step_log + ((long)g * cfg.max_steps + steps) * 3;
      lg[0] = tprev;
      lg[1] = tnext;
      lg[2] = en;
    }
    if (!isfinite(en)) { st = 2; break; }
    const bool keep = en < 1.0;
    const double factor = pid.adapt(en, keep, cfg);
    if (keep) {
      if (isave < nout && tout[isave] <= tnext) {  // outputs in (tprev, tnext]: interpolate every unit
        int isv = isave;
        for (int j = 0; j < batch; ++j) {
          const double k = unit_k[u0 + j];
          const double *ya = st0 + 2L * n * j + (long)cur * n;
          const double *yb = st0 + 2L * n * j + (long)(1 - cur) * n;
          for (int i = tid; i < n; i += NT) { W.y0[i] = ya[i]; W.y1[i] = yb[i]; }
          __syncthreads();
          isv = isave;
          if constexpr (SRC) {
            while (isv < nout && tout[isv] <= tnext) ++isv;
            emit_sources<NQ, NT>(b, L, k, tprev, tnext, W.y0, W.y1, tout, isave, isv, u0 + j, nout, *so, W, tid);
            __syncthreads();  // W.y0, W.y1 are read before the next unit's states replace them
          } else {
            while (isv < nout && tout[isv] <= tnext) {
              double th = (tout[isv] - tprev) / (tnext - tprev);
              for (int i = tid; i < n; i += NT) W.u[i] = W.y0[i] + th * (W.y1[i] - W.y0[i]);
              __syncthreads();
              emit_output<NT>(b, L, k, tout[isv], W.u, yo + ((long)j * nout + isv) * wout, prob.return_full, tid);
              __syncthreads();
              ++isv;
            }
          }
        }
        isave = isv;
      }
      cur = 1 - cur;
      ++na;
    } else {
      ++nr;
    }
    double nt0 = keep ? tnext : tprev;
    tprev = fmin(nt0, t1);
    double tn = nt0 + dt * factor;
    tnext = (tn > t1 - 1e-10) ? t1 : tn;
    ++steps;
  }
  if (st != 0) {
    for (int j = 0; j < batch; ++j) {
      if constexpr (SRC) nan_sources<NT>(u0 + j, isave, nout, *so, tid);
      else
        for (int r = isave; r < nout; ++r)
          for (int i = tid; i < wout; i += NT) yo[((long)j * nout + r) * wout + i] = __builtin_nan("");
    }
  }
  for (int j = tid; j < batch; j += NT) {
    status[u0 + j] = st;
    n_acc[u0 + j] = na;
    n_rej[u0 + j] = nr;
  }
}

}  // namespace
}  // namespace deb
