// deb_adjoint.hip -- discrete adjoint of the frozen-mesh Rodas5 replay: d<gbar, outputs>/d theta for
// the six parameters that enter the background and the perturbations (omega_c, omega_b, h, m_nu,
// w0, wa), one workgroup (four wavefronts) per (cosmology, k) unit.
//
// The map differentiated is the replay the forward kernel runs on a fixed accepted-step mesh
// (steps ending at t0 + s_i (t1 - t0), s_last = 1): adiabatic initial state at t0, Rodas5 steps
// (ode_integrators_stiff.py:776-892), outputs as linear interpolants between step ends, RSA
// projection and conversion to the 20 output fields (perturbations.py:904-925, 528-677).  The
// reference differentiates its pipeline in forward mode (jax.jacfwd through diffeqsolve with
// DirectAdjoint, perturbations.py:919; ForwardMode in RECFAST, thermodynamics_recfast.py:295);
// this kernel computes the same frozen-mesh derivative in reverse mode:
//   1. forward replay of the unit, checkpointing every accepted state to HBM;
//   2. for each step, from the last: recompute the step from its checkpoint (Jacobian point,
//      structured factorisation, eight stages), then sweep the stages in reverse:
//        r_s = W^-T kbar_s by the transposed structured solve (chains swept the other way, the
//            transpose of the bordered 19 x 19 system through the inverse the forward step keeps),
//        the vector-Jacobian product of the right-hand side at the stage point: its linear part
//            A^T r column by column (unit-vector evaluation of the dense rows, chain coefficients),
//            its scale-factor, time and parameter parts by dual numbers,
//        the stage-input, C/h, h d_s and 1/(gamma h) terms;
//      then the dependence of J(t0, y0) and df/dtau(t0, y0) on the scale factor, the time, the
//      parameters and (through the scale-factor column) on the state;
//   3. the initial state, the start and end times and the output times by central differences
//      over the twelve shifted cosmologies the caller provides (as the finite-difference
//      gradient takes them).
// Parameter derivatives of the background come from derivative tables: the central difference
// of the shifted cosmologies' tables, knot by knot (the thermal history's log a knots and the
// neutrino tables' knots do not move with the parameters, and a natural spline is linear in its
// knot values, so these are the splines of the derivatives).
#include "deb_perturb_dev.hpp"

namespace deb {
namespace {

constexpr int NDIR = 8;      // derivative directions: a, tau, then the six parameters
constexpr int NPAR = 6;      // parameters differentiated through the ODE
constexpr int NSHIFT = 2 * NPAR;
constexpr int MAXOUT = 64;   // output times per unit supported by the adjoint

// Rodas5 stage-input coefficients a_{s+1, j+1} (stage s <= 5 takes y0 + sum_j a k_j) and the
// C_{s+1, j+1} of stage s's right-hand side, selected without run-time-indexed arrays
__device__ __forceinline__ double rodas_A(int s, int j) {
  using namespace r5;
  switch (s * 8 + j) {
    case 8: return a21;
    case 16: return a31; case 17: return a32;
    case 24: return a41; case 25: return a42; case 26: return a43;
    case 32: return a51; case 33: return a52; case 34: return a53; case 35: return a54;
    case 40: return a61; case 41: return a62; case 42: return a63; case 43: return a64; case 44: return a65;
    default: return 0.0;
  }
}
__device__ __forceinline__ double rodas_C(int s, int j) {
  using namespace r5;
  switch (s * 8 + j) {
    case 8: return C21;
    case 16: return C31; case 17: return C32;
    case 24: return C41; case 25: return C42; case 26: return C43;
    case 32: return C51; case 33: return C52; case 34: return C53; case 35: return C54;
    case 40: return C61; case 41: return C62; case 42: return C63; case 43: return C64; case 44: return C65;
    case 48: return C71; case 49: return C72; case 50: return C73; case 51: return C74; case 52: return C75; case 53: return C76;
    case 56: return C81; case 57: return C82; case 58: return C83; case 59: return C84; case 60: return C85; case 61: return C86;
    case 62: return C87;
    default: return 0.0;
  }
}

template <class T> struct IsDouble { static constexpr bool value = false; };
template <> struct IsDouble<double> { static constexpr bool value = true; };

// ---------------------------------------------------------------------------------------
// dual-number view of the background
// ---------------------------------------------------------------------------------------
template <int ND> struct BgD {
  Dn<ND> grhom, Omegam, Omegab, OmegaDE, Omegak, w0, wa, amnu, akthom;
  double grhog, grhor, Neff, Nmnu, cs2DE;
  bool lambda_de;
};

// dt[l]: derivative table of local direction l (NULL: no parameter part)
template <int ND> __device__ __forceinline__ Dn<ND> dual_at(double v, const double *const *dt, long off) {
  Dn<ND> r(v);
#pragma unroll
  for (int l = 0; l < ND; ++l) r.d[l] = dt[l] ? dt[l][off] : 0.0;
  return r;
}

template <int ND> __device__ __forceinline__ void make_bgd(const BgView &b, const double *const *dt, BgD<ND> &d) {
  d.grhom = dual_at<ND>(b.grhom, dt, DEB_HDR_GRHOM);
  d.Omegam = dual_at<ND>(b.Omegam, dt, DEB_P_OMEGAM);
  d.Omegab = dual_at<ND>(b.Omegab, dt, DEB_P_OMEGAB);
  d.OmegaDE = dual_at<ND>(b.OmegaDE, dt, DEB_HDR_OMEGADE);
  d.Omegak = dual_at<ND>(b.Omegak, dt, DEB_P_OMEGAK);
  d.w0 = dual_at<ND>(b.w0, dt, DEB_P_W0);
  d.wa = dual_at<ND>(b.wa, dt, DEB_P_WA);
  d.amnu = dual_at<ND>(b.amnu, dt, DEB_HDR_AMNU);
  d.akthom = dual_at<ND>(b.akthom, dt, DEB_HDR_AKTHOM);
  d.grhog = b.grhog; d.grhor = b.grhor; d.Neff = b.Neff; d.Nmnu = b.Nmnu; d.cs2DE = b.cs2DE; d.lambda_de = b.lambda_de;
}

template <class T> struct PBgX : PBg<T> {
  T rhonu, pnu;  // massive-neutrino density and pressure ratios (for the output fields)
};

template <class T, class K>
__device__ __forceinline__ T cubic_g(double x0, double x1, const K &y0, const K &y1, const K &S0, const K &S1, const T &xq) {
  const double hl = x1 - x0;
  const T t = (xq - x0) / hl;
  const T A = 1.0 - t;
  const T B = t;
  return A * y0 + B * y1 + ((A * A * A - A) * S0 + (B * B * B - B) * S1) * (hl * hl / 6.0);
}
template <class T, class K>
__device__ __forceinline__ T cubic_deriv_g(double x0, double x1, const K &y0, const K &y1, const K &S0, const K &S1, const T &xq) {
  const double hl = x1 - x0;
  const T d = xq - x0;
  const K bb = (y1 - y0) / hl - hl * (S1 + 2.0 * S0) / 6.0;
  return bb + 2.0 * (S0 / 2.0) * d + 3.0 * ((S1 - S0) / (6.0 * hl)) * d * d;
}

// Background at (tau, a) with derivatives in ND directions (pbg_eval's formulas, every table
// value and density parameter carrying the derivative-table parts of its direction).  Switch
// decisions use the values.  Every lane of the wave takes part (per-bin factors on lanes
// VQ_LANE0..).
template <int ND>
__device__ __forceinline__ void pbg_eval_ad(const BgView &b, const BgD<ND> &bd, const double *const *dt, const PLay &L, double k,
                                            const Dn<ND> &tau, const Dn<ND> &a, PBgX<Dn<ND>> &g, int lane) {
  typedef Dn<ND> D;
  g.a = a;
  g.tau = tau;
  const D loga = dlog(a);
  const int N = b.N;
  const int i = spl_index(b.loga, N, loga.v);
  const int j = spl_index_uniform(b.nu_x0, b.nu_dxinv, DEB_NNU, loga.v);
  auto kn = [&](const double *arr, int slot, int idx) { return dual_at<ND>(arr[idx], dt, deb_array_offset(slot, N) + idx); };
  const double x0 = b.loga[i], x1 = b.loga[i + 1];
  const D cs2c = cubic_g(x0, x1, kn(b.cs2a, DEB_A_CS2A, i), kn(b.cs2a, DEB_A_CS2A, i + 1), kn(b.S_cs2a_of_loga, DEB_A_S_CS2A_OF_LOGA, i),
                         kn(b.S_cs2a_of_loga, DEB_A_S_CS2A_OF_LOGA, i + 1), loga);
  const D ye0 = kn(b.xe, DEB_A_XE, i), ye1 = kn(b.xe, DEB_A_XE, i + 1);
  const D Se0 = kn(b.S_xe_of_loga, DEB_A_S_XE_OF_LOGA, i), Se1 = kn(b.S_xe_of_loga, DEB_A_S_XE_OF_LOGA, i + 1);
  g.cs2 = cs2c / a;
  g.xe = cubic_g(x0, x1, ye0, ye1, Se0, Se1, loga);
  const D dxe = cubic_deriv_g(x0, x1, ye0, ye1, Se0, Se1, loga);
  g.pb43 = 4.0 / 3.0 * (bd.grhog / (bd.grhom * bd.Omegab * a));
  const double nx0 = b.nu_loga[j], nx1 = b.nu_loga[j + 1];
  const D rhonu = dexp(cubic_g(nx0, nx1, kn(b.nu_lr, DEB_A_NU_LOGRHO, j), kn(b.nu_lr, DEB_A_NU_LOGRHO, j + 1),
                               kn(b.nu_lrS, DEB_A_NU_LOGRHO_S, j), kn(b.nu_lrS, DEB_A_NU_LOGRHO_S, j + 1), loga));
  const D pnu = dexp(cubic_g(nx0, nx1, kn(b.nu_lp, DEB_A_NU_LOGP, j), kn(b.nu_lp, DEB_A_NU_LOGP, j + 1), kn(b.nu_lpS, DEB_A_NU_LOGP_S, j),
                             kn(b.nu_lpS, DEB_A_NU_LOGP_S, j + 1), loga));
  g.rhonu = rhonu;
  g.pnu = pnu;
  g.w_Q = bd.w0 + bd.wa * (1.0 - a);
  g.rho_Q = dexp(-3.0 * (1.0 + bd.w0 + bd.wa) * loga + 3.0 * (a - 1.0) * bd.wa);
  const D a2 = a * a;
  const D grho = bd.grhom * bd.Omegam / a + (bd.grhog + bd.grhor * (bd.Neff + bd.Nmnu * rhonu)) / a2 + bd.grhom * bd.OmegaDE * g.rho_Q * a2 +
                 bd.grhom * bd.Omegak;
  g.aH = dsqrt(grho / 3.0);
  g.xeprime = dxe * g.aH;
  const D gpres = ((bd.grhog + bd.grhor * bd.Neff) / 3.0 + bd.grhor * bd.Nmnu * pnu) / a2 + g.w_Q * bd.grhom * bd.OmegaDE * g.rho_Q * a2;
  g.aHH = 0.5 * (g.aH * g.aH - gpres);
  const D wprime = -bd.wa * g.aH * a;
  g.ca2_Q = g.w_Q - wprime / 3.0 / ((1.0 + g.w_Q) + 1e-6) / g.aH;
  g.opac = g.xe * bd.akthom / a2;
  g.tauc = 1.0 / dmax(g.opac, 1e-30);
  g.taucprime = g.tauc * (2.0 * g.aH - g.xeprime / dmax(g.xe, 1e-30));
  const double tv = tau.v, tc = g.tauc.v, th = 1.0 / fmax(g.aH.v, 1e-30), tk = 1.0 / fmax(k, 1e-30);
  g.rsa = b.use_rsa && (tc / fmax(tv, 1e-30) > b.rsa_c) && (tv * k > b.rsa_k);
  g.ufa = !g.rsa && b.use_ufa && (tv * k > b.ufa_k);
  const bool uncoupled = (tc / fmax(tk, 1e-30) > b.tca_k) || ((tc / fmax(th, 1e-30) > b.tca_h) && (tc / fmax(tk, 1e-30) > 0.1 * b.tca_k));
  g.tca = !g.rsa && b.use_tca && !uncoupled;
  g.amnu_a = a * bd.amnu;
  g.ia = 1.0 / a;
  g.ia2 = 1.0 / a2;
  g.iaH = 1.0 / dmax(g.aH, 1e-30);
  g.ipb = 1.0 / (1.0 + g.pb43);
  g.ipbm = 1.0 / dmax(g.pb43, 1e-30);
  const int qb = lane - VQ_LANE0;
  const D aq = g.amnu_a / L.q[(qb >= 0 && qb < L.nq) ? qb : 0];
  g.ivl = dsqrt(1.0 + aq * aq);
  g.vl = 1.0 / g.ivl;
}

// ---------------------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------------------
// v of the lane CTRL / ROWS name (a DPP move of both halves); a lane the move does not write,
// and a lane whose source lies outside its row, receives 0
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_or_zero(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = (int)(unsigned)u, hi = (int)(unsigned)(u >> 32);
  const unsigned tlo = (unsigned)__builtin_amdgcn_update_dpp(0, lo, CTRL, ROWS, 0xf, false);
  const unsigned thi = (unsigned)__builtin_amdgcn_update_dpp(0, hi, CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)thi << 32) | tlo));
}
// Wave-wide sum, returned to every lane (all 64 lanes active): an inclusive scan within each row
// of 16 lanes by row shifts of 1, 2, 4 and 8, then lane 15's total into rows 1 and 3 and lane
// 31's into rows 2 and 3, so that lane 63 holds the sum, which readlane returns.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_or_zero<0x111, 0xf>(v);  // row_shr:1
  v += dpp_or_zero<0x112, 0xf>(v);  // row_shr:2
  v += dpp_or_zero<0x114, 0xf>(v);  // row_shr:4
  v += dpp_or_zero<0x118, 0xf>(v);  // row_shr:8
  v += dpp_or_zero<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_or_zero<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return lane_bcast(v, WAVE - 1);
}
// block-wide sums of K values per thread; red: LDS scratch of (NT / 64) K doubles; every thread
// receives the totals
template <int NT, int K> __device__ __forceinline__ void block_sums(double (&v)[K], double *red, int tid) {
  const int w = tid / WAVE, lane = tid & (WAVE - 1);
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < K; ++q) red[w * K + q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < K; ++q) {
    double s = 0.0;
    for (int ww = 0; ww < NT / WAVE; ++ww) s += red[ww * K + q];
    v[q] = s;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------------
// adjoint workspace (LDS, after the forward solver's)
// ---------------------------------------------------------------------------------------
struct AdjWs {
  double *kb;    // 8 n: stage cotangents kbar_s, replaced by r_s = W^-T kbar_s
  double *lam;   // n: cotangent of the step's end state
  double *ucar;  // n: cotangent of the stage inputs 6..8 (carried down)
  double *yb0;   // n: cotangent of the step's start state
  double *ust;   // n: stage input (recomputed)
  double *usb;   // n: vector-Jacobian product at the stage point
  double *dTb;   // n: cotangent of df/dtau at the Jacobian point
  double *rho;   // n: sum_s k_s[0] r_s
  double *pend;  // n: output cotangents of the step's start state
  double *yn;    // n: the step's end state (output interpolation)
  double *part;  // (NT / 64) x 128: per-row-group column sums
  double *red;   // reduction scratch, 64 doubles
  double *cb;    // NQMAX + 3: per-chain head terms of the transposed solve
  double *xd2;   // 3 NQMAX + 17: dense right-hand side of the transposed solve
  double *z19;   // 2 NRED: reduced right-hand side and solution
  double *sc;    // scalars: see SC_*
  double *taub;  // MAXOUT: cotangents of the output times
  double *gb;    // 20: output-field cotangent of the current output
  int *perm64;   // n + 256: the right-hand side rows in class-grouped order for one wavefront (rows_dot)
};
enum { SC_T0B, SC_HB, SC_TSB, SC_TEB, SC_T0BAR, SC_T1BAR, SC_G0, SC_COUNT = SC_G0 + NPAR };

__host__ __device__ constexpr long adj_extra_doubles(long n, int nt) {
  return 18 * n + (nt / WAVE) * 128 + 64 + (NQMAX + 3) + (3 * NQMAX + 17) + 2 * NRED + SC_COUNT + MAXOUT + 20 + 8 + (n + 256 + 1) / 2;
}

// ---------------------------------------------------------------------------------------
// per-step record of the forward replay: everything the reverse sweep reads of a step's
// stages and factorisation -- K (8 n), J0, dT, the chains' rtd, tlo, tmul (5 n), the chain
// heads' parent entries of Jh (nch) and the dense part: with a compile-time bin count the
// reduced inverse (NRED^2), Q and the neutrino columns of G (2 x 2 NN), the core rows' neutrino
// columns of the Schur block (NCORE NN), the 3 x 3 bin inverses (9 NQ) and the pivot maps
// (2 NRED); otherwise the m x m inverse and its pivot map (m^2 + m).  Stored during the forward
// replay and read back in the reverse sweep instead of recomputing the step (the same values).
// ---------------------------------------------------------------------------------------
__host__ __device__ constexpr long adj_rec_doubles(long n, long m, long nq, long nch, bool reduced) {
  const long dense = reduced ? NRED * NRED + 4 * 3 * nq + NCORE * 3 * nq + 9 * nq + 2 * NRED : m * m + m;
  return (13 * n + nch + dense + 3) / 4 * 4;
}
// the bin counts with a compile-time instantiation (the bordered reduction)
__host__ __device__ constexpr bool adj_reduced(int nq) { return nq == 3 || nq == 5 || nq == 8; }

template <int NQ, int NT> __device__ __forceinline__ void save_record(const PLay &L, const Ws &W, double *__restrict__ rec, int tid) {
  const int n = L.n, m = L.m, nch = L.nch;
  for (int p = tid; p < 8 * n; p += NT) rec[p] = W.K[p];
  double *r = rec + 8 * n;
  for (int p = tid; p < n; p += NT) {
    r[p] = W.J0[p];
    r[n + p] = W.dT[p];
    r[2 * n + p] = W.rtd[p];
    r[3 * n + p] = W.tlo[p];
    r[4 * n + p] = W.tmul[p];
  }
  r += 5 * n;
  for (int c = tid; c < nch; c += NT) r[c] = W.Jh[c * m + W.cpar[c]];
  r += nch;
  if constexpr (NQ > 0) {
    constexpr int NN = 3 * NQ;
    for (int p = tid; p < NRED * NRED; p += NT) r[p] = W.Sinv[p];
    r += NRED * NRED;
    for (int p = tid; p < 2 * NN; p += NT) {
      r[p] = W.Q[p];
      r[2 * NN + p] = W.G[30 + p];  // G[2 (15 + jn) + t]
    }
    r += 4 * NN;
    for (int p = tid; p < NCORE * NN; p += NT) {
      const int i = p / NN, jn = p - i * NN;
      r[p] = W.S[core_index(i, NQ) * m + 15 + jn];
    }
    r += NCORE * NN;
    for (int p = tid; p < 9 * NQ; p += NT) r[p] = W.Minv[p];
    r += 9 * NQ;
    for (int p = tid; p < NRED; p += NT) {
      r[p] = (double)W.rmap[p];
      r[NRED + p] = (double)W.pi[p];
    }
  } else {
    for (int p = tid; p < m * m; p += NT) r[p] = W.Sinv[p];
    for (int p = tid; p < m; p += NT) r[m * m + p] = (double)W.pi[p];
  }
}

template <int NQ, int NT> __device__ __forceinline__ void load_record(const PLay &L, Ws &W, const double *__restrict__ rec, int tid) {
  const int n = L.n, m = L.m, nch = L.nch;
  for (int p = tid; p < 8 * n; p += NT) W.K[p] = rec[p];
  const double *r = rec + 8 * n;
  for (int p = tid; p < n; p += NT) {
    W.J0[p] = r[p];
    W.dT[p] = r[n + p];
    W.rtd[p] = r[2 * n + p];
    W.tlo[p] = r[3 * n + p];
    W.tmul[p] = r[4 * n + p];
  }
  r += 5 * n;
  for (int c = tid; c < nch; c += NT) W.Jh[c * m + W.cpar[c]] = r[c];
  r += nch;
  if constexpr (NQ > 0) {
    constexpr int NN = 3 * NQ;
    W.Sinv = W.R;
    for (int p = tid; p < NRED * NRED; p += NT) W.Sinv[p] = r[p];
    r += NRED * NRED;
    for (int p = tid; p < 2 * NN; p += NT) {
      W.Q[p] = r[p];
      W.G[30 + p] = r[2 * NN + p];
    }
    r += 4 * NN;
    for (int p = tid; p < NCORE * NN; p += NT) {
      const int i = p / NN, jn = p - i * NN;
      W.S[core_index(i, NQ) * m + 15 + jn] = r[p];
    }
    r += NCORE * NN;
    for (int p = tid; p < 9 * NQ; p += NT) W.Minv[p] = r[p];
    r += 9 * NQ;
    for (int p = tid; p < NRED; p += NT) {
      W.rmap[p] = (int)r[p];
      W.pi[p] = (int)r[NRED + p];
    }
  } else {
    W.Sinv = W.S;
    for (int p = tid; p < m * m; p += NT) W.Sinv[p] = r[p];
    for (int p = tid; p < m; p += NT) W.pi[p] = (int)r[m * m + p];
  }
}

// ---------------------------------------------------------------------------------------
// transposed structured solve W^T x = b (v in LDS, in place), wavefront 0 alone (tid = lane);
// the factorisation of the forward step (factorise / reduce_dense / gauss_jordan_wave) is read:
// W = I/(gamma h) - J with row 0 [W00, 0...], column 0 -J0, the tail chains T = U L (bottom-up
// elimination: diagonal td, multipliers tmul), the Schur-corrected dense block and, with a
// compile-time bin count, its bordered reduction (R^-1 kept in W.Sinv with rmap / pi).
// ---------------------------------------------------------------------------------------
template <int MC>
__device__ __forceinline__ void wsolveT(const PLay &L, Ws &W, AdjWs &A, double W00, double *v, int lane) {
  const int n = L.n, m = L.m, nch = L.nch;
  // 1. chains: L^T w = b from the bottom; the head's row keeps beta = b_h + lo_1 w_1
  for (int c = lane; c < nch; c += WAVE) {
    const int h = W.chead[c], s = W.cstride[c], len = W.clen[c];
    int ii = h + (len - 1) * s;
    double w = 0.0;
    if (len > 1) {
      w = v[ii] * W.rtd[ii];
      v[ii] = w;
      int j = len - 2;
      for (; j >= 4; j -= 4) {  // four elements: every operand loaded first, then the serial recurrence
        const int i0 = ii - s, i1 = ii - 2 * s, i2 = ii - 3 * s, i3 = ii - 4 * s;
        const double a0 = v[i0], a1 = v[i1], a2 = v[i2], a3 = v[i3];
        const double l0 = W.tlo[ii], l1 = W.tlo[i0], l2 = W.tlo[i1], l3 = W.tlo[i2];
        const double r0 = W.rtd[i0], r1 = W.rtd[i1], r2 = W.rtd[i2], r3 = W.rtd[i3];
        const double w0 = (a0 + l0 * w) * r0, w1 = (a1 + l1 * w0) * r1, w2 = (a2 + l2 * w1) * r2;
        w = (a3 + l3 * w2) * r3;
        v[i0] = w0;
        v[i1] = w1;
        v[i2] = w2;
        v[i3] = w;
        ii = i3;
      }
      for (; j >= 1; --j) {
        ii -= s;
        w = (v[ii] + W.tlo[ii + s] * w) * W.rtd[ii];
        v[ii] = w;
      }
    }
    A.cb[c] = v[h] + (len > 1 ? W.tlo[h + s] * w : 0.0);
  }
  wave_sync();
  // dense right-hand side: b_r, plus lo_h beta_c / td_h at a chain's parent row
  for (int r = lane; r < m; r += WAVE) {
    const int c = W.pch[r];
    double x = v[W.dmap[r]];
    if (c >= 0) {
      const int h = W.chead[c];
      x += W.tlo[h] * A.cb[c] * W.rtd[h];
    }
    A.xd2[r] = x;
  }
  wave_sync();
  // 2. dense block
  if constexpr (MC > 0) {
    constexpr int NQ = (MC - NCORE) / 3;
    constexpr int NN = 3 * NQ;
    // reduced right-hand side [b_c; Q^T b_n], scattered by the pivot rows: z19[NRED + rmap[p]] = rhs_p
    for (int i = lane; i < NRED; i += WAVE) {
      double x;
      if (i < NCORE) {
        x = A.xd2[core_index(i, NQ)];
      } else {
        x = 0.0;
        for (int jn = 0; jn < NN; ++jn) x += W.Q[2 * jn + (i - NCORE)] * A.xd2[15 + jn];
      }
      A.z19[NRED + W.rmap[i]] = x;
    }
    wave_sync();
    // [x_c; z'] = R^-T rhs: x_r = sum_row Ainv[row][pi[r]] rhs'[row]
    double zr = 0.0;
    if (lane < NRED) {
      const int col = W.pi[lane];
      for (int row = 0; row < NRED; ++row) zr += W.Sinv[row * NRED + col] * A.z19[NRED + row];
    }
    wave_sync();
    if (lane < NRED) A.z19[lane] = zr;
    wave_sync();
    for (int i = lane; i < NCORE; i += WAVE) v[W.dmap[core_index(i, NQ)]] = A.z19[i];
    // x_n = M^-T (b_n - W_cn^T x_c + G_n z'), bin by bin
    for (int jn = lane; jn < NN; jn += WAVE) {
      double t = A.xd2[15 + jn] + W.G[2 * (15 + jn)] * A.z19[NCORE] + W.G[2 * (15 + jn) + 1] * A.z19[NCORE + 1];
      for (int i = 0; i < NCORE; ++i) t -= W.S[core_index(i, NQ) * m + 15 + jn] * A.z19[i];
      W.f[jn] = t;                // W.f is free during the reverse sweep
    }
    wave_sync();
    for (int jn = lane; jn < NN; jn += WAVE) {
      const int l = jn / NQ, q = jn - l * NQ;
      const double *mi = W.Minv + 9 * q;
      v[W.dmap[15 + jn]] = mi[l] * W.f[q] + mi[3 + l] * W.f[NQ + q] + mi[6 + l] * W.f[2 * NQ + q];
    }
  } else {
    // generic bin count: x_c = sum_r Sinv[r][pi[c]] b_r
    for (int c = lane; c < m; c += WAVE) {
      const int col = W.pi[c];
      double x = 0.0;
      for (int r = 0; r < m; ++r) x += W.Sinv[r * m + col] * A.xd2[r];
      v[W.dmap[c]] = x;
    }
  }
  wave_sync();
  // 3. chains from the head: x_h = (beta + Jh x_parent) / td_h, x_{j+1} = w_{j+1} - tmul_j x_j
  for (int c = lane; c < nch; c += WAVE) {
    const int h = W.chead[c], s = W.cstride[c], len = W.clen[c], pr = W.cpar[c];
    double x = (A.cb[c] + W.Jh[c * m + pr] * v[W.dmap[pr]]) * W.rtd[h];
    v[h] = x;
    int ii = h;
    int j = 1;
    for (; j + 3 < len; j += 4) {  // four elements at a time, as above
      const int i0 = ii + s, i1 = ii + 2 * s, i2 = ii + 3 * s, i3 = ii + 4 * s;
      const double a0 = v[i0], a1 = v[i1], a2 = v[i2], a3 = v[i3];
      const double m0 = W.tmul[ii], m1 = W.tmul[i0], m2 = W.tmul[i1], m3 = W.tmul[i2];
      const double x0 = a0 - m0 * x, x1 = a1 - m1 * x0, x2 = a2 - m2 * x1;
      x = a3 - m3 * x2;
      v[i0] = x0;
      v[i1] = x1;
      v[i2] = x2;
      v[i3] = x;
      ii = i3;
    }
    for (; j < len; ++j) {
      const int in = ii + s;
      x = v[in] - W.tmul[ii] * x;
      v[in] = x;
      ii = in;
    }
  }
  wave_sync();
  // 4. row 0: W00 x0 - sum_p J0[p] x_p = b0
  double acc = 0.0;
  for (int p = 1 + lane; p < n; p += WAVE) acc += W.J0[p] * v[p];
  acc = wave_sum(acc);
  if (lane == 0) v[0] = (v[0] + acc) / W00;
  wave_sync();
}

// ---------------------------------------------------------------------------------------
// Transposed linear right-hand side, dense rows: sum_i c_i A_iq over the dense-block rows i for
// one column q (a dense-block entry or a chain head), where A = d f / d y at the background g
// (f is linear in y_1..).  Transposed by hand from metric() and core_row(): the rows' cotangents
// are carried back to the metric's building blocks (h', eta', alpha, the density, pressure and
// velocity sums, the photon and neutrino fields with their radiation-streaming and
// tight-coupling forms), and each column collects its direct terms and its share of those.  The
// same values as evaluating the rows on the unit vector e_q, up to rounding.
// ---------------------------------------------------------------------------------------
// vj, ivj: the velocity factor of column q's momentum bin and its reciprocal (read from the bin
// lanes by the caller, in control flow every lane of the wave takes; any value for other columns)
template <class T, class BV>
__device__ __forceinline__ T dense_colT(const BV &b, const PLay &L, double k, const PBg<T> &g, const double *c, int q, const T &vj,
                                        const T &ivj) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0, idxgp = L.idxgp, idxr = L.idxr;
  const double k2 = k * k, ik2 = 1.0 / (k * k), sk2 = fmax(k * k, 1e-30);
  const auto Omegac = b.Omegam - b.Omegab;
  const double cs2_Q = b.cs2DE;
  const bool rsa = g.rsa, tca = g.tca, ufa = g.ufa;
  const bool gam = !rsa && !tca;  // photon l <= 2 and polarisation rows in their full form
  // 1. rows -> the metric's fields
  const double c1 = c[1], c2 = c[2], c3 = c[3], c5 = c[5], c6 = c[6], c7 = c[7], c8 = c[8], c9 = c[9];
  T hp(0.0), ep(0.0), al(0.0), dgrho(0.0), dgpres(0.0), dg(0.0), tg(0.0), sg(0.0), dr(0.0), tr(0.0), sr(0.0);
  T tgp(0.0), tbp(0.0), dgp(0.0), dbp(0.0), tbp0(0.0), slip(0.0);
  dgrho = dgrho - g.a * c1;
  dgpres = dgpres - 3.0 * g.a * c1;
  ep = ep + c2;
  hp = hp - 0.5 * c3 - 0.5 * c5;
  if (tca) tbp = tbp + c6;
  else tg = tg + g.pb43 * g.opac * c6;
  if (!rsa) {
    if (tca) {
      dgp = dgp + c7;
      tgp = tgp + c8;
    } else {
      tg = tg - 4.0 / 3.0 * c7 + 8.0 / 15.0 * c9 - g.opac * c8;
      hp = hp - 2.0 / 3.0 * c7;
      dg = dg + 0.25 * k2 * c8;
      sg = sg - k2 * c8;
      al = al + 8.0 / 15.0 * k2 * c9;
    }
    const double cr0 = c[idxr], cr1 = c[idxr + 1], cr2 = c[idxr + 2];
    tr = tr - 4.0 / 3.0 * cr0;
    hp = hp - 2.0 / 3.0 * cr0;
    dr = dr + 0.25 * k2 * cr1;
    sr = sr - k2 * cr1;
    if (ufa) {
      sr = sr - 3.0 / dmax(g.tau, 1e-30) * cr2;
      tr = tr + 4.0 / 15.0 * cr2;
      hp = hp + 2.0 / 15.0 * cr2;
    } else {
      tr = tr + 8.0 / 15.0 * cr2;
      al = al + 8.0 / 15.0 * k2 * cr2;
    }
  }
  for (int j = 0; j < nq; ++j) {
    const double dl = L.dlfdlq[j], c0q = c[iq0 + j], c2q = c[iq0 + 2 * nq + j];
    hp = hp + dl * (1.0 / 6.0) * c0q - dl * (1.0 / 15.0) * c2q;
    ep = ep - 0.4 * dl * c2q;
  }
  if (!b.lambda_de) hp = hp - 0.5 * (1.0 + g.w_Q) * c[n - 2];
  if (tca) {  // m.tgp_tca = (-tbp - aH tb + k2 cs2 db) ipbm + ...; tbp = tbp0 + pb43 ipb slip; slip (.. dbp, dgp ..)
    tbp = tbp - g.ipbm * tgp;
    tbp0 = tbp;
    slip = g.pb43 * g.ipb * tbp;
    dbp = dbp + g.tauc * k2 * g.cs2 * g.ipb * slip;
    dgp = dgp - 0.25 * g.tauc * k2 * g.ipb * slip;
    hp = hp - 2.0 / 3.0 * dgp - 0.5 * dbp;
  }
  // 2. alpha = (h' + 6 eta') / (2 k^2); eta' = dgtheta / (2 k^2)
  hp = hp + 0.5 * ik2 * al;
  ep = ep + 3.0 * ik2 * al;
  const T dgth = 0.5 * ik2 * ep;
  tg = tg + 4.0 / 3.0 * b.grhog * g.ia2 * dgth;
  tr = tr + 4.0 / 3.0 * b.Neff * b.grhor * g.ia2 * dgth;
  // 3. pressure: dg, dr, and the neutrino, dark-energy density and velocity parts
  dg = dg + b.grhog * g.ia2 * (1.0 / 3.0) * dgpres;
  dr = dr + b.grhor * b.Neff * g.ia2 * (1.0 / 3.0) * dgpres;
  // 4. h': from the full density sum, or under RSA from its form without photons and neutrinos
  T hprsa(0.0), c_eta(0.0);
  if (rsa) {
    hprsa = hp;
    const T stc = dmax(g.tauc, 1e-30);
    const T f3 = 3.0 / (sk2 * stc);
    hprsa = hprsa + (4.0 / sk2 * g.aH + 4.0 / (sk2 * stc) * 0.5) * dg + (-0.5 + f3 * (-g.taucprime / stc) * 0.5) * tg + 4.0 / sk2 * g.aH * dr -
            0.5 * tr;
    c_eta = c_eta - 4.0 * dg + f3 * sk2 * tg - 4.0 * dr;
  } else {
    dgrho = dgrho + g.iaH * hp;
    c_eta = c_eta + 2.0 * k2 * g.iaH * hp;
  }
  const T wo = g.iaH * hprsa;
  c_eta = c_eta + 2.0 * k2 * g.iaH * hprsa;
  const T dens = dgrho + wo;  // the matter and neutrino density terms enter both sums alike
  // 5. the column
  if (q == 2) return c_eta;
  if (q == 3) return b.grhom * Omegac * g.ia * dens;
  if (q == 4) return -c3 - g.aH * c[4] + b.grhom * Omegac * g.ia * dgth;
  if (q == 5) {
    T v = b.grhom * b.Omegab * g.ia * dens;
    if (!tca) v = v + k2 * g.cs2 * c6;
    if (rsa) v = v + 3.0 / (sk2 * dmax(g.tauc, 1e-30)) * g.cs2 * sk2 * tg;
    if (tca) v = v + k2 * g.cs2 * g.ipbm * tgp + k2 * g.cs2 * g.ipb * tbp0;
    return v;
  }
  if (q == 6) {
    T v = -c5 + b.grhom * b.Omegab * g.ia * dgth;
    if (!tca) v = v + (-g.aH - g.pb43 * g.opac) * c6;
    if (gam) v = v + g.opac * c8;
    if (rsa) {
      const T stc = dmax(g.tauc, 1e-30);
      v = v + 4.0 / (sk2 * stc) * dg + 3.0 / (sk2 * stc) * (-g.taucprime / stc - g.aH) * tg;
    }
    if (tca) v = v - g.aH * g.ipbm * tgp - g.aH * g.ipb * tbp0 + (2.0 * g.pb43 * g.ipb * g.aH - g.tauc * g.aHH * g.ipb) * slip - dbp;
    return v;
  }
  if (q == 7) {
    T v = b.grhog * g.ia2 * dgrho;
    if (!rsa) v = v + dg;
    if (tca) v = v + 0.25 * k2 * tgp + 0.25 * k2 * g.pb43 * g.ipb * tbp0 - 0.5 * g.tauc * g.aH * k2 * g.ipb * slip;
    return v;
  }
  if (q == 8) {
    T v(0.0);
    if (!rsa) v = v + tg;
    if (tca) v = v - 2.0 * g.pb43 * g.ipb * g.aH * slip - 4.0 / 3.0 * dgp;
    return v;
  }
  const double cg0 = c[idxgp], cg1 = c[idxgp + 1], cg2 = c[idxgp + 2];
  if (q == 9) {
    T v(0.0);
    if (gam) v = v + (-g.opac + 0.1 * g.opac) * c9 + g.opac * 0.5 * cg0 + 0.1 * g.opac * cg2;
    if (!rsa) v = v + 0.5 * sg;
    if (tca) v = v - 0.5 * k2 * tgp - 0.5 * k2 * g.pb43 * g.ipb * tbp0;
    return v;
  }
  if (q == 10) return gam ? T(-3.0 / 5.0 * k * c9) : T(0.0);
  if (q >= idxgp && q <= idxgp + 3) {
    if (!gam) return T(0.0);
    const int l = q - idxgp;
    if (l == 0) return (-g.opac + g.opac * 0.5) * cg0 + k * (1.0 / 3.0) * cg1 + 0.1 * g.opac * cg2 + 0.1 * g.opac * c9;
    if (l == 1) return T(-k * cg0 + k * 0.2 * 2.0 * cg2) - g.opac * cg1;
    if (l == 2) return 0.1 * g.opac * c9 + g.opac * 0.5 * cg0 - 2.0 * k * (1.0 / 3.0) * cg1 + (-g.opac + 0.1 * g.opac) * cg2;
    return T(-k * 0.2 * 3.0 * cg2);
  }
  if (q >= idxr && q <= idxr + 3) {
    const int l = q - idxr;
    if (l == 0) {
      T v = b.grhor * b.Neff * g.ia2 * dgrho;
      if (!rsa) v = v + dr;
      return v;
    }
    if (rsa) return T(0.0);
    if (l == 1) return tr;
    if (l == 2) return 0.5 * sr;
    return ufa ? T(0.0) : T(-0.6 * k * c[idxr + 2]);
  }
  if (q >= iq0 && q < iq0 + 4 * nq) {
    const int l = (q - iq0) / nq, j = (q - iq0) - l * nq;
    const T v = vj, kv = k * v;
    if (l == 0)
      return kv * (1.0 / 3.0) * c[iq0 + nq + j] + L.w[j] * ivj * (b.grhor * b.Nmnu * g.ia2 * dens) +
             L.w[j] * v * (1.0 / 3.0) * (b.grhor * b.Nmnu * g.ia2 * dgpres);
    if (l == 1) return -kv * c[iq0 + j] + kv * 0.2 * 2.0 * c[iq0 + 2 * nq + j] + L.w[j] * (b.Nmnu * b.grhor * k * g.ia2 * dgth);
    if (l == 2) return -kv * (2.0 / 3.0) * c[iq0 + nq + j];
    return -kv * 0.2 * 3.0 * c[iq0 + 2 * nq + j];
  }
  const T a2 = g.a * g.a;
  if (q == n - 2) {
    T v = b.grhom * b.OmegaDE * g.rho_Q * a2 * (dens + cs2_Q * dgpres);
    if (!b.lambda_de) v = v - 3.0 * (cs2_Q - g.w_Q) * g.aH * c[n - 2] + cs2_Q / (1.0 + g.w_Q) * k2 * c[n - 1];
    return v;
  }
  if (q == n - 1) {
    T v = (1.0 + g.w_Q) * g.rho_Q * b.grhom * b.OmegaDE * a2 * (dgth + (cs2_Q - g.ca2_Q) * (3.0 * g.aH * ik2) * dgpres);
    if (!b.lambda_de)
      v = v + (-(1.0 + g.w_Q) - 9.0 * (1.0 + g.w_Q) * (cs2_Q - g.ca2_Q) * g.aH * g.aH / k2) * c[n - 2] - (1.0 - 3.0 * cs2_Q) * g.aH * c[n - 1];
    return v;
  }
  return T(0.0);
}

// ---------------------------------------------------------------------------------------
// out[q] += sum_r rb[r] A_rq (q >= 1) for the linear part of the right-hand side at the
// background g: dense columns by unit-vector evaluation of the dense rows (row group per
// wavefront), tail-chain entries from the chain coefficients.  All threads.
// With a dual background (T = Dn<2>, directions a and tau) it accumulates instead
// sum_r (wa[r] dA_rq/da + wt[r] dA_rq/dtau) (wa = rb, wt = rb2).
// ---------------------------------------------------------------------------------------
template <int NT, class T, class BV>
__device__ __forceinline__ void colT(const BV &b, const PLay &L, double k, const PBg<T> &g, const double *rb, const double *rb2, double *out,
                                     Ws &W, AdjWs &A, int tid) {
  constexpr int NR = NT / WAVE;
  const int m = L.m, nch = L.nch, n = L.n;
  const int rg = __builtin_amdgcn_readfirstlane(tid / WAVE), lane = tid & (WAVE - 1);
  auto wsum = [&](const T &v, int p) -> double {
    if constexpr (IsDouble<T>::value) return rb[p] * val(v);
    else return rb[p] * v.d[0] + rb2[p] * v.d[1];
  };
  // dense-block and chain-head columns by the transposed rows (dense_colT), one column per
  // thread, the columns spread over the wavefronts; with a dual background the two directions
  // take their own cotangents (rb for a, rb2 for tau)
  for (int c0 = 0; c0 < m + nch; c0 += NT) {  // a uniform trip count: the bin factors are read with every lane active
    const int c = c0 + lane * NR + rg;
    const int q = (c < m + nch) ? ((c < m) ? W.dmap[c] : W.chead[c - m]) : 0;
    const int j = (q >= L.iq0 && q < L.iq0 + 4 * L.nq) ? (q - L.iq0) % L.nq : 0;
    T vj = lane_bcast(g.vl, VQ_LANE0), ivj = lane_bcast(g.ivl, VQ_LANE0);
    for (int jj = 1; jj < L.nq; ++jj) {
      const T t = lane_bcast(g.vl, VQ_LANE0 + jj), ti = lane_bcast(g.ivl, VQ_LANE0 + jj);
      if (j == jj) { vj = t; ivj = ti; }
    }
    if (c < m + nch) {
      double sum;
      if constexpr (IsDouble<T>::value) {
        sum = dense_colT<T>(b, L, k, g, rb, q, vj, ivj);
      } else {
        sum = dense_colT<T>(b, L, k, g, rb, q, vj, ivj).d[0] + dense_colT<T>(b, L, k, g, rb2, q, vj, ivj).d[1];
      }
      A.part[c] = sum;
    }
  }
  __syncthreads();
  for (int c = tid; c < m + nch; c += NT) {
    const int idx = (c < m) ? W.dmap[c] : W.chead[c - m];
    out[idx] += A.part[c];
  }
  __syncthreads();
  // tail chains: f_i = lo_i y_{i-s} + di_i y_i + up_i y_{i+s}
  for (int i = tid; i < n; i += NT) {
    const int rk = W.rk[i];
    if (rk < 0) continue;
    const int ch = rk >> 8, l = rk & 255, s = W.rs[i];
    auto vsel = [&](int q) { return g.vq_l(L, q); };
    T lo, di, up;
    tail_coef<T>(ch, l, L, k, g, W.tca[i], W.tcb[i], vsel, lo, di, up);
    double acc = wsum(di, i);
    const bool head = (i == W.chead[ch]);
    if (!head) {  // row i - s is a tail row: its up coefficient
      T lo2, di2, up2;
      tail_coef<T>(ch, l - 1, L, k, g, W.tca[i - s], W.tcb[i - s], vsel, lo2, di2, up2);
      acc += wsum(up2, i - s);
    }
    if (i + s < n && W.rk[i + s] >= 0 && (W.rk[i + s] >> 8) == ch) {  // row i + s: its lo coefficient
      T lo3, di3, up3;
      tail_coef<T>(ch, l + 1, L, k, g, W.tca[i + s], W.tcb[i + s], vsel, lo3, di3, up3);
      acc += wsum(lo3, i + s);
    }
    out[i] += acc;
    if (head) out[W.dmap[W.cpar[ch]]] += wsum(lo, i);  // the head row reads its parent
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------------
// Directional contraction: wave w evaluates the right-hand side rows lo.. at (tau, a, y) with
// the dual background of its directions gdir = w ND + l and returns sum_rows r[i] df_i/d(dir)
// (in acc[l], summed over the calls that share one accumulator).  ND = 8 / (NT / 64).
// ---------------------------------------------------------------------------------------
template <int ND> struct DirCtx {
  const double *dt[ND];  // derivative table of each local direction (NULL for a and tau)
  int ia, it;            // local slot of the a and tau directions (-1: not in this wave)
  BgD<ND> bd;
};

template <int ND>
__device__ __forceinline__ void dir_ctx(const BgView &b, const double *dtab_c, long tstride, int wave, DirCtx<ND> &c) {
  c.ia = -1;
  c.it = -1;
#pragma unroll
  for (int l = 0; l < ND; ++l) {
    const int gd = wave * ND + l;
    c.dt[l] = (gd >= 2) ? dtab_c + (long)(gd - 2) * tstride : nullptr;
    if (gd == 0) c.ia = l;
    if (gd == 1) c.it = l;
  }
  make_bgd<ND>(b, c.dt, c.bd);
}

template <int ND>
__device__ __forceinline__ void dual_point(const DirCtx<ND> &c, double tau, double a, Dn<ND> &td, Dn<ND> &ad) {
  td = Dn<ND>(tau);
  ad = Dn<ND>(a);
#pragma unroll
  for (int l = 0; l < ND; ++l) {
    td.d[l] = (l == c.it) ? 1.0 : 0.0;
    ad.d[l] = (l == c.ia) ? 1.0 : 0.0;
  }
}

template <int ND, class YF>
__device__ __forceinline__ void rows_dot(const BgView &b, const DirCtx<ND> &c, const PLay &L, double k, const PBgX<Dn<ND>> &g, const YF &yv,
                                         const double *r, int row_lo, Ws &W, const int *perm64, int lane, double (&acc)[ND]) {
  typedef Dn<ND> D;
  Met<D> m;
  metric<D>(c.bd, L, k, g, yv, m);
  const int n = L.n;
  const int nit = (n + WAVE - 1) / WAVE;
  for (int it = 0; it < nit; ++it) {  // uniform trip count: every lane takes part in the shuffles
    const int i = perm64[lane + it * WAVE];  // rows grouped by formula class, so that a wave takes few branches
    const int ic = i >= 0 ? i : 0;
    const int src = VQ_LANE0 + W.rbin[ic];
    D vsh;
    vsh.v = __shfl(g.vl.v, src);
#pragma unroll
    for (int l = 0; l < ND; ++l) vsh.d[l] = __shfl(g.vl.d[l], src);
    auto vrow = [&](int) { return vsh; };
    if (i >= 0 && i >= row_lo) {
      const int rk = W.rk[i];
      D f;
      if (rk < 0) {
        f = core_row<D>(i, c.bd, L, k, g, m, yv, vrow);
      } else {
        const int ch = rk >> 8, l = rk & 255, s = W.rs[i];
        D lo, di, up;
        tail_coef<D>(ch, l, L, k, g, W.tca[i], W.tcb[i], vrow, lo, di, up);
        const double yn = (i + s < n) ? yv(i + s) : 0.0;
        f = lo * yv(i - s) + di * yv(i) + up * yn;
      }
      const double ri = r[i];
#pragma unroll
      for (int l = 0; l < ND; ++l) acc[l] += ri * f.d[l];
    }
  }
}

// rows_dot over the dense-block rows alone (one slot per dense row, in block order)
template <int ND, class YF>
__device__ __forceinline__ void rows_dot_dense(const DirCtx<ND> &c, const PLay &L, double k, const PBgX<Dn<ND>> &g, const YF &yv,
                                               const double *r, Ws &W, int lane, double (&acc)[ND]) {
  typedef Dn<ND> D;
  Met<D> m;
  metric<D>(c.bd, L, k, g, yv, m);
  const int nit = (L.m + WAVE - 1) / WAVE;
  for (int it = 0; it < nit; ++it) {  // uniform trip count: every lane takes part in the shuffles
    const int slot = lane + it * WAVE;
    const int i = slot < L.m ? W.dmap[slot] : -1;
    const int src = VQ_LANE0 + W.rbin[i >= 0 ? i : 0];
    D vsh;
    vsh.v = __shfl(g.vl.v, src);
#pragma unroll
    for (int l = 0; l < ND; ++l) vsh.d[l] = __shfl(g.vl.d[l], src);
    auto vrow = [&](int) { return vsh; };
    if (i >= 0) {
      const D f = core_row<D>(i, c.bd, L, k, g, m, yv, vrow);
      const double ri = r[i];
#pragma unroll
      for (int l = 0; l < ND; ++l) acc[l] += ri * f.d[l];
    }
  }
}

// The tail rows' part of sum_s r_s^T dA k_s: a tail row's coefficients (lo, di, up) do not depend on
// the stage, so its share is lo' pm_i + di' p0_i + up' pp_i with the stage sums pm_i = sum_s r_s,i
// k_s,i-st, p0_i = sum_s r_s,i k_s,i, pp_i = sum_s r_s,i k_s,i+st formed beforehand
template <int ND>
__device__ __forceinline__ void rows_dot_tails(const PLay &L, double k, const PBgX<Dn<ND>> &g, const double *pm, const double *p0,
                                               const double *pp, Ws &W, int lane, double (&acc)[ND]) {
  typedef Dn<ND> D;
  const int n = L.n;
  const int nit = (n + WAVE - 1) / WAVE;
  for (int it = 0; it < nit; ++it) {  // uniform trip count: every lane takes part in the shuffles
    const int i = lane + it * WAVE;
    const int ic = i < n ? i : 0;
    const int src = VQ_LANE0 + W.rbin[ic];
    D vsh;
    vsh.v = __shfl(g.vl.v, src);
#pragma unroll
    for (int l = 0; l < ND; ++l) vsh.d[l] = __shfl(g.vl.d[l], src);
    auto vrow = [&](int) { return vsh; };
    const int rk = i < n ? W.rk[i] : -1;
    if (rk >= 0) {
      const int ch = rk >> 8, l = rk & 255;
      D lo, di, up;
      tail_coef<D>(ch, l, L, k, g, W.tca[i], W.tcb[i], vrow, lo, di, up);
#pragma unroll
      for (int d = 0; d < ND; ++d) acc[d] += lo.d[d] * pm[i] + di.d[d] * p0[i] + up.d[d] * pp[i];
    }
  }
}

// ---------------------------------------------------------------------------------------
// The output map: 20 fields of an interpolated state, with the RSA projection
// (approximations.py:84-176, perturbations.py:528-677), contracted with gb.
// ---------------------------------------------------------------------------------------
template <class T, class BV, class Y>
__device__ __forceinline__ T out_dot(const BV &b, const PLay &L, double k, const PBgX<T> &g, bool rsa, const Y &y, const double *gb) {
  const int n = L.n, nq = L.nq, iq0 = L.iq0;
  const auto Omegac = b.Omegam - b.Omegab;
  const T a = g.a, ia = g.ia, ia2 = g.ia2, a2 = a * a;
  const double eta = y(2), deltac = y(3), thetac = y(4), deltab = y(5), thetab = y(6);
  T drhonu(0.0), fnu(0.0);
  for (int q = 0; q < nq; ++q) {
    drhonu += L.w[q] * y(iq0 + q) * g.ivq_u(q);
    fnu += T(L.w[q] * y(iq0 + nq + q));
  }
  T deltag, thetag, deltar, thetar;
  if (rsa) {
    const double sk2 = fmax(k * k, 1e-30);
    const T dgw = b.grhom * (Omegac * deltac + b.Omegab * deltab) * ia + b.grhor * b.Nmnu * drhonu * ia2 +
                  b.grhom * b.OmegaDE * y(n - 2) * g.rho_Q * a2;
    const T hp = (2.0 * k * k * eta + dgw) / dmax(g.aH, 1e-30);
    const T stc = dmax(g.tauc, 1e-30);
    deltag = 4.0 / sk2 * (g.aH * hp - sk2 * eta) + 4.0 / (sk2 * stc) * (thetab + 0.5 * hp);
    thetag = -0.5 * hp + 3.0 / (sk2 * stc) * (-g.taucprime / stc * (thetab + 0.5 * hp) + (-g.aH * thetab + g.cs2 * sk2 * deltab + sk2 * eta));
    deltar = 4.0 / sk2 * (g.aH * hp - sk2 * eta);
    thetar = -0.5 * hp;
  } else {
    deltag = T(y(7)); thetag = T(y(8)); deltar = T(y(L.idxr)); thetar = T(y(L.idxr + 1));
  }
  const double deltaq = y(n - 2), thetaq = y(n - 1);
  const T rpptq = (1.0 + g.w_Q) * g.rho_Q * b.grhom * b.OmegaDE * thetaq * a2;
  const T dgrho = b.grhom * (Omegac * deltac + b.Omegab * deltab) * ia + (b.grhog * deltag + b.grhor * (b.Neff * deltar + b.Nmnu * drhonu)) * ia2 +
                  b.grhom * b.OmegaDE * deltaq * g.rho_Q * a2;
  const T dgtheta = b.grhom * (Omegac * thetac + b.Omegab * thetab) * ia + 4.0 / 3.0 * (b.grhog * thetag + b.Neff * b.grhor * thetar) * ia2 +
                    b.Nmnu * b.grhor * k * fnu * ia2 + rpptq;
  const T aH = g.aH;
  const T hprime = (2.0 * k * k * eta + dgrho) / aH;
  const T etaprime = 0.5 * dgtheta / (k * k);
  const T alpha = (hprime + 6. * etaprime) / 2. / (k * k);
  const T mden = b.grhom * b.Omegam * ia + (b.grhor * b.Nmnu * g.rhonu) * ia2;
  const T cdm = b.grhom * (Omegac * deltac + b.Omegab * deltab) * ia;
  const T cth = b.grhom * (Omegac * thetac + b.Omegab * thetab) * ia;
  const T deltam = (cdm + (b.grhor * b.Nmnu * drhonu) * ia2) / mden;
  T thetam = (cth + b.Nmnu * b.grhor * k * fnu * ia2) / (3.0 * (b.grhom * b.Omegam * ia + b.grhor * b.Nmnu * g.rhonu * ia2));
  const T deltabc = cdm / (b.grhom * b.Omegam * ia);
  T thetabc = cth / (3.0 * (b.grhom * b.Omegam * ia) * ia2);
  thetam = thetam + alpha * k * k;
  thetabc = thetabc + alpha * k * k;
  const T deltanu = drhonu / g.rhonu, thetanu = k * fnu / (g.rhonu + g.pnu);
  const T iaH = 1.0 / aH;
  T s = gb[0] * T(eta) + gb[1] * etaprime + gb[2] * hprime + gb[3] * alpha + gb[4] * deltam + gb[5] * (thetam * iaH) + gb[6] * deltabc +
        gb[7] * (thetabc * iaH) + gb[8] * T(deltac) + gb[9] * (thetac * iaH) + gb[10] * T(deltab) + gb[11] * (thetab * iaH);
  s = s + gb[12] * deltag + gb[13] * (thetag * iaH) + gb[14] * deltar + gb[15] * (thetar * iaH) + gb[16] * deltanu + gb[17] * (thetanu * iaH) +
      gb[18] * T(deltaq) + gb[19] * (thetaq * iaH);
  return s;
}

// double background with the neutrino density / pressure ratios, lane-parallel evaluation
__device__ __forceinline__ void pbg_eval_out(const BgView &b, const PLay &L, double k, double tau, double a, PBgX<double> &g, SplCache *sc,
                                             int tid) {
  pbg_eval_lanes(b, L, k, tau, a, g, sc, tid);
  const double la = log(a);
  g.rhonu = exp(b.logrhonu(la));
  g.pnu = exp(b.logpnu(la));
}

// emit_output's RSA decision at (tau, a)
__device__ __forceinline__ bool out_rsa(const BgView &b, double k, double tau, double tauc) {
  return b.use_rsa && (tauc / fmax(tau, 1e-30) > b.rsa_c) && (tau * k > b.rsa_k);
}

// ---------------------------------------------------------------------------------------
// the adjoint of one unit, in two parts with their own launches: REPLAY the forward replay of the
// mesh, which writes each step's start state and record to the unit's checkpoint slot; otherwise
// the reverse sweep and the shifted-cosmology terms, which read them.  The replay needs the
// forward solver's LDS and registers only, so it runs at more units per compute unit.
// ---------------------------------------------------------------------------------------
template <int NQ, int NT, bool REPLAY>
__device__ __forceinline__ void adjoint_unit(int unit, const double *__restrict__ tables, long tstride, const double *__restrict__ shtab,
                                             const double *__restrict__ dtab, const double *__restrict__ dth, const int *__restrict__ unit_cosmo,
                                             const double *__restrict__ unit_k, const double *__restrict__ tau_out,
                                             const double *__restrict__ tau_out_sh, int nout, const double *__restrict__ mesh, long mstride,
                                             const int *__restrict__ mesh_len, const double *__restrict__ gbar, double *__restrict__ ckpt,
                                             long cstride, double *__restrict__ grad, double *__restrict__ lam0, double *__restrict__ tbar,
                                             int *__restrict__ status, const PLay &L, Ws &W, AdjWs &A, int tid) {
  using namespace r5;
  constexpr int MC = NQ > 0 ? 17 + 3 * NQ : 0;
  constexpr int ND = NDIR * WAVE / NT;
  const int cosmo = unit_cosmo[unit];
  const double k = unit_k[unit];
  const int n = L.n;
  const int lane = tid & (WAVE - 1), wave = __builtin_amdgcn_readfirstlane(tid / WAVE);
  __shared__ BgView s_bg;
  __syncthreads();
  if (tid == 0) { W.sc->i = -1; W.sc->j = -1; s_bg.init(tables + (long)cosmo * tstride); }
  __syncthreads();
  const BgView &b = s_bg;
  double *gu = grad + (long)unit * NPAR;
  if (!b.lambda_de && (1.0 + b.w0) * (1.0 + b.w0 + b.wa) <= 0.0) {  // the phantom divide (as the forward kernel)
    if (tid < NPAR) gu[tid] = __builtin_nan("");
    if (tid == 0) status[unit] = 4;
    return;
  }
  const double *tout = tau_out + (long)cosmo * nout;
  double tmin = tout[0], t1 = tout[0];
  for (int j = 1; j < nout; ++j) { tmin = fmin(tmin, tout[j]); t1 = fmax(t1, tout[j]); }
  const double t0 = 0.99 * fmin(tmin, starting_time(b, k));
  const double *rs = mesh + (long)unit * mstride;
  const int N = mesh_len[unit];
  const long rsz = adj_rec_doubles(n, L.m, L.nq, L.nch, NQ > 0);
  if (N < 1 || N > mstride || (long)(N + 1) * n + (long)N * rsz > cstride) {  // the mesh row or the checkpoint slot would overflow
    if (tid < NPAR) gu[tid] = __builtin_nan("");
    if (tid == 0) status[unit] = 7;
    return;
  }
  auto mesh_s = [&](int i) { return i == 0 ? 0.0 : (rs[i - 1] >= 1.0 ? 1.0 : rs[i - 1]); };  // point i of 0..N
  auto mesh_t = [&](int i) { return i == 0 ? t0 : (rs[i - 1] >= 1.0 ? t1 : t0 + rs[i - 1] * (t1 - t0)); };
  double *ck = ckpt + (long)unit * cstride;
  double *rec0 = ck + (long)(N + 1) * n;  // step i's record at rec0 + i rsz
  // 1. forward replay with checkpoints: the state at every step start and the step's record
  if constexpr (REPLAY) {
    if (tid == 0) initial_conditions(b, L, t0, k, W.y0);
    __syncthreads();
    int gsrc = 0;
    for (int i = 0; i < N; ++i) {
      for (int p = tid; p < n; p += NT) ck[(long)i * n + p] = W.y0[p];
      const double tp = mesh_t(i), tn = mesh_t(i + 1);
      rodas5_attempt<NQ, NT>(b, L, k, tp, tn - tp, W, tid, gsrc, tn);
      save_record<NQ, NT>(L, W, rec0 + (long)i * rsz, tid);
      for (int p = tid; p < n; p += NT) W.y0[p] = W.y1[p];
      __syncthreads();
      gsrc = (NT > WAVE) ? 1 : 0;
    }
    for (int p = tid; p < n; p += NT) ck[(long)N * n + p] = W.y0[p];
    return;
  }
  // output j lies in the first step whose end reaches it
  // 2. reverse sweep
  for (int p = tid; p < n; p += NT) A.lam[p] = 0.0;
  if (tid < SC_COUNT) A.sc[tid] = 0.0;
  if (tid < MAXOUT) A.taub[tid] = 0.0;
  const double *dtab_c = dtab + (long)cosmo * NPAR * tstride;
  // each wave's directions (derivative-table pointers, dual density parameters) in LDS: read at
  // each use, so that they do not occupy registers across the sweep
  __shared__ DirCtx<ND> s_dc[NT / WAVE];
  if (lane == 0) dir_ctx<ND>(b, dtab_c, tstride, wave, s_dc[wave]);
  const DirCtx<ND> &dc = s_dc[wave];
  __syncthreads();
  int jout = nout - 1;  // outputs are handled from the last
  for (int i = N - 1; i >= 0; --i) {
    const double T0 = mesh_t(i), T1 = mesh_t(i + 1), h = T1 - T0;
    // the start and end states of the step
    for (int p = tid; p < n; p += NT) {
      W.y0[p] = ck[(long)i * n + p];
      A.yn[p] = ck[(long)(i + 1) * n + p];
      A.pend[p] = 0.0;
    }
    if (tid == 0) { A.sc[SC_TSB] = 0.0; A.sc[SC_TEB] = 0.0; }
    __syncthreads();
    // outputs of this step: tout[j] in (T0, T1] (the first step reaching it), the last first
    while (jout >= 0 && tout[jout] > T0) {
      const double tj = tout[jout], th = (tj - T0) / h;
      if (tid < 20) A.gb[tid] = gbar[((long)unit * nout + jout) * 20 + tid];
      for (int p = tid; p < n; p += NT) A.ust[p] = W.y0[p] + th * (A.yn[p] - W.y0[p]);
      __syncthreads();
      PBgX<double> go;
      pbg_eval_out(b, L, k, tj, A.ust[0], go, W.sc, tid);
      const bool rsa = out_rsa(b, k, tj, go.tauc);
      // linear part, q >= 1: thread q evaluates the fields on e_q
      const int nitq = (n - 1 + NT - 1) / NT;
      for (int it = 0; it < nitq; ++it) {
        const int q = 1 + tid + it * NT;
        const int qq = q < n ? q : 1;
        auto ev = [&](int ii) { return ii == qq ? 1.0 : 0.0; };
        const double v = out_dot<double>(b, L, k, go, rsa, ev, A.gb);
        if (q < n) A.usb[q] = v;
      }
      // scale factor and parameters: dual evaluation on the state
      {
        Dn<ND> td, ad;
        dual_point<ND>(dc, tj, A.ust[0], td, ad);
        PBgX<Dn<ND>> gd;
        pbg_eval_ad<ND>(b, dc.bd, dc.dt, L, k, td, ad, gd, lane);
        auto yv = [&](int ii) { return A.ust[ii]; };
        const Dn<ND> s = out_dot<Dn<ND>>(dc.bd, L, k, gd, rsa, yv, A.gb);
        if (lane == 0)
#pragma unroll
          for (int l = 0; l < ND; ++l) A.red[wave * ND + l] = s.d[l];
      }
      __syncthreads();
      if (tid == 0) {
        A.usb[0] = A.red[0];  // direction 0: a
        for (int p = 0; p < NPAR; ++p) A.sc[SC_G0 + p] += A.red[2 + p];
      }
      __syncthreads();
      double thb[1] = {0.0};
      for (int p = tid; p < n; p += NT) {
        const double yb = A.usb[p];
        A.lam[p] += th * yb;
        A.pend[p] += (1.0 - th) * yb;
        thb[0] += yb * (A.yn[p] - W.y0[p]);
      }
      block_sums<NT, 1>(thb, A.red, tid);
      if (tid == 0) {
        A.sc[SC_TSB] += -thb[0] * (1.0 - th) / h;
        A.sc[SC_TEB] += -thb[0] * th / h;
        A.taub[jout] += thb[0] / h;
      }
      __syncthreads();
      --jout;
    }
    // the step's stages and factorisation from its record
    __syncthreads();
    load_record<NQ, NT>(L, W, rec0 + (long)i * rsz, tid);
    __syncthreads();
    const double W00 = 1.0 / (h * gamma) - W.J0[0];
    double *K = W.K;
    for (int p = tid; p < n; p += NT) {
      for (int s = 0; s < 7; ++s) A.kb[s * n + p] = 0.0;
      A.kb[7 * n + p] = A.lam[p];
      A.ucar[p] = A.lam[p];
      A.yb0[p] = 0.0;
      A.dTb[p] = 0.0;
    }
    if (tid == 0) { A.sc[SC_T0B] = 0.0; A.sc[SC_HB] = 0.0; }
    __syncthreads();
    for (int s = 7; s >= 0; --s) {
      double *rb = A.kb + s * n;
      const double ts = stage_time(s, T0, h);
      // r_s = W^-T kbar_s (wave 0) while the others form the stage input u_s
      if (tid < WAVE) {
        wsolveT<MC>(L, W, A, W00, rb, tid);
      } else {
        for (int p = tid - WAVE; p < n; p += NT - WAVE) {
          double u5 = stage_in(s < 5 ? s : 5, W.y0[p], 0.0, [&](int j) { return K[j * n + p]; });
          if (s >= 6) u5 = u5 + K[5 * n + p];
          if (s >= 7) u5 = u5 + K[6 * n + p];
          A.ust[p] = u5;
        }
      }
      __syncthreads();
      for (int p = tid; p < n; p += NT) A.usb[p] = 0.0;
      __syncthreads();
      // linear part of the stage point's vector-Jacobian product
      {
        PBg<double> gs;
        pbg_eval_lanes(b, L, k, ts, A.ust[0], gs, W.sc, tid);
        colT<NT, double>(b, L, k, gs, rb, nullptr, A.usb, W, A, tid);
      }
      // scale factor, time and parameters at the stage point
      {
        Dn<ND> td, ad;
        dual_point<ND>(dc, ts, A.ust[0], td, ad);
        PBgX<Dn<ND>> gd;
        pbg_eval_ad<ND>(b, dc.bd, dc.dt, L, k, td, ad, gd, lane);
        double acc[ND];
#pragma unroll
        for (int l = 0; l < ND; ++l) acc[l] = 0.0;
        auto yv = [&](int ii) { return A.ust[ii]; };
        rows_dot<ND>(b, dc, L, k, gd, yv, rb, 0, W, A.perm64, lane, acc);
#pragma unroll
        for (int l = 0; l < ND; ++l) acc[l] = wave_sum(acc[l]);
        if (lane == 0)
#pragma unroll
          for (int l = 0; l < ND; ++l) A.red[wave * ND + l] = acc[l];
      }
      __syncthreads();
      // dot products of r_s with dT and the stages
      double dots[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) dots[q] = 0.0;
      for (int p = tid; p < n; p += NT) {
        const double r = rb[p];
        dots[8] += r * W.dT[p];
#pragma unroll
        for (int j = 0; j < 8; ++j) dots[j] += r * K[j * n + p];
      }
      const double ua = A.red[0], tsb = A.red[1];
      double gpar[NPAR];
#pragma unroll
      for (int p = 0; p < NPAR; ++p) gpar[p] = A.red[2 + p];
      block_sums<NT, 9>(dots, A.red + NDIR, tid);
      if (tid == 0) {
        A.usb[0] += ua;
        for (int p = 0; p < NPAR; ++p) A.sc[SC_G0 + p] += gpar[p];
        const double cs = (s == 0) ? 0.0 : (s == 1) ? c2 : (s == 2) ? c3 : (s == 3) ? c4 : (s == 4) ? c5 : 1.0;
        double dks = 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j == s) dks = dots[j];
        double hb = cs * tsb + dks / (gamma * h * h);
        if (s <= 4) {
          const double ds = (s == 0) ? d1 : (s == 1) ? d2 : (s == 2) ? d3 : (s == 3) ? d4 : d5;
          hb += ds * dots[8];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (j < s) hb -= rodas_C(s, j) / (h * h) * dots[j];
        A.sc[SC_T0B] += tsb;
        A.sc[SC_HB] += hb;
      }
      __syncthreads();
      // vector updates: dT, C/h, stage inputs
      const double hds = (s <= 4) ? h * ((s == 0) ? d1 : (s == 1) ? d2 : (s == 2) ? d3 : (s == 3) ? d4 : d5) : 0.0;
      for (int p = tid; p < n; p += NT) {
        const double r = rb[p], ub = A.usb[p];
        if (s <= 4) A.dTb[p] += hds * r;
        for (int j = 0; j < s; ++j) A.kb[j * n + p] += rodas_C(s, j) / h * r;
        if (s >= 5) {
          const double uc = A.ucar[p] + ub;
          A.ucar[p] = uc;
          if (s >= 6) {
            A.kb[(s - 1) * n + p] += uc;
          } else {
            A.yb0[p] += uc;
            for (int j = 0; j < 5; ++j) A.kb[j * n + p] += rodas_A(5, j) * uc;
          }
        } else {
          A.yb0[p] += ub;
          for (int j = 0; j < s; ++j) A.kb[j * n + p] += rodas_A(s, j) * ub;
        }
      }
      __syncthreads();
    }
    // J(t0, y0) and df/dtau(t0, y0): rho = sum_s k_s[0] r_s; and the tail rows' stage sums
    for (int p = tid; p < n; p += NT) {
      double r = 0.0;
      for (int s = 0; s < 8; ++s) r += K[s * n] * A.kb[s * n + p];
      A.rho[p] = r;
      double pm = 0.0, p0 = 0.0, pp = 0.0;
      const int rk = W.rk[p];
      if (rk >= 0) {
        const int st = W.rs[p];
        for (int s = 0; s < 8; ++s) {
          const double rs = A.kb[s * n + p];
          pm += rs * K[s * n + p - st];
          p0 += rs * K[s * n + p];
          if (p + st < n) pp += rs * K[s * n + p + st];
        }
      }
      A.ust[p] = pm;
      A.usb[p] = p0;
      A.yn[p] = pp;
    }
    __syncthreads();
    const double a0 = W.y0[0];
    {
      double acc[ND];
#pragma unroll
      for (int l = 0; l < ND; ++l) acc[l] = 0.0;
      Dn<ND> td, ad;
      // (b) sum_s r_s^T d(A(t0, a0)) k_s, rows >= 1: dense rows stage by stage, tail rows through
      // their stage sums (A.ust, A.usb, A.yn are free here)
      dual_point<ND>(dc, T0, a0, td, ad);
      PBgX<Dn<ND>> gd;
      pbg_eval_ad<ND>(b, dc.bd, dc.dt, L, k, td, ad, gd, lane);
      for (int s = 0; s < 8; ++s) {
        auto yv = [&](int ii) { return K[s * n + ii]; };
        rows_dot_dense<ND>(dc, L, k, gd, yv, A.kb + s * n, W, lane, acc);
      }
      rows_dot_tails<ND>(L, k, gd, A.ust, A.usb, A.yn, W, lane, acc);
      // (c) d/da of rho^T df/d(dir) at (t0, (a, y0[1:])), central difference in a
      const double da = 1e-5 * a0;
      double accp[ND], accm[ND];
#pragma unroll
      for (int l = 0; l < ND; ++l) { accp[l] = 0.0; accm[l] = 0.0; }
      auto yv0 = [&](int ii) { return W.y0[ii]; };
      dual_point<ND>(dc, T0, a0 + da, td, ad);
      pbg_eval_ad<ND>(b, dc.bd, dc.dt, L, k, td, ad, gd, lane);
      rows_dot<ND>(b, dc, L, k, gd, yv0, A.rho, 0, W, A.perm64, lane, accp);
      dual_point<ND>(dc, T0, a0 - da, td, ad);
      pbg_eval_ad<ND>(b, dc.bd, dc.dt, L, k, td, ad, gd, lane);
      rows_dot<ND>(b, dc, L, k, gd, yv0, A.rho, 0, W, A.perm64, lane, accm);
#pragma unroll
      for (int l = 0; l < ND; ++l) acc[l] += (accp[l] - accm[l]) / (2.0 * da);
#pragma unroll
      for (int l = 0; l < ND; ++l) acc[l] = wave_sum(acc[l]);
      if (lane == 0)
#pragma unroll
        for (int l = 0; l < ND; ++l) A.red[wave * ND + l] = acc[l];
    }
    __syncthreads();
    {
      const double ua = A.red[0], tb = A.red[1];
      double gpar[NPAR];
#pragma unroll
      for (int p = 0; p < NPAR; ++p) gpar[p] = A.red[2 + p];
      __syncthreads();
      if (tid == 0) {
        A.yb0[0] += ua;
        A.sc[SC_T0B] += tb;
        for (int p = 0; p < NPAR; ++p) A.sc[SC_G0 + p] += gpar[p];
      }
    }
    // (d) y0[1:]: sum_r (rho_r dA_rq/da + dTb_r dA_rq/dtau), a two-direction dual background
    {
      const double *dnull[2] = {nullptr, nullptr};
      BgD<2> bd2;
      make_bgd<2>(b, dnull, bd2);
      Dn<2> td(T0), ad(a0);
      td.d[0] = 0.0; td.d[1] = 1.0;
      ad.d[0] = 1.0; ad.d[1] = 0.0;
      PBgX<Dn<2>> g2;
      pbg_eval_ad<2>(b, bd2, dnull, L, k, td, ad, g2, lane);
      colT<NT, Dn<2>>(bd2, L, k, g2, A.rho, A.dTb, A.yb0, W, A, tid);
    }
    // (e) second tau derivative of dTb^T f at (t0, y0): truncation rows and the UFA row
    {
      double acc[1] = {0.0};
      PBg<double> g0;
      pbg_eval_lanes(b, L, k, T0, a0, g0, W.sc, tid);
      const double it3 = 1.0 / (T0 * T0 * T0);
      for (int c = tid; c < L.nch; c += NT) {
        const bool off = (c <= 1) ? (g0.rsa || g0.tca) : (c == 2) ? (g0.rsa || g0.ufa) : false;
        if (off) continue;
        const int lmax = (c == 0) ? L.lmaxg : (c == 1) ? L.lmaxgp : (c == 2) ? L.lmaxr : L.lmaxnu;
        const int last = W.chead[c] + (W.clen[c] - 1) * W.cstride[c];
        acc[0] += A.dTb[last] * (-2.0 * (double)(lmax + 1) * it3) * W.y0[last];
      }
      if (tid == 0 && g0.ufa && !g0.rsa) acc[0] += A.dTb[L.idxr + 2] * (-6.0 * it3) * (0.5 * W.y0[L.idxr + 2]);
      block_sums<NT, 1>(acc, A.red, tid);
      if (tid == 0) A.sc[SC_T0B] += acc[0];
    }
    __syncthreads();
    // the step's times: t_i = T0, t_{i+1} = T0 + h
    if (tid == 0) {
      const double tsb = A.sc[SC_TSB] + A.sc[SC_T0B] - A.sc[SC_HB];
      const double teb = A.sc[SC_TEB] + A.sc[SC_HB];
      const double si = mesh_s(i), se = mesh_s(i + 1);
      A.sc[SC_T0BAR] += (1.0 - si) * tsb + (1.0 - se) * teb;
      A.sc[SC_T1BAR] += si * tsb + se * teb;
    }
    for (int p = tid; p < n; p += NT) A.lam[p] = A.yb0[p] + A.pend[p];
    __syncthreads();
  }
  // 3. initial state, start / end times and output times over the shifted cosmologies
  {
    __shared__ BgView s_sh[NSHIFT];
    if (tid < NSHIFT) s_sh[tid].init(shtab + ((long)cosmo * NSHIFT + tid) * tstride);
    __syncthreads();
    double *ysh = W.K;  // 12 n doubles: K (8 n) and the chain tables after it (tlo .. td), no longer read
    if (tid < NSHIFT) {
      const BgView &bs = s_sh[tid];
      const double *to = tau_out_sh + ((long)cosmo * NSHIFT + tid) * nout;
      double tm = to[0];
      for (int j = 1; j < nout; ++j) tm = fmin(tm, to[j]);
      const double t0s = 0.99 * fmin(tm, starting_time(bs, k));
      initial_conditions(bs, L, t0s, k, ysh + (long)tid * n);
      A.red[tid] = t0s;
    }
    __syncthreads();
    double acc[NPAR];
#pragma unroll
    for (int p = 0; p < NPAR; ++p) acc[p] = 0.0;
    for (int q = tid; q < n; q += NT) {
      const double l = A.lam[q];
#pragma unroll
      for (int p = 0; p < NPAR; ++p) acc[p] += l * (ysh[(long)(2 * p) * n + q] - ysh[(long)(2 * p + 1) * n + q]);
    }
    block_sums<NT, NPAR>(acc, A.red + 16, tid);
    if (tid == 0) {
      for (int p = 0; p < NPAR; ++p) {
        const double d2 = 2.0 * dth[(long)cosmo * NPAR + p];
        const double *top = tau_out_sh + ((long)cosmo * NSHIFT + 2 * p) * nout, *tom = top + nout;
        double t1p = top[0], t1m = tom[0];
        for (int j = 1; j < nout; ++j) { t1p = fmax(t1p, top[j]); t1m = fmax(t1m, tom[j]); }
        double gm = A.sc[SC_T0BAR] * (A.red[2 * p] - A.red[2 * p + 1]) / d2 + A.sc[SC_T1BAR] * (t1p - t1m) / d2;
        for (int j = 0; j < nout; ++j) gm += A.taub[j] * (top[j] - tom[j]) / d2;
        gu[p] = A.sc[SC_G0 + p] + acc[p] / d2 + gm;
      }
      status[unit] = 0;
      if (tbar) { tbar[2 * (long)unit] = A.sc[SC_T0BAR]; tbar[2 * (long)unit + 1] = A.sc[SC_T1BAR]; }
    }
    if (lam0)
      for (int q = tid; q < n; q += NT) lam0[(long)unit * n + q] = A.lam[q];
  }
}

template <int NQ, int NT>
__global__ __launch_bounds__(NT) void adjoint_kernel(const double *__restrict__ tables, long tstride, const double *__restrict__ shtab,
                                                      const double *__restrict__ dtab, const double *__restrict__ dth,
                                                      const int *__restrict__ unit_cosmo, const double *__restrict__ unit_k, int n_units,
                                                      const double *__restrict__ tau_out, const double *__restrict__ tau_out_sh, int nout,
                                                      const double *__restrict__ nu_qw, DebProblem prob, long lds_bytes,
                                                      const double *__restrict__ mesh, long mstride, const int *__restrict__ mesh_len,
                                                      const double *__restrict__ gbar, double *__restrict__ ckpt, long cstride,
                                                      double *__restrict__ grad, double *__restrict__ lam0, double *__restrict__ tbar,
                                                      int *__restrict__ status, const int *__restrict__ order) {
  if ((int)blockIdx.x >= n_units) return;
  const int tid = threadIdx.x;
  PLay L;
  Ws W;
  const long fwd = deb_perturb_lds_bytes_dev(prob);
  if (!setup_solver<NQ, NT>(prob, nu_qw, fwd, L, W, tid)) {
    for (int u = blockIdx.x * NT + tid; u < n_units; u += gridDim.x * NT) status[u] = 3;
    return;
  }
  extern __shared__ double smem[];
  double *p = smem + (fwd + 7) / 8;
  const int n = L.n;
  AdjWs A;
  A.kb = p; p += 8 * n; A.lam = p; p += n; A.ucar = p; p += n; A.yb0 = p; p += n; A.ust = p; p += n; A.usb = p; p += n;
  A.dTb = p; p += n; A.rho = p; p += n; A.pend = p; p += n; A.yn = p; p += n;
  A.part = p; p += (NT / WAVE) * 128; A.red = p; p += 64; A.cb = p; p += NQMAX + 3; A.xd2 = p; p += 3 * NQMAX + 17; A.z19 = p; p += 2 * NRED;
  A.sc = p; p += SC_COUNT; A.taub = p; p += MAXOUT; A.gb = p; p += 20;
  A.perm64 = (int *)p; p += (n + 256 + 1) / 2;
  if ((long)((char *)p - (char *)smem) > lds_bytes) {
    for (int u = blockIdx.x * NT + tid; u < n_units; u += gridDim.x * NT) status[u] = 3;
    return;
  }
  if (tid == 0) build_perm(L, W, A.perm64, WAVE);
  __syncthreads();
  const int unit = order ? order[blockIdx.x] : (int)blockIdx.x;
  adjoint_unit<NQ, NT, false>(unit, tables, tstride, shtab, dtab, dth, unit_cosmo, unit_k, tau_out, tau_out_sh, nout, mesh, mstride, mesh_len, gbar,
                       ckpt, cstride, grad, lam0, tbar, status, L, W, A, tid);
}

// The forward replay part of the adjoint (adjoint_unit<..., true>): the forward solver's LDS only.
template <int NQ, int NT>
__global__ __launch_bounds__(NT) void adjoint_replay_kernel(const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                                             const double *__restrict__ unit_k, int n_units, const double *__restrict__ tau_out,
                                                             int nout, const double *__restrict__ nu_qw, DebProblem prob, long lds_bytes,
                                                             const double *__restrict__ mesh, long mstride, const int *__restrict__ mesh_len,
                                                             double *__restrict__ ckpt, long cstride, double *__restrict__ grad,
                                                             int *__restrict__ status, const int *__restrict__ order) {
  if ((int)blockIdx.x >= n_units) return;
  const int tid = threadIdx.x;
  PLay L;
  Ws W;
  if (!setup_solver<NQ, NT>(prob, nu_qw, lds_bytes, L, W, tid)) {
    for (int u = blockIdx.x * NT + tid; u < n_units; u += gridDim.x * NT) status[u] = 3;
    return;
  }
  AdjWs A{};  // not read by the replay
  const int unit = order ? order[blockIdx.x] : (int)blockIdx.x;
  adjoint_unit<NQ, NT, true>(unit, tables, tstride, nullptr, nullptr, nullptr, unit_cosmo, unit_k, tau_out, nullptr, nout, mesh, mstride,
                             mesh_len, nullptr, ckpt, cstride, grad, nullptr, nullptr, status, L, W, A, tid);
}

// derivative tables: dtab[c, p] = (shtab[c, 2p] - shtab[c, 2p + 1]) / (2 dth[c, p]), element-wise
__global__ void dtab_kernel(const double *__restrict__ shtab, const double *__restrict__ dth, int n_cosmo, long tstride,
                            double *__restrict__ dtab) {
  const long total = (long)n_cosmo * NPAR * tstride;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long cp = e / tstride, off = e - cp * tstride;
    const long c = cp / NPAR, pp = cp - c * NPAR;
    const double *hi = shtab + (c * NSHIFT + 2 * pp) * tstride, *lo = hi + tstride;
    dtab[e] = (hi[off] - lo[off]) / (2.0 * dth[cp]);
  }
}

}  // namespace
}  // namespace deb

extern "C" long deb_adjoint_record_doubles(const DebProblem *prob) {
  const long n = 7 + (prob->lmaxg + 1) + (prob->lmaxgp + 1) + (prob->lmaxr + 1) + (long)prob->nqmax * (prob->lmaxnu + 1) + 2;
  return deb::adj_rec_doubles(n, 17 + 3L * prob->nqmax, prob->nqmax, 3 + prob->nqmax, deb::adj_reduced(prob->nqmax));
}

extern "C" long deb_adjoint_lds_bytes(const DebProblem *prob, int nt) {
  const long n = 7 + (prob->lmaxg + 1) + (prob->lmaxgp + 1) + (prob->lmaxr + 1) + (long)prob->nqmax * (prob->lmaxnu + 1) + 2;
  const long fwd = deb::deb_perturb_lds_bytes_dev(*prob);
  return ((fwd + 7) / 8) * 8 + 8 * deb::adj_extra_doubles(n, nt);
}

namespace {
typedef decltype(&deb::adjoint_kernel<0, 256>) AdjFn;
typedef decltype(&deb::adjoint_replay_kernel<0, 256>) ReplayFn;
template <int NT> ReplayFn pick_replay_nt(int nq) {
  switch (nq) {
    case 3: return deb::adjoint_replay_kernel<3, NT>;
    case 5: return deb::adjoint_replay_kernel<5, NT>;
    case 8: return deb::adjoint_replay_kernel<8, NT>;
    default: return deb::adjoint_replay_kernel<0, NT>;
  }
}
// compute units of the current device (once per device)
int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int v = 0;
    cus[dev] = (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ? v : 256;
  }
  return cus[dev];
}
AdjFn pick_adjoint(int nq) {
  switch (nq) {
    case 3: return deb::adjoint_kernel<3, 256>;
    case 5: return deb::adjoint_kernel<5, 256>;
    case 8: return deb::adjoint_kernel<8, 256>;
    default: return deb::adjoint_kernel<0, 256>;
  }
}
}  // namespace

extern "C" int deb_launch_adjoint(const double *d_tables, int num_thermo, const double *d_shtab, double *d_dtab, const double *d_dth, int n_cosmo,
                                  const int *d_unit_cosmo, const double *d_unit_k, int n_units, const double *d_tau_out,
                                  const double *d_tau_out_sh, int nout, const double *d_nu_qw, const DebProblem *prob, const double *d_mesh,
                                  long mstride, const int *d_mesh_len, const double *d_gbar, double *d_ckpt, long cstride, double *d_grad,
                                  double *d_lam0, double *d_tbar, int *d_status, const int *d_order, hipStream_t stream) {
  if (nout > deb::MAXOUT) return -5;
  const long stride = deb_table_doubles(num_thermo);
  const long total = (long)n_cosmo * deb::NPAR * stride;
  const int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL(deb::dtab_kernel, dim3(blocks), dim3(256), 0, stream, d_shtab, d_dth, n_cosmo, stride, d_dtab);
  // the forward replay: 256 threads while every unit has a compute unit to itself, else 128 (two
  // units per compute unit)
  const long flds = deb::deb_perturb_lds_bytes_dev(*prob);
  if (flds > 160 * 1024) return -2;
  const int rnt = (n_units <= device_cus()) ? 256 : 128;
  ReplayFn rk = (rnt == 256) ? pick_replay_nt<256>(prob->nqmax) : pick_replay_nt<128>(prob->nqmax);
  if (flds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)rk, hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds);
    if (e != hipSuccess) return (int)e;
  }
  hipLaunchKernelGGL(rk, dim3(n_units), dim3(rnt), (size_t)flds, stream, d_tables, stride, d_unit_cosmo, d_unit_k, n_units, d_tau_out, nout,
                     d_nu_qw, *prob, flds, d_mesh, mstride, d_mesh_len, d_ckpt, cstride, d_grad, d_status, d_order);
  const long lds = deb_adjoint_lds_bytes(prob, 256);
  if (lds > 160 * 1024) return -2;
  AdjFn kern = pick_adjoint(prob->nqmax);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
  }
  hipLaunchKernelGGL(kern, dim3(n_units), dim3(256), (size_t)lds, stream, d_tables, stride, d_shtab, d_dtab, d_dth, d_unit_cosmo, d_unit_k,
                     n_units, d_tau_out, d_tau_out_sh, nout, d_nu_qw, *prob, lds, d_mesh, mstride, d_mesh_len, d_gbar, d_ckpt, cstride,
                     d_grad, d_lam0, d_tbar, d_status, d_order);
  return (int)hipGetLastError();
}
