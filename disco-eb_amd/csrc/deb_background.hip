// deb_background.hip -- background evolution and RECFAST thermal history on the GPU.
//
// One workgroup per cosmology writes that cosmology's table block (deb_layout.h).  Follows
// the reference's evolve_background(thermo_module='RECFAST') (src/discoeb/background.py:
// 141-345) and thermodynamics_recfast.py:124-492:
//   * neutrino background on 512 log-spaced a (8-node generalized Gauss-Laguerre bins),
//   * Omega_DE closure, tau_min = a_min / adotrad,
//   * RECFAST on the adaptive a-grid: the interval loop is sequential by nature and runs on
//     lane 0; each interval is an adaptive GRKT4 (Kaps-Rentrop) solve whose 3x3 Jacobian is
//     taken exactly with 3-direction dual numbers,
//   * per-knot work (He II Saha and its a-derivative, sound speed, Romberg increments of
//     tau(a), opacity and visibility) spread over the workgroup,
//   * natural cubic splines (Thomas algorithm, one lane per spline).
#include <hip/hip_runtime.h>

#include "deb_cosmo.hpp"

namespace deb {
namespace {

constexpr double PI = 3.141592653589793;

// RECFAST constants (thermodynamics_recfast.py:37-122): physical constants and fit parameters
struct RfConst {
  double G, mH, me, mHe_mH, c, h, kB, sigma, arad, bigH, dens_fac, EionHe12s;
  double Lambda, Lambda_He, L_H_ion, L_H_alpha, L_He1_ion, L_He_2s, L_He_2p, A2P_s, A2P_t, L_He_2Pt, L_He_2St, L_He2St_ion;
  double sigma_He_2Ps, sigma_He_2Pt, AGauss1, AGauss2, zGauss1, zGauss2, wGauss1, wGauss2, H_frac;
  double CDB, CDB_He, CB1, CB1_He1, CR, CK, CK_He, CL, CL_He, CT, Bfact, CL_PSt, hcL2St_kB, hcL2Stion_kB;
  double a_PPB, b_PPB, c_PPB, d_PPB, a_VF, b_VF, T_0, T_1, a_trip, b_trip;
  __device__ void init() {
    G = 6.67430e-11; mH = 1.67353284e-27; me = 9.1093837015e-31; mHe_mH = 3.97146570884; c = 2.99792458e+08;
    h = 6.62607015e-34; kB = 1.380649e-23; sigma = 6.6524587321e-29; arad = 4 * 5.670374419e-8 / c;
    bigH = 3.2407792902755102e-18; dens_fac = 11.223810928601939; EionHe12s = 6.314878282674e5;
    Lambda = 8.2245809; Lambda_He = 51.3; L_H_ion = 1.096787737e7; L_H_alpha = 8.225916453e6; L_He1_ion = 1.98310772e7;
    L_He_2s = 1.66277434e7; L_He_2p = 1.71134891e7; A2P_s = 1.798287e9; A2P_t = 177.58; L_He_2Pt = 1.690871466e7;
    L_He_2St = 1.5985597526e7; L_He2St_ion = 3.8454693845e6; sigma_He_2Ps = 1.436289e-22; sigma_He_2Pt = 1.484872e-22;
    AGauss1 = -0.14; AGauss2 = 0.05; zGauss1 = 7.28; zGauss2 = 6.75; wGauss1 = 0.18; wGauss2 = 0.33; H_frac = 1.e-3;
    double Lalpha = 1.0 / L_H_alpha, Lalpha_He = 1.0 / L_He_2p;
    CDB = h * c * (L_H_ion - L_H_alpha) / kB;
    CDB_He = h * c * (L_He1_ion - L_He_2s) / kB;
    CB1 = h * c * L_H_ion / kB;
    CB1_He1 = h * c * L_He1_ion / kB;
    CR = 2.0 * PI * (me / h) * (kB / h);
    CK = Lalpha * Lalpha * Lalpha / (8.0 * PI);
    CK_He = Lalpha_He * Lalpha_He * Lalpha_He / (8.0 * PI);
    CL = c * h / (kB * Lalpha);
    CL_He = c * h / (kB / L_He_2s);
    CT = (8.0 / 3.0) * (sigma / (me * c)) * arad;
    Bfact = h * c * (L_He_2p - L_He_2s) / kB;
    CL_PSt = h * c * (L_He_2Pt - L_He_2St) / kB;
    hcL2St_kB = h * c * L_He_2St / kB;
    hcL2Stion_kB = h * c * L_He2St_ion / kB;
    a_PPB = 4.309; b_PPB = -0.6166; c_PPB = 0.6703; d_PPB = 0.5300;
    a_VF = pow(10., -16.744); b_VF = 0.711; T_0 = pow(10., 0.477121); T_1 = pow(10., 5.114);
    a_trip = pow(10., -16.306); b_trip = 0.761;
  }
};

// d(x_H, x_He, T_m)/da  (thermodynamics_recfast.py:124-281)
template <class T>
__device__ void ionization(const BgView &b, const RfConst &K, const T &a, const T *y, T *f) {
  T z = 1.0 / a - 1.0;
  const double H = b.H0 / 100.0, HO = H * K.bigH, fu = 1.105;
  const double mu_H = 1.0 / (1.0 - b.YHe);
  const double fHe = b.YHe / (K.mHe_mH * (1.0 - b.YHe));
  const double Nnow = K.dens_fac * H * H * b.Omegab / mu_H;
  T x_H = y[0], x_He = y[1];
  T x = x_H + fHe * x_He;
  T Tmat = dabs(y[2]);
  T zt = 1.0 + z;
  T n = Nnow * zt * zt * zt;
  T n_He = fHe * n;
  T Trad = b.Tcmb * zt;
  T Hz = (1e-5 * b.aprimeoa(a)) / a * K.c * K.bigH;
  T Tmat_1e4 = Tmat / 1e4;
  T CR_Tmat_15 = dpow(K.CR * Tmat, 1.5);
  T Rdown = 1e-19 * K.a_PPB * dpow(Tmat_1e4, K.b_PPB) / (1.0 + K.c_PPB * dpow(Tmat_1e4, K.d_PPB));
  T Rup = Rdown * CR_Tmat_15 * dexp(-K.CDB / Tmat);
  T sq_0 = dsqrt(Tmat / K.T_0), sq_1 = dsqrt(Tmat / K.T_1);
  T sq_0_term = 1.0 + sq_0, sq_1_term = 1.0 + sq_1;
  T Rdown_He = K.a_VF / (sq_0 * dpow(sq_0_term, 1 - K.b_VF) * dpow(sq_1_term, 1 + K.b_VF));
  T Rup_He = 4.0 * Rdown_He * CR_Tmat_15 * dexp(-K.CDB_He / Tmat);
  T He_Boltz = dexp(dmin(K.Bfact / Tmat, 680.0));
  T Rdown_trip = K.a_trip / (sq_0 * dpow(sq_0_term, 1 - K.b_trip) * dpow(sq_1_term, 1 + K.b_trip));
  T Rup_trip = Rdown_trip * dexp(-K.hcL2Stion_kB / Tmat) * CR_Tmat_15 * (4.0 / 3.0);
  T lz = dlog(zt);
  T g1 = (lz - K.zGauss1) / K.wGauss1, g2 = (lz - K.zGauss2) / K.wGauss2;
  T K_gauss = 1.0 + K.AGauss1 * dexp(-(g1 * g1)) + K.AGauss2 * dexp(-(g2 * g2));
  T Kp = K.CK / Hz * K_gauss;
  T omxHe = 1.0 - x_He, omxH = 1.0 - x_H;
  T nHe1mx = n_He * omxHe;
  T tauHe_s = K.A2P_s * K.CK_He * 3.0 * nHe1mx / Hz;
  T pHe_s = (1.0 - dexp(-tauHe_s)) / tauHe_s;
  T Dop = dsqrt(2.0 * K.kB * Tmat / (K.mH * K.mHe_mH * K.c * K.c));
  T Dop_2p = K.c * K.L_He_2p * Dop;
  T Dop_2Pt = K.c * K.L_He_2Pt * Dop;
  const double cL2p = K.c * K.L_He_2p, cL2Pt = K.c * K.L_He_2Pt;
  T gamma_2Ps = (3.0 * K.A2P_s * fHe * omxHe * K.c * K.c) / (sqrt(PI) * K.sigma_He_2Ps * 8.0 * PI * Dop_2p * omxH * (cL2p * cL2p));
  T AHcon = K.A2P_s / (1.0 + 0.36 * dpow(gamma_2Ps, 0.86));
  bool c1 = val(x_He) < 5e-9 || val(x_He) > 0.98;
  bool c2 = val(x_H) < 0.9999999;
  T K_He = c1 ? T(K.CK_He / Hz) : (c2 ? T(1.0 / ((K.A2P_s * pHe_s + AHcon) * 3.0 * nHe1mx)) : T(1.0 / (K.A2P_s * pHe_s * 3.0 * nHe1mx)));
  T tauHe_t = K.A2P_t * nHe1mx * 3.0 / (8.0 * PI * Hz * (K.L_He_2Pt * K.L_He_2Pt * K.L_He_2Pt));
  T pHe_t = (1.0 - dexp(-tauHe_t)) / tauHe_t;
  T gamma_2Pt = (3.0 * K.A2P_t * fHe * omxHe * K.c * K.c) / (sqrt(PI) * K.sigma_He_2Pt * 8.0 * PI * Dop_2Pt * omxH * (cL2Pt * cL2Pt));
  T AHcon_t = K.A2P_t / (1.0 + 0.66 * dpow(gamma_2Pt, 0.9)) / 3.0;
  T eCL = dexp(-K.CL_PSt / Tmat);
  T CfHe_t = val(x_H) > 0.99999 ? T(K.A2P_t * pHe_t * eCL) : T((K.A2P_t * pHe_t + AHcon_t) * eCL);
  CfHe_t = CfHe_t / (Rup_trip + CfHe_t);
  T Trad2 = Trad * Trad;
  T timeTh = (1.0 / (K.CT * Trad2 * Trad2)) * (1.0 + x + fHe) / x;
  T timeH = 2.0 / (3.0 * HO * dpow(zt, 1.5));
  T Hz_zt = Hz * zt;
  T rate_diff = x * x_H * n * Rdown - Rup * omxH * dexp(-K.CL / Tmat);
  T KLn = Kp * K.Lambda * n * omxH;
  T denom = 1.0 / fu + KLn / fu + Kp * Rup * n * omxH;
  T f0;
  if (val(x_H) > 0.99) f0 = T(0.0);
  else if (val(x_H) > 0.985) f0 = rate_diff / Hz_zt;
  else f0 = (rate_diff * (1.0 + KLn)) / (Hz_zt * denom);
  T rdHe = x * x_He * n * Rdown_He - Rup_He * omxHe * dexp(-K.CL_He / Tmat);
  T KHe_t = K_He * K.Lambda_He * nHe1mx * He_Boltz;
  T He_den = 1.0 + K_He * (K.Lambda_He + Rup_He) * nHe1mx * He_Boltz;
  T f1_main = (rdHe * (1.0 + KHe_t)) / (Hz_zt * He_den);
  T trd = x * x_He * n * Rdown_trip - omxHe * 3.0 * Rup_trip * dexp(-K.hcL2St_kB / Tmat);
  T tcontrib = trd * CfHe_t / Hz_zt;
  T f1 = val(x_He) < 1e-8 ? T(0.0) : (c1 ? f1_main : f1_main + tcontrib);
  T f2;
  if (val(timeTh) < K.H_frac * val(timeH)) f2 = Tmat / zt;
  else f2 = K.CT * (Trad2 * Trad2) * x / (1.0 + x + fHe) * (Tmat - Trad) / Hz_zt + 2.0 * Tmat / zt;
  T dzda = -1.0 / (a * a);
  f[0] = f0 * dzda;
  f[1] = f1 * dzda;
  f[2] = f2 * dzda;
}

// value of x on lane src (uniform src) of the calling wavefront
__device__ __forceinline__ double lane_read(double x, int src) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, src), hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// One GRKT4 step (ode_integrators_stiff.py:166-281, GRK4A coefficients); every stage uses the
// vector field at t0.  Called by a whole wavefront with the same arguments on every lane: lane
// j < 3 carries the dual direction of y_j, so one dual evaluation gives the Jacobian (column j
// read from lane j) and, in its value part, the first stage's vector field.
__device__ void grkt4_step(const BgView &b, const RfConst &K, double t0, double dt, const double *y0, double *y1, double *yerr, int lane) {
  const double gamma = 0.395, g21 = -0.767672395484, g31 = -0.851675323742, g32 = 0.522967289188;
  const double g41 = 0.288463109545, g42 = 0.880214273381e-1, g43 = -0.337389840627;
  const double a21 = 0.438, a31 = 0.796920457938, a32 = 0.730795420615e-1;
  const double ch1 = 0.346325833758, ch2 = 0.285693175712, ch3 = 0.367980990530;
  const double c1 = 0.199293275701, c2 = 0.482645235674, c3 = 0.680614886256e-1, c4 = 0.25;
  double J[3][3], r[3];
  {
    Dual yd[3], fd[3];
    for (int i = 0; i < 3; ++i) {
      yd[i] = Dual(y0[i]);
      yd[i].d[0] = (lane == i) ? 1.0 : 0.0;
    }
    ionization<Dual>(b, K, Dual(t0), yd, fd);
    for (int i = 0; i < 3; ++i) {
      r[i] = fd[i].v * dt;
      for (int j = 0; j < 3; ++j) J[i][j] = dt * lane_read(fd[i].d[0], j);
    }
  }
  double A00 = 1.0 - gamma * J[0][0], A01 = -gamma * J[0][1], A02 = -gamma * J[0][2];
  double A10 = -gamma * J[1][0], A11 = 1.0 - gamma * J[1][1], A12 = -gamma * J[1][2];
  double A20 = -gamma * J[2][0], A21 = -gamma * J[2][1], A22 = 1.0 - gamma * J[2][2];
  double u00 = A00, u01 = A01, u02 = A02;
  double l10 = A10 / u00, u11 = A11 - l10 * u01, u12 = A12 - l10 * u02;
  double l20 = A20 / u00, r21 = A21 - l20 * u01, r22 = A22 - l20 * u02;
  double l21 = r21 / u11, u22 = r22 - l21 * u12;
  auto solve = [&](const double *r, double *x) {
    double z0 = r[0], z1 = r[1] - l10 * z0, z2 = r[2] - l20 * z0 - l21 * z1;
    x[2] = z2 / u22;
    x[1] = (z1 - u12 * x[2]) / u11;
    x[0] = (z0 - u01 * x[1] - u02 * x[2]) / u00;
  };
  auto vf = [&](const double *y, double *f) {
    ionization<double>(b, K, t0, y, f);
    f[0] *= dt; f[1] *= dt; f[2] *= dt;
  };
  double k1[3], k2[3], k3[3], k4[3], yy[3], f[3], Jk1[3], Jk2[3], Jk3[3];
  solve(r, k1);
  for (int i = 0; i < 3; ++i) yy[i] = y0[i] + a21 * k1[i];
  vf(yy, f);
  for (int i = 0; i < 3; ++i) Jk1[i] = J[i][0] * k1[0] + J[i][1] * k1[1] + J[i][2] * k1[2];
  for (int i = 0; i < 3; ++i) r[i] = f[i] + g21 * Jk1[i];
  solve(r, k2);
  for (int i = 0; i < 3; ++i) yy[i] = y0[i] + a31 * k1[i] + a32 * k2[i];
  vf(yy, f);
  for (int i = 0; i < 3; ++i) Jk2[i] = J[i][0] * k2[0] + J[i][1] * k2[1] + J[i][2] * k2[2];
  for (int i = 0; i < 3; ++i) r[i] = f[i] + g31 * Jk1[i] + g32 * Jk2[i];
  solve(r, k3);
  for (int i = 0; i < 3; ++i) Jk3[i] = J[i][0] * k3[0] + J[i][1] * k3[1] + J[i][2] * k3[2];
  for (int i = 0; i < 3; ++i) r[i] = f[i] + g41 * Jk1[i] + g42 * Jk2[i] + g43 * Jk3[i];
  solve(r, k4);
  for (int i = 0; i < 3; ++i) {
    y1[i] = y0[i] + c1 * k1[i] + c2 * k2[i] + c3 * k3[i] + c4 * k4[i];
    yerr[i] = (c1 - ch1) * k1[i] + (c2 - ch2) * k2[i] + (c3 - ch3) * k3[i] + c4 * k4[i];
  }
}

// Adaptive GRKT4 over [astart, aend] with the default I-controller (safety 0.9, factor clip
// [0.2 (1 on accept), 10], error order 3), rtol 1e-3, atol 1e-6, at most 128 steps, RMS norm.
// Run by a whole wavefront (see grkt4_step).
__device__ bool solve_ionization(const BgView &b, const RfConst &K, double astart, double aend, const double *ys, double *ye, int lane) {
  const double rtol = 1e-3, atol = 1e-6;
  double y[3] = {ys[0], ys[1], ys[2]};
  double tprev = astart, t1 = aend;
  double tnext = fmin(astart + fabs(astart * 1e-3), t1);
  int steps = 0;
  while (tprev < t1 && steps < 128) {
    double y1[3], err[3];
    grkt4_step(b, K, tprev, tnext - tprev, y, y1, err, lane);
    bool nan = isnan(y1[0]) || isnan(y1[1]) || isnan(y1[2]);
    double s = 0.0;
    for (int i = 0; i < 3; ++i) {
      double yy = fmax(fabs(y[i]), fabs(nan ? y[i] : y1[i]));
      double e = err[i] / (atol + yy * rtol);
      s += e * e;
    }
    double en = sqrt(s / 3.0);
    bool keep = en < 1.0;
    double factor = 0.9 * pow(1.0 / en, 1.0 / 3.0);
    factor = fmin(fmax(factor, keep ? 1.0 : 0.2), 10.0);
    double dt = (tnext - tprev) * factor;
    double nt0 = keep ? tnext : tprev;
    if (keep) { y[0] = y1[0]; y[1] = y1[1]; y[2] = y1[2]; }
    tprev = fmin(nt0, t1);
    double tn = nt0 + dt;
    tnext = (tn > t1 - 1e-10) ? t1 : tn;
    ++steps;
    if (!(en == en)) break;  // non-finite error: stop (reported through the failure count)
  }
  ye[0] = y[0]; ye[1] = y[1]; ye[2] = y[2];
  return tprev >= t1;
}

// He II Saha equation (thermodynamics_recfast.py:303-322)
template <class T> __device__ T saha_HeII(const BgView &b, const RfConst &K, const T &a) {
  const double Hfac = 1.0 / (1.0e+06 * 3.0856775807e+13);
  const double mu_H = 1.0 / (1.0 - b.YHe);
  const double rho_c = 3.0 * (b.H0 * Hfac) * (b.H0 * Hfac) / (8.0 * PI * K.G);
  T NH = rho_c * b.Omegab / (K.mH * mu_H) / (a * a * a);
  T Temp = b.Tcmb / a;
  T betaE = K.EionHe12s / Temp;
  const double A = 1.0 + b.fHe, B = 1.0 + 2.0 * b.fHe;
  T R = dpow(2.0 * PI * K.me * K.kB / (K.h * K.h) * Temp, 1.5) / NH * dexp(-betaE);
  if (val(R) > 1e5) return b.fHe * (1.0 - B / R + (1.0 + 5.0 * b.fHe + 6.0 * b.fHe * b.fHe) / (R * R));
  T RA = R - A;
  return -RA / 2.0 + dsqrt(RA * RA / 4.0 + R * B) - A;
}

// 10**(log10(start)*(1-s) + log10(stop)*s), the arithmetic of jax.numpy.geomspace
__device__ __forceinline__ double geom_point(double start, double stop, int i, int num, bool endpoint) {
  double l0 = log10(start), l1 = log10(stop);
  int div = endpoint ? num - 1 : num;
  if (endpoint && i == num - 1) return pow(10.0, l1);
  double s = (double)i / (double)div;
  return pow(10.0, l0 * (1.0 - s) + l1 * s);
}

// jax_cosmo romb with divmax = 6 (Romberg R(6,6), 64 panels) of dtau/da
__device__ double romb_dtauda(const BgView &b, double a0, double a1) {
  const int divmax = 6;
  double range = a1 - a0;
  double ordsum = 0.5 * (b.dtauda(a0) + b.dtauda(a1));
  double prev[divmax + 1], cur[divmax + 1];
  prev[0] = range * ordsum;
  for (int k = 1; k <= divmax; ++k) prev[k] = prev[0];
  for (int i = 1; i <= divmax; ++i) {
    int nn = 1 << i, nsum = nn / 2;
    double h = range / nsum, lox = a0 + 0.5 * h, s = 0.0;
    for (int j = 0; j < nsum; ++j) s += b.dtauda(lox + h * j);
    ordsum += s;
    cur[0] = range * ordsum / nn;
    double t4 = 1.0;
    for (int k = 0; k < divmax; ++k) {
      t4 *= 4.0;
      cur[k + 1] = (t4 * cur[k] - prev[k]) / (t4 - 1.0);
    }
    for (int k = 0; k <= divmax; ++k) prev[k] = cur[k];
  }
  return prev[divmax];
}

// The 8-node momentum grid of the neutrino background, once computed on this device (each
// device holds its own copy; launches that compute it at once write the same values).
__device__ double g_bg_nu[16];
__device__ int g_bg_nu_ready;

// thermal history of the MB95 module (thermodynamics_mb95.py:6-90): semi-implicit hydrogen step
__device__ double mb95_ionize(double tempb, double a, double adot, double dtau, double xe, double YHe, double H0, double Omegab) {
  const double iswitch = 0.5, tion = 1.5789e5, beta0 = 43.082, dec2g = 8.468e14;
  const double alpha0 = 2.3866e-6 * (1.0 - YHe) * Omegab * H0 * H0;
  const double crec = 8.0138e-26 * (1.0 - YHe) * Omegab * H0 * H0;
  const double phi2 = fmax(0.448 * log(tion / tempb), 0.0);
  const double alpha = alpha0 / sqrt(tempb) * phi2 / (a * a * a);
  const double beta = tempb * phi2 * exp(beta0 - tion / tempb);
  double cpeebles = 1.0;
  if (!(tempb <= 200.0)) {
    const double cp1 = crec * dec2g * (1.0 - xe) / (a * adot);
    const double cp2 = crec * tempb * phi2 * exp(beta0 - 0.25 * tion / tempb) * (1.0 - xe) / (a * adot);
    cpeebles = (1.0 + cp1) / (1.0 + cp1 + cp2);
  }
  const double aa = a * dtau * alpha * cpeebles, bb = a * dtau * beta * cpeebles;
  const double b1 = 1.0 + iswitch * bb;
  const double bbxe = bb + xe - (1.0 - iswitch) * (bb * xe + aa * xe * xe);
  const double rat = iswitch * aa * bbxe / (b1 * b1);
  return rat < 5e-5 ? bbxe / b1 * (1.0 - rat) : b1 / (2.0 * iswitch * aa) * (sqrt(4.0 * rat + 1.0) - 1.0);
}

// helium Saha equations, six fixed-point iterations (thermodynamics_mb95.py:47-89)
__device__ void mb95_ionHe(double tempb, double a, double x0, double &x1, double &x2, double YHe, double H0, double Omegab) {
  const double tion1 = 2.855e5, tion2 = 6.313e5;
  const double b0 = 2.150e24 / ((1.0 - YHe) * Omegab * H0 * H0);
  const double bp = b0 * a * a * a * tempb * sqrt(tempb);
  const double r1 = 4.0 * bp * exp(-tion1 / tempb), r2 = bp * exp(-tion2 / tempb);
  const double c = 0.25 * YHe / (1.0 - YHe);
  for (int it = 0; it < 6; ++it) {
    const double xe = x0 + c * (x1 + 2.0 * x2);
    const double den = r1 * r2 + xe * r1 + xe * xe;
    const double x2new = r1 * r2 / den;
    x1 = xe * r1 / den;
    x2 = x2new;
  }
}

// the MB95 scan (thermodynamics_mb95.py:92-212) on N points log-uniform in tau over [taumin, taumax]
__device__ void mb95_scan(const BgView &b, int N, double taumax, double *aexp, double *tauv, double *xeo, double *xHo, double *xHe1o,
                          double *xHe2o, double *tbo, double *cs2o) {
  const double Tcmb = b.Tcmb, YHe = b.YHe, H0 = b.H0, Omegab = b.Omegab;
  const double thomc0 = 5.0577e-8 * Tcmb * Tcmb * Tcmb * Tcmb;
  const double barssc_raw = 9.1820e-14, fx = 0.25 * YHe / (1.0 - YHe);
  const double taumin = b.taumin, dlntau = log(taumax / taumin) / (N - 1);
  double adot = b.adotrad, tau = taumin, a = adot * tau, tb = Tcmb / a;
  double xHII = 1.0, xHeII = 0.0, xHeIII = 1.0;
  double xe = xHII + fx * (xHeII + 2.0 * xHeIII);
  double cs2 = 4.0 / 3.0 * barssc_raw * (1.0 - 0.75 * YHe + (1.0 - YHe) * xe) * tb;
  for (int i = 0;; ++i) {
    aexp[i] = a; tauv[i] = tau; xeo[i] = xe; xHo[i] = xHII; xHe1o[i] = xHeII; xHe2o[i] = xHeIII; tbo[i] = tb; cs2o[i] = cs2;
    if (i == N - 1) break;
    const double new_tau = taumin * exp((i + 1) * dlntau);
    const double dtau = new_tau - tau;
    double new_a = a + adot * dtau;
    const double new_adot = b.aprimeoa(new_a) * new_a;
    new_a = a + 2.0 * dtau / (1.0 / adot + 1.0 / new_adot);
    const double tg0 = Tcmb / a, ahalf = 0.5 * (a + new_a), adothalf = 0.5 * (adot + new_adot);
    const double fe = (1.0 - YHe) * xe / (1.0 - 0.75 * YHe + (1.0 - YHe) * xe);
    const double thomc = thomc0 * fe / adothalf / (ahalf * ahalf * ahalf);
    double a2t;
    if (fe < 1e-3) {
      a2t = (a - new_a) * Tcmb + a * a * (tb - tg0) + (0.5 * (a - new_a) * (a - new_a) * Tcmb + a * a * (a - new_a) * (tb - tg0)) * thomc;
    } else {
      const double etc = exp(-thomc * (new_a - a));
      a2t = a * a * (tb - tg0) * etc - Tcmb / thomc * (1.0 - etc);
    }
    const double new_tb = Tcmb / new_a + a2t / (new_a * new_a);
    const double tbhalf = 0.5 * (tb + new_tb);
    const double new_xHII = mb95_ionize(tbhalf, ahalf, adothalf, dtau, xHII, YHe, H0, Omegab);
    mb95_ionHe(tb, new_a, new_xHII, xHeII, xHeIII, YHe, H0, Omegab);
    const double new_xe = new_xHII + fx * (xHeII + 2.0 * xHeIII);
    const double dtbdla = -2.0 * new_tb - thomc * a2t / new_a;
    const double barssc = barssc_raw * (1.0 - 0.75 * YHe + (1.0 - YHe) * new_xe);
    cs2 = barssc * new_tb * (1.0 - dtbdla / new_tb / 3.0);
    a = new_a; adot = new_adot; tau = new_tau; tb = new_tb; xHII = new_xHII; xe = new_xe;
  }
}

// Reionization (no counterpart in the reference, whose thermal histories have none): the tanh
// model of Lewis (2008, Phys. Rev. D 78, 023002), as CAMB uses it -- hydrogen and the first
// helium ionisation
//   x_e = x_rec + (f_re - x_rec) F,  F = (1 + tanh((y_re - y) / dy)) / 2,  y = (1 + z)^(3/2),
//   y_re = (1 + z_re)^(3/2),  dy = (3/2) (1 + z_re)^(1/2) dz,  dz = 0.5,  f_re = 1 + f_He,
// and the second helium ionisation, f_He (1 + tanh((3.5 - z) / 0.5)) / 2.  z_re is chosen so that
// the optical depth these add to the recombination history, akthom int (x_e - x_rec) / a^2 dtau
// from today to z = 150, equals the requested tau_reion.
struct Reion {
  double yre, dy, fre, fHe;
  __device__ void set(double zre, double fhe) {
    yre = pow(1.0 + zre, 1.5);
    dy = 1.5 * sqrt(1.0 + zre) * 0.5;
    fre = 1.0 + fhe;
    fHe = fhe;
  }
  // x_e - x_rec at scale factor a, and (into d) its a-derivative at fixed x_rec; FH the hydrogen fraction F
  __device__ double dxe(double a, double xrec, double &d, double &FH) const {
    const double y = pow(a, -1.5), z = 1.0 / a - 1.0;
    const double th = tanh((yre - y) / dy), te = tanh((3.5 - z) / 0.5);
    FH = 0.5 * (1.0 + th);
    const double Fe = 0.5 * (1.0 + te);
    d = (fre - xrec) * 0.5 * (1.0 - th * th) * 1.5 * pow(a, -2.5) / dy + fHe * 0.5 * (1.0 - te * te) / (0.5 * a * a);
    return (fre - xrec) * FH + fHe * Fe;
  }
};
constexpr int REION_NQ = 4096;  // Simpson intervals in ln a over z in [0, 150]

// the optical depth added by reionization at z_re, by Simpson's rule in ln a (block-wide sum;
// every thread returns it); xrec from the natural spline S of the recombination x_e in ln a
__device__ double reion_depth(const BgView &b, const double *loga, const double *xe, const double *S, int N, double zre, double *red,
                              int tid, int nth) {
  Reion r;
  r.set(zre, b.fHe);
  const double l0 = -log(151.0), h = -l0 / REION_NQ;
  double acc = 0.0;
  for (int i = tid; i <= REION_NQ; i += nth) {
    const double la = l0 + h * i, a = exp(la);
    const int j = spl_index(loga, N, la);
    const double xrec = spl_eval_at(loga, xe, S, j, la);
    double d, FH;
    const double w = (i == 0 || i == REION_NQ) ? 1.0 : ((i & 1) ? 4.0 : 2.0);
    acc += w * r.dxe(a, xrec, d, FH) * b.akthom / a * b.dtauda(a);  // (x_e - x_rec) akthom / a^2 dtau/da da, da = a dln a
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nth / 64; ++w) s += red[w];
  return s * h / 3.0;
}

// thermo: 0 RECFAST, 1 MB95, 2 the thermal history given in thermo_in (per cosmology 5 x N
// doubles: a ascending, tau, x_e, T_b, c_b^2 -- the CLASS input of background.py:271-285)
__global__ __launch_bounds__(256) void background_kernel(const double *__restrict__ params, int N, double *__restrict__ tables,
                                                         long stride, int thermo, const double *__restrict__ thermo_in) {
  const int cosmo = blockIdx.x;
  const int tid = threadIdx.x, nth = blockDim.x;
  double *t = tables + (long)cosmo * stride;
  const double *p = params + (long)cosmo * DEB_NPARAM;
  __shared__ double s_q[8], s_w[8], s_tau_pre, s_xe_full, s_tau0, s_odt;
  __shared__ RfConst K;
  double *A = t;  // alias for readability
  auto arr = [&](int slot) { return t + deb_array_offset(slot, N); };
  const double *tin = thermo == 2 ? thermo_in + (long)cosmo * 5 * N : nullptr;
  // the a-range of the background (background.py:230-236): the given table's, else [1e-9, 1.01]
  const double amin = tin ? tin[0] : 1e-9, amax = tin ? tin[N - 1] : 1.01;

  if (tid == 0) {
    K.init();
    for (int i = 0; i < DEB_HDR_SIZE; ++i) t[i] = 0.0;
    for (int i = 0; i < DEB_NPARAM; ++i) t[i] = p[i];
    const double H0 = p[DEB_P_H0], Tc = p[DEB_P_TCMB];
    const double grhom = 3.33795017e-11 * H0 * H0;
    const double grhog = 1.49594245e-13 * Tc * Tc * Tc * Tc;
    const double grhor = 3.39739477e-14 * Tc * Tc * Tc * Tc;
    t[DEB_HDR_GRHOM] = grhom;
    t[DEB_HDR_GRHOG] = grhog;
    t[DEB_HDR_GRHOR] = grhor;
    t[DEB_HDR_ADOTRAD] = sqrt((grhog + grhor * (p[DEB_P_NEFF] + p[DEB_P_NMNU])) / 3.0);
    t[DEB_HDR_AMNU] = p[DEB_P_MNU] * 1.62581581e4 / Tc;
    t[DEB_HDR_FHE] = p[DEB_P_YHE] / (K.mHe_mH * (1.0 - p[DEB_P_YHE]));
    t[DEB_HDR_AKTHOM] = 2.3038921003709498e-9 * (1.0 - p[DEB_P_YHE]) * p[DEB_P_OMEGAB] * H0 * H0;
    t[DEB_HDR_LAMBDA_DE] = (p[DEB_P_W0] == -1.0 && p[DEB_P_WA] == 0.0) ? 1.0 : 0.0;
    t[DEB_HDR_N] = (double)N;
    t[DEB_HDR_TAUMIN] = amin / t[DEB_HDR_ADOTRAD];
    t[DEB_HDR_AMIN] = amin;
    t[DEB_HDR_AMAX] = amax;
    t[DEB_HDR_THERMO] = (double)thermo;
    if (g_bg_nu_ready) {
      for (int i = 0; i < 8; ++i) { s_q[i] = g_bg_nu[i]; s_w[i] = g_bg_nu[8 + i]; }
    } else {
      gauss_laguerre_fd(8, s_q, s_w);
      for (int i = 0; i < 8; ++i) { g_bg_nu[i] = s_q[i]; g_bg_nu[8 + i] = s_w[i]; }
      __threadfence();
      g_bg_nu_ready = 1;
    }
  }
  __syncthreads();
  // neutrino background on 512 log-spaced a in [0.9 amin, 1.1 amax] (background.py:158-168)
  {
    const double amnu = t[DEB_HDR_AMNU];
    double *la = arr(DEB_A_NU_LOGA), *lr = arr(DEB_A_NU_LOGRHO), *lp = arr(DEB_A_NU_LOGP), *lpp = arr(DEB_A_NU_LOGPP);
    for (int i = tid; i < DEB_NNU; i += nth) {
      double a = geom_point(amin * 0.9, amax * 1.1, i, DEB_NNU, true);
      double r = 0.0, ps = 0.0, pps = 0.0;
      for (int j = 0; j < 8; ++j) {
        double aq = a * amnu / s_q[j];
        double v = 1.0 / sqrt(1.0 + aq * aq);
        r += s_w[j] / v;
        ps += s_w[j] * v / 3.0;
        pps += s_w[j] * v * v * v / 3.0;
      }
      la[i] = log(a);
      lr[i] = log(r);
      lp[i] = log(ps);
      lpp[i] = log(pps);
    }
  }
  __syncthreads();
  if (tid < 3) {
    const int ys[3] = {DEB_A_NU_LOGRHO, DEB_A_NU_LOGP, DEB_A_NU_LOGPP};
    spl_build(arr(DEB_A_NU_LOGA), arr(ys[tid]), DEB_NNU, arr(ys[tid] + 1), arr(DEB_A_SCR0 + tid));
  }
  if (tid == 0) {
    const double *la = arr(DEB_A_NU_LOGA);
    t[DEB_HDR_NU_X0] = la[0];
    t[DEB_HDR_NU_DXINV] = 1.0 / (la[1] - la[0]);
  }
  __syncthreads();
  BgView b;
  if (tid == 0) {
    b.init(t);
    double rhonu1 = exp(b.logrhonu(0.0));
    t[DEB_HDR_OMEGAMNU] = b.grhor * rhonu1 / b.grhom;
    double Omegar = (b.Neff + b.Nmnu * rhonu1) * b.grhor / b.grhom;
    double Omegag = b.grhog / b.grhom;
    t[DEB_HDR_OMEGADE] = 1.0 - b.Omegak - Omegar - Omegag - b.Omegam;
  }
  __syncthreads();
  b.init(t);
  // adaptive a-grid with N+1 points (thermodynamics_recfast.py:325-361); a[0] -> header,
  // a[1..N] -> aexp
  if (thermo == 0) {
    const int Np = N + 1;
    const int n1 = max(8, (int)(Np * 0.05)), n2 = max(8, (int)(Np * 0.10)), n3 = (int)(Np * 0.50);
    const int n4 = Np - n1 - n2 - n3;
    const double a0 = 1e-9, a1 = 1.01;
    const double ab1 = fmax(1.0 / (1.0 + 10000.0), a0), ab2 = 1.0 / (1.0 + 1400.0), ab3 = 1.0 / (1.0 + 600.0);
    double *aexp = arr(DEB_A_AEXP);
    for (int i = tid; i < Np; i += nth) {
      double a;
      if (i < n1) a = geom_point(a0, ab1, i, n1, false);
      else if (i < n1 + n2) a = geom_point(ab1, ab2, i - n1, n2, false);
      else if (i < n1 + n2 + n3) a = geom_point(ab2, ab3, i - n1 - n2, n3, false);
      else a = geom_point(ab3, a1, i - n1 - n2 - n3, n4, true);
      if (i == 0) t[DEB_HDR_A_START] = a;
      else aexp[i - 1] = a;
    }
  }
  __syncthreads();
  // tau_max by Romberg over [amin, amax] (background.py:181-187), beside the RECFAST loop
  if (tid == 64) t[DEB_HDR_TAUMAX] = b.taumin + romb_dtauda(b, amin, amax);
  if (thermo != 0) {
    __syncthreads();  // tau_max first: the MB95 grid ends there
    if (tid == 0) {
      double *xe = arr(DEB_A_XE), *xH = arr(DEB_A_XEHI), *xHe1 = arr(DEB_A_XEHEI), *xHe2 = arr(DEB_A_XEHEII), *Tm = arr(DEB_A_TM);
      double *cs2 = arr(DEB_A_CS2), *aexp = arr(DEB_A_AEXP), *tau = arr(DEB_A_TAU);
      if (thermo == 1) {
        mb95_scan(b, N, t[DEB_HDR_TAUMAX], aexp, tau, xe, xH, xHe1, xHe2, Tm, cs2);
      } else {
        for (int i = 0; i < N; ++i) {
          aexp[i] = tin[i]; tau[i] = tin[N + i]; xe[i] = tin[2 * N + i]; Tm[i] = tin[3 * N + i]; cs2[i] = tin[4 * N + i];
          xH[i] = 0.0; xHe1[i] = 0.0; xHe2[i] = 0.0;
        }
      }
      t[DEB_HDR_RECFAST_FAIL] = 0.0;
      t[DEB_HDR_A_START] = aexp[0];
    }
    __syncthreads();
    double *aexp = arr(DEB_A_AEXP), *mu = arr(DEB_A_MU), *cs2a = arr(DEB_A_CS2A), *tempba = arr(DEB_A_TEMPBA);
    double *xpr = arr(DEB_A_XEPRIME_RECF), *loga = arr(DEB_A_LOGA);
    const double *cs2 = arr(DEB_A_CS2), *Tm = arr(DEB_A_TM);
    for (int i = tid; i < N; i += nth) {
      const double a = aexp[i];
      cs2a[i] = a * cs2[i];
      tempba[i] = a * Tm[i];
      loga[i] = log(a);
      mu[i] = 0.0;  // not produced by these modules
      xpr[i] = 0.0;
    }
  }
  if (thermo == 0 && tid < 64) {
    // RECFAST interval loop (thermodynamics_recfast.py:392-452) on wavefront 0 (every lane runs
    // the same sequence; lanes 0-2 carry the Jacobian directions), lane 0 writes
    const double *aexp = arr(DEB_A_AEXP);
    double *xH = arr(DEB_A_XEHI), *xHe = arr(DEB_A_XEHEI), *Tm = arr(DEB_A_TM);
    double *d0 = arr(DEB_A_SCR0), *d1 = arr(DEB_A_SCR1), *d2 = arr(DEB_A_SCR2);
    const double Tcmb = b.Tcmb, Tcmb2 = Tcmb * Tcmb, Tcmb3 = Tcmb2 * Tcmb;
    const double HH = b.H0 / 100.0, HO = HH * K.bigH, mu_H = 1.0 / (1.0 - b.YHe);
    const double Nnow = 3.0 * HO * HO * b.Omegab / (8.0 * PI * K.G * mu_H * K.mH);
    const double fHe = b.fHe, CR = K.CR, CB1 = K.CB1, CB1_He1 = K.CB1_He1;
    double yp[6];
    int fails = 0;
    for (int i = 0; i < N; ++i) {
      double astart = (i == 0) ? t[DEB_HDR_A_START] : aexp[i - 1];
      double aend = aexp[i];
      double zstart = 1.0 / astart - 1.0, zend = 1.0 / aend - 1.0, zz = 1.0 + zend;
      double dzda = -1.0 / (aend * aend);
      if (i == 0) { yp[0] = 1.0; yp[1] = 1.0; yp[2] = Tcmb * (1.0 + zstart); yp[3] = 0.0; yp[4] = 0.0; yp[5] = -Tcmb * (1.0 + zstart); }
      double out[6];
      if (zend > 3500.0) {
        out[0] = 1.0; out[1] = 1.0; out[2] = Tcmb * zz; out[3] = 0.0; out[4] = 0.0; out[5] = -Tcmb * zz;
      } else if (i > 0 && yp[1] > 0.99) {
        double rhs = exp(1.5 * log(CR * Tcmb / zz) - CB1_He1 / (Tcmb * zz)) / Nnow * 4.0;
        double xHe0 = 0.5 * (sqrt((rhs - 1.0) * (rhs - 1.0) + 4.0 * (1.0 + fHe) * rhs) - (rhs - 1.0));
        double qv = -((CB1_He1 - CR * Tcmb2) / (Tcmb * zz));
        double cb = -CB1_He1 + CR * Tcmb2;
        double sq = sqrt(Nnow * Nnow + (16.0 * cb * cb * cb) / (Tcmb3 * zz * zz * zz) + 8.0 * (1.0 + 2.0 * fHe) * Nnow * pow(qv, 1.5));
        double dxHeIdz = (-3.0 * pow(-(CB1_He1 / Tcmb) + CR * Tcmb, 1.5) * (Nnow + 2.0 * fHe * Nnow + 4.0 * pow(qv, 1.5) - sq)) /
                         (Nnow * pow(zz, 2.5) * sq);
        out[0] = 1.0; out[1] = (xHe0 - 1.0) / fHe; out[2] = Tcmb * zz; out[3] = 0.0; out[4] = dxHeIdz * dzda; out[5] = -Tcmb * zz;
      } else {
        double ye[3], f[3];
        if (!solve_ionization(b, K, astart, aend, yp, ye, tid)) ++fails;
        ionization<double>(b, K, aend, ye, f);
        out[0] = ye[0]; out[1] = ye[1]; out[2] = ye[2]; out[3] = f[0]; out[4] = f[1]; out[5] = f[2];
        if (i > 0 && yp[0] > 0.99) {  // hydrogen Saha (f_case3)
          double rhs = exp(1.5 * log(CR * Tcmb / zz) - CB1 / (Tcmb * zz)) / Nnow;
          double xH0 = 0.5 * (sqrt(rhs * rhs + 4.0 * rhs) - rhs);
          double qv = -((CB1 - CR * Tcmb2) / (Tcmb * zz));
          double sqq = sqrt(qv);
          double inner = CB1 * CB1 - 2.0 * CB1 * CR * Tcmb2 + Tcmb2 * (CR * CR * Tcmb2 + Nnow * zz * zz * sqq);
          double num = (CB1 - CR * Tcmb2) * (2.0 * CB1 * CB1 - 4.0 * CB1 * CR * Tcmb2 + Tcmb2 * (2.0 * CR * CR * Tcmb2 + Nnow * zz * zz * sqq));
          double den = pow(Tcmb, 1.5) * pow(zz, 2.5) * sqrt((-CB1 + CR * Tcmb2) * inner);
          double dxHdz = (3.0 * ((2.0 * pow(qv, 1.5)) / zz + num / den)) / (2.0 * Nnow);
          out[0] = xH0; out[3] = dxHdz * dzda;
        }
      }
      if (tid == 0) { xH[i] = out[0]; xHe[i] = out[1]; Tm[i] = out[2]; d0[i] = out[3]; d1[i] = out[4]; d2[i] = out[5]; }
      for (int j = 0; j < 6; ++j) yp[j] = out[j];
    }
    if (tid == 0) t[DEB_HDR_RECFAST_FAIL] = (double)fails;
  }
  __syncthreads();
  // evaluate_thermo per knot (thermodynamics_recfast.py:456-492)
  if (thermo == 0) {
    const double *aexp = arr(DEB_A_AEXP), *xH = arr(DEB_A_XEHI), *xHe = arr(DEB_A_XEHEI), *Tm = arr(DEB_A_TM);
    const double *d0 = arr(DEB_A_SCR0), *d1 = arr(DEB_A_SCR1), *d2 = arr(DEB_A_SCR2);
    double *xe = arr(DEB_A_XE), *xeHeII = arr(DEB_A_XEHEII), *mu = arr(DEB_A_MU), *cs2 = arr(DEB_A_CS2), *cs2a = arr(DEB_A_CS2A);
    double *tempba = arr(DEB_A_TEMPBA), *xpr = arr(DEB_A_XEPRIME_RECF), *loga = arr(DEB_A_LOGA), *inc = arr(DEB_A_SCR3);
    for (int i = tid; i < N; i += nth) {
      double a = aexp[i];
      Dual ad(a);
      ad.d[0] = 1.0;
      Dual sh = saha_HeII<Dual>(b, K, ad);
      xeHeII[i] = sh.v;
      double xev = xH[i] + b.fHe * xHe[i] + sh.v;
      xe[i] = xev;
      double muv = 1.0 / (1.0 + (1.0 / K.mHe_mH - 1.0) * b.YHe + (1.0 - b.YHe) * xev);
      mu[i] = muv;
      double daTmda = Tm[i] + a * d2[i];
      double c2v = K.kB / K.mH / (K.c * K.c) / muv * Tm[i] * (4.0 - daTmda / Tm[i]) / 3.0;
      cs2[i] = c2v;
      cs2a[i] = a * c2v;
      tempba[i] = a * Tm[i];
      xpr[i] = (d0[i] + b.fHe * d1[i] + sh.d[0]) * (1.0 / b.dtauda(a));
      loga[i] = log(a);
      inc[i] = (i > 0) ? romb_dtauda(b, aexp[i - 1], a) : 0.0;
    }
  }
  __syncthreads();
  if (thermo == 0 && tid == 0) {  // tau = taumin + concatenate([0], cumsum(increments))
    double *tau = arr(DEB_A_TAU);
    const double *inc = arr(DEB_A_SCR3);
    double cum = 0.0;
    tau[0] = b.taumin;
    for (int i = 1; i < N; ++i) { cum += inc[i]; tau[i] = b.taumin + cum; }
  }
  __syncthreads();
  // reionization on the RECFAST history (tau_reion > 0): z_re by bisection on the added optical
  // depth, then x_e, its conformal-time derivative, mu and the sound speed at every knot
  if (thermo == 0 && p[DEB_P_TAU_REION] > 0.0) {
    const double tre = p[DEB_P_TAU_REION];
    const double *loga = arr(DEB_A_LOGA), *aexp = arr(DEB_A_AEXP), *Tm = arr(DEB_A_TM);
    double *xe = arr(DEB_A_XE), *xpr = arr(DEB_A_XEPRIME_RECF), *mu = arr(DEB_A_MU), *cs2 = arr(DEB_A_CS2), *cs2a = arr(DEB_A_CS2A);
    double *S = arr(DEB_A_SCR4);
    __shared__ double s_red[4], s_lo, s_hi;
    if (tid == 0) spl_build(loga, xe, N, S, arr(DEB_A_SCR5));
    __syncthreads();
    const double dlo = reion_depth(b, loga, xe, S, N, 0.0, s_red, tid, nth);
    const double dhi = reion_depth(b, loga, xe, S, N, 100.0, s_red, tid, nth);
    if (tid == 0) { s_lo = 0.0; s_hi = 100.0; }
    __syncthreads();
    if (tre > dlo && tre < dhi) {
      for (int it = 0; it < 60; ++it) {  // the added depth increases with z_re
        const double zm = 0.5 * (s_lo + s_hi);
        const double dm = reion_depth(b, loga, xe, S, N, zm, s_red, tid, nth);
        __syncthreads();
        if (tid == 0) {
          if (dm < tre) s_lo = zm; else s_hi = zm;
        }
        __syncthreads();
      }
    } else if (tid == 0) {  // outside what z_re in [0, 100] gives: the nearest end
      s_lo = s_hi = (tre <= dlo) ? 0.0 : 100.0;
    }
    __syncthreads();
    const double zre = 0.5 * (s_lo + s_hi);
    const double dfin = reion_depth(b, loga, xe, S, N, zre, s_red, tid, nth);
    if (tid == 0) { t[DEB_HDR_Z_REION] = zre; t[DEB_HDR_TAU_REION_OUT] = dfin; }
    Reion r;
    r.set(zre, b.fHe);
    for (int i = tid; i < N; i += nth) {
      const double a = aexp[i], xrec = xe[i];
      double d, FH;
      const double dx = r.dxe(a, xrec, d, FH);
      const double xn = xrec + dx;
      // d/dtau [x_rec + (f_re - x_rec) F_H + ...] = x_rec' (1 - F_H) + (d/da at fixed x_rec) / (dtau/da)
      xpr[i] = xpr[i] * (1.0 - FH) + d / b.dtauda(a);
      const double mun = 1.0 / (1.0 + (1.0 / K.mHe_mH - 1.0) * b.YHe + (1.0 - b.YHe) * xn);
      cs2[i] *= mun / mu[i];  // c_s^2 is proportional to 1 / mu at a given T_m
      cs2a[i] = a * cs2[i];
      mu[i] = mun;
      xe[i] = xn;
      (void)Tm;
    }
  }
  __syncthreads();
  // splines of background.py:248-256
  if (tid < 7) {
    int xs, ys, ss;
    switch (tid) {
      case 0: xs = DEB_A_AEXP; ys = DEB_A_TAU; ss = DEB_A_S_TAU_OF_A; break;
      case 1: xs = DEB_A_TAU; ys = DEB_A_AEXP; ss = DEB_A_S_A_OF_TAU; break;
      case 2: xs = DEB_A_TAU; ys = DEB_A_XE; ss = DEB_A_S_XE_OF_TAU; break;
      case 3: xs = DEB_A_TAU; ys = DEB_A_CS2A; ss = DEB_A_S_CS2A_OF_TAU; break;
      case 4: xs = DEB_A_TAU; ys = DEB_A_TEMPBA; ss = DEB_A_S_TEMPBA_OF_TAU; break;
      case 5: xs = DEB_A_LOGA; ys = DEB_A_XE; ss = DEB_A_S_XE_OF_LOGA; break;
      default: xs = DEB_A_LOGA; ys = DEB_A_CS2A; ss = DEB_A_S_CS2A_OF_LOGA; break;
    }
    spl_build(arr(xs), arr(ys), N, arr(ss), arr(DEB_A_SCR0 + tid));
  }
  __syncthreads();
  if (tid == 0) {
    s_tau_pre = b.tau_of_a(1e-4);
    s_xe_full = 1.0 + b.YHe / (1.0 - b.YHe);
  }
  __syncthreads();
  // opacity (background.py:309-319)
  {
    const double *tau = arr(DEB_A_TAU), *xe = arr(DEB_A_XE), *aexp = arr(DEB_A_AEXP), *Sx = arr(DEB_A_S_XE_OF_TAU);
    double *xv = arr(DEB_A_XE_VIS), *xp = arr(DEB_A_XEPRIME), *op = arr(DEB_A_OPAC);
    for (int i = tid; i < N; i += nth) {
      double tt = tau[i];
      bool pre = tt <= s_tau_pre;
      int j = spl_index(tau, N, tt);
      double xev = pre ? s_xe_full : spl_eval_at(tau, xe, Sx, j, tt);
      xv[i] = xev;
      xp[i] = pre ? 0.0 : spl_deriv_at(tau, xe, Sx, j, tt);
      op[i] = xev * b.akthom / (aexp[i] * aexp[i]);
    }
  }
  __syncthreads();
  if (tid == 0) {
    const double *tau = arr(DEB_A_TAU), *op = arr(DEB_A_OPAC);
    double *S = arr(DEB_A_S_OPAC), *I = arr(DEB_A_I_OPAC);
    spl_build(tau, op, N, S, arr(DEB_A_SCR0));
    spl_cumint(tau, op, S, N, I);
    double tau0 = b.tau_of_a(1.0);
    int j = spl_index(tau, N, tau0);
    s_tau0 = tau0;
    s_odt = spl_integral_at(tau, op, S, I, j, tau0);
    t[DEB_HDR_TAU0] = tau0;
    t[DEB_HDR_OPTDEPTH_TODAY] = s_odt;
  }
  __syncthreads();
  // optical depth and visibility (background.py:321-335)
  {
    const double *tau = arr(DEB_A_TAU), *op = arr(DEB_A_OPAC), *S = arr(DEB_A_S_OPAC), *I = arr(DEB_A_I_OPAC);
    double *od = arr(DEB_A_OPTDEPTH), *g = arr(DEB_A_GVIS), *g1 = arr(DEB_A_GVISP), *g2 = arr(DEB_A_GVISPP);
    for (int i = tid; i < N; i += nth) {
      double tt = tau[i];
      int j = spl_index(tau, N, tt);
      double o = op[i];
      double o1 = spl_deriv_at(tau, op, S, j, tt);
      double o2 = spl_deriv2_at(tau, S, j, tt);
      double odv = s_odt - spl_integral_at(tau, op, S, I, j, tt);
      double em = exp(-odv);
      od[i] = odv;
      g[i] = o * em;
      g1[i] = (o1 + o * o) * em;
      g2[i] = (o2 + 3.0 * o * o1 + o * o * o) * em;
    }
  }
  __syncthreads();
  if (tid == 0) {  // tau at the first maximum of the visibility
    const double *g = arr(DEB_A_GVIS), *tau = arr(DEB_A_TAU);
    int im = 0;
    for (int i = 1; i < N; ++i)
      if (g[i] > g[im]) im = i;
    t[DEB_HDR_TAU_MAXVIS] = tau[im];
  }
  (void)A;
}

}  // namespace

}  // namespace deb

// host launcher (called by the C ABI)
extern "C" int deb_launch_background(const double *d_params, int n_cosmo, int num_thermo, int thermo, const double *d_thermo_in,
                                     double *d_tables, hipStream_t stream) {
  long stride = deb_table_doubles(num_thermo);
  hipLaunchKernelGGL(deb::background_kernel, dim3(n_cosmo), dim3(256), 0, stream, d_params, num_thermo, d_tables, stride, thermo,
                     d_thermo_in);
  return (int)hipGetLastError();
}
