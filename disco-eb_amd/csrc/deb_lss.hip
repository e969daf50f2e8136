// deb_lss.hip -- large-scale-structure observables on top of the solver's output
// (perturbations.py:1409-1569, util.py:291-328): Savitzky-Golay taps and filtering, the smoothed
// power spectrum, the Kaiser anisotropic spectrum and its multipoles, and the FFTLog correlation
// function.  Inputs are a few hundred to tens of thousands of k-modes: every kernel is a short
// elementwise or reduction pass over HBM-resident vectors, one thread per output element.
//
// The correlation function needs an rfft / irfft of the (arbitrary, even) length n of the k-grid;
// it is done here as two direct transforms (one thread per output coefficient, O(n^2) multiply-adds,
// twiddles from the exact integer phase (j m mod n) / n), which for n <= 2^15 costs well under a
// millisecond and needs no power-of-two length or planning.
#include <hip/hip_runtime.h>

#include "deb_complex.hpp"

namespace deb {
namespace {
using namespace cplx;

constexpr int kMaxPolyorder = 8;

// Taps of util.py:291-328: the minimum-norm solution of A c = e_0, A[i][j] = x_j^i, x decreasing
// window positions about the centre.  With x scaled by s (A_s = D A, D = diag(s^-i)), the
// minimum-norm solution is c = A_s^T (A_s A_s^T)^-1 e_0 (D e_0 = e_0), which keeps the normal
// matrix well conditioned for long windows.  One thread; Gaussian elimination with partial pivots.
__global__ void savgol_coef_kernel(int w, int p, double *__restrict__ c) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int half = w / 2;
  const double pos = (w % 2 == 0) ? half - 0.5 : (double)half;
  const double s = pos > 1.0 ? pos : 1.0;
  const int m = p + 1;
  double G[kMaxPolyorder + 1][kMaxPolyorder + 2];
  for (int i = 0; i < m; ++i)
    for (int j = 0; j <= m; ++j) G[i][j] = 0.0;
  for (int t = 0; t < w; ++t) {
    const double x = (pos - t) / s;  // decreasing positions
    double xi = 1.0;
    double pw[2 * kMaxPolyorder + 1];
    for (int e = 0; e <= 2 * p; ++e) { pw[e] = xi; xi *= x; }
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) G[i][j] += pw[i + j];
  }
  G[0][m] = 1.0;
  for (int col = 0; col < m; ++col) {
    int piv = col;
    for (int r = col + 1; r < m; ++r)
      if (fabs(G[r][col]) > fabs(G[piv][col])) piv = r;
    if (piv != col)
      for (int j = 0; j <= m; ++j) { const double t = G[col][j]; G[col][j] = G[piv][j]; G[piv][j] = t; }
    for (int r = col + 1; r < m; ++r) {
      const double f = G[r][col] / G[col][col];
      for (int j = col; j <= m; ++j) G[r][j] -= f * G[col][j];
    }
  }
  double z[kMaxPolyorder + 1];
  for (int i = m - 1; i >= 0; --i) {
    double v = G[i][m];
    for (int j = i + 1; j < m; ++j) v -= G[i][j] * z[j];
    z[i] = v / G[i][i];
  }
  for (int t = 0; t < w; ++t) {
    const double x = (pos - t) / s;
    double v = 0.0, xi = 1.0;
    for (int i = 0; i < m; ++i) { v += z[i] * xi; xi *= x; }
    c[t] = v;
  }
}

// np.convolve(y, c, 'same'): out[i] = sum_m c[m] y[i + (w - 1) / 2 - m], zero beyond the ends
__device__ __forceinline__ double conv_same(const double *__restrict__ y, int n, const double *__restrict__ c, int w, int i) {
  const int h = (w - 1) / 2;
  double v = 0.0;
  for (int m = 0; m < w; ++m) {
    const int j = i + h - m;
    if (j >= 0 && j < n) v += c[m] * y[j];
  }
  return v;
}

__global__ void savgol_apply_kernel(const double *__restrict__ y, int n, const double *__restrict__ c, int w, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = conv_same(y, n, c, w, i);
}

__device__ __forceinline__ double primordial(double k, double A_s, double n_s, double k_p) {
  return 2.0 * kPiC * kPiC * A_s * pow(k / k_p, n_s - 1.0) / (k * k * k);
}

__global__ void log_power_kernel(const double *__restrict__ k, const double *__restrict__ y, int n, int ystride, int idx, double A_s, double n_s,
                                 double k_p, double *__restrict__ lp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = y[(long)i * ystride + idx];
  lp[i] = log(primordial(k[i], A_s, n_s, k_p) * v * v);
}

// exp of the filtered log P; the first w / 2 and the last (w + 1) / 2 points keep P itself
// (the reference's slices [:w // 2] and [-w // 2:], perturbations.py:1501-1503)
__global__ void smooth_power_kernel(const double *__restrict__ lp, int n, const double *__restrict__ c, int w, double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool keep = i < w / 2 || i >= n - (w + 1) / 2;
  out[i] = keep ? exp(lp[i]) : exp(conv_same(lp, n, c, w, i));
}

// P(k, mu) = (b delta - mu^2 theta)^2 (perturbations.py:1507-1543); pd, pt non-NULL: delta =
// sqrt(pd), theta = -sqrt(pt) (the smoothed spectra), else the scaled transfer functions y[4], y[5]
__global__ void kaiser_kernel(const double *__restrict__ k, const double *__restrict__ y, int n, int ystride, const double *__restrict__ pd,
                              const double *__restrict__ pt, double b, double A_s, double n_s, double k_p, int nmu, int mu_sampling,
                              double *__restrict__ pkmu, double *__restrict__ mu_out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)n * nmu) return;
  const int i = (int)(t / nmu), j = (int)(t % nmu);
  const double frac = nmu > 1 ? (double)j / (double)(nmu - 1) : 0.0;
  double mu;
  if (mu_sampling) {
    mu = (nmu > 1 && j == nmu - 1) ? 1.0 : -1.0 + 2.0 * frac;
  } else {
    mu = cospi(frac);
  }
  double dm, tm;
  if (pd != nullptr) {
    dm = sqrt(pd[i]);
    tm = -sqrt(pt[i]);
  } else {
    const double amp = sqrt(primordial(k[i], A_s, n_s, k_p));
    dm = amp * y[(long)i * ystride + 4];
    tm = amp * y[(long)i * ystride + 5];
  }
  const double v = b * dm - mu * mu * tm;
  pkmu[t] = v * v;
  if (i == 0) mu_out[j] = mu;
}

// monopole, quadrupole, hexadecapole of the Kaiser spectrum (perturbations.py:1547-1569)
__global__ void multipoles_kernel(const double *__restrict__ k, const double *__restrict__ y, int n, int ystride, double b, double A_s,
                                  double n_s, double k_p, double *__restrict__ p0, double *__restrict__ p2, double *__restrict__ p4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double amp = sqrt(primordial(k[i], A_s, n_s, k_p));
  const double dm = amp * y[(long)i * ystride + 4], tm = amp * y[(long)i * ystride + 5];
  p0[i] = b * b * dm * dm - 2.0 * b / 3.0 * dm * tm + tm * tm / 5.0;
  p2[i] = -4.0 * b / 3.0 * dm * tm + 4.0 / 7.0 * tm * tm;
  p4[i] = 8.0 / 35.0 * tm * tm;
}

// F_m = sum_j P_j k_j^1.5 e^{-2 pi i j m / n} (m = 0..n/2), then times
// exp(2 i (Im lnGamma((1.5 + l) / 2 + i kappa_m) - kappa_m ln pi)), kappa_m = pi m / ln(k_max / k_min)
// (perturbations.py:1427-1436)
__global__ void xi_forward_kernel(const double *__restrict__ k, const double *__restrict__ pk, int n, int ell, double2 *__restrict__ F) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m > n / 2) return;
  double re = 0.0, im = 0.0;
  long ph = 0;  // j m mod n
  for (int j = 0; j < n; ++j) {
    const double f = pk[j] * k[j] * sqrt(k[j]);
    double s, c;
    sincospi(-2.0 * (double)ph / (double)n, &s, &c);
    re += f * c;
    im += f * s;
    ph += m;
    if (ph >= n) ph -= n;
  }
  const double L = log(k[n - 1] / k[0]);
  const double kap = kPiC * (double)m / L;
  const double th = lngamma_complex(make_double2((1.5 + ell) / 2.0, kap)).y;
  const double ang = 2.0 * (th - log(kPiC) * kap);
  double s, c;
  sincos(ang, &s, &c);
  F[m] = cmul(make_double2(re, im), make_double2(c, s));
}

// x_j = irfft(F)_j, xi_j = Re(i^l) x_j / (2 pi r_j)^1.5 with r_j = 2 pi / k_j, written in reverse
// order (increasing r) as the reference returns them (perturbations.py:1438-1443)
__global__ void xi_inverse_kernel(const double *__restrict__ k, const double2 *__restrict__ F, int n, int ell, double *__restrict__ xi,
                                  double *__restrict__ r) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int M = n / 2;
  double v = F[0].x + ((j & 1) ? -F[M].x : F[M].x);
  long ph = j;  // m j mod n
  for (int m = 1; m < M; ++m) {
    double s, c;
    sincospi(2.0 * (double)ph / (double)n, &s, &c);
    v += 2.0 * (F[m].x * c - F[m].y * s);
    ph += j;
    if (ph >= n) ph -= n;
  }
  v /= (double)n;
  const double sgn = (ell % 4 == 0) ? 1.0 : ((ell % 4 == 2) ? -1.0 : 0.0);  // Re(i^l)
  const double rj = 2.0 * kPiC / k[j];
  const double d = 2.0 * kPiC * rj;
  xi[n - 1 - j] = sgn * v / (d * sqrt(d));
  r[n - 1 - j] = rj;
}

}  // namespace
}  // namespace deb

extern "C" int deb_launch_savgol_coefficients(int w, int p, double *d_c, hipStream_t stream) {
  if (p > deb::kMaxPolyorder) return -1;
  hipLaunchKernelGGL(deb::savgol_coef_kernel, dim3(1), dim3(64), 0, stream, w, p, d_c);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_savgol_apply(const double *d_y, int n, const double *d_c, int w, double *d_out, hipStream_t stream) {
  hipLaunchKernelGGL(deb::savgol_apply_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_y, n, d_c, w, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_power_smoothed(const double *d_k, const double *d_y, int n, int ystride, int idx, double A_s, double n_s, double k_p,
                                         int w, double *d_out, double *d_work, hipStream_t stream) {
  double *c = d_work, *lp = d_work + w;
  hipLaunchKernelGGL(deb::savgol_coef_kernel, dim3(1), dim3(64), 0, stream, w, 3, c);
  hipLaunchKernelGGL(deb::log_power_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_k, d_y, n, ystride, idx, A_s, n_s, k_p, lp);
  hipLaunchKernelGGL(deb::smooth_power_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, lp, n, c, w, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_kaiser(const double *d_k, const double *d_y, int n, int ystride, const double *d_pd, const double *d_pt, double b,
                                 double A_s, double n_s, double k_p, int nmu, int mu_sampling, double *d_pkmu, double *d_mu, hipStream_t stream) {
  const long tot = (long)n * nmu;
  hipLaunchKernelGGL(deb::kaiser_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream, d_k, d_y, n, ystride, d_pd, d_pt, b, A_s,
                     n_s, k_p, nmu, mu_sampling, d_pkmu, d_mu);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_multipoles(const double *d_k, const double *d_y, int n, int ystride, double b, double A_s, double n_s, double k_p,
                                     double *d_p0, double *d_p2, double *d_p4, hipStream_t stream) {
  hipLaunchKernelGGL(deb::multipoles_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d_k, d_y, n, ystride, b, A_s, n_s, k_p, d_p0, d_p2,
                     d_p4);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_xi(const double *d_k, const double *d_pk, int n, int ell, double *d_xi, double *d_r, double *d_work, hipStream_t stream) {
  double2 *F = reinterpret_cast<double2 *>(d_work);
  hipLaunchKernelGGL(deb::xi_forward_kernel, dim3((n / 2 + 1 + 63) / 64), dim3(64), 0, stream, d_k, d_pk, n, ell, F);
  hipLaunchKernelGGL(deb::xi_inverse_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, d_k, F, n, ell, d_xi, d_r);
  return (int)hipGetLastError();
}
