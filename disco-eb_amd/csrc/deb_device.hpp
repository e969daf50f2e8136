// deb_device.hpp -- device-side building blocks shared by the background and perturbation
// kernels: forward-mode dual numbers, natural-cubic-spline evaluation on HBM tables, and the
// per-cosmology table layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "deb_layout.h"

namespace deb {

// ---------------------------------------------------------------------------------------
// Forward-mode dual numbers with N derivative directions (N=1: d/da or d/dtau; N=3: the
// RECFAST 3x3 Jacobian in one pass).
// ---------------------------------------------------------------------------------------
template <int N>
struct Dn {
  double v;
  double d[N];
  __device__ __host__ Dn() {}
  __device__ __host__ Dn(double x) : v(x) {
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
  }
};
typedef Dn<1> Dual;

__device__ __forceinline__ double val(double x) { return x; }
template <int N> __device__ __forceinline__ double val(const Dn<N> &x) { return x.v; }

template <int N> __device__ __forceinline__ Dn<N> operator+(const Dn<N> &a, const Dn<N> &b) {
  Dn<N> r; r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator-(const Dn<N> &a, const Dn<N> &b) {
  Dn<N> r; r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator-(const Dn<N> &a) {
  Dn<N> r; r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator*(const Dn<N> &a, const Dn<N> &b) {
  Dn<N> r; r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator/(const Dn<N> &a, const Dn<N> &b) {
  Dn<N> r; double ib = 1.0 / b.v; r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator+(const Dn<N> &a, double b) { Dn<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ Dn<N> operator+(double b, const Dn<N> &a) { Dn<N> r = a; r.v += b; return r; }
template <int N> __device__ __forceinline__ Dn<N> operator-(const Dn<N> &a, double b) { Dn<N> r = a; r.v -= b; return r; }
template <int N> __device__ __forceinline__ Dn<N> operator-(double b, const Dn<N> &a) {
  Dn<N> r; r.v = b - a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator*(const Dn<N> &a, double b) {
  Dn<N> r; r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> operator*(double b, const Dn<N> &a) { return a * b; }
template <int N> __device__ __forceinline__ Dn<N> operator/(const Dn<N> &a, double b) { return a * (1.0 / b); }
template <int N> __device__ __forceinline__ Dn<N> operator/(double b, const Dn<N> &a) {
  Dn<N> r; double ia = 1.0 / a.v; r.v = b * ia;
  double s = -r.v * ia;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
  return r;
}
template <int N> __device__ __forceinline__ Dn<N> &operator+=(Dn<N> &a, const Dn<N> &b) { a = a + b; return a; }
template <int N> __device__ __forceinline__ Dn<N> &operator+=(Dn<N> &a, double b) { a.v += b; return a; }

template <int N> __device__ __forceinline__ Dn<N> chain(const Dn<N> &a, double fv, double dfdx) {
  Dn<N> r; r.v = fv;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = dfdx * a.d[i];
  return r;
}
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double dpow(double x, double p) { return pow(x, p); }
template <int N> __device__ __forceinline__ Dn<N> dexp(const Dn<N> &a) { double e = exp(a.v); return chain(a, e, e); }
template <int N> __device__ __forceinline__ Dn<N> dlog(const Dn<N> &a) { return chain(a, log(a.v), 1.0 / a.v); }
template <int N> __device__ __forceinline__ Dn<N> dsqrt(const Dn<N> &a) { double s = sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int N> __device__ __forceinline__ Dn<N> dpow(const Dn<N> &a, double p) {
  double f = pow(a.v, p);
  return chain(a, f, p * pow(a.v, p - 1.0));
}
// |x| along the real line, max/min against a constant (branch on the value)
template <class T> __device__ __forceinline__ T dabs(const T &x) { return val(x) < 0 ? T(-x) : x; }
template <class T> __device__ __forceinline__ T dmax(const T &x, double b) { return val(x) > b ? x : T(b); }
template <class T> __device__ __forceinline__ T dmin(const T &x, double b) { return val(x) < b ? x : T(b); }

// ---------------------------------------------------------------------------------------
// Natural cubic spline evaluation (tables built by the background kernel).  Interval index:
// number of knots strictly below x, minus one, clipped to [0, n-2]; on uniform grids
// floor((x - x0) / dx) clipped to the same range.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int spl_index(const double *__restrict__ x, int n, double xq) {
  // lower_bound: first knot >= xq
  int lo = 0, len = n;
  while (len > 0) {
    int half = len >> 1;
    if (x[lo + half] < xq) { lo += half + 1; len -= half + 1; }
    else len = half;
  }
  int i = lo - 1;
  i = i < 0 ? 0 : i;
  return i > n - 2 ? n - 2 : i;
}
__device__ __forceinline__ int spl_index_uniform(double x0, double dxinv, int n, double xq) {
  double f = floor((xq - x0) * dxinv);
  if (!(f > 0.0)) f = 0.0;
  if (f > (double)(n - 2)) f = (double)(n - 2);
  return (int)f;
}
template <class T>
__device__ __forceinline__ T spl_eval_at(const double *__restrict__ x, const double *__restrict__ y,
                                         const double *__restrict__ S, int i, const T &xq) {
  double x0 = x[i], x1 = x[i + 1], y0 = y[i], y1 = y[i + 1], S0 = S[i], S1 = S[i + 1];
  double hl = x1 - x0;
  T t = (xq - x0) / hl;
  T A = 1.0 - t;
  T B = t;
  return A * y0 + B * y1 + ((A * A * A - A) * S0 + (B * B * B - B) * S1) * (hl * hl / 6.0);
}
template <class T>
__device__ __forceinline__ T spl_deriv_at(const double *__restrict__ x, const double *__restrict__ y,
                                          const double *__restrict__ S, int i, const T &xq) {
  double x0 = x[i], x1 = x[i + 1], y0 = y[i], y1 = y[i + 1], S0 = S[i], S1 = S[i + 1];
  double hl = x1 - x0;
  T d = xq - x0;
  double b = (y1 - y0) / hl - hl * (S1 + 2.0 * S0) / 6.0;
  double c = S0 / 2.0;
  double dd = (S1 - S0) / (6.0 * hl);
  return b + 2.0 * c * d + 3.0 * dd * d * d;
}
__device__ __forceinline__ double spl_deriv2_at(const double *__restrict__ x, const double *__restrict__ S, int i, double xq) {
  double hl = x[i + 1] - x[i];
  double d = xq - x[i];
  return 2.0 * (S[i] / 2.0) + 6.0 * ((S[i + 1] - S[i]) / (6.0 * hl)) * d;
}
__device__ __forceinline__ double spl_integral_at(const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ S, const double *__restrict__ Icum, int i,
                                                  double xq) {
  double hl = x[i + 1] - x[i];
  double d = xq - x[i];
  double a = y[i];
  double b = (y[i + 1] - y[i]) / hl - hl * (S[i + 1] + 2.0 * S[i]) / 6.0;
  double c = S[i] / 2.0;
  double dd = (S[i + 1] - S[i]) / (6.0 * hl);
  return Icum[i] + a * d + b * d * d / 2.0 + c * d * d * d / 3.0 + dd * d * d * d * d / 4.0;
}

// Second derivatives of a natural cubic spline by the Thomas algorithm (single thread).
// dp is scratch of length n.  S[0] = S[n-1] = 0.
__device__ inline void spl_build(const double *__restrict__ x, const double *__restrict__ y, int n, double *__restrict__ S,
                                 double *__restrict__ dp) {
  S[0] = 0.0;
  S[n - 1] = 0.0;
  int m = n - 2;
  if (m <= 0) return;
  // cp is kept in S[i+1] until the back substitution overwrites it in place
  double h0 = x[1] - x[0], h1 = x[2] - x[1];
  double bd = 2.0 * (h0 + h1);
  double d0 = 6.0 * ((y[2] - y[1]) / h1 - (y[1] - y[0]) / h0);
  S[1] = (m > 1) ? h1 / bd : 0.0;
  dp[0] = d0 / bd;
  for (int i = 1; i < m; ++i) {
    double hi = x[i + 1] - x[i], hi1 = x[i + 2] - x[i + 1];
    double denom = 2.0 * (hi + hi1) - hi * S[i];
    double di = 6.0 * ((y[i + 2] - y[i + 1]) / hi1 - (y[i + 1] - y[i]) / hi);
    S[i + 1] = (i < m - 1) ? hi1 / denom : 0.0;
    dp[i] = (di - hi * dp[i - 1]) / denom;
  }
  S[m] = dp[m - 1];
  for (int i = m - 2; i >= 0; --i) S[i + 1] = dp[i] - S[i + 1] * S[i + 2];
}
// cumulative integral from the first knot to each knot
__device__ inline void spl_cumint(const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ S,
                                  int n, double *__restrict__ Icum) {
  Icum[0] = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    double h = x[i + 1] - x[i];
    double a = y[i];
    double b = (y[i + 1] - y[i]) / h - h * (S[i + 1] + 2.0 * S[i]) / 6.0;
    double c = S[i] / 2.0;
    double d = (S[i + 1] - S[i]) / (6.0 * h);
    Icum[i + 1] = Icum[i] + (a * h + b * h * h / 2 + c * h * h * h / 3 + d * h * h * h * h / 4);
  }
}

// Nodes and Fermi-Dirac weights of n-point generalized Gauss-Laguerre quadrature (alpha = 1),
// background.py:40-45 / util.py:83-124: eigen-decomposition of the Golub-Welsch Jacobi matrix
// (diag 2i, off-diag sqrt(i(i+1))) by cyclic Jacobi rotations; weight = first eigenvector
// component squared (Gamma(2) = 1), then w *= q^3/(1+exp(-q))/q / (7 pi^4/120).  n <= 16.
__device__ inline void gauss_laguerre_fd(int n, double *q, double *w) {
  double A[16][16], V[16][16];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) { A[i][j] = 0.0; V[i][j] = (i == j) ? 1.0 : 0.0; }
  for (int i = 1; i <= n; ++i) A[i - 1][i - 1] = 2.0 * i;
  for (int i = 1; i < n; ++i) { double bb = sqrt((double)i * (i + 1.0)); A[i - 1][i] = bb; A[i][i - 1] = bb; }
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int r = p + 1; r < n; ++r) off += A[p][r] * A[p][r];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int r = p + 1; r < n; ++r) {
        double apq = A[p][r];
        if (fabs(apq) < 1e-300) continue;
        double theta = (A[r][r] - A[p][p]) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double cc = 1.0 / sqrt(t * t + 1.0), s = t * cc;
        for (int k = 0; k < n; ++k) { double akp = A[k][p], akq = A[k][r]; A[k][p] = cc * akp - s * akq; A[k][r] = s * akp + cc * akq; }
        for (int k = 0; k < n; ++k) { double apk = A[p][k], aqk = A[r][k]; A[p][k] = cc * apk - s * aqk; A[r][k] = s * apk + cc * aqk; }
        for (int k = 0; k < n; ++k) { double vkp = V[k][p], vkq = V[k][r]; V[k][p] = cc * vkp - s * vkq; V[k][r] = s * vkp + cc * vkq; }
      }
  }
  double ev[16], wt[16];
  for (int j = 0; j < n; ++j) { ev[j] = A[j][j]; wt[j] = V[0][j] * V[0][j]; }
  for (int i = 1; i < n; ++i)
    for (int j = i; j > 0 && ev[j - 1] > ev[j]; --j) {
      double t = ev[j]; ev[j] = ev[j - 1]; ev[j - 1] = t;
      t = wt[j]; wt[j] = wt[j - 1]; wt[j - 1] = t;
    }
  const double fdc = 5.682196976983475;
  for (int j = 0; j < n; ++j) {
    q[j] = ev[j];
    w[j] = wt[j] * ev[j] * ev[j] * ev[j] / (1.0 + exp(-ev[j])) / ev[j] / fdc;
  }
}

}  // namespace deb
