// deb_misc.hip -- small elementwise kernels: spline queries on the background tables and P(k).
#include <hip/hip_runtime.h>

#include "deb_cosmo.hpp"

namespace deb {
namespace {

// spline ids follow include/discoeb_amd.h (DEB_SPL_*)
__global__ void spline_eval_kernel(const double *__restrict__ t, int N, int sid, int kind, const double *__restrict__ x, int nx,
                                   double *__restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx) return;
  const int xs[11] = {DEB_A_AEXP, DEB_A_TAU, DEB_A_TAU, DEB_A_TAU, DEB_A_TAU, DEB_A_LOGA, DEB_A_LOGA, DEB_A_NU_LOGA, DEB_A_NU_LOGA,
                      DEB_A_NU_LOGA, DEB_A_TAU};
  const int ys[11] = {DEB_A_TAU, DEB_A_AEXP, DEB_A_XE, DEB_A_CS2A, DEB_A_TEMPBA, DEB_A_XE, DEB_A_CS2A, DEB_A_NU_LOGRHO, DEB_A_NU_LOGP,
                      DEB_A_NU_LOGPP, DEB_A_OPAC};
  const int ss[11] = {DEB_A_S_TAU_OF_A, DEB_A_S_A_OF_TAU, DEB_A_S_XE_OF_TAU, DEB_A_S_CS2A_OF_TAU, DEB_A_S_TEMPBA_OF_TAU,
                      DEB_A_S_XE_OF_LOGA, DEB_A_S_CS2A_OF_LOGA, DEB_A_NU_LOGRHO_S, DEB_A_NU_LOGP_S, DEB_A_NU_LOGPP_S, DEB_A_S_OPAC};
  const double *X = t + deb_array_offset(xs[sid], N), *Y = t + deb_array_offset(ys[sid], N), *S = t + deb_array_offset(ss[sid], N);
  const bool uni = (sid >= 7 && sid <= 9);
  const int n = uni ? DEB_NNU : N;
  double xq = x[i];
  int j = uni ? spl_index_uniform(t[DEB_HDR_NU_X0], t[DEB_HDR_NU_DXINV], n, xq) : spl_index(X, n, xq);
  double r;
  if (kind == 0) r = spl_eval_at(X, Y, S, j, xq);
  else if (kind == 1) r = spl_deriv_at(X, Y, S, j, xq);
  else if (kind == 2) r = spl_deriv2_at(X, S, j, xq);
  else r = spl_integral_at(X, Y, S, t + deb_array_offset(DEB_A_I_OPAC, N), j, xq);
  out[i] = r;
}

__global__ void power_kernel(const double *__restrict__ k, int krep, const double *__restrict__ y, long n, int ystride, int idx,
                             double A_s, double n_s, double k_p, double *__restrict__ pk) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double kv = k[i / krep];
  double d = y[i * ystride + idx];
  pk[i] = 2.0 * 3.141592653589793 * 3.141592653589793 * A_s * pow(kv / k_p, n_s - 1.0) * pow(kv, -3.0) * d * d;
}

// P(k) of every (unit, output time) row, A_s, n_s and k_p taken from the unit's cosmology
// (the parameter copy at the head of its table block)
__global__ void power_units_kernel(const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                   const double *__restrict__ unit_k, long n_rows, int nout, const double *__restrict__ y, int ystride,
                                   int idx, double *__restrict__ pk) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_rows) return;
  long u = i / nout;
  const double *hdr = tables + (long)unit_cosmo[u] * tstride;
  double kv = unit_k[u];
  double d = y[i * ystride + idx];
  pk[i] = 2.0 * 3.141592653589793 * 3.141592653589793 * hdr[DEB_P_AS] * pow(kv / hdr[DEB_P_KP], hdr[DEB_P_NS] - 1.0) * pow(kv, -3.0) * d *
          d;
}

}  // namespace
}  // namespace deb

extern "C" int deb_launch_power_units(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                                      int nout, const double *d_y, int ystride, int idx, double *d_pk, hipStream_t stream) {
  long rows = (long)n_units * nout;
  if (rows <= 0) return 0;
  hipLaunchKernelGGL(deb::power_units_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, stream, d_tables,
                     deb_table_doubles(num_thermo), d_unit_cosmo, d_unit_k, rows, nout, d_y, ystride, idx, d_pk);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_spline_eval(const double *d_tables, int N, int sid, int kind, const double *d_x, int nx, double *d_out,
                                      hipStream_t stream) {
  if (nx <= 0) return 0;
  hipLaunchKernelGGL(deb::spline_eval_kernel, dim3((nx + 255) / 256), dim3(256), 0, stream, d_tables, N, sid, kind, d_x, nx, d_out);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_power(const double *d_k, int krep, const double *d_y, long n, int ystride, int idx, double A_s, double n_s,
                                double k_p, double *d_pk, hipStream_t stream) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(deb::power_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, d_k, krep, d_y, n, ystride, idx, A_s,
                     n_s, k_p, d_pk);
  return (int)hipGetLastError();
}
