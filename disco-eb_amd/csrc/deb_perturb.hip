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
#include "deb_perturb_dev.hpp"

namespace deb {
namespace {

// neutrino momentum bins (background.py:14-45): CAMB tables for 3-5 bins, else generalized
// Gauss-Laguerre (alpha = 1) reweighted to the Fermi-Dirac integrand; qw[0..nq) nodes,
// qw[NQMAX..NQMAX+nq) weights
// Nodes and weights already computed on this device, per bin count (each device holds its own
// copy).  Two launches that compute the same count at once write the same values.
__device__ double g_nu_cache[NQMAX + 1][2 * NQMAX];
__device__ int g_nu_ready[NQMAX + 1];

__global__ void nu_bins_kernel(int nq, double *qw) {
  if (g_nu_ready[nq]) {  // computed by an earlier launch: copy
    for (int i = threadIdx.x; i < 2 * NQMAX; i += blockDim.x) qw[i] = g_nu_cache[nq][i];
    return;
  }
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double fdc = 5.682196976983475;
  if (nq >= 3 && nq <= 5) {
    const double q3[] = {0.913201, 3.37517, 7.79184}, w3[] = {0.0687359, 3.31435, 2.29911};
    const double q4[] = {0.7, 2.62814, 5.90428, 12.0}, w4[] = {0.0200251, 1.84539, 3.52736, 0.289427};
    const double q5[] = {0.583165, 2.0, 4.0, 7.26582, 13.0}, w5[] = {0.0081201, 0.689407, 2.8063, 2.05156, 0.12681};
    const double *qq = nq == 3 ? q3 : nq == 4 ? q4 : q5;
    const double *ww = nq == 3 ? w3 : nq == 4 ? w4 : w5;
    for (int i = 0; i < nq; ++i) {
      double dl = -qq[i] / (1.0 + exp(-qq[i]));
      qw[i] = qq[i];
      qw[NQMAX + i] = ww[i] / (-0.25 * dl) / fdc;
    }
  } else {
    gauss_laguerre_fd(nq, qw, qw + NQMAX);
  }
  for (int i = 0; i < nq; ++i) {
    g_nu_cache[nq][i] = qw[i];
    g_nu_cache[nq][NQMAX + i] = qw[NQMAX + i];
  }
  __threadfence();
  g_nu_ready[nq] = 1;
}

// Work distribution: workgroup g solves unit order[g] (g when order is NULL); the grid covers
// n_units.  The dispatcher starts workgroups in index order as slots free up, so an order that
// lists the most expensive units first keeps the tail short.
template <int NQ, int NT>
__global__ __launch_bounds__(NT) void perturb_kernel(const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                                       const double *__restrict__ unit_k, int n_units, const double *__restrict__ tau_out,
                                                       int nout, const double *__restrict__ nu_qw, DebProblem prob, DebSolverCfg cfg,
                                                       long lds_bytes, double *__restrict__ step_log, double *__restrict__ y_out,
                                                       int *__restrict__ status, int *__restrict__ n_acc, int *__restrict__ n_rej,
                                                       const int *__restrict__ order, const double *__restrict__ replay,
                                                       long replay_stride, const int *__restrict__ replay_len) {
  if ((int)blockIdx.x >= n_units) return;
  const int tid = threadIdx.x;
  PLay L;
  Ws W;
  if (!setup_solver<NQ, NT>(prob, nu_qw, lds_bytes, L, W, tid)) {  // launch/carve-up mismatch: refuse to run
    for (int u = blockIdx.x * NT + tid; u < n_units; u += gridDim.x * NT) { status[u] = 3; n_acc[u] = 0; n_rej[u] = 0; }
    return;
  }
  const int unit = order ? order[blockIdx.x] : (int)blockIdx.x;
  solve_unit<NQ, NT>(unit, tables, tstride, unit_cosmo, unit_k, tau_out, nout, prob, cfg, step_log, y_out, status, n_acc, n_rej, replay,
                 replay_stride, replay_len, L, W, tid);
}


// Batched shared-step evolution: one workgroup per batch (solve_batch).
template <int NQ, int NT>
__global__ __launch_bounds__(NT) void perturb_batched_kernel(const double *__restrict__ tables, long tstride,
                                                               const int *__restrict__ unit_cosmo, const double *__restrict__ unit_k,
                                                               int n_units, int batch, const double *__restrict__ tau_out, int nout,
                                                               const double *__restrict__ nu_qw, DebProblem prob, DebSolverCfg cfg,
                                                               long lds_bytes, double *__restrict__ step_log, double *__restrict__ y_out,
                                                               int *__restrict__ status, int *__restrict__ n_acc, int *__restrict__ n_rej,
                                                               double *__restrict__ state) {
  solve_batch<NQ, NT>(tables, tstride, unit_cosmo, unit_k, n_units, batch, tau_out, nout, nu_qw, prob, cfg, lds_bytes, step_log, y_out, status,
                      n_acc, n_rej, state);
}


// dy/dtau at saved states (compute_time_derivatives, perturbations.py:1339-1406): one wave per
// k-mode, looping over the output times
__global__ __launch_bounds__(WAVE) void time_deriv_kernel(const double *__restrict__ tables, const double *__restrict__ yin,
                                                          const double *__restrict__ tau, int n_tau, const double *__restrict__ kk,
                                                          const double *__restrict__ nu_qw, DebProblem prob, double *__restrict__ dy) {
  const int ik = blockIdx.x, lane = threadIdx.x;
  const double k = kk[ik];
  BgView b;
  b.init(tables);
  PLay L;
  L.lmaxg = prob.lmaxg; L.lmaxgp = prob.lmaxgp; L.lmaxr = prob.lmaxr; L.lmaxnu = prob.lmaxnu; L.nq = prob.nqmax;
  L.n = 7 + (L.lmaxg + 1) + (L.lmaxgp + 1) + (L.lmaxr + 1) + L.nq * (L.lmaxnu + 1) + 2;
  L.idxg = 7; L.idxgp = 7 + (L.lmaxg + 1); L.idxr = 9 + L.lmaxg + L.lmaxgp; L.iq0 = 10 + L.lmaxg + L.lmaxgp + L.lmaxr;
  L.m = 17 + 3 * L.nq;
  L.nch = 3 + L.nq;
  const int n = L.n, m = L.m, nch = L.nch;
  extern __shared__ double smem[];
  Ws W;
  double *p = smem;
  W.u = p; p += n; W.f = p; p += n; W.tca = p; p += n; W.tcb = p; p += n;
  double *sq = p; p += NQMAX; double *sw = p; p += NQMAX; double *sdl = p; p += NQMAX;
  for (int q = lane; q < L.nq; q += WAVE) {
    sq[q] = nu_qw[q];
    sw[q] = nu_qw[NQMAX + q];
    sdl[q] = -nu_qw[q] / (1.0 + exp(-nu_qw[q]));
  }
  L.q = sq; L.w = sw; L.dlfdlq = sdl;
  int *ip = (int *)p;
  W.dmap = ip; ip += m; W.rk = ip; ip += n; W.chead = ip; ip += nch; W.cstride = ip; ip += nch; W.clen = ip; ip += nch;
  W.cpar = ip; ip += nch; W.rs = ip; ip += n; W.rbin = ip; ip += n; W.perm = ip; ip += n + 256;
  ip += (2 - ((long)((char *)ip - (char *)smem) / 4) % 2) % 2;
  W.sc = (SplCache *)ip;
  if (lane == 0) { W.sc->i = -1; W.sc->j = -1; }
  setup_maps<WAVE>(L, W, lane);
  for (int it = 0; it < n_tau; ++it) {
    const double *src = yin + ((long)ik * n_tau + it) * n;
    for (int i = lane; i < n; i += WAVE) W.u[i] = src[i];
    __syncthreads();
    rhs_eval<WAVE>(b, L, k, tau[it], W.u, W.f, W, lane);
    __syncthreads();
    double *dst = dy + ((long)ik * n_tau + it) * n;
    for (int i = lane; i < n; i += WAVE) dst[i] = W.f[i];
    __syncthreads();
  }
}

// tau_out[c, j] = tau_of_a(aexp_out[j]) for every cosmology
__global__ void tau_out_kernel(const double *__restrict__ tables, long tstride, int n_cosmo, const double *__restrict__ aexp_out,
                               int nout, double *__restrict__ tau_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cosmo * nout) return;
  int c = i / nout, j = i - c * nout;
  BgView b;
  b.init(tables + (long)c * tstride);
  tau_out[i] = b.tau_of_a(aexp_out[j]);
}

}  // namespace
}  // namespace deb

extern "C" long deb_perturb_lds_bytes(int lmaxg, int lmaxgp, int lmaxr, int lmaxnu, int nqmax) {
  const DebProblem p{lmaxg, lmaxgp, lmaxr, lmaxnu, nqmax, 0};
  return deb::deb_perturb_lds_bytes_dev(p);
}

extern "C" int deb_launch_nu_bins(int nqmax, double *d_qw, hipStream_t stream) {
  if (nqmax < 1 || nqmax > deb::NQMAX) return -1;
  hipLaunchKernelGGL(deb::nu_bins_kernel, dim3(1), dim3(64), 0, stream, nqmax, d_qw);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_time_deriv(const double *d_tables, const double *d_y, const double *d_tau, int n_tau, const double *d_k,
                                     int n_k, const double *d_nu_qw, const DebProblem *prob, double *d_dy, hipStream_t stream) {
  long n = 7 + (prob->lmaxg + 1) + (prob->lmaxgp + 1) + (prob->lmaxr + 1) + (long)prob->nqmax * (prob->lmaxnu + 1) + 2;
  long m = 17 + 3 * (long)prob->nqmax, nch = 3 + prob->nqmax;
  long lds = (4 * n + 3 * deb::NQMAX) * 8 + (m + 4 * n + 256 + 4 * nch + 1) * 4 + (long)sizeof(deb::SplCache);
  hipLaunchKernelGGL(deb::time_deriv_kernel, dim3(n_k), dim3(deb::WAVE), (size_t)lds, stream, d_tables, d_y, d_tau, n_tau, d_k, d_nu_qw,
                     *prob, d_dy);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_tau_out(const double *d_tables, int num_thermo, int n_cosmo, const double *d_aexp_out, int nout,
                                  double *d_tau_out, hipStream_t stream) {
  long stride = deb_table_doubles(num_thermo);
  int total = n_cosmo * nout;
  hipLaunchKernelGGL(deb::tau_out_kernel, dim3((total + 255) / 256), dim3(256), 0, stream, d_tables, stride, n_cosmo, d_aexp_out, nout,
                     d_tau_out);
  return (int)hipGetLastError();
}

namespace {
typedef decltype(&deb::perturb_kernel<0, deb::WAVE>) PerturbFn;
typedef decltype(&deb::perturb_batched_kernel<0, deb::WAVE>) PerturbBatchedFn;

// Instantiation for nq momentum bins (3, 5 and 8 at compile time, others generic) and
// threads_per_unit threads (64, 128 or 256) per unit.
template <int NQ> PerturbFn pick_nt(int nt) {
  return nt == 256 ? deb::perturb_kernel<NQ, 256> : (nt == 128 ? deb::perturb_kernel<NQ, 128> : deb::perturb_kernel<NQ, 64>);
}
template <int NQ> PerturbBatchedFn pick_nt_batched(int nt) {
  return nt == 256 ? deb::perturb_batched_kernel<NQ, 256>
                   : (nt == 128 ? deb::perturb_batched_kernel<NQ, 128> : deb::perturb_batched_kernel<NQ, 64>);
}
PerturbFn pick_perturb(int nq, int nt) {
  switch (nq) {
    case 3: return pick_nt<3>(nt);
    case 5: return pick_nt<5>(nt);
    case 8: return pick_nt<8>(nt);
    default: return pick_nt<0>(nt);
  }
}
PerturbBatchedFn pick_perturb_batched(int nq, int nt) {
  switch (nq) {
    case 3: return pick_nt_batched<3>(nt);
    case 5: return pick_nt_batched<5>(nt);
    case 8: return pick_nt_batched<8>(nt);
    default: return pick_nt_batched<0>(nt);
  }
}
}  // namespace

extern "C" int deb_launch_perturb(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k, int n_units,
                                  const double *d_tau_out, int nout, const double *d_nu_qw, const DebProblem *prob,
                                  const DebSolverCfg *cfg, double *d_step_log, double *d_y_out, int *d_status, int *d_nacc, int *d_nrej,
                                  const int *d_order, int threads_per_unit, const double *d_replay, long replay_stride,
                                  const int *d_replay_len, hipStream_t stream) {
  long stride = deb_table_doubles(num_thermo);
  long lds = deb_perturb_lds_bytes(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax);
  if (lds > 160 * 1024) return -2;
  const int nt = threads_per_unit > 0 ? threads_per_unit : DEB_DEFAULT_THREADS_PER_UNIT;
  if (nt != 64 && nt != 128 && nt != 256) return -4;
  PerturbFn kern = pick_perturb(prob->nqmax, nt);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
  }
  hipLaunchKernelGGL(kern, dim3(n_units), dim3(nt), (size_t)lds, stream, d_tables, stride, d_unit_cosmo, d_unit_k, n_units, d_tau_out,
                     nout, d_nu_qw, *prob, *cfg, lds, d_step_log, d_y_out, d_status, d_nacc, d_nrej, d_order, d_replay, replay_stride,
                     d_replay_len);
  return (int)hipGetLastError();
}

extern "C" int deb_launch_perturb_batched(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k,
                                          int n_units, int batch, const double *d_tau_out, int nout, const double *d_nu_qw,
                                          const DebProblem *prob, const DebSolverCfg *cfg, double *d_step_log, double *d_y_out,
                                          int *d_status, int *d_nacc, int *d_nrej, double *d_state, int threads_per_unit,
                                          hipStream_t stream) {
  long stride = deb_table_doubles(num_thermo);
  long lds = deb_perturb_lds_bytes(prob->lmaxg, prob->lmaxgp, prob->lmaxr, prob->lmaxnu, prob->nqmax);
  if (lds > 160 * 1024) return -2;
  if (batch <= 0 || n_units % batch != 0) return -3;
  const int nt = threads_per_unit > 0 ? threads_per_unit : DEB_DEFAULT_THREADS_PER_UNIT;
  if (nt != 64 && nt != 128 && nt != 256) return -4;
  PerturbBatchedFn kern = pick_perturb_batched(prob->nqmax, nt);
  if (lds > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return (int)e;
  }
  hipLaunchKernelGGL(kern, dim3(n_units / batch), dim3(nt), (size_t)lds, stream, d_tables, stride, d_unit_cosmo, d_unit_k, n_units,
                     batch, d_tau_out, nout, d_nu_qw, *prob, *cfg, lds, d_step_log, d_y_out, d_status, d_nacc, d_nrej, d_state);
  return (int)hipGetLastError();
}
