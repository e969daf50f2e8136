// deb_sources.hip -- the CMB chain perturbations -> time derivatives -> source function in one
// pass.  The per-mode and the batched perturbation kernels run with the source epilogue
// (solve_unit / solve_batch with SRC = true): the outputs of each accepted step are formed one
// per thread -- the interpolated state projected (RSA), the right-hand side rows the source terms
// read evaluated there, the source terms formed from the two -- so that only the (k, tau) source
// planes reach HBM -- never the states and their
// derivatives, 2 x nvar doubles per point, that the staged chain of the reference
// (cmb.py:1030-1073: evolve_perturbations_batched with return_full, compute_time_derivatives,
// compute_source_function) materialises.
#include "deb_perturb_dev.hpp"

namespace deb {
namespace {

template <int NQ, int NT>
__global__ __launch_bounds__(NT) void perturb_sources_kernel(const double *__restrict__ tables, long tstride, const int *__restrict__ unit_cosmo,
                                                               const double *__restrict__ unit_k, int n_units, const double *__restrict__ tau_out,
                                                               int nout, const double *__restrict__ nu_qw, DebProblem prob, DebSolverCfg cfg,
                                                               long lds_bytes, SrcOut so, int *__restrict__ status, int *__restrict__ n_acc,
                                                               int *__restrict__ n_rej, const int *__restrict__ order, double *__restrict__ step_log,
                                                               const double *__restrict__ replay, long replay_stride,
                                                               const int *__restrict__ replay_len) {
  if ((int)blockIdx.x >= n_units) return;
  const int tid = threadIdx.x;
  PLay L;
  Ws W;
  if (!setup_solver<NQ, NT>(prob, nu_qw, lds_bytes, L, W, tid)) {  // launch/carve-up mismatch: refuse to run
    for (int u = blockIdx.x * NT + tid; u < n_units; u += gridDim.x * NT) { status[u] = 3; n_acc[u] = 0; n_rej[u] = 0; }
    return;
  }
  const int unit = order ? order[blockIdx.x] : (int)blockIdx.x;
  solve_unit<NQ, NT, true>(unit, tables, tstride, unit_cosmo, unit_k, tau_out, nout, prob, cfg, step_log, nullptr, status, n_acc, n_rej, replay,
                           replay_stride, replay_len, L, W, tid, &so);
}

template <int NQ, int NT>
__global__ __launch_bounds__(NT) void perturb_batched_sources_kernel(const double *__restrict__ tables, long tstride,
                                                                       const int *__restrict__ unit_cosmo, const double *__restrict__ unit_k,
                                                                       int n_units, int batch, const double *__restrict__ tau_out, int nout,
                                                                       const double *__restrict__ nu_qw, DebProblem prob, DebSolverCfg cfg,
                                                                       long lds_bytes, SrcOut so, int *__restrict__ status,
                                                                       int *__restrict__ n_acc, int *__restrict__ n_rej,
                                                                       double *__restrict__ state, double *__restrict__ step_log) {
  solve_batch<NQ, NT, true>(tables, tstride, unit_cosmo, unit_k, n_units, batch, tau_out, nout, nu_qw, prob, cfg, lds_bytes, step_log, nullptr,
                            status, n_acc, n_rej, state, &so);
}

__global__ void fill_int_kernel(int *__restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

typedef decltype(&perturb_sources_kernel<0, WAVE>) SrcFn;
typedef decltype(&perturb_batched_sources_kernel<0, WAVE>) SrcBatchedFn;

// the instantiations of the state-output kernels (compile-time bin counts 3, 5 and 8, generic
// otherwise), so that both forms take the same linear-algebra route and round alike
template <int NQ> SrcFn pick_src_nt(int nt) {
  return nt == 256 ? perturb_sources_kernel<NQ, 256> : (nt == 128 ? perturb_sources_kernel<NQ, 128> : perturb_sources_kernel<NQ, 64>);
}
template <int NQ> SrcBatchedFn pick_src_batched_nt(int nt) {
  return nt == 256 ? perturb_batched_sources_kernel<NQ, 256>
                   : (nt == 128 ? perturb_batched_sources_kernel<NQ, 128> : perturb_batched_sources_kernel<NQ, 64>);
}
SrcFn pick_src(int nq, int nt) {
  switch (nq) {
    case 3: return pick_src_nt<3>(nt);
    case 5: return pick_src_nt<5>(nt);
    case 8: return pick_src_nt<8>(nt);
    default: return pick_src_nt<0>(nt);
  }
}
SrcBatchedFn pick_src_batched(int nq, int nt) {
  switch (nq) {
    case 3: return pick_src_batched_nt<3>(nt);
    case 5: return pick_src_batched_nt<5>(nt);
    case 8: return pick_src_batched_nt<8>(nt);
    default: return pick_src_batched_nt<0>(nt);
  }
}

}  // namespace
}  // namespace deb

extern "C" int deb_launch_nu_bins(int nqmax, double *d_qw, hipStream_t stream);
extern "C" int deb_launch_visibility(const double *d_t, int num_thermo, const double *d_tau, const double *d_aexp, int n_tau, double *d_V,
                                     double *d_work, hipStream_t stream);
extern "C" int deb_launch_cmb_isw_finish(const double *d_tau, int n_tau, int n_k, const double *d_pot, const double *d_V, double *d_out,
                                         double *d_work, hipStream_t stream);

// workspace (doubles): momentum bins 32 | visibility build 12 N | visibility table NV n_tau |
// ISW potential n_tau n_k | ISW spline n_tau n_k + n_tau | unit cosmology indices (n_k ints) |
// batched states 2 nvar n_k (batch > 0)
extern "C" long deb_evolve_sources_work_doubles(int num_thermo, int n_k, int n_tau, int batch, int nvar) {
  long w = 32 + 12L * num_thermo + (long)deb::NV * n_tau + 2L * n_tau * n_k + n_tau + (n_k + 1) / 2 + 8;
  if (batch > 0) w += 2L * nvar * n_k;
  return w;
}

// d_t: the tables of the one cosmology every unit belongs to; d_tau / d_aexp: its n_tau output
// times and scale factors.  batch 0: per-mode adaptive steps (d_order, threads_per_unit, the step
// log and the replay mesh as deb_launch_perturb); batch > 0: batched shared steps (n_k a multiple
// of batch; a step log per batch, no replay).
extern "C" int deb_launch_evolve_sources(const double *d_t, int num_thermo, const double *d_k, int n_k, int batch, const double *d_aexp,
                                         const double *d_tau, int n_tau, const DebProblem *prob, const DebSolverCfg *cfg, int isw,
                                         double *d_out, int *d_status, int *d_nacc, int *d_nrej, double *d_work, const int *d_order,
                                         int threads_per_unit, double *d_step_log, const double *d_replay, long replay_stride,
                                         const int *d_replay_len, hipStream_t stream) {
  const long lds = deb::deb_perturb_lds_bytes_dev(*prob);
  if (lds > 160 * 1024) return -2;
  if (batch < 0 || (batch > 0 && n_k % batch != 0)) return -3;
  if (batch > 0 && d_replay != nullptr) return -5;
  const int nt = threads_per_unit > 0 ? threads_per_unit : DEB_DEFAULT_THREADS_PER_UNIT;
  if (nt != 64 && nt != 128 && nt != 256) return -4;
  double *qw = d_work, *vw = qw + 32, *V = vw + 12L * num_thermo, *pot = V + (long)deb::NV * n_tau;
  double *iswk = pot + (long)n_tau * n_k, *ic = iswk + (long)n_tau * n_k + n_tau;
  int *unit_cosmo = (int *)ic;
  double *state = ic + (n_k + 1) / 2 + 8;
  int rc = deb_launch_nu_bins(prob->nqmax, qw, stream);
  if (rc) return rc;
  rc = deb_launch_visibility(d_t, num_thermo, d_tau, d_aexp, n_tau, V, vw, stream);
  if (rc) return rc;
  hipLaunchKernelGGL(deb::fill_int_kernel, dim3((n_k + 255) / 256), dim3(256), 0, stream, unit_cosmo, n_k, 0);
  deb::SrcOut so;
  so.V = V;
  so.aexp = d_aexp;
  so.qw = qw;
  so.planes = d_out;
  so.pot = pot;
  so.isw = isw;
  so.n_units = n_k;
  DebProblem pr = *prob;
  pr.return_full = 1;
  const long stride = deb_table_doubles(num_thermo);
  if (batch == 0) {
    deb::SrcFn kern = deb::pick_src(prob->nqmax, nt);
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(kern, dim3(n_k), dim3(nt), (size_t)lds, stream, d_t, stride, unit_cosmo, d_k, n_k, d_tau, n_tau, qw, pr, *cfg, lds, so,
                       d_status, d_nacc, d_nrej, d_order, d_step_log, d_replay, replay_stride, d_replay_len);
  } else {
    deb::SrcBatchedFn kern = deb::pick_src_batched(prob->nqmax, nt);
    if (lds > 64 * 1024) {
      hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return (int)e;
    }
    hipLaunchKernelGGL(kern, dim3(n_k / batch), dim3(nt), (size_t)lds, stream, d_t, stride, unit_cosmo, d_k, n_k, batch, d_tau, n_tau, qw, pr,
                       *cfg, lds, so, d_status, d_nacc, d_nrej, state, d_step_log);
  }
  rc = (int)hipGetLastError();
  if (rc) return rc;
  if (isw == 0) rc = deb_launch_cmb_isw_finish(d_tau, n_tau, n_k, pot, V, d_out, iswk, stream);
  return rc;
}
