// deb_evolve_batched without a workspace argument and without any device allocation: the
// momentum bins and the two state vectors per unit live in a file-scope device array.
//
// Each device holds its own copy of the array (DEB_NOALLOC_SCRATCH_DOUBLES doubles, 8 MiB)
// until the process ends.  Its address on a device is looked up once, with that device current,
// on the first call there, and kept per device: the array's name in host code is a host-side
// stand-in and is never passed to a kernel.  Every call on a device shares that device's array,
// so calls that overlap on one device (two streams, or two host threads) race; calls ordered on
// one stream do not.  The batched kernel writes every word of its state vectors (initial
// conditions) and the momentum-bin kernel every word of its bins before they are read, so
// nothing is cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/discoeb_amd.h"

namespace {

constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[DEB_NOALLOC_SCRATCH_DOUBLES];

std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// The device address of this device's scratch array (nullptr and an error code on failure).
int scratch_on_current_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

namespace deb {
int capi_fail(int code, const char *msg);
}

extern "C" long deb_noalloc_scratch_doubles(void) { return DEB_NOALLOC_SCRATCH_DOUBLES; }

extern "C" int deb_evolve_batched_noalloc(const double *d_tables, int num_thermo, const int *d_unit_cosmo, const double *d_unit_k,
                                          int n_units, int batch_size, const double *d_tau_out, int nout, const DebProblem *prob,
                                          const DebSolverCfg *cfg, double *d_y_out, int *d_status, int *d_n_accept, int *d_n_reject,
                                          double *d_step_log, const DebSchedule *sched, void *stream) {
  if (!prob) return deb::capi_fail(-1, "problem descriptor is NULL");
  if (n_units < 0) return deb::capi_fail(-1, "n_units must be >= 0");
  const long need = deb_evolve_batched_workspace_doubles(n_units, prob);  // an invalid problem is rejected below
  if (need > DEB_NOALLOC_SCRATCH_DOUBLES)
    return deb::capi_fail(-1, "deb_evolve_batched_noalloc: n_units x 2 nvar + 32 doubles exceed the static scratch of "
                              "deb_noalloc_scratch_doubles() doubles; use deb_evolve_batched with a workspace");
  double *ws = nullptr;
  int rc = scratch_on_current_device(&ws);
  if (rc) return deb::capi_fail(rc, "deb_evolve_batched_noalloc: no device address for the static scratch on the current device");
  return deb_evolve_batched(d_tables, num_thermo, d_unit_cosmo, d_unit_k, n_units, batch_size, d_tau_out, nout, prob, cfg, d_y_out,
                            d_status, d_n_accept, d_n_reject, ws, d_step_log, sched, stream);
}
