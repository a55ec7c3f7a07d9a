// spss_noalloc.hip -- an entry point that allocates nothing, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc): the order parameter of a device state array, with the block partials of its
// reduction in a fixed device array of this file instead of a handle's scratch.
//
// The array `scratch` is one per GPU and lives until the process ends.  Its device address is
// looked up once per device (hipGetSymbolAddress, with that device current) on the first call
// made on that device, and only that address is handed to kernels: the array's name in host code
// is a host-side stand-in, not a device address.  Every call on a GPU shares that GPU's array, so
// calls that overlap on one GPU (two streams, two threads) race; calls in one stream are ordered.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/spss_noalloc.h"
#include "spss_launch.hpp"

using namespace spss;

namespace {

// block partials of the order-parameter reduction: 4 doubles (two sums and their compensations)
// per workgroup of kBlock particles
constexpr long kScratchBlocks = SPSS_NOALLOC_MAX_PARTICLES / kBlock;
static_assert(SPSS_NOALLOC_MAX_PARTICLES % kBlock == 0, "whole workgroups");

__device__ __align__(256) double scratch[4 * kScratchBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device, once looked up

// the device address of this device's `scratch` (the device current on entry)
int scratch_of_current_device(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return set_error(SPSS_ERR_UNSUPPORTED, "device index beyond 64");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(SPSS_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return SPSS_OK;
}

}  // namespace

extern "C" {

int spss_order_parameter_noalloc(const double *d_state, int n, int s, int slot, double *d_out, void *hip_stream) {
  if (!d_state || !d_out) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (n <= 0) return set_error(SPSS_ERR_INVALID, "n must be positive");
  if (s != 3 && s != 4 && s != 6) return set_error(SPSS_ERR_INVALID, "s must be 3, 4 or 6");
  if (slot < 0 || slot >= s) return set_error(SPSS_ERR_INVALID, "slot out of range");
  if ((long)n > SPSS_NOALLOC_MAX_PARTICLES)
    return set_error(SPSS_ERR_CAPACITY, "spss_order_parameter_noalloc takes at most SPSS_NOALLOC_MAX_PARTICLES "
                                        "(16,777,216) particles: its fixed scratch holds that many block partials");
  double *partial = nullptr;
  const int rc = scratch_of_current_device(&partial);
  if (rc != SPSS_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(hip_stream);
  // k_order_partial writes every partial it uses before k_order_final reads it: nothing to clear
  launch_order_parameter(st, n, s, slot, d_state, partial, d_out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string("order-parameter launch: ") + hipGetErrorString(e));
  return SPSS_OK;
}

}  // extern "C"
