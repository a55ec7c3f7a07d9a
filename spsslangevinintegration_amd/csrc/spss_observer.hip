// spss_observer.hip -- the reference's trajectory and summary outputs and its restart from the
// last saved frame (the spss_frame*, spss_summary*, spss_*_last_frame part of the C ABI).
//
// Frame (BinaryObserverFirstOrderPtr.cpp:108-121, BinaryObserverFirstOrder.cpp:101-122): a float
// time, then the s*n state values as floats.  Models 3 and 4 wrap the float copy: x and y by the
// float overload of PeriodicBoundaryConditions::ApplyPeriodicBoundaryConditions
// (PeriodicBoundaryConditions.cpp:137-145), phi by phi -= floor(phi / 2 pi) * 2 pi; each of these
// runs in double on the float value and is rounded back to float, as the float -= double of the
// reference does.  Model 6 writes the raw float copy (BinaryObserverSecondOrderPtr.cpp:108-113).
//
// Summary (BinaryObserverFirstOrderPtr.cpp:126-146, BinaryObserverSecondOrderPtr.cpp:175-193):
// Z = (1/n) sum_i ((float) cos phi_i + i (float) sin phi_i), accumulated in complex<float> in
// particle order; the line "t \t |Z| \t arg Z" with the default stream formatting.  The phasors
// are computed in parallel and summed in particle order by one lane, so the float sums are the
// reference's; |Z| and arg Z are taken on the host with std::abs / std::arg of complex<float>.
//
// Restart (SimulationEngineFirstOrderPtr.cpp:70-94): the last complete frame of a file, found
// from the file size, its floats widened to double.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/spss_hip.h"
#include "spss_launch.hpp"

using namespace spss;

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

__global__ void ob_frame(const double *__restrict__ state, long m, int s, int wrap, double xs, double ys, double xrs,
                         double yrs, float *__restrict__ out) {
  const long k = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (k >= m) return;
  float v = (float)state[k];
  if (wrap) {
    const int q = (int)(k % s);
    if (q == 0)
      v = (float)((double)v - floor((double)v * xrs) * xs);
    else if (q == 1)
      v = (float)((double)v - floor((double)v * yrs) * ys);
    else if (q == 2)
      v = (float)((double)v - floor((double)v / kTwoPi) * kTwoPi);
  }
  out[k] = v;
}

__global__ void ob_phasors(const double *__restrict__ state, int n, int s, int slot, float2 *__restrict__ ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double phi = state[(long)i * s + slot];
  ph[i] = make_float2((float)cos(phi), (float)sin(phi));
}

// the complex<float> accumulation of the reference, in particle order (one lane; every add
// depends on the previous one)
__global__ void ob_sequential_sum(const float2 *__restrict__ ph, int n, float *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float re = 0.f, im = 0.f;
  for (int i = 0; i < n; ++i) {
    const float2 v = ph[i];
    re = re + v.x;
    im = im + v.y;
  }
  out[0] = re;
  out[1] = im;
}

int check(hipError_t e, const char *what) {
  return e == hipSuccess ? SPSS_OK : set_error(SPSS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define OCHK(call)                                \
  do {                                            \
    int r_ = check((call), #call);                \
    if (r_ != SPSS_OK) return r_;                 \
  } while (0)

// the 1 + s*n floats of a frame into host memory
int frame_to_host(spss_system *sys, double t, float *host) {
  SystemView v{};
  int rc = system_view(sys, &v);
  if (rc != SPSS_OK) return rc;
  const long m = (long)v.s * v.n;
  float *d = nullptr;
  OCHK(hipMalloc(&d, sizeof(float) * (m > 0 ? m : 1)));
  const unsigned nb = (unsigned)((m + 255) / 256);
  if (m > 0)
    hipLaunchKernelGGL(ob_frame, dim3(nb), dim3(256), 0, v.stream, v.state, m, v.s, v.s == 6 ? 0 : 1, v.x_size,
                       v.y_size, v.x_rsize, v.y_rsize, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && m > 0) e = hipMemcpyAsync(host + 1, d, sizeof(float) * m, hipMemcpyDeviceToHost, v.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  hipFree(d);
  OCHK(e);
  host[0] = (float)t;
  return SPSS_OK;
}

int summary_sums(spss_system *sys, float *re, float *im, int *n_out) {
  SystemView v{};
  int rc = system_view(sys, &v);
  if (rc != SPSS_OK) return rc;
  if (v.n < 1) return set_error(SPSS_ERR_INVALID, "no particles");
  const int slot = v.s == 6 ? 4 : 2;
  float2 *ph = nullptr;
  float *out = nullptr;
  OCHK(hipMalloc(&ph, sizeof(float2) * v.n));
  hipError_t e = hipMalloc(&out, 2 * sizeof(float));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(ob_phasors, dim3((v.n + 255) / 256), dim3(256), 0, v.stream, v.state, v.n, v.s, slot, ph);
    hipLaunchKernelGGL(ob_sequential_sum, dim3(1), dim3(64), 0, v.stream, ph, v.n, out);
    e = hipGetLastError();
  }
  float h[2] = {0.f, 0.f};
  if (e == hipSuccess) e = hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, v.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(v.stream);
  hipFree(ph);
  hipFree(out);
  OCHK(e);
  *re = h[0];
  *im = h[1];
  *n_out = v.n;
  return SPSS_OK;
}

}  // namespace

extern "C" {

int spss_frame(spss_system *sys, double t, float *host_frame) {
  if (!sys || !host_frame) return set_error(SPSS_ERR_INVALID, "NULL argument");
  return frame_to_host(sys, t, host_frame);
}

int spss_frame_append(spss_system *sys, const char *path, double t) {
  if (!sys || !path) return set_error(SPSS_ERR_INVALID, "NULL argument");
  SystemView v{};
  int rc = system_view(sys, &v);
  if (rc != SPSS_OK) return rc;
  std::vector<float> buf(1 + (size_t)v.s * v.n);
  if ((rc = frame_to_host(sys, t, buf.data())) != SPSS_OK) return rc;
  FILE *f = std::fopen(path, "ab");
  if (!f) return set_error(SPSS_ERR_INVALID, std::string("cannot open ") + path);
  const size_t w = std::fwrite(buf.data(), sizeof(float), buf.size(), f);
  const int closed = std::fclose(f);
  if (w != buf.size() || closed != 0) return set_error(SPSS_ERR_INVALID, std::string("cannot write ") + path);
  return SPSS_OK;
}

int spss_summary(spss_system *sys, float *abs_z, float *arg_z) {
  if (!sys || !abs_z || !arg_z) return set_error(SPSS_ERR_INVALID, "NULL argument");
  float re = 0.f, im = 0.f;
  int n = 0;
  int rc = summary_sums(sys, &re, &im, &n);
  if (rc != SPSS_OK) return rc;
  std::complex<float> z(re, im);
  z /= (float)n;
  *abs_z = std::abs(z);
  *arg_z = std::arg(z);
  return SPSS_OK;
}

int spss_summary_append(spss_system *sys, const char *path, double t) {
  if (!sys || !path) return set_error(SPSS_ERR_INVALID, "NULL argument");
  float a = 0.f, g = 0.f;
  int rc = spss_summary(sys, &a, &g);
  if (rc != SPSS_OK) return rc;
  FILE *f = std::fopen(path, "a");
  if (!f) return set_error(SPSS_ERR_INVALID, std::string("cannot open ") + path);
  // operator<< of a float with the default stream settings prints it as %g of its double value
  const int w = std::fprintf(f, "%g\t%g\t%g\n", (double)(float)t, (double)a, (double)g);
  const int closed = std::fclose(f);
  if (w < 0 || closed != 0) return set_error(SPSS_ERR_INVALID, std::string("cannot write ") + path);
  return SPSS_OK;
}

int spss_read_last_frame(const char *path, int s, int n, float *t, double *state) {
  if (!path || !t || !state) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (s < 1 || n < 1) return set_error(SPSS_ERR_INVALID, "s and n must be >= 1");
  FILE *f = std::fopen(path, "rb");
  if (!f) return set_error(SPSS_ERR_INVALID, std::string("cannot open ") + path);
  const long frame = (1L + (long)s * n) * (long)sizeof(float);
  long size = -1;
  if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
  const long t_max = size > 0 ? size / frame : 0;
  if (t_max < 1) {
    std::fclose(f);
    return set_error(SPSS_ERR_INVALID, std::string("no complete frame in ") + path);
  }
  std::vector<float> buf((size_t)s * n);
  bool ok = std::fseek(f, frame * (t_max - 1), SEEK_SET) == 0;
  ok = ok && std::fread(t, sizeof(float), 1, f) == 1;
  ok = ok && std::fread(buf.data(), sizeof(float), buf.size(), f) == buf.size();
  std::fclose(f);
  if (!ok) return set_error(SPSS_ERR_INVALID, std::string("cannot read ") + path);
  for (size_t k = 0; k < buf.size(); ++k) state[k] = (double)buf[k];
  return SPSS_OK;
}

int spss_load_last_frame(spss_system *sys, const char *path, double *t) {
  if (!sys || !path) return set_error(SPSS_ERR_INVALID, "NULL argument");
  SystemView v{};
  int rc = system_view(sys, &v);
  if (rc != SPSS_OK) return rc;
  std::vector<double> st((size_t)v.s * v.n);
  float tf = 0.f;
  if ((rc = spss_read_last_frame(path, v.s, v.n, &tf, st.data())) != SPSS_OK) return rc;
  if ((rc = spss_set_state(sys, st.data())) != SPSS_OK) return rc;
  if (t) *t = (double)tf;
  return SPSS_OK;
}

}  // extern "C"
