// spss_mlp.hip -- the network of the learned sweep screen (the spss_mlp_* part of the C ABI).
//
// A fully connected FP32 network: GELU (erf form) after every hidden layer, a linear output
// layer, mean-squared-error loss and Adam.  It maps the parameters of a replica (coupling,
// noise, inertia, frequency spread, size, initial coherence) to its steady-state order
// parameter statistics, so that a parameter sweep can spend full simulations only near the
// transition.  There is no counterpart in the reference.
//
// Kernels (spss_mlp_kernels.hpp):
//   ml_fwd        Z = A W^T + b, H = gelu(Z)
//   ml_mse        loss and dL/dY
//   ml_bwd_weight dW = D^T A, db = sum over the batch of D
//   ml_bwd_data   D_prev = (D W) * gelu'(Z_prev)
//   ml_adam       the Adam update of every parameter
// Parameters are one flat array, layer after layer: W_l [d_{l+1} x d_l] row-major, then b_l.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/spss_hip.h"
#include "spss_device.hpp"
#include "spss_launch.hpp"
#include "spss_mlp_kernels.hpp"

using namespace spss;
using namespace spss::mlp;

namespace {

constexpr int kMaxLayers = 8;
constexpr int kMaxDim = 4096;
constexpr int kLossCap = 1 << 16;

int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace

struct spss_mlp {
  int L = 0;                    // layers
  int dims[kMaxLayers + 1] = {};  // widths, input first
  int max_batch = 0;
  long nparam = 0;
  long woff[kMaxLayers] = {}, boff[kMaxLayers] = {};
  hipStream_t st = nullptr;
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  float *xin = nullptr, *tin = nullptr;  // staging of inputs and targets [max_batch rows]
  float *act[kMaxLayers + 1] = {};       // act[l], l >= 1: gelu(Z_{l-1}) [max_batch x d_l]
  float *z[kMaxLayers] = {};             // Z_l [max_batch x d_{l+1}]
  float *dz[2] = {};                     // gradients w.r.t. Z, ping-pong [max_batch x max width]
  float *data_x = nullptr, *data_t = nullptr;
  long ndata = 0;
  float *loss = nullptr;  // [kLossCap] per-step losses of one train call
  long adam_t = 0;        // Adam steps taken
};

namespace {

#define MCHK(call)                                                                                \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

void mlp_free(spss_mlp *m) {
  float **bufs[] = {&m->p, &m->g, &m->m1, &m->m2, &m->xin, &m->tin, &m->dz[0], &m->dz[1], &m->data_x, &m->data_t,
                    &m->loss};
  for (float **b : bufs)
    if (*b) hipFree(*b), *b = nullptr;
  for (int l = 0; l <= kMaxLayers; ++l)
    if (m->act[l]) hipFree(m->act[l]), m->act[l] = nullptr;
  for (int l = 0; l < kMaxLayers; ++l)
    if (m->z[l]) hipFree(m->z[l]), m->z[l] = nullptr;
  if (m->st) hipStreamDestroy(m->st), m->st = nullptr;
}

// forward pass over B rows of X (rows (x_off + b) mod x_mod); the output is z[L-1]
int forward(spss_mlp *m, const float *X, long x_off, long x_mod, int B) {
  for (int l = 0; l < m->L; ++l) {
    const int in = m->dims[l], out = m->dims[l + 1];
    const float *A = l == 0 ? X : m->act[l];
    const long off = l == 0 ? x_off : 0, mod = l == 0 ? x_mod : B;
    float *H = l + 1 < m->L ? m->act[l + 1] : nullptr;
    hipLaunchKernelGGL(ml_fwd, dim3(cdiv(B, kT), cdiv(out, kT)), dim3(kTB), 0, m->st, A, off, mod, B, in, out,
                       m->p + m->woff[l], m->p + m->boff[l], m->z[l], H);
  }
  MCHK(hipGetLastError());
  return SPSS_OK;
}

// loss of the last forward pass and the gradients of every parameter into m->g
int backward(spss_mlp *m, const float *X, long x_off, long x_mod, const float *T, int B, float *loss_dev) {
  const int L = m->L;
  hipLaunchKernelGGL(ml_mse, dim3(1), dim3(kTB), 0, m->st, m->z[L - 1], T, x_off, x_mod, B, m->dims[L], m->dz[0],
                     loss_dev);
  int cur = 0;
  for (int l = L - 1; l >= 0; --l) {
    const int in = m->dims[l], out = m->dims[l + 1];
    const float *A = l == 0 ? X : m->act[l];
    const long off = l == 0 ? x_off : 0, mod = l == 0 ? x_mod : B;
    hipLaunchKernelGGL(ml_bwd_weight, dim3(cdiv(out, kT), cdiv(in + 1, kT)), dim3(kTB), 0, m->st, m->dz[cur], B, out,
                       A, off, mod, in, m->g + m->woff[l], m->g + m->boff[l]);
    if (l > 0) {
      hipLaunchKernelGGL(ml_bwd_data, dim3(cdiv(B, kT), cdiv(in, kT)), dim3(kTB), 0, m->st, m->dz[cur], B, out, in,
                         m->p + m->woff[l], m->z[l - 1], m->dz[cur ^ 1]);
      cur ^= 1;
    }
  }
  MCHK(hipGetLastError());
  return SPSS_OK;
}

int check_batch(const spss_mlp *m, long B) {
  if (B <= 0 || B > m->max_batch)
    return set_error(SPSS_ERR_INVALID, "batch must be in [1, " + std::to_string(m->max_batch) + "]");
  return SPSS_OK;
}

}  // namespace

extern "C" {

int spss_mlp_create(spss_mlp **out, const int *dims, int ndims, int max_batch, uint64_t seed) {
  if (!out || !dims) return set_error(SPSS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (ndims < 2 || ndims > kMaxLayers + 1) return set_error(SPSS_ERR_INVALID, "ndims must be in [2, 9]");
  if (max_batch < 1 || max_batch > (1 << 20)) return set_error(SPSS_ERR_INVALID, "max_batch must be in [1, 2^20]");
  for (int l = 0; l < ndims; ++l)
    if (dims[l] < 1 || dims[l] > kMaxDim) return set_error(SPSS_ERR_INVALID, "layer widths must be in [1, 4096]");
  spss_mlp *m = new spss_mlp;
  m->L = ndims - 1;
  m->max_batch = max_batch;
  int widest = 0;
  for (int l = 0; l < ndims; ++l) m->dims[l] = dims[l], widest = std::max(widest, dims[l]);
  for (int l = 0; l < m->L; ++l) {
    m->woff[l] = m->nparam;
    m->nparam += (long)dims[l + 1] * dims[l];
    m->boff[l] = m->nparam;
    m->nparam += dims[l + 1];
  }
  // torch.nn.Linear's default initialisation, U(-1/sqrt(in), 1/sqrt(in)) for W and b, from
  // Philox4x32-10 keyed by the seed with counter (parameter index, 0x4D4C5030, 0, 0)
  std::vector<float> init(m->nparam);
  for (int l = 0; l < m->L; ++l) {
    const float bound = 1.f / std::sqrt((float)dims[l]);
    for (long e = m->woff[l]; e < m->boff[l] + dims[l + 1]; ++e) {
      const U4 r = philox4x32_10(U4{(uint32_t)e, 0x4D4C5030u, 0u, 0u}, (uint32_t)seed, (uint32_t)(seed >> 32));
      const float u = (float)(r.x >> 8) * (1.f / 16777216.f);
      init[e] = (2.f * u - 1.f) * bound;
    }
  }
  auto fail = [&](hipError_t e, const char *what) {
    mlp_free(m);
    delete m;
    return set_error(SPSS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e = hipStreamCreate(&m->st);
  if (e != hipSuccess) return fail(e, "hipStreamCreate");
  const size_t pb = m->nparam * sizeof(float);
  for (float **b : {&m->p, &m->g, &m->m1, &m->m2})
    if ((e = hipMalloc(b, pb)) != hipSuccess) return fail(e, "hipMalloc");
  const size_t rows = (size_t)max_batch;
  if ((e = hipMalloc(&m->xin, rows * dims[0] * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&m->tin, rows * dims[m->L] * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
  for (int l = 0; l < m->L; ++l) {
    if ((e = hipMalloc(&m->z[l], rows * dims[l + 1] * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
    if (l + 1 < m->L && (e = hipMalloc(&m->act[l + 1], rows * dims[l + 1] * sizeof(float))) != hipSuccess)
      return fail(e, "hipMalloc");
  }
  for (int k = 0; k < 2; ++k)
    if ((e = hipMalloc(&m->dz[k], rows * widest * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMalloc(&m->loss, kLossCap * sizeof(float))) != hipSuccess) return fail(e, "hipMalloc");
  if ((e = hipMemcpyAsync(m->p, init.data(), pb, hipMemcpyHostToDevice, m->st)) != hipSuccess)
    return fail(e, "hipMemcpyAsync");
  if ((e = hipMemsetAsync(m->g, 0, pb, m->st)) != hipSuccess) return fail(e, "hipMemsetAsync");
  if ((e = hipMemsetAsync(m->m1, 0, pb, m->st)) != hipSuccess) return fail(e, "hipMemsetAsync");
  if ((e = hipMemsetAsync(m->m2, 0, pb, m->st)) != hipSuccess) return fail(e, "hipMemsetAsync");
  if ((e = hipStreamSynchronize(m->st)) != hipSuccess) return fail(e, "hipStreamSynchronize");
  *out = m;
  return SPSS_OK;
}

int spss_mlp_destroy(spss_mlp *m) {
  if (!m) return SPSS_OK;
  hipStreamSynchronize(m->st);
  mlp_free(m);
  delete m;
  return SPSS_OK;
}

int spss_mlp_param_count(const spss_mlp *m, long *count) {
  if (!m || !count) return set_error(SPSS_ERR_INVALID, "NULL argument");
  *count = m->nparam;
  return SPSS_OK;
}

int spss_mlp_set_params(spss_mlp *m, const float *params, long count) {
  if (!m || !params) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (count != m->nparam) return set_error(SPSS_ERR_INVALID, "count must equal the parameter count");
  MCHK(hipMemcpyAsync(m->p, params, count * sizeof(float), hipMemcpyHostToDevice, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return SPSS_OK;
}

int spss_mlp_get_params(spss_mlp *m, float *params, long count) {
  if (!m || !params) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (count != m->nparam) return set_error(SPSS_ERR_INVALID, "count must equal the parameter count");
  MCHK(hipMemcpyAsync(params, m->p, count * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return SPSS_OK;
}

int spss_mlp_reset_optimizer(spss_mlp *m) {
  if (!m) return set_error(SPSS_ERR_INVALID, "NULL argument");
  const size_t pb = m->nparam * sizeof(float);
  MCHK(hipMemsetAsync(m->m1, 0, pb, m->st));
  MCHK(hipMemsetAsync(m->m2, 0, pb, m->st));
  MCHK(hipStreamSynchronize(m->st));
  m->adam_t = 0;
  return SPSS_OK;
}

int spss_mlp_predict(spss_mlp *m, const float *x, long n, float *y) {
  if (!m || (n > 0 && (!x || !y))) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (n < 0) return set_error(SPSS_ERR_INVALID, "n must be >= 0");
  const int din = m->dims[0], dout = m->dims[m->L];
  for (long r0 = 0; r0 < n; r0 += m->max_batch) {
    const int B = (int)std::min<long>(m->max_batch, n - r0);
    MCHK(hipMemcpyAsync(m->xin, x + r0 * din, (size_t)B * din * sizeof(float), hipMemcpyHostToDevice, m->st));
    int rc = forward(m, m->xin, 0, B, B);
    if (rc != SPSS_OK) return rc;
    MCHK(hipMemcpyAsync(y + r0 * dout, m->z[m->L - 1], (size_t)B * dout * sizeof(float), hipMemcpyDeviceToHost,
                        m->st));
  }
  MCHK(hipStreamSynchronize(m->st));
  return SPSS_OK;
}

int spss_mlp_gradients(spss_mlp *m, const float *x, const float *t, int batch, float *grad, float *loss) {
  if (!m || !x || !t || !grad) return set_error(SPSS_ERR_INVALID, "NULL argument");
  int rc = check_batch(m, batch);
  if (rc != SPSS_OK) return rc;
  const int din = m->dims[0], dout = m->dims[m->L];
  MCHK(hipMemcpyAsync(m->xin, x, (size_t)batch * din * sizeof(float), hipMemcpyHostToDevice, m->st));
  MCHK(hipMemcpyAsync(m->tin, t, (size_t)batch * dout * sizeof(float), hipMemcpyHostToDevice, m->st));
  if ((rc = forward(m, m->xin, 0, batch, batch)) != SPSS_OK) return rc;
  if ((rc = backward(m, m->xin, 0, batch, m->tin, batch, m->loss)) != SPSS_OK) return rc;
  MCHK(hipMemcpyAsync(grad, m->g, m->nparam * sizeof(float), hipMemcpyDeviceToHost, m->st));
  float l = 0.f;
  MCHK(hipMemcpyAsync(&l, m->loss, sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  if (loss) *loss = l;
  return SPSS_OK;
}

int spss_mlp_set_data(spss_mlp *m, const float *x, const float *t, long n) {
  if (!m || !x || !t) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (n < 1) return set_error(SPSS_ERR_INVALID, "n must be >= 1");
  if (m->data_x) hipFree(m->data_x), m->data_x = nullptr;
  if (m->data_t) hipFree(m->data_t), m->data_t = nullptr;
  m->ndata = 0;
  const int din = m->dims[0], dout = m->dims[m->L];
  MCHK(hipMalloc(&m->data_x, (size_t)n * din * sizeof(float)));
  MCHK(hipMalloc(&m->data_t, (size_t)n * dout * sizeof(float)));
  MCHK(hipMemcpyAsync(m->data_x, x, (size_t)n * din * sizeof(float), hipMemcpyHostToDevice, m->st));
  MCHK(hipMemcpyAsync(m->data_t, t, (size_t)n * dout * sizeof(float), hipMemcpyHostToDevice, m->st));
  MCHK(hipStreamSynchronize(m->st));
  m->ndata = n;
  return SPSS_OK;
}

int spss_mlp_train(spss_mlp *m, int nsteps, int batch, float lr, float beta1, float beta2, float eps,
                   float *losses) {
  if (!m) return set_error(SPSS_ERR_INVALID, "NULL argument");
  if (!m->data_x) return set_error(SPSS_ERR_INVALID, "no training data (spss_mlp_set_data)");
  int rc = check_batch(m, batch);
  if (rc != SPSS_OK) return rc;
  if (batch > m->ndata) return set_error(SPSS_ERR_INVALID, "batch exceeds the number of training rows");
  if (nsteps < 0 || nsteps > kLossCap) return set_error(SPSS_ERR_INVALID, "nsteps must be in [0, 65536]");
  if (!(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) || !(beta2 >= 0.f && beta2 < 1.f) || !(eps > 0.f))
    return set_error(SPSS_ERR_INVALID, "bad Adam hyper-parameters");
  const int adam_grid = std::min(cdiv(m->nparam, kTB), 1024);
  for (int s = 0; s < nsteps; ++s) {
    // mini-batch: the next `batch` rows of the resident set, wrapping around
    const long off = (m->adam_t * (long)batch) % m->ndata;
    if ((rc = forward(m, m->data_x, off, m->ndata, batch)) != SPSS_OK) return rc;
    if ((rc = backward(m, m->data_x, off, m->ndata, m->data_t, batch, m->loss + s)) != SPSS_OK) return rc;
    ++m->adam_t;
    const double bc1 = 1.0 - std::pow((double)beta1, (double)m->adam_t);
    const double bc2 = 1.0 - std::pow((double)beta2, (double)m->adam_t);
    hipLaunchKernelGGL(ml_adam, dim3(adam_grid), dim3(kTB), 0, m->st, m->p, m->g, m->m1, m->m2, m->nparam, beta1,
                       beta2, (float)(lr / bc1), (float)(1.0 / std::sqrt(bc2)), eps);
  }
  MCHK(hipGetLastError());
  if (losses && nsteps > 0)
    MCHK(hipMemcpyAsync(losses, m->loss, nsteps * sizeof(float), hipMemcpyDeviceToHost, m->st));
  MCHK(hipStreamSynchronize(m->st));
  return SPSS_OK;
}

int spss_mlp_adam_steps(const spss_mlp *m, long *steps) {
  if (!m || !steps) return set_error(SPSS_ERR_INVALID, "NULL argument");
  *steps = m->adam_t;
  return SPSS_OK;
}

}  // extern "C"
