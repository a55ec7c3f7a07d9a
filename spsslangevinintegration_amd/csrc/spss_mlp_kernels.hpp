// spss_mlp_kernels.hpp -- device code of the sweep-screen network (included by spss_mlp.hip).
//
// One launch per layer and direction, 256 threads, 32 x 32 output tiles; reductions are staged
// through LDS in chunks of 64 and summed in ascending index order with FMAs, without atomics,
// so that every result is reproducible bit for bit.  Matrices are row-major FP32.
#pragma once

#include <hip/hip_runtime.h>

namespace spss {
namespace mlp {

constexpr int kT = 32;    // edge of an output tile
constexpr int kK = 64;    // reduction chunk staged in LDS
constexpr int kTB = 256;  // threads per block: 32 columns x 8 row groups, 4 rows each

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }

// d/dz [z Phi(z)] = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu_grad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return __builtin_fmaf(z, pdf, cdf);
}

// Z[b, o] = sum_i A[row(b), i] W[o, i] + bias[o]; H[b, o] = gelu(Z[b, o]) when H is given.
// row(b) = (a_off + b) mod a_mod: a ring over the resident training set for the input layer.
__global__ __launch_bounds__(kTB) void ml_fwd(const float *A, long a_off, long a_mod, int B, int in, int out,
                                              const float *W, const float *bias, float *Z, float *H) {
  __shared__ float sa[kT][kK + 1];
  __shared__ float sw[kT][kK + 1];
  const int b0 = blockIdx.x * kT, o0 = blockIdx.y * kT;
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int i0 = 0; i0 < in; i0 += kK) {
    const int kk = min(kK, in - i0);
    for (int e = threadIdx.x; e < kT * kK; e += kTB) {
      const int r = e / kK, c = e % kK;
      float va = 0.f, vw = 0.f;
      if (c < kk) {
        if (b0 + r < B) va = A[((a_off + b0 + r) % a_mod) * in + i0 + c];
        if (o0 + r < out) vw = W[(long)(o0 + r) * in + i0 + c];
      }
      sa[r][c] = va;
      sw[r][c] = vw;
    }
    __syncthreads();
    for (int c = 0; c < kk; ++c) {
      const float w = sw[tx][c];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_fmaf(sa[ty + 8 * q][c], w, acc[q]);
    }
    __syncthreads();
  }
  const int o = o0 + tx;
  if (o >= out) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = b0 + ty + 8 * q;
    if (b < B) {
      const float z = acc[q] + bias[o];
      Z[(long)b * out + o] = z;
      if (H) H[(long)b * out + o] = gelu(z);
    }
  }
}

// mean squared error over the B x m outputs: D = 2 (Y - T) / (B m), loss = sum (Y - T)^2 / (B m)
// (FP64 partial sums in a fixed tree); targets of row b at T[(t_off + b) mod t_mod]
__global__ __launch_bounds__(kTB) void ml_mse(const float *Y, const float *T, long t_off, long t_mod, int B, int m,
                                              float *D, float *loss) {
  __shared__ double red[kTB];
  const float scale = 2.f / (float)(B * m);
  double s = 0.0;
  for (int e = threadIdx.x; e < B * m; e += kTB) {
    const int b = e / m, j = e % m;
    const float d = Y[e] - T[((t_off + b) % t_mod) * m + j];
    D[e] = scale * d;
    s += (double)d * (double)d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kTB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = (float)(red[0] / (double)(B * m));
}

// dW[o, i] = sum_b D[b, o] A[row(b), i], b ascending; column i == in is the bias (A = 1)
__global__ __launch_bounds__(kTB) void ml_bwd_weight(const float *D, int B, int out, const float *A, long a_off,
                                                     long a_mod, int in, float *dW, float *db) {
  __shared__ float sd[kK][kT + 1];
  __shared__ float sa[kK][kT + 1];
  const int o0 = blockIdx.x * kT, i0 = blockIdx.y * kT;
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int bb = 0; bb < B; bb += kK) {
    const int kk = min(kK, B - bb);
    for (int e = threadIdx.x; e < kK * kT; e += kTB) {
      const int r = e / kT, c = e % kT;
      float vd = 0.f, va = 0.f;
      if (r < kk) {
        if (o0 + c < out) vd = D[(long)(bb + r) * out + o0 + c];
        const int i = i0 + c;
        if (i < in) va = A[((a_off + bb + r) % a_mod) * in + i];
        else if (i == in) va = 1.f;
      }
      sd[r][c] = vd;
      sa[r][c] = va;
    }
    __syncthreads();
    for (int r = 0; r < kk; ++r) {
      const float a = sa[r][tx];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_fmaf(sd[r][ty + 8 * q], a, acc[q]);
    }
    __syncthreads();
  }
  const int i = i0 + tx;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int o = o0 + ty + 8 * q;
    if (o >= out) continue;
    if (i < in) dW[(long)o * in + i] = acc[q];
    else if (i == in) db[o] = acc[q];
  }
}

// G[b, i] = (sum_o D[b, o] W[o, i]) gelu'(Zp[b, i]), o ascending
__global__ __launch_bounds__(kTB) void ml_bwd_data(const float *D, int B, int out, int in, const float *W,
                                                   const float *Zp, float *G) {
  __shared__ float sd[kT][kK + 1];
  __shared__ float sw[kK][kT + 1];
  const int b0 = blockIdx.x * kT, i0 = blockIdx.y * kT;
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  for (int o0 = 0; o0 < out; o0 += kK) {
    const int kk = min(kK, out - o0);
    for (int e = threadIdx.x; e < kT * kK; e += kTB) {
      const int r = e / kK, c = e % kK;
      sd[r][c] = (c < kk && b0 + r < B) ? D[(long)(b0 + r) * out + o0 + c] : 0.f;
      const int rr = e / kT, cc = e % kT;
      sw[rr][cc] = (rr < kk && i0 + cc < in) ? W[(long)(o0 + rr) * in + i0 + cc] : 0.f;
    }
    __syncthreads();
    for (int c = 0; c < kk; ++c) {
      const float w = sw[c][tx];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_fmaf(sd[ty + 8 * q][c], w, acc[q]);
    }
    __syncthreads();
  }
  const int i = i0 + tx;
  if (i >= in) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = b0 + ty + 8 * q;
    if (b < B) G[(long)b * in + i] = acc[q] * gelu_grad(Zp[(long)b * in + i]);
  }
}

// Adam (Kingma & Ba 2015) in the form of torch.optim.Adam without weight decay:
//   m = b1 m + (1 - b1) g,  v = b2 v + (1 - b2) g^2,
//   p -= (lr / (1 - b1^t)) m / (sqrt(v) / sqrt(1 - b2^t) + eps)
__global__ __launch_bounds__(kTB) void ml_adam(float *p, const float *g, float *m, float *v, long n, float b1,
                                               float b2, float step_size, float inv_sqrt_bc2, float eps) {
  for (long e = blockIdx.x * (long)kTB + threadIdx.x; e < n; e += (long)gridDim.x * kTB) {
    const float gr = g[e];
    const float mm = __builtin_fmaf(b1, m[e], (1.f - b1) * gr);
    const float vv = __builtin_fmaf(b2, v[e], (1.f - b2) * gr * gr);
    m[e] = mm;
    v[e] = vv;
    p[e] -= step_size * mm / (sqrtf(vv) * inv_sqrt_bc2 + eps);
  }
}

}  // namespace mlp
}  // namespace spss
