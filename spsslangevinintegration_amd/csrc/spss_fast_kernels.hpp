// spss_fast_kernels.hpp -- device code of the FP32-compensated cell-list path
// (included by spss_kernels_fast.hip; host logic lives there).
#pragma once

#include <stdint.h>

#include "spss_device.hpp"

namespace spss {

namespace {
constexpr int kFB = 256;       // threads per pair workgroup
constexpr int kStageMax = 1536; // staged particles per workgroup
constexpr int kMaxRows = 7;    // 2k+1 with k <= 3
constexpr int kMaxCols = 64 + 6;
}  // namespace


struct FastArgs {
  int n, ncf, k, W, chunks, ntiles;
  double H;
  float r2lo, r2hi;
  float alpha_rev;
};

// fused stochastic update (MODE 3: Euler, MODE 4: SRK); the state is double-buffered:
// the kernel reads `stage` (= the current state, k_coef = 0) and writes `state_out`
struct StepArgs {
  double *state_out;
  double dt, sqrt_dt, int_phi, c15, is3, sq2d, friction;
  int noise_mode;
  const double *noise;
  uint64_t seed, step;
};

// ---------------------------------------------------------------------------
template <int S>
__global__ void fk_keys(DevParams P, const double *__restrict__ stage, int ncf, int *__restrict__ key,
                        int *__restrict__ slot, int *__restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double xw, yw;
  apply_pbc(P, stage[(long)i * S], stage[(long)i * S + 1], xw, yw);
  const int c = cell_coord(yw, ncf) * ncf + cell_coord(xw, ncf);
  key[i] = c;
  slot[i] = atomicAdd(count + c, 1);
}

// two-level exclusive scan of count[0..m): 2048 counts per block
constexpr int kScanItems = 8;
__global__ __launch_bounds__(256) void fk_scan_local(const int *__restrict__ count, int m, int *__restrict__ start,
                                                     int *__restrict__ blocksum) {
  __shared__ int sh[256];
  const int base = blockIdx.x * 256 * kScanItems + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = (base + q < m) ? count[base + q] : 0;
    s += v[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int t = (threadIdx.x >= off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - s;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < m) start[base + q] = run;
    run += v[q];
  }
  if (threadIdx.x == 255) blocksum[blockIdx.x] = sh[255];
}

// adds the scanned block totals, zeroes the counts for the next build, writes start[m]
__global__ __launch_bounds__(256) void fk_scan_add(int *__restrict__ count, int m, int *__restrict__ start,
                                                   const int *__restrict__ blocksum, int nblocks) {
  __shared__ int off;
  if (threadIdx.x == 0) {
    int s = 0;
    for (int b = 0; b < (int)blockIdx.x; ++b) s += blocksum[b];
    off = s;
    if (blockIdx.x == nblocks - 1) start[m] = s + blocksum[blockIdx.x];
  }
  __syncthreads();
  const int base = blockIdx.x * 256 * kScanItems;
  for (int q = threadIdx.x; q < 256 * kScanItems; q += 256) {
    const int c = base + q;
    if (c < m) {
      start[c] += off;
      count[c] = 0;
    }
  }
}

__global__ void fk_scatter(int n, const int *__restrict__ key, const int *__restrict__ slot,
                           const int *__restrict__ start, int *__restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[start[key[i]] + slot[i]] = i;
}

// stable rank + snapshot in sorted order
template <int S>
__global__ void fk_rank(DevParams P, const double *__restrict__ stage, int ncf, double H,
                        const int *__restrict__ key, const int *__restrict__ start, const int *__restrict__ tmp,
                        int *__restrict__ sorted, float4 *__restrict__ snap, float *__restrict__ omega_f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int c = key[i];
  const int b = start[c], e = start[c + 1];
  int r = 0;
  for (int t = b; t < e; ++t) r += (tmp[t] < i) ? 1 : 0;
  const int p = b + r;
  sorted[p] = i;
  const double *st = stage + (long)i * S;
  double xw, yw;
  apply_pbc(P, st[0], st[1], xw, yw);
  const int cx = c % ncf, cy = c / ncf;
  const double offx = xw - (double)cx * H;
  const double offy = yw - (double)cy * H;
  // phase in revolutions, reduced to [0, 1) in FP64 before rounding to FP32
  const double rev = st[2] * (1.0 / (2.0 * M_PI));
  const double red = rev - floor(rev);
  snap[p] = make_float4((float)offx, (float)offy, (float)red, __int_as_float(i));
  if (omega_f) omega_f[p] = (S == 4) ? (float)st[3] : 0.f;
}

// ---------------------------------------------------------------------------
// the pair kernel

__device__ __forceinline__ double philox_z(uint64_t seed, int i, uint32_t block, uint64_t step) {
  double z0, z1;
  philox_normal2(seed, (uint32_t)i, block, step, z0, z1);
  return z0;
}

__device__ __forceinline__ void two_sum_f(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// FP64 re-test of a guard-band pair with the reference's formula
// (x_i wrapped, x_j as stored, ClassCEffectiveParticleDistanceUnsigned, dr2 <= rho^2)
template <int S>
__device__ __forceinline__ bool fp64_in_range(const DevParams &P, const double *stage, double x_i, double y_i,
                                              int j) {
  double dx, dy;
  class_c_unsigned(P, x_i, y_i, stage[(long)j * S], stage[(long)j * S + 1], dx, dy);
  double dr2 = dx * dx + dy * dy;
  return dr2 <= P.rho2;
}

// MODE 0: RHS; MODE 1: count neighbours; MODE 2: write neighbour lists
template <int S, bool SRK, int MODE>
__global__ __launch_bounds__(kFB) void fk_pairs(DevParams P, FastArgs A, const double *__restrict__ stage,
                                                const int *__restrict__ start, const float4 *__restrict__ snap,
                                                const float *__restrict__ omega_f, const double *__restrict__ nat,
                                                double *__restrict__ k_next, double *__restrict__ compl_out,
                                                int *__restrict__ pcount, const int *__restrict__ pstart,
                                                int *__restrict__ plist, StepArgs U) {
  __shared__ float2 spos[kStageMax];               // x re-based to staged column 0, y offset in its cell
  __shared__ float2 saux[kStageMax];               // phase in revolutions, particle index (bits)
  __shared__ float somega[SRK ? kStageMax : 1];
  __shared__ uint8_t scol[kStageMax];              // staged column of each staged particle
  __shared__ int cstart[kMaxRows][kMaxCols + 1];   // global sorted start of each staged cell; [ncols] = end
  __shared__ int cpos[kMaxRows * kMaxCols + 1];    // LDS start of staged cell (r, q) at r*ncols+q; [nrows*ncols] = total
  __shared__ int s_total;

  // XCD-aware tile order: consecutive tiles (neighbouring rows) share an XCD's L2
  const int per = gridDim.x / 8;  // the grid is a multiple of 8
  const int tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= A.ntiles) return;
  const int row = tile / A.chunks;
  const int ca = (tile % A.chunks) * A.W;
  const int cb = min(ca + A.W, A.ncf);
  const int W = cb - ca;
  const int k = A.k;
  const int nrows = 2 * k + 1;
  const int ncols = W + 2 * k;
  const int ncf = A.ncf;
  const double H = A.H;

  // ---- staging plan, all threads: global extent of every staged cell
  const int ncells = nrows * ncols;
  for (int e = threadIdx.x; e < ncells; e += kFB) {
    const int r = e / ncols, q = e % ncols;
    const int rrow = ((row + r - k) % ncf + ncf) % ncf;
    const int col = ((ca - k + q) % ncf + ncf) % ncf;
    const int c = rrow * ncf + col;
    const int g0 = start[c];
    cstart[r][q] = g0;
    cpos[e] = start[c + 1] - g0;  // count for now
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive prefix over at most 7 x 70 cells
    int base = 0;
    for (int e = 0; e < ncells; ++e) {
      const int c = cpos[e];
      cpos[e] = base;
      base += c;
    }
    cpos[ncells] = base;
    s_total = base;
  }
  __syncthreads();
  const int total = s_total;
  const bool fits = total <= kStageMax;

  if (fits) {
    // copy: staged entry e belongs to the last cell whose LDS start is <= e
    for (int e = threadIdx.x; e < total; e += kFB) {
      int lo = 0, hi = ncells - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpos[mid] <= e) lo = mid; else hi = mid - 1;
      }
      const int r = lo / ncols, q = lo % ncols;
      const int g = cstart[r][q] + (e - cpos[lo]);
      const float4 v = snap[g];
      const float xt = (float)((double)q * H + (double)v.x);
      spos[e] = make_float2(xt, v.y);
      saux[e] = make_float2(v.z, v.w);
      scol[e] = (uint8_t)q;
      if (SRK) somega[e] = omega_f[g];
    }
  }
  __syncthreads();

  // particles of the tile: staged row k (dr = 0), columns k .. k+W-1
  const int rowk = k * ncols;
  const int gtb = cstart[k][k];
  const int ntile = start[row * ncf + cb] - start[row * ncf + ca];
  const int ltb = fits ? cpos[rowk + k] : 0;
  const float r2lo = A.r2lo, r2hi = A.r2hi;

  for (int u = threadIdx.x; u < ntile; u += kFB) {
    const int li = ltb + u;      // LDS index (fits)
    const int gp = gtb + u;      // global sorted position
    float xi_t, yi, phi_i;
    int i, qi;
    if (fits) {
      const float2 mp = spos[li], ma = saux[li];
      xi_t = mp.x;
      yi = mp.y;
      phi_i = ma.x;
      i = __float_as_int(ma.y);
      qi = scol[li];
    } else {
      const float4 mg = snap[gp];
      xi_t = mg.x;  // offset in its own cell
      yi = mg.y;
      phi_i = mg.z;
      i = __float_as_int(mg.w);
      // own cell column: last tile cell whose start <= gp
      qi = 0;
      for (int q = k; q < k + W; ++q)
        if (cstart[k][q] <= gp) qi = q;
    }
    const float om_i = SRK ? (fits ? somega[li] : omega_f[gp]) : 0.f;
    bool have64 = false;
    double x_i64 = 0.0, y_i64 = 0.0;
    float acc = 0.f, accc = 0.f, cs = 0.f, csc = 0.f;
    int nnb = 0;
    int w = 0;
    const int wbase = (MODE == 2) ? pstart[i] : 0;

    auto use_pair = [&](float d2, int j, float phij, float omj) {
      if (d2 >= r2lo) {
        if (!have64) {
          apply_pbc(P, stage[(long)i * S], stage[(long)i * S + 1], x_i64, y_i64);
          have64 = true;
        }
        if (!fp64_in_range<S>(P, stage, x_i64, y_i64, j)) return;
      }
      ++nnb;
      if (MODE == 0 || MODE >= 3) {
        const float drev = phij - phi_i - A.alpha_rev;
        float s, e;
        two_sum_f(acc, __builtin_amdgcn_sinf(drev), s, e);
        acc = s;
        accc += e;
        if (SRK) {
          two_sum_f(cs, (omj - om_i) * __builtin_amdgcn_cosf(drev), s, e);
          cs = s;
          csc += e;
        }
      } else if (MODE == 2) {
        plist[wbase + w++] = j;
      }
    };

    if (fits) {
      for (int r = 0; r < nrows; ++r) {
        const float syi = (float)((double)yi - (double)(r - k) * H);
        const int lo = cpos[r * ncols + qi - k];
        const int hi = cpos[r * ncols + qi + k + 1];
        for (int t = lo; t < hi; ++t) {
          const float2 o = spos[t];
          const float dx = o.x - xi_t;
          const float dy = o.y - syi;
          const float d2 = __builtin_fmaf(dy, dy, dx * dx);
          if (d2 <= r2hi && t != li) {
            const float2 a2 = saux[t];
            use_pair(d2, __float_as_int(a2.y), a2.x, SRK ? somega[t] : 0.f);
          }
        }
      }
    } else {
      // global path (staging would overflow LDS): per stencil cell, own offsets shifted
      for (int r = 0; r < nrows; ++r) {
        const int dr = r - k;
        const float syi = (float)((double)yi - (double)dr * H);
        for (int dc = -k; dc <= k; ++dc) {
          const float sxi = (float)((double)xi_t - (double)dc * H);
          const int g0 = cstart[r][qi + dc];
          const int rrow = ((row + dr) % ncf + ncf) % ncf;
          const int col = ((ca - k + qi + dc) % ncf + ncf) % ncf;
          const int g1 = start[rrow * ncf + col + 1];
          for (int g = g0; g < g1; ++g) {
            if (g == gp) continue;
            const float4 o = snap[g];
            const float dx = o.x - sxi;
            const float dy = o.y - syi;
            const float d2 = __builtin_fmaf(dy, dy, dx * dx);
            if (d2 <= r2hi) use_pair(d2, __float_as_int(o.w), o.z, SRK ? omega_f[g] : 0.f);
          }
        }
      }
    }

    if (MODE == 1) {
      pcount[i] = nnb;
    } else if (MODE == 0) {
      const long ii = (long)i * S;
      const double *sti = stage + ii;
      const double phi64 = sti[2];
      double sn, cn;
      sincos(phi64, &sn, &cn);
      const double A64 = (double)acc + (double)accc;
      const double card = (double)nnb;
      const double nat_i = nat ? nat[i] : 0.0;
      k_next[ii] = P.v0 * cn;
      k_next[ii + 1] = P.v0 * sn;
      if (S == 4) {
        const double omega_i = sti[3];
        k_next[ii + 2] = omega_i;
        if (P.natfreq_local) {
          if (nnb) k_next[ii + 3] = -P.xi * omega_i + nat_i + P.sigma * A64 / card;
          else k_next[ii + 3] = -P.xi * omega_i + nat_i;
        } else if (nnb) {
          k_next[ii + 3] = -P.xi * omega_i + P.sigma * A64 / card;
        }
        if (SRK && compl_out) {
          double c = (double)cs + (double)csc;
          if (nnb) c *= P.sigma / card;
          compl_out[i] = c;
        }
      } else {
        k_next[ii + 2] = nnb ? nat_i + P.sigma * A64 / card : nat_i;
      }
    } else if (MODE >= 3) {
      // fused stepper: the derivative of this particle, as the zero-filled derivative
      // vector of StochasticEulerStepperPtr.hpp:42-47 would hold it, then the update
      const long ii = (long)i * S;
      const double *sti = stage + ii;
      double *out = U.state_out + ii;
      const double phi64 = sti[2];
      double sn, cn;
      sincos(phi64, &sn, &cn);
      const double A64 = (double)acc + (double)accc;
      const double card = (double)nnb;
      const double nat_i = nat ? nat[i] : 0.0;
      const double f0 = P.v0 * cn, f1 = P.v0 * sn;
      if (S == 4) {
        const double omega_i = sti[3];
        const double f2 = omega_i;
        double f3 = 0.0;
        if (P.natfreq_local) {
          f3 = nnb ? -P.xi * omega_i + nat_i + P.sigma * A64 / card : -P.xi * omega_i + nat_i;
        } else if (nnb) {
          f3 = -P.xi * omega_i + P.sigma * A64 / card;
        }
        if (MODE == 3) {  // StochasticEulerStepperPtr.hpp:50-67
          const double z = (U.noise_mode == 1) ? U.noise[i] : (U.noise_mode == 3 ? philox_z(U.seed, i, 0, U.step) : 0.0);
          const double dW = U.sqrt_dt * z;
          out[0] = sti[0] + f0 * U.dt;
          out[1] = sti[1] + f1 * U.dt;
          out[2] = sti[2] + f2 * U.dt;
          out[3] = sti[3] + (f3 * U.dt + U.int_phi * dW);
        } else {  // StochasticRungeKuttaStepperPtr.hpp:50-71
          double C = (double)cs + (double)csc;
          if (nnb) C *= P.sigma / card;
          if (compl_out) compl_out[i] = C;
          double U1 = 0.0, U2 = 0.0;
          if (U.noise_mode == 1) {
            U1 = U.noise[2L * i];
            U2 = U.noise[2L * i + 1];
          } else if (U.noise_mode == 3) {
            philox_normal2(U.seed, (uint32_t)i, 0, U.step, U1, U2);
          }
          const double dt = U.dt;
          const double dW = U1 * U.sqrt_dt;
          const double dZ = U.c15 * (U1 + U.is3 * U2);
          out[0] = sti[0] + (f0 * dt - 0.5 * f2 * f1 * dt * dt);
          out[1] = sti[1] + (f1 * dt + 0.5 * f2 * f0 * dt * dt);
          out[2] = sti[2] + (f2 * dt + 0.5 * f3 * dt * dt + U.sq2d * dZ);
          out[3] = sti[3] + (f3 * dt + 0.5 * (C - U.friction * f3) * dt * dt + U.sq2d * (dW - U.friction * dZ));
        }
      } else {  // model 3 Euler
        const double f2 = nnb ? nat_i + P.sigma * A64 / card : nat_i;
        const double z = (U.noise_mode == 1) ? U.noise[i] : (U.noise_mode == 3 ? philox_z(U.seed, i, 0, U.step) : 0.0);
        out[0] = sti[0] + f0 * U.dt;
        out[1] = sti[1] + f1 * U.dt;
        out[2] = sti[2] + (f2 * U.dt + U.int_phi * (U.sqrt_dt * z));
      }
    }
  }
}

}  // namespace spss
