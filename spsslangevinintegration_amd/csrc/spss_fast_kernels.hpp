// spss_fast_kernels.hpp -- device code of the FP32-compensated cell-list path
// (included by spss_kernels_fast.hip; host logic lives there).
//
// Two ways to feed the pair kernel:
//  * rebuilt (MODE 0 RHS, MODE 1/2 neighbour export): every call sorts the particles
//    of an FP64 stage state (original order) into cells -- fk_keys, a two-level scan,
//    fk_scatter, fk_rank (stable: ascending particle index inside a cell);
//  * resident (MODE 3 Euler, MODE 4 SRK): the FP64 state, the particle ids and the
//    natural frequencies stay in cell order between steps.  The pair kernel writes the
//    new state and each particle's move (one of the 3x3 neighbour cells) in place of
//    the old order; fr_cell_counts, a two-level scan of the arrival counts and
//    fr_scatter then move every particle to its new slot: inside a cell, particles that
//    stayed come first in their previous order, then arrivals grouped by the direction
//    they came from, each group in its previous order, then the rare particles that jumped
//    farther than the 3x3 block around their cell, by previous position (fr_far_bucket).
//    No atomic decides an order, so results are deterministic.
#pragma once

#include <stdint.h>

#include "spss_device.hpp"

namespace spss {

namespace {
constexpr int kFB = 256;        // threads per pair workgroup
constexpr int kStageMax = 1536; // staged particles per workgroup
constexpr int kMaxRows = 7;     // 2k+1 with k <= 3
constexpr int kMaxCols = 64 + 6;
constexpr int kScanItems = 8;   // counts per thread in the two-level scans
}  // namespace

struct FastArgs {
  int n, ncf, k, W, chunks, ntiles;
  double H;
  float r2lo, r2hi;
  float alpha_rev;
};

// buffers of one pair-kernel launch (unused ones are null)
struct PairBufs {
  const double *stage;   // rebuilt modes: stage state, original order (AoS FP64)
  const int *sorted;     // rebuilt modes: particle id of each sorted position
  const double *rstate;  // resident: FP64 state in cell order
  const int *rid;        // resident: particle id of each sorted position
  const double *rnat;    // resident: natural frequency of each sorted position (or null)
  const int *start;      // [cells + 1] first sorted position of each cell
  const float4 *snap;    // FP32 snapshot in cell order: offsets in the cell, phase in revolutions
  const float *omega_f;  // SRK: omega of each sorted position
  const double *nat;     // rebuilt modes: natural frequency by particle id (or null)
  double *k_next;        // MODE 0, by particle id
  double *compl_out;     // complementary force by particle id (SRK)
  int *pcount;           // MODE 1, by particle id
  const int *pstart;     // MODE 2, by particle id
  int *plist;            // MODE 2
  double *tmp_state;     // MODE 3/4: new state of each (old) sorted position
  int *mv;               // MODE 3/4: old cell * 16 + move of each (old) sorted position
  int *farcnt;           // MODE 3/4: per cell, arrivals from farther than the 3x3 block (zeroed by the scan)
  int *farslot;          // MODE 3/4: arrival slot of such a far move (unordered)
};

// the stochastic update of MODE 3 (Euler) and MODE 4 (SRK)
struct StepArgs {
  double dt, sqrt_dt, int_phi, c15, is3, sq2d, friction;
  int noise_mode;  // 0 none, 1 device array indexed by particle id, 3 Philox
  const double *noise;
  uint64_t seed, step;
};

// ---------------------------------------------------------------------------
// moves of the resident re-sort: d = (dy+1)*3 + (dx+1) for a move into one of the 3x3 cells
// around the old cell, kFar for a longer jump.  Inside a destination cell the stayers (d = 4)
// come first, then moves 0, 1, 2, 3, 5, 6, 7, 8, then the far arrivals by old position.
constexpr int kFar = 9;

__device__ __forceinline__ int move_order(int d) { return d == 4 ? 0 : (d < 4 ? d + 1 : d); }

// source cell of an arrival into cell c by move d
__device__ __forceinline__ int move_source(int c, int d, int ncf) {
  const int cx = c % ncf, cy = c / ncf;
  const int sx = (cx - (d % 3 - 1) + ncf) % ncf;
  const int sy = (cy - (d / 3 - 1) + ncf) % ncf;
  return sy * ncf + sx;
}

__device__ __forceinline__ int new_count_of(const int *cnt9, int c, int ncf) {
  int s = 0;
#pragma unroll
  for (int d = 0; d < 9; ++d) s += cnt9[(long)move_source(c, d, ncf) * 9 + d];
  return s;
}

// ---------------------------------------------------------------------------
// rebuilt cell list

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

// block-local exclusive scan of 2048 values per block; a value is count[c] plus, for the
// resident re-sort, the near arrivals of cell c (either source may be null)
__global__ __launch_bounds__(256) void fk_scan_local(const int *__restrict__ count, const int *__restrict__ cnt9,
                                                     int ncf, int m, int *__restrict__ start,
                                                     int *__restrict__ blocksum) {
  __shared__ int sh[256];
  const int base = blockIdx.x * 256 * kScanItems + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = 0;
    if (base + q < m) {
      if (count) v[q] += count[base + q];
      if (cnt9) v[q] += new_count_of(cnt9, base + q, ncf);
    }
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

// adds the scanned block totals, writes start[m], zeroes count[] (when given) for the next build
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
      if (count) count[c] = 0;
    }
  }
}

__global__ void fk_scatter(int n, const int *__restrict__ key, const int *__restrict__ slot,
                           const int *__restrict__ start, int *__restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[start[key[i]] + slot[i]] = i;
}

// snapshot entry of a particle: offsets inside cell (cx, cy) in FP64, rounded once;
// phase in revolutions reduced to [0, 1) in FP64 before rounding
template <int S>
__device__ __forceinline__ float4 snap_entry(const DevParams &P, const double *st, int cx, int cy, double H,
                                             int id) {
  double xw, yw;
  apply_pbc(P, st[0], st[1], xw, yw);
  const double offx = xw - (double)cx * H;
  const double offy = yw - (double)cy * H;
  const double rev = st[2] * (1.0 / (2.0 * M_PI));
  const double red = rev - floor(rev);
  return make_float4((float)offx, (float)offy, (float)red, __int_as_float(id));
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
  snap[p] = snap_entry<S>(P, st, c % ncf, c / ncf, H, i);
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
__device__ __forceinline__ bool fp64_in_range(const DevParams &P, const double *pj, double x_i, double y_i) {
  double dx, dy;
  class_c_unsigned(P, x_i, y_i, pj[0], pj[1], dx, dy);
  double dr2 = dx * dx + dy * dy;
  return dr2 <= P.rho2;
}

// MODE 0: RHS; 1: count neighbours; 2: write neighbour lists; 3: fused Euler; 4: fused SRK
template <int S, bool SRK, int MODE>
__global__ __launch_bounds__(kFB) void fk_pairs(DevParams P, FastArgs A, PairBufs B, StepArgs U) {
  constexpr bool RES = MODE >= 3;
  __shared__ float2 spos[kStageMax];               // x re-based to staged column 0, y offset in its cell
  __shared__ float2 saux[kStageMax];               // phase in revolutions, global sorted position (bits)
  __shared__ float somega[SRK ? kStageMax : 1];
  __shared__ uint8_t scol[kStageMax];              // staged column of each staged particle
  __shared__ int cstart[kMaxRows][kMaxCols + 1];   // global sorted start of each staged cell
  __shared__ int cpos[kMaxRows * kMaxCols + 1];    // LDS start of staged cell (r, q) at r*ncols+q
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
  const int *start = B.start;

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
      const float4 v = B.snap[g];
      const float xt = (float)((double)q * H + (double)v.x);
      spos[e] = make_float2(xt, v.y);
      saux[e] = make_float2(v.z, __int_as_float(g));
      scol[e] = (uint8_t)q;
      if (SRK) somega[e] = B.omega_f[g];
    }
  }
  __syncthreads();

  // particles of the tile: staged row k (dr = 0), columns k .. k+W-1
  const int gtb = cstart[k][k];
  const int ntile = start[row * ncf + cb] - start[row * ncf + ca];
  const int ltb = fits ? cpos[k * ncols + k] : 0;
  const float r2lo = A.r2lo, r2hi = A.r2hi;

  for (int u = threadIdx.x; u < ntile; u += kFB) {
    const int li = ltb + u;  // LDS index (fits)
    const int gp = gtb + u;  // global sorted position
    const int i = RES ? B.rid[gp] : B.sorted[gp];
    const double *pi = RES ? B.rstate + (long)gp * S : B.stage + (long)i * S;
    double x_i64, y_i64;
    apply_pbc(P, pi[0], pi[1], x_i64, y_i64);
    float xi_t, yi, phi_i;
    int qi;
    if (fits) {
      const float2 mp = spos[li], ma = saux[li];
      xi_t = mp.x;
      yi = mp.y;
      phi_i = ma.x;
      qi = scol[li];
    } else {
      const float4 mg = B.snap[gp];
      xi_t = mg.x;  // offset in its own cell
      yi = mg.y;
      phi_i = mg.z;
      qi = k;  // own cell column: last tile cell whose start <= gp
      for (int q = k; q < k + W; ++q)
        if (cstart[k][q] <= gp) qi = q;
    }
    const float phia = phi_i + A.alpha_rev;
    const float om_i = SRK ? (fits ? somega[li] : B.omega_f[gp]) : 0.f;
    float acc = 0.f, accc = 0.f, cs = 0.f, csc = 0.f;
    int nnb = 0;
    int w = 0;
    const int wbase = (MODE == 2) ? B.pstart[i] : 0;

    // a candidate within r2hi: guard-band re-test in FP64, then the coupling terms
    auto use_pair = [&](float d2, int g, float phij, float omj) {
      if (d2 >= r2lo) {
        const double *pj = RES ? B.rstate + (long)g * S : B.stage + (long)B.sorted[g] * S;
        if (!fp64_in_range(P, pj, x_i64, y_i64)) return;
      }
      ++nnb;
      if (MODE == 0 || MODE >= 3) {
        const float drev = phij - phia;
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
        B.plist[wbase + w++] = B.sorted[g];
      }
    };

    if (fits) {
      if (MODE == 1 || MODE == 2) {
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
              use_pair(d2, __float_as_int(a2.y), a2.x, 0.f);
            }
          }
        }
      } else {
        // the particle itself is among its candidates exactly once (distinct stencil
        // columns since int(1/rho) >= 3): it is counted with d2 = 0 and removed below
        for (int r = 0; r < nrows; ++r) {
          const float syi = (float)((double)yi - (double)(r - k) * H);
          const int lo = cpos[r * ncols + qi - k];
          const int hi = cpos[r * ncols + qi + k + 1];
#pragma unroll 2
          for (int t = lo; t < hi; ++t) {
            const float2 o = spos[t];
            const float dx = o.x - xi_t;
            const float dy = o.y - syi;
            const float d2 = __builtin_fmaf(dy, dy, dx * dx);
            if (d2 <= r2hi) {
              const float2 a2 = saux[t];
              use_pair(d2, __float_as_int(a2.y), a2.x, SRK ? somega[t] : 0.f);
            }
          }
        }
        --nnb;
        float s, e;
        two_sum_f(acc, -__builtin_amdgcn_sinf(phi_i - phia), s, e);  // the self term added above
        acc = s;
        accc += e;
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
            const float4 o = B.snap[g];
            const float dx = o.x - sxi;
            const float dy = o.y - syi;
            const float d2 = __builtin_fmaf(dy, dy, dx * dx);
            if (d2 <= r2hi) use_pair(d2, g, o.z, SRK ? B.omega_f[g] : 0.f);
          }
        }
      }
    }

    const double A64 = (double)acc + (double)accc;
    const double card = (double)nnb;
    if (MODE == 1) {
      B.pcount[i] = nnb;
    } else if (MODE == 0) {
      const long ii = (long)i * S;
      double sn, cn;
      sincos(pi[2], &sn, &cn);
      const double nat_i = B.nat ? B.nat[i] : 0.0;
      B.k_next[ii] = P.v0 * cn;
      B.k_next[ii + 1] = P.v0 * sn;
      if (S == 4) {
        const double omega_i = pi[3];
        B.k_next[ii + 2] = omega_i;
        if (P.natfreq_local) {
          if (nnb) B.k_next[ii + 3] = -P.xi * omega_i + nat_i + P.sigma * A64 / card;
          else B.k_next[ii + 3] = -P.xi * omega_i + nat_i;
        } else if (nnb) {
          B.k_next[ii + 3] = -P.xi * omega_i + P.sigma * A64 / card;
        }
        if (SRK && B.compl_out) {
          double c = (double)cs + (double)csc;
          if (nnb) c *= P.sigma / card;
          B.compl_out[i] = c;
        }
      } else {
        B.k_next[ii + 2] = nnb ? nat_i + P.sigma * A64 / card : nat_i;
      }
    } else if (MODE >= 3) {
      // fused stepper: the derivative of this particle as the zero-filled derivative vector
      // of StochasticEulerStepperPtr.hpp:42-47 would hold it, then the update
      double *out = B.tmp_state + (long)gp * S;
      double sn, cn;
      sincos(pi[2], &sn, &cn);
      const double nat_i = B.rnat ? B.rnat[gp] : 0.0;
      const double f0 = P.v0 * cn, f1 = P.v0 * sn;
      if (S == 4) {
        const double omega_i = pi[3];
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
          out[0] = pi[0] + f0 * U.dt;
          out[1] = pi[1] + f1 * U.dt;
          out[2] = pi[2] + f2 * U.dt;
          out[3] = pi[3] + (f3 * U.dt + U.int_phi * dW);
        } else {  // StochasticRungeKuttaStepperPtr.hpp:50-71
          double C = (double)cs + (double)csc;
          if (nnb) C *= P.sigma / card;
          if (B.compl_out) B.compl_out[i] = C;
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
          out[0] = pi[0] + (f0 * dt - 0.5 * f2 * f1 * dt * dt);
          out[1] = pi[1] + (f1 * dt + 0.5 * f2 * f0 * dt * dt);
          out[2] = pi[2] + (f2 * dt + 0.5 * f3 * dt * dt + U.sq2d * dZ);
          out[3] = pi[3] + (f3 * dt + 0.5 * (C - U.friction * f3) * dt * dt + U.sq2d * (dW - U.friction * dZ));
        }
      } else {  // model 3 Euler
        const double f2 = nnb ? nat_i + P.sigma * A64 / card : nat_i;
        const double z = (U.noise_mode == 1) ? U.noise[i] : (U.noise_mode == 3 ? philox_z(U.seed, i, 0, U.step) : 0.0);
        out[0] = pi[0] + f0 * U.dt;
        out[1] = pi[1] + f1 * U.dt;
        out[2] = pi[2] + (f2 * U.dt + U.int_phi * (U.sqrt_dt * z));
      }
      // the particle's move from its cell (row, ca + qi - k)
      double xw, yw;
      apply_pbc(P, out[0], out[1], xw, yw);
      const int nx = cell_coord(xw, ncf), ny = cell_coord(yw, ncf);
      const int ocol = ca + qi - k;
      int dxc = nx - ocol, dyc = ny - row;
      if (dxc == ncf - 1) dxc = -1;
      if (dxc == 1 - ncf) dxc = 1;
      if (dyc == ncf - 1) dyc = -1;
      if (dyc == 1 - ncf) dyc = 1;
      int d;
      if (dxc >= -1 && dxc <= 1 && dyc >= -1 && dyc <= 1) {
        d = (dyc + 1) * 3 + (dxc + 1);
      } else {
        d = kFar;
        B.farslot[gp] = atomicAdd(B.farcnt + ny * ncf + nx, 1);
      }
      B.mv[gp] = ((row * ncf + ocol) << 4) | d;
    }
  }
}

// ---------------------------------------------------------------------------
// resident re-sort

// per cell: how many of its particles take each of the nine near moves
__global__ void fr_cell_counts(int cells, const int *__restrict__ start, const int *__restrict__ mv,
                               int *__restrict__ cnt9) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  int v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int p = start[c]; p < start[c + 1]; ++p) {
    const int d = mv[p] & 15;
#pragma unroll
    for (int e = 0; e < 9; ++e) v[e] += (d == e) ? 1 : 0;
  }
#pragma unroll
  for (int d = 0; d < 9; ++d) cnt9[(long)c * 9 + d] = v[d];
}

template <int S>
__device__ __forceinline__ int cell_of_state(const DevParams &P, const double *st, int ncf) {
  double xw, yw;
  apply_pbc(P, st[0], st[1], xw, yw);
  return cell_coord(yw, ncf) * ncf + cell_coord(xw, ncf);
}

// far arrivals: unordered bucket behind the near arrivals of their destination cell
template <int S>
__global__ void fr_far_bucket(DevParams P, int n, int ncf, const int *__restrict__ mv,
                              const double *__restrict__ tmp_state, const int *__restrict__ nstart,
                              const int *__restrict__ cnt9, const int *__restrict__ farslot,
                              int *__restrict__ bucket) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || (mv[p] & 15) != kFar) return;
  const int cd = cell_of_state<S>(P, tmp_state + (long)p * S, ncf);
  bucket[nstart[cd] + new_count_of(cnt9, cd, ncf) + farslot[p]] = p;
}

// move every particle to its new slot; write the new resident arrays and the snapshot
template <int S>
__global__ void fr_scatter(DevParams P, int n, int ncf, double H, const int *__restrict__ start,
                           const int *__restrict__ nstart, const int *__restrict__ mv,
                           const int *__restrict__ cnt9, const int *__restrict__ bucket,
                           const double *__restrict__ tmp_state, const int *__restrict__ rid,
                           const double *__restrict__ rnat, double *__restrict__ nstate, int *__restrict__ nrid,
                           double *__restrict__ nnat, float4 *__restrict__ snap, float *__restrict__ omega_f) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int m = mv[p];
  const int cs = m >> 4, d = m & 15;
  const double *src = tmp_state + (long)p * S;
  int cd, pos;
  if (d == kFar) {
    cd = cell_of_state<S>(P, src, ncf);
    const int b = nstart[cd] + new_count_of(cnt9, cd, ncf);
    const int e = nstart[cd + 1];
    int rank = 0;
    for (int q = b; q < e; ++q) rank += (bucket[q] < p) ? 1 : 0;
    pos = b + rank;
  } else {
    int rank = 0;
    for (int q = start[cs]; q < p; ++q) rank += ((mv[q] & 15) == d) ? 1 : 0;
    const int dcx = d % 3 - 1, dcy = d / 3 - 1;
    cd = ((cs / ncf + dcy + ncf) % ncf) * ncf + (cs % ncf + dcx + ncf) % ncf;
    int base = 0;
    const int od = move_order(d);
    for (int e = 0; e < 9; ++e)
      if (move_order(e) < od) base += cnt9[(long)move_source(cd, e, ncf) * 9 + e];
    pos = nstart[cd] + base + rank;
  }
  double *dst = nstate + (long)pos * S;
#pragma unroll
  for (int q = 0; q < S; ++q) dst[q] = src[q];
  const int id = rid[p];
  nrid[pos] = id;
  if (rnat) nnat[pos] = rnat[p];
  snap[pos] = snap_entry<S>(P, src, cd % ncf, cd / ncf, H, id);
  if (omega_f) omega_f[pos] = (S == 4) ? (float)src[3] : 0.f;
}

// entry into resident mode: gather the state (original order) into cell order
template <int S>
__global__ void fr_gather(int n, const double *__restrict__ state, const int *__restrict__ sorted,
                          const double *__restrict__ nat, double *__restrict__ rstate, int *__restrict__ rid,
                          double *__restrict__ rnat) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int id = sorted[p];
#pragma unroll
  for (int q = 0; q < S; ++q) rstate[(long)p * S + q] = state[(long)id * S + q];
  rid[p] = id;
  if (nat) rnat[p] = nat[id];
}

// leave resident mode: the state back into original order
template <int S>
__global__ void fr_unpermute(int n, const double *__restrict__ rstate, const int *__restrict__ rid,
                             double *__restrict__ state) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int id = rid[p];
#pragma unroll
  for (int q = 0; q < S; ++q) state[(long)id * S + q] = rstate[(long)p * S + q];
}

}  // namespace spss
