// spss_kernels_fp64.hip -- FP64 kernels that follow the reference's operation order.
//
// Every pair sum is accumulated in the order the reference visits the pairs
// (j ascending for all-pairs; stencil order then ascending j for cells; list
// order for Verlet), so results agree with the reference to the last few ulps
// of the device sin/cos.  Built with -ffp-contract=off.
#pragma clang fp contract(off)

#include "spss_device.hpp"
#include "spss_launch.hpp"
#include "spss_scan.hpp"

namespace spss {

// ---------------------------------------------------------------------------
// stage state y + (k_coef*dt)*k_prev, ParticleSystemFirstOrderPtr.cpp:234-241
__global__ void k_stage(long m, const double *__restrict__ state, const double *__restrict__ kprev,
                        double kc_dt, double *__restrict__ stage) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) stage[k] = state[k] + kc_dt * kprev[k];
}

// the stage of an all-zero k_prev: state + 0.0 is what state + kc_dt * 0.0 gives for any finite
// kc_dt (-0.0 becomes +0.0), without reading a zero vector
__global__ void k_stage_zero(long m, const double *__restrict__ state, double *__restrict__ stage) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) stage[k] = state[k] + 0.0;
}

// ---------------------------------------------------------------------------
// all-pairs, ParticleSystemFirstOrderPtr.cpp:225-316 / ParticleSystemSecondOrderPtr.cpp:155-230.
// One thread per i (BLOCK particles per workgroup); j-tiles of JT particles of the stage state are
// staged through LDS and visited in ascending j, the reference's summation order, so every
// (BLOCK, JT) gives the same sums bit for bit: the autotuner picks the fastest.
template <int S, int BLOCK, int JT>
__global__ __launch_bounds__(BLOCK) void k_all_pairs(DevParams P, const double *__restrict__ stage,
                                                     const double *__restrict__ nat, double *__restrict__ k_next,
                                                     double *__restrict__ compl_out) {
  __shared__ double tile[JT * S];
  const int n = P.n;
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  double st[S];
  double x_i = 0.0, y_i = 0.0, phi_i = 0.0, omega_i = 0.0;
  if (i < n) {
#pragma unroll
    for (int q = 0; q < S; ++q) st[q] = stage[(long)i * S + q];
    x_i = st[0];
    y_i = st[1];
    phi_i = st[Slots<S>::phi];
    omega_i = (S == 3) ? 0.0 : st[Slots<S>::omega];
    if (S != 6) apply_pbc(P, x_i, y_i, x_i, y_i);  // :270 (model 6 wraps nothing, :188-193)
  }
  PairAcc a{0.0, 0.0, 0.0};
  for (int j0 = 0; j0 < n; j0 += JT) {
    const int cnt = min(JT, n - j0);
    __syncthreads();
    for (int q = threadIdx.x; q < cnt * S; q += BLOCK) tile[q] = stage[(long)j0 * S + q];
    __syncthreads();
    if (i < n) {
      for (int t = 0; t < cnt; ++t) {
        const int j = j0 + t;
        if (j == i) continue;
        const double *sj = tile + t * S;
        pair_term<S>(P, x_i, y_i, phi_i, omega_i, sj[0], sj[1], sj[Slots<S>::phi],
                     (S == 3) ? 0.0 : sj[Slots<S>::omega], a);
      }
    }
  }
  if (i < n) write_k<S>(P, st, i, a, nat ? nat[i] : 0.0, k_next, compl_out);
}

// ---------------------------------------------------------------------------
// Cell list by a stable device counting sort over the cell keys.
// Keys: ParticleSystemFirstOrderPtr.cpp:364-381.  Buckets in ascending particle
// index (:388-391) come from the stable rank pass below.

template <int S>
__global__ void k_cell_keys(DevParams P, const double *__restrict__ stage, int ncx, int ncy,
                            int *__restrict__ key, int *__restrict__ slot, int *__restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double xw, yw;
  key_wrap<S>(P, stage[(long)i * S], stage[(long)i * S + 1], xw, yw);
  int c = cell_coord(yw, ncy) * ncx + cell_coord(xw, ncx);
  key[i] = c;
  slot[i] = atomicAdd(count + c, 1);
}

__global__ void k_scatter(int n, const int *__restrict__ key, const int *__restrict__ slot,
                          const int *__restrict__ start, int *__restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[start[key[i]] + slot[i]] = i;
}

// Each cell's members in ascending particle index (the order in which the FP64 pair loop sums a
// cell's members).  Cells of at most 64 members (every cell at the densities of the workloads) are
// sorted by one wave each (wave_sort64, grid-stride over cells); the members of denser cells are
// then ranked one thread each by counting the smaller indices of their cell (grid-stride over the
// particles, so a single cell holding most particles is spread over the whole grid).
__global__ __launch_bounds__(256) void k_stable_rank(int n, int ncells, const int *__restrict__ key,
                                                     const int *__restrict__ start, const int *__restrict__ tmp,
                                                     int *__restrict__ sorted) {
  const int lane = threadIdx.x & 63;
  for (int c = blockIdx.x * 4 + (threadIdx.x >> 6); c < ncells; c += gridDim.x * 4) {
    const int b = start[c], m = start[c + 1] - b;
    if (m > 64) continue;  // the whole wave
    const int v = wave_sort64(lane < m ? tmp[b + lane] : 0x7fffffff, lane);
    if (lane < m) sorted[b + lane] = v;
  }
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int c = key[i];
    const int b = start[c], e = start[c + 1];
    if (e - b <= 64) continue;
    int r = 0;
    for (int t = b; t < e; ++t) r += tmp[t] < i ? 1 : 0;
    sorted[b + r] = i;
  }
}

// linked-list pair loop, ParticleSystemFirstOrderPtr.cpp:398-466 (x_i wrapped, x_j as stored) and
// ParticleSystemSecondOrderPtr.cpp:318-381 (model 6: both wrapped when outside the box, ClassA).
// Threads walk particles in cell order for locality; sums follow the stencil order of :83-88.
template <int S>
__global__ __launch_bounds__(kBlock) void k_linked_list(DevParams P, const double *__restrict__ stage, int ncx,
                                                        int ncy, const int *__restrict__ start,
                                                        const int *__restrict__ sorted,
                                                        const double *__restrict__ nat, double *__restrict__ k_next,
                                                        double *__restrict__ compl_out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P.n) return;
  const int i = sorted[t];
  double st[S];
#pragma unroll
  for (int q = 0; q < S; ++q) st[q] = stage[(long)i * S + q];
  double x_i, y_i;
  key_wrap<S>(P, st[0], st[1], x_i, y_i);
  const double phi_i = st[Slots<S>::phi];
  const double omega_i = (S == 3) ? 0.0 : st[Slots<S>::omega];
  const int cx = cell_coord(x_i, ncx), cy = cell_coord(y_i, ncy);
  PairAcc a{0.0, 0.0, 0.0};
  for (int dyc = -1; dyc <= 1; ++dyc) {
    const int jy = adjust_cell(cy + dyc, ncy);
    for (int dxc = -1; dxc <= 1; ++dxc) {
      const int jx = adjust_cell(cx + dxc, ncx);
      const int c = jy * ncx + jx;
      const int e = start[c + 1];
      for (int u = start[c]; u < e; ++u) {
        const int j = sorted[u];
        if (j == i) continue;
        const double *sj = stage + (long)j * S;
        double x_j = sj[0], y_j = sj[1];
        if (S == 6) key_wrap<6>(P, x_j, y_j, x_j, y_j);  // :356-359
        pair_term<S>(P, x_i, y_i, phi_i, omega_i, x_j, y_j, sj[Slots<S>::phi],
                     (S == 3) ? 0.0 : sj[Slots<S>::omega], a);
      }
    }
  }
  write_k<S>(P, st, i, a, nat ? nat[i] : 0.0, k_next, compl_out);
}

// ---------------------------------------------------------------------------
// Verlet lists.  Build over all pairs (ParticleSystemFirstOrderPtr.cpp:497-543: x_i as stored)
// or from the cell stencil (:643-743: x_i wrapped).  Two passes: counts, then lists written in
// the reference's visiting order after a scan of the counts.

template <int S, bool FILL>
__global__ void k_verlet_build_all(DevParams P, const double *__restrict__ stage, double rmax2,
                                   int *__restrict__ vcount, const int *__restrict__ vstart,
                                   int *__restrict__ vlist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const double x_i = stage[(long)i * S], y_i = stage[(long)i * S + 1];
  int c = 0;
  int w = FILL ? vstart[i] : 0;
  for (int j = 0; j < P.n; ++j) {
    if (j == i) continue;
    double dx, dy;
    class_c_unsigned(P, x_i, y_i, stage[(long)j * S], stage[(long)j * S + 1], dx, dy);
    double dr2 = dx * dx + dy * dy;
    if (dr2 <= rmax2) {
      if (FILL) vlist[w++] = j;
      ++c;
    }
  }
  if (!FILL) vcount[i] = c;
}

template <int S, bool FILL>
__global__ void k_verlet_build_cells(DevParams P, const double *__restrict__ stage, int ncx, int ncy,
                                     const int *__restrict__ start, const int *__restrict__ sorted, double rmax2,
                                     int *__restrict__ vcount, const int *__restrict__ vstart,
                                     int *__restrict__ vlist) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double x_i, y_i;
  key_wrap<S>(P, stage[(long)i * S], stage[(long)i * S + 1], x_i, y_i);
  const int cx = cell_coord(x_i, ncx), cy = cell_coord(y_i, ncy);
  int cnt = 0;
  int w = FILL ? vstart[i] : 0;
  for (int dyc = -1; dyc <= 1; ++dyc) {
    const int jy = adjust_cell(cy + dyc, ncy);
    for (int dxc = -1; dxc <= 1; ++dxc) {
      const int c = jy * ncx + adjust_cell(cx + dxc, ncx);
      const int e = start[c + 1];
      for (int u = start[c]; u < e; ++u) {
        const int j = sorted[u];
        if (j == i) continue;
        double dx, dy;
        class_c_unsigned(P, x_i, y_i, stage[(long)j * S], stage[(long)j * S + 1], dx, dy);
        double dr2 = dx * dx + dy * dy;
        if (dr2 <= rmax2) {
          if (FILL) vlist[w++] = j;
          ++cnt;
        }
      }
    }
  }
  if (!FILL) vcount[i] = cnt;
}

// Verlet pair loop, ParticleSystemFirstOrderPtr.cpp:553-613 (x_i as stored, list order)
template <int S>
__global__ __launch_bounds__(kBlock) void k_verlet_pairs(DevParams P, const double *__restrict__ stage,
                                                         const int *__restrict__ vstart, const int *__restrict__ vlist,
                                                         const double *__restrict__ nat, double *__restrict__ k_next,
                                                         double *__restrict__ compl_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double st[S];
#pragma unroll
  for (int q = 0; q < S; ++q) st[q] = stage[(long)i * S + q];
  const double phi_i = st[Slots<S>::phi];
  const double omega_i = (S == 3) ? 0.0 : st[Slots<S>::omega];
  PairAcc a{0.0, 0.0, 0.0};
  const int e = vstart[i + 1];
  for (int u = vstart[i]; u < e; ++u) {
    const int j = vlist[u];
    if (j == i) continue;
    const double *sj = stage + (long)j * S;
    pair_term<S>(P, st[0], st[1], phi_i, omega_i, sj[0], sj[1], sj[Slots<S>::phi],
                 (S == 3) ? 0.0 : sj[Slots<S>::omega], a);
  }
  write_k<S>(P, st, i, a, nat ? nat[i] : 0.0, k_next, compl_out);
}

// ---------------------------------------------------------------------------
// global polar order field, ParticleSystemFirstOrderPtr.cpp:177-223.
// Sums (cos, sin, w cos, w sin) per block in a fixed tree, then in block order.

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// the final reductions: loads issued together per lane
constexpr int kFinalBatch = 16;

// the 64 compensated lane pairs (sum, correction) of a wave combined pairwise in a fixed tree
// (lane l with lane l + 32, then + 16, ...); lane 0 returns sum + correction
__device__ __forceinline__ double wave_combine(double s, double c) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double s2 = __shfl_down(s, off, 64), c2 = __shfl_down(c, off, 64);
    double t, e;
    two_sum(s, s2, t, e);
    s = t;
    c = c + c2 + e;
  }
  return s + c;
}

struct Sum4 {
  double v[4];
  double c[4];
};

__global__ __launch_bounds__(kBlock) void k_polar_partial(int n, int s, int slot_phi, int slot_omega,
                                                          const double *__restrict__ stage,
                                                          double *__restrict__ partial) {
  __shared__ double sh[4][kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < n) {
    double phi = stage[(long)i * s + slot_phi];
    double c = cos(phi), sn = sin(phi);
    v[0] = c;
    v[1] = sn;
    if (slot_omega >= 0) {
      double om = stage[(long)i * s + slot_omega];
      v[2] = om * c;
      v[3] = om * sn;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
#pragma unroll
      for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) partial[(long)blockIdx.x * 4 + q] = sh[q][0];
  }
}

// the four block-partial columns summed with compensation (two_sum), one wave per column: each
// lane takes every 64th block in block order, then the wave combines its 64 compensated pairs in
// a fixed tree; the sums are divided by n (the field Z and W).  The order is fixed, so the result
// does not vary between runs.
__global__ __launch_bounds__(256) void k_polar_final(int nblocks, const double *__restrict__ partial,
                                                     double *__restrict__ sums, double n) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s = 0.0, c = 0.0;
  // eight loads in flight per lane, then the sums in block order (the partials were written by
  // other CUs: one load at a time would wait out that latency once per block)
  for (int b0 = lane; b0 < nblocks; b0 += 64 * kFinalBatch) {
    double x[kFinalBatch];
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) {
      const int b = b0 + 64 * k;
      x[k] = b < nblocks ? partial[(long)b * 4 + q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) {
      if (b0 + 64 * k < nblocks) {
        double t, e;
        two_sum(s, x[k], t, e);
        s = t;
        c += e;
      }
    }
  }
  const double total = wave_combine(s, c);
  if (lane == 0) sums[q] = total / n;
}

// k of the global branch; fields = (Re Z, Im Z, Re W, Im W) already divided by N
template <int S>
__global__ void k_global_rhs(DevParams P, const double *__restrict__ stage, const double *__restrict__ fields,
                             const double *__restrict__ nat, double *__restrict__ k_next,
                             double *__restrict__ compl_out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const double absZ = hypot(fields[0], fields[1]), argZ = atan2(fields[1], fields[0]);
  const long ii = (long)i * S;
  const double phi_i = stage[ii + 2];
  const double nat_i = nat ? nat[i] : 0.0;
  if (S == 4) {
    const double omega_i = stage[ii + 3];
    k_next[ii] = P.v0 * cos(phi_i);
    k_next[ii + 1] = P.v0 * sin(phi_i);
    k_next[ii + 2] = omega_i;
    k_next[ii + 3] =
        -P.xi * omega_i + nat_i + P.sigma * absZ * (sin(argZ - phi_i) * P.cos_a - cos(argZ - phi_i) * P.sin_a);
    if (P.higher_order) {
      const double absW = hypot(fields[2], fields[3]), argW = atan2(fields[3], fields[2]);
      // :215-219, with arg(W) in the last sine as the reference has it
      compl_out[i] = P.sigma * (absW * (cos(argW - phi_i) * P.cos_a + sin(argW - phi_i) * P.sin_a) -
                                omega_i * absZ * (cos(argZ - phi_i) * P.cos_a + sin(argW - phi_i) * P.sin_a));
    }
  } else {
    k_next[ii] = P.v0 * cos(phi_i);
    k_next[ii + 1] = P.v0 * sin(phi_i);
    k_next[ii + 2] = nat_i + P.sigma * absZ * (sin(argZ - phi_i) * P.cos_a - cos(argZ - phi_i) * P.sin_a);
  }
}


// ---------------------------------------------------------------------------
// steppers.  Each writes the step's total displacement only as the per-block
// max of dx^2 + dy^2 that CalculateMaxDisplacement needs (:841-850).

__device__ __forceinline__ void block_max_store(double v, double *partial) {
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = sh[0];
}

// RungeKutta4StepperPtr.cpp:78-87
template <int S>
__global__ __launch_bounds__(kBlock) void k_rk4_combine(int n, double *__restrict__ state, const double *__restrict__ k1,
                                                        const double *__restrict__ k2, const double *__restrict__ k3,
                                                        const double *__restrict__ k4, double dt,
                                                        double *__restrict__ max_partial) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double m2 = 0.0;
  if (i < n) {
    double tot[S];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const long k = (long)i * S + q;
      tot[q] = (k1[k] + 2.0 * k2[k] + 2.0 * k3[k] + k4[k]) * dt / 6.0;
      state[k] += tot[q];
    }
    m2 = tot[0] * tot[0] + tot[1] * tot[1];
  }
  block_max_store(m2, max_partial);
}

__device__ __forceinline__ double draw(int mode, const double *noise, long idx, uint64_t seed, uint32_t particle,
                                       uint32_t which, uint64_t step) {
  if (mode == 1) return noise[idx];  // injected (host array already copied to the device)
  if (mode == 3) {
    double z0, z1;
    philox_normal2(seed, particle, which >> 1, step, z0, z1);
    return (which & 1) ? z1 : z0;
  }
  return 0.0;
}

// StochasticEulerStepperPtr.hpp:40-73 (model 4), StochasticEulerStepper.hpp:33-51 +
// ParticleSystemSecondOrder.cpp:439-451 (model 6, draws for slots 2, 3, 5), model 3 on phi.
template <int S>
__global__ __launch_bounds__(kBlock) void k_euler(int n, double *__restrict__ state, const double *__restrict__ f,
                                                  double dt, double sqrt_dt, double int_phi, double int_r,
                                                  int noise_mode, const double *__restrict__ noise, uint64_t seed,
                                                  uint64_t step, double *__restrict__ max_partial) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double m2 = 0.0;
  if (i < n) {
    const long ii = (long)i * S;
    if (S == 4) {
      double dW = sqrt_dt * draw(noise_mode, noise, i, seed, i, 0, step);
      double t0 = f[ii] * dt, t1 = f[ii + 1] * dt, t2 = f[ii + 2] * dt;
      double t3 = f[ii + 3] * dt + int_phi * dW;
      state[ii] += t0;
      state[ii + 1] += t1;
      state[ii + 2] += t2;
      state[ii + 3] += t3;
      m2 = t0 * t0 + t1 * t1;
    } else if (S == 3) {
      double dW = sqrt_dt * draw(noise_mode, noise, i, seed, i, 0, step);
      double t0 = f[ii] * dt, t1 = f[ii + 1] * dt;
      double t2 = f[ii + 2] * dt + int_phi * dW;
      state[ii] += t0;
      state[ii + 1] += t1;
      state[ii + 2] += t2;
      m2 = t0 * t0 + t1 * t1;
    } else {
      double g2 = int_r * draw(noise_mode, noise, 3L * i, seed, i, 0, step);
      double g3 = int_r * draw(noise_mode, noise, 3L * i + 1, seed, i, 1, step);
      double g5 = int_phi * draw(noise_mode, noise, 3L * i + 2, seed, i, 2, step);
      const double g[6] = {0.0, 0.0, g2, g3, 0.0, g5};
#pragma unroll
      for (int q = 0; q < 6; ++q) state[ii + q] += dt * f[ii + q] + sqrt_dt * g[q];
    }
  }
  block_max_store(m2, max_partial);
}

// StochasticRungeKuttaStepperPtr.hpp:50-76 (model 4).  c15 = 0.5*pow(dt,1.5), is3 = 1/sqrt(3),
// sq2d = sqrt(2*D_phi) come from the host so they round exactly as there.
__global__ __launch_bounds__(kBlock) void k_srk(int n, double *__restrict__ state, const double *__restrict__ f,
                                                const double *__restrict__ C, double dt, double sqrt_dt,
                                                double c15, double is3, double sq2d, double friction,
                                                int noise_mode, const double *__restrict__ noise, uint64_t seed,
                                                uint64_t step, double *__restrict__ max_partial) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double m2 = 0.0;
  if (i < n) {
    double U_1 = draw(noise_mode, noise, 2L * i, seed, i, 0, step);
    double U_2 = draw(noise_mode, noise, 2L * i + 1, seed, i, 1, step);
    double delta_W_i = U_1 * sqrt_dt;
    double delta_Z_i = c15 * (U_1 + is3 * U_2);
    const long ii = (long)i * 4;
    const double f0 = f[ii], f1 = f[ii + 1], f2 = f[ii + 2], f3 = f[ii + 3];
    double t0 = f0 * dt - 0.5 * f2 * f1 * dt * dt;
    double t1 = f1 * dt + 0.5 * f2 * f0 * dt * dt;
    double t2 = f2 * dt + 0.5 * f3 * dt * dt + sq2d * delta_Z_i;
    double t3 = f3 * dt + 0.5 * (C[i] - friction * f3) * dt * dt + sq2d * (delta_W_i - friction * delta_Z_i);
    state[ii] += t0;
    state[ii + 1] += t1;
    state[ii + 2] += t2;
    state[ii + 3] += t3;
    m2 = t0 * t0 + t1 * t1;
  }
  block_max_store(m2, max_partial);
}

// max of the block partials
__global__ void k_max_final(int nb, const double *__restrict__ partial, double *__restrict__ out) {
  __shared__ double sh[kBlock];
  double m = 0.0;
  for (int b0 = threadIdx.x; b0 < nb; b0 += kBlock * kFinalBatch) {
    double x[kFinalBatch];
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) x[k] = b0 + kBlock * k < nb ? partial[b0 + kBlock * k] : 0.0;
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) m = fmax(m, x[k]);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// max of dx^2+dy^2 over a total-displacement vector (CalculateMaxDisplacement on a given vector)
__global__ __launch_bounds__(kBlock) void k_disp_max(int n, int s, const double *__restrict__ tot,
                                                     double *__restrict__ max_partial) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  double m2 = 0.0;
  if (i < n) {
    double dx = tot[(long)i * s], dy = tot[(long)i * s + 1];
    m2 = dx * dx + dy * dy;
  }
  block_max_store(m2, max_partial);
}

// ---------------------------------------------------------------------------
// observers and utilities

// order parameter sums: per-block compensated (cos, sin) sums of state[s*i + slot]
__global__ __launch_bounds__(kBlock) void k_order_partial(int n, int s, int slot, const double *__restrict__ state,
                                                          double *__restrict__ partial) {
  __shared__ double sh[2][kBlock];
  __shared__ double shc[2][kBlock];
  int i = blockIdx.x * kBlock + threadIdx.x;
  double c = 0.0, sn = 0.0;
  if (i < n) {
    double phi = state[(long)i * s + slot];
    c = cos(phi);
    sn = sin(phi);
  }
  sh[0][threadIdx.x] = c;
  sh[1][threadIdx.x] = sn;
  shc[0][threadIdx.x] = 0.0;
  shc[1][threadIdx.x] = 0.0;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        double t, e;
        two_sum(sh[q][threadIdx.x], sh[q][threadIdx.x + off], t, e);
        sh[q][threadIdx.x] = t;
        shc[q][threadIdx.x] += shc[q][threadIdx.x + off] + e;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[(long)blockIdx.x * 4 + 0] = sh[0][0];
    partial[(long)blockIdx.x * 4 + 1] = sh[1][0];
    partial[(long)blockIdx.x * 4 + 2] = shc[0][0];
    partial[(long)blockIdx.x * 4 + 3] = shc[1][0];
  }
}

// the (cos, sin) block partials summed with compensation, one wave per column: each lane takes
// every 64th block, then the wave combines its 64 compensated pairs in a fixed tree
__global__ __launch_bounds__(128) void k_order_final(int nb, const double *__restrict__ partial, double n,
                                                     double *__restrict__ out) {
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double s = 0.0, c = 0.0;
  for (int b0 = lane; b0 < nb; b0 += 64 * kFinalBatch) {
    double x[kFinalBatch], y[kFinalBatch];
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) {
      const int b = b0 + 64 * k;
      x[k] = b < nb ? partial[(long)b * 4 + q] : 0.0;
      y[k] = b < nb ? partial[(long)b * 4 + 2 + q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kFinalBatch; ++k) {
      if (b0 + 64 * k < nb) {
        double t, e;
        two_sum(s, x[k], t, e);
        s = t;
        c += e + y[k];
      }
    }
  }
  const double total = wave_combine(s, c);
  if (lane == 0) out[q] = total / n;
}

// ApplyPeriodicBoundaryConditions(Real*, long), PeriodicBoundaryConditions.cpp:147-155
__global__ void k_apply_pbc(int n, int s, double *__restrict__ state, double xs, double ys, double xrs, double yrs) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long ii = (long)i * s;
  state[ii] -= floor(state[ii] * xrs) * xs;
  state[ii + 1] -= floor(state[ii + 1] * yrs) * ys;
}

// AveragingObserver::operator(), AveragingObserver.cpp:29-37
__global__ void k_accumulate(long m, const double *__restrict__ state, double *__restrict__ acc) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) acc[k] = state[k] + acc[k];
}

__global__ void k_fill_double(long m, double *__restrict__ p, double v) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) p[k] = v;
}

// plain streaming copy, one 16-byte vector per thread: measures attainable HBM bandwidth and
// calibrates the FETCH_SIZE / WRITE_SIZE counters against a kernel of known traffic.  Without a
// grid-stride loop every load is issued as soon as its wave starts, which keeps more bytes in
// flight than a loop over a fixed grid (on MI355X at 1 GiB: ~6.2 TB/s against ~4.6 TB/s for 4096
// workgroups striding)
__global__ __launch_bounds__(256) void k_copy16(long n16, const uint4 *__restrict__ src, uint4 *__restrict__ dst) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n16) dst[k] = src[k];
}

__global__ void k_fill_int(long m, int *__restrict__ p, int v) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) p[k] = v;
}

// ---------------------------------------------------------------------------
// host-side launchers

static inline unsigned nblk(long m, int b = kBlock) { return (unsigned)((m + b - 1) / b); }

void launch_stage(hipStream_t st, long m, const double *state, const double *kprev, double kc_dt, double *stage) {
  if (kprev)
    k_stage<<<dim3(nblk(m)), dim3(kBlock), 0, st>>>(m, state, kprev, kc_dt, stage);
  else
    k_stage_zero<<<dim3(nblk(m)), dim3(kBlock), 0, st>>>(m, state, stage);
}

#define SPSS_DISPATCH_S(S, ...) \
  switch (S) {                   \
    case 3: { constexpr int SS = 3; __VA_ARGS__; } break; \
    case 4: { constexpr int SS = 4; __VA_ARGS__; } break; \
    case 6: { constexpr int SS = 6; __VA_ARGS__; } break; \
    default: break;              \
  }

const int kAllPairsConfigs[kAllPairsConfigCount][2] = {{64, 64}, {64, 256}, {128, 128}, {128, 512},
                                                        {256, 256}, {256, 512}};

template <int BLOCK, int JT>
static void all_pairs_as(hipStream_t st, const DevParams &P, const double *stage, const double *nat, double *k_next,
                         double *compl_out) {
  const unsigned nb = (unsigned)((P.n + BLOCK - 1) / BLOCK);
  SPSS_DISPATCH_S(P.s, k_all_pairs<SS, BLOCK, JT><<<dim3(nb), dim3(BLOCK), 0, st>>>(P, stage, nat, k_next, compl_out));
}

void launch_all_pairs(hipStream_t st, const DevParams &P, const double *stage, const double *nat, double *k_next,
                      double *compl_out, int config) {
  switch (config) {
    case 0: all_pairs_as<64, 64>(st, P, stage, nat, k_next, compl_out); break;
    case 1: all_pairs_as<64, 256>(st, P, stage, nat, k_next, compl_out); break;
    case 2: all_pairs_as<128, 128>(st, P, stage, nat, k_next, compl_out); break;
    case 3: all_pairs_as<128, 512>(st, P, stage, nat, k_next, compl_out); break;
    case 5: all_pairs_as<256, 512>(st, P, stage, nat, k_next, compl_out); break;
    default: all_pairs_as<256, 256>(st, P, stage, nat, k_next, compl_out); break;
  }
}

void launch_cell_build(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy, CellBuffers &cb) {
  const int m = ncx * ncy;
  SPSS_DISPATCH_S(P.s, k_cell_keys<SS><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, ncx, ncy, cb.key, cb.slot, cb.count));
  launch_scan(st, cb.count, cb.start, m, cb.blocksum);
  k_scatter<<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P.n, cb.key, cb.slot, cb.start, cb.tmp);
  k_stable_rank<<<dim3(std::max((m + 3) / 4, (int)nblk(P.n))), dim3(256), 0, st>>>(P.n, m, cb.key, cb.start, cb.tmp,
                                                                                 cb.sorted);
}

void launch_linked_list(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                        const CellBuffers &cb, const double *nat, double *k_next, double *compl_out) {
  SPSS_DISPATCH_S(P.s, k_linked_list<SS><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, ncx, ncy, cb.start, cb.sorted, nat, k_next, compl_out));
}

void launch_scan(hipStream_t st, int *count, int *start, int m, int *blocksum) {
  const int nb = scan_blocks(m);
  scan_local<<<dim3(nb), dim3(256), 0, st>>>(count, nullptr, m, start, blocksum);
  scan_add<<<dim3(nb), dim3(256), 0, st>>>(count, nullptr, m, start, blocksum, nb, nullptr);
}

void launch_verlet_count_all(hipStream_t st, const DevParams &P, const double *stage, double rmax2, int *vcount) {
  SPSS_DISPATCH_S(P.s, k_verlet_build_all<SS, false><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, rmax2, vcount, nullptr, nullptr));
}
void launch_verlet_fill_all(hipStream_t st, const DevParams &P, const double *stage, double rmax2,
                            const int *vstart, int *vlist) {
  SPSS_DISPATCH_S(P.s, k_verlet_build_all<SS, true><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, rmax2, nullptr, vstart, vlist));
}
void launch_verlet_count_cells(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                               const CellBuffers &cb, double rmax2, int *vcount) {
  SPSS_DISPATCH_S(P.s, k_verlet_build_cells<SS, false><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, ncx, ncy, cb.start, cb.sorted, rmax2, vcount, nullptr, nullptr));
}
void launch_verlet_fill_cells(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                              const CellBuffers &cb, double rmax2, const int *vstart, int *vlist) {
  SPSS_DISPATCH_S(P.s, k_verlet_build_cells<SS, true><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, ncx, ncy, cb.start, cb.sorted, rmax2, nullptr, vstart, vlist));
}
void launch_verlet_pairs(hipStream_t st, const DevParams &P, const double *stage, const int *vstart,
                         const int *vlist, const double *nat, double *k_next, double *compl_out) {
  SPSS_DISPATCH_S(P.s, k_verlet_pairs<SS><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, vstart, vlist, nat, k_next, compl_out));
}

void launch_global(hipStream_t st, const DevParams &P, const double *stage, double *partial, double *fields,
                   const double *nat, double *k_next, double *compl_out) {
  const int s = P.s;
  const int nb = (int)nblk(P.n);
  k_polar_partial<<<dim3(nb), dim3(kBlock), 0, st>>>(P.n, s, Slots<4>::phi, (s == 4) ? 3 : -1, stage, partial);
  k_polar_final<<<dim3(1), dim3(256), 0, st>>>(nb, partial, fields, (double)P.n);
  if (s == 4)
    k_global_rhs<4><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, fields, nat, k_next, compl_out);
  else
    k_global_rhs<3><<<dim3(nblk(P.n)), dim3(kBlock), 0, st>>>(P, stage, fields, nat, k_next, compl_out);
}

void launch_rk4_combine(hipStream_t st, int s, int n, double *state, const double *k1, const double *k2,
                        const double *k3, const double *k4, double dt, double *max_partial) {
  SPSS_DISPATCH_S(s, k_rk4_combine<SS><<<dim3(nblk(n)), dim3(kBlock), 0, st>>>(n, state, k1, k2, k3, k4, dt, max_partial));
}

void launch_euler(hipStream_t st, int s, int n, double *state, const double *f, double dt, double sqrt_dt,
                  double int_phi, double int_r, int noise_mode, const double *noise, uint64_t seed, uint64_t step,
                  double *max_partial) {
  SPSS_DISPATCH_S(s, k_euler<SS><<<dim3(nblk(n)), dim3(kBlock), 0, st>>>(n, state, f, dt, sqrt_dt, int_phi, int_r, noise_mode, noise, seed, step, max_partial));
}

void launch_srk(hipStream_t st, int n, double *state, const double *f, const double *C, double dt, double sqrt_dt,
                double c15, double is3, double sq2d, double friction, int noise_mode, const double *noise,
                uint64_t seed, uint64_t step, double *max_partial) {
  k_srk<<<dim3(nblk(n)), dim3(kBlock), 0, st>>>(n, state, f, C, dt, sqrt_dt, c15, is3, sq2d, friction, noise_mode, noise, seed, step, max_partial);
}

void launch_max_final(hipStream_t st, int nb, const double *partial, double *out) {
  k_max_final<<<dim3(1), dim3(kBlock), 0, st>>>(nb, partial, out);
}

void launch_disp_max(hipStream_t st, int n, int s, const double *tot, double *partial) {
  k_disp_max<<<dim3(nblk(n)), dim3(kBlock), 0, st>>>(n, s, tot, partial);
}

void launch_order_parameter(hipStream_t st, int n, int s, int slot, const double *state, double *partial,
                            double *out) {
  const int nb = (int)nblk(n);
  k_order_partial<<<dim3(nb), dim3(kBlock), 0, st>>>(n, s, slot, state, partial);
  k_order_final<<<dim3(1), dim3(128), 0, st>>>(nb, partial, (double)n, out);
}

void launch_apply_pbc(hipStream_t st, int n, int s, double *state, double xs, double ys, double xrs, double yrs) {
  k_apply_pbc<<<dim3(nblk(n)), dim3(kBlock), 0, st>>>(n, s, state, xs, ys, xrs, yrs);
}

void launch_accumulate(hipStream_t st, long m, const double *state, double *acc) {
  k_accumulate<<<dim3(nblk(m)), dim3(kBlock), 0, st>>>(m, state, acc);
}

void launch_fill_double(hipStream_t st, long m, double *p, double v) {
  if (m > 0) k_fill_double<<<dim3(nblk(m)), dim3(kBlock), 0, st>>>(m, p, v);
}

void launch_copy16(hipStream_t st, long n16, const void *src, void *dst) {
  if (n16 > 0)
    k_copy16<<<dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, st>>>(n16, (const uint4 *)src, (uint4 *)dst);
}

void launch_fill_int(hipStream_t st, long m, int *p, int v) {
  if (m > 0) k_fill_int<<<dim3(nblk(m)), dim3(kBlock), 0, st>>>(m, p, v);
}

}  // namespace spss
