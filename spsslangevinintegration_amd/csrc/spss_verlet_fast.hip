// spss_verlet_fast.hip -- Verlet neighbour lists in FP32C precision (exact mode): models 3 and 4,
// stochastic Euler and Runge-Kutta steps, rebuilds decided on the device.
//
// Reference: EvaluateInteractionsWithVerletNeighborList and ...WithLinkedListAndVerletNeighborList
// (ParticleSystemFirstOrderPtr.cpp:470-815) with CalculateMaxDisplacement (:836-861), in this
// library's exact mode: r_max = rho + skin, candidates from cells of side >= r_max, so every list
// holds every j within r_max and each step's neighbour set is {j : |r_ij| <= rho} whatever the
// rebuild schedule.  The schedule itself follows the exact mode of the FP64 path: after every
// step acc += sqrt(max_i |dr_i|^2), rebuild when acc + that step's value > (r_max - rho) / 2.
//
// Layout (the resident copy): particles in the cell order of the last rebuild on a grid of
// side Hv = 1/int(1/r_max); between rebuilds they keep their slots.  Per slot: FP64 state
// (double-buffered by step), id, natural frequency, FP32 snapshot {x, y wrapped into the box,
// cos phi, sin phi} and FP32 omega (double-buffered with the state), and an ELL list of width K
// with its length, in blocks of 256 slots and chunks of four entries (lslot).  Grids of at most
// kRowTileCols columns also have row tiles of the rebuild's cell order (spans of whole cells of
// one row, as for the cell-list pair kernel) and a 16-bit copy of every list: the entries' LDS
// positions in the tile's stage, tile-local in the same chunked form (l16chunk).  Either way a
// wave's loads of one chunk are contiguous and a particle's entries are not on pages of their own.
//
// One step is a rebuild -- launches that return at once unless the device flag asks for one:
// cell keys and histogram, the two-level scan, bucket fill, stable rank and gather into scratch,
// back into the resident arrays with the new snapshot, the row tiles, the lists -- and the step:
// one workgroup per row tile stages the snapshots (and omega) of the 3 x (W + 2) cells around its
// W cells into LDS; each lane walks its list from LDS, sixteen entries per pass (the next pass's
// entries loaded meanwhile): FP32 distance (the stage holds each position as its periodic image
// nearest its staged cell, so separations are plain differences; the walks through global memory
// take x_j - x_i minus its nearest integer), a guard band around rho^2 decided
// with the reference's FP64 ClassC formula on the stored positions (so neighbour sets are exact),
// phasor sums in four FP32 partials flushed into FP64 every 16 entries; then the Euler/SRK update
// of step_particle, the new snapshot and the step's largest squared displacement (maxima by
// group of workgroups); a one-wave launch after the step (vc_flag) turns it into the next step's
// flag on the device.  No host synchronisation: steps replay from HIP graphs.  Without tiles
// (wider grids), vr_lists and vs_step do the same per particle through global memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/spss_hip.h"
#include "spss_launch.hpp"
#include "spss_scan.hpp"
#include "spss_step.hpp"
#include "spss_tiles.hpp"
#include "spss_verlet_fast.hpp"

namespace spss {

namespace {
thread_local std::string g_verr;
constexpr int kVB = 256;         // threads per workgroup
constexpr int kVGraphSteps = 10;  // steps of the bulk graph (even: the state buffers end where they started)
constexpr int kGateGrid = 1024;   // workgroups of the launches that run on rebuilds (and vt_rest) only
}  // namespace

// device control block
constexpr int kGroups = 64, kLineU64 = 16;  // workgroup groups of the displacement maximum, a 128-byte line each
struct VCtl {
  int flag;                     // rebuild before the next step (should_update_lists_)
  int pad0;
  double acc;                   // accumulated_displacement_
  unsigned long long overflow;  // lists longer than the width K since entry (walked through their cells)
  unsigned long long step;      // Philox step of graph-replayed steps
  int big_tiles;                // tiles of the last rebuild holding more particles than a workgroup
  int pad1;
  // the step's largest |dr|^2 by group of workgroups (workgroup b in group b % kGroups), as the
  // bits of a non-negative double, one word per cache line so that the groups' atomics run apart
  unsigned long long gmax[kGroups * kLineU64];
};

struct VArgs {
  int n, ncv, cells, K;
  double rmax2;           // FP64 list radius squared
  double half_skin;       // (r_max - rho) / 2
  float r2c, r2h, r2k;    // guard band around rho^2: |d^2 - r2c| <= r2h is decided in FP64
  float m2c, m2h;         // guard band around r_max^2 (list build)
  VCtl *ctl;
};

__device__ __forceinline__ bool gate_off(const VArgs &A) { return A.ctl->flag == 0; }

// device addresses of fields of the control block, formed on the host without dereferencing it
static int *ctl_flag(VCtl *c) { return reinterpret_cast<int *>(reinterpret_cast<char *>(c) + offsetof(VCtl, flag)); }
static unsigned long long *ctl_step(VCtl *c) {
  return reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(c) + offsetof(VCtl, step));
}

// snapshot entry: wrapped position and unit phasor, each computed in FP64 and rounded once
__device__ __forceinline__ float4 vsnap_of(const DevParams &P, const double *st) {
  double xw, yw, sn, cn;
  apply_pbc(P, st[0], st[1], xw, yw);
  sincos(st[2], &sn, &cn);
  return make_float4((float)xw, (float)yw, (float)cn, (float)sn);
}

// FP32 squared minimum-image distance minus c2
__device__ __forceinline__ float dd_of(float4 o, float xi, float yi, float c2) {
  float dx = o.x - xi;
  dx -= __builtin_rintf(dx);
  float dy = o.y - yi;
  dy -= __builtin_rintf(dy);
  return __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, -c2));
}

// entry t of slot p in the ELL lists of width K (a multiple of 16): blocks of 256 slots, inside a
// block chunks of 4 consecutive entries of one slot, slot-major (the chunks of entries 4c..4c+3 of
// the block's 256 slots side by side).  A builder writes a slot's entries four at a time, the loads
// of a chunk by a wave's lanes are one contiguous run, and a particle's entries lie 4 KB apart
// (with one n-long array per entry they would each sit on a page of their own).
__host__ __device__ __forceinline__ long lslot(long p, long t, int K) {
  return ((((p >> 8) * (K >> 2) + (t >> 2)) * 256 + (p & 255)) << 2) + (t & 3);
}
// the 16-bit copy, tile-local in the same chunked form: chunk c of the tile's lane u at
// (tile * K / 4 + c) * 256 + u, four 16-bit entries (uint2) each
__device__ __forceinline__ long l16chunk(long tile, int u, int c, int K) { return (tile * (K >> 2) + c) * 256 + u; }

// ---------------------------------------------------------------------------
// rebuild (every launch returns at once when the flag is clear)

template <int S>
__global__ __launch_bounds__(kVB) void vr_keys(DevParams P, VArgs A, const double *__restrict__ state,
                                               int *__restrict__ key, int *__restrict__ slotc,
                                               int *__restrict__ count) {
  if (gate_off(A)) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) A.ctl->big_tiles = 0;  // counted again by vt_lists
  // the particles are in the last rebuild's cell order and have moved less than half the skin
  // since, so a wave holds a few runs of equal keys: one atomic per run, the run's lanes take
  // consecutive places after it (the order inside a cell is settled by vr_gather)
  const int lane = threadIdx.x & 63;
  const unsigned long long below = (1ull << lane) - 1ull;
  const int nwave = (A.n + 63) & ~63;  // every lane of a wave takes the same trips
  for (int p = blockIdx.x * kVB + threadIdx.x; p < nwave; p += gridDim.x * kVB) {
    const bool valid = p < A.n;
    int c = -1;
    if (valid) {
      double xw, yw;
      apply_pbc(P, state[(long)p * S], state[(long)p * S + 1], xw, yw);
      c = cell_coord(yw, A.ncv) * A.ncv + cell_coord(xw, A.ncv);
      key[p] = c;
    }
    const int pc = __shfl_up(c, 1, 64);
    const unsigned long long heads = __ballot(lane == 0 || pc != c);
    const int seg = 63 - __clzll((long long)(heads & (below | (1ull << lane))));  // this run's first lane
    int base = 0;
    if (valid && seg == lane) {
      const unsigned long long later = lane == 63 ? 0ull : heads >> (lane + 1);
      const int len = later ? __ffsll((long long)later) : 64 - lane;
      base = atomicAdd(count + c, len);
    }
    base = __shfl(base, seg, 64);
    if (valid) slotc[p] = base + (lane - seg);
  }
}

__global__ __launch_bounds__(kVB) void vr_scatter(VArgs A, const int *__restrict__ key, const int *__restrict__ slotc,
                                                  const int *__restrict__ start, int *__restrict__ tmp) {
  if (gate_off(A)) return;
  for (int p = blockIdx.x * kVB + threadIdx.x; p < A.n; p += gridDim.x * kVB) tmp[start[key[p]] + slotc[p]] = p;
}

// each cell's particles in ascending previous slot, gathered into the new order in scratch: cells
// of at most 64 particles one wave each (wave_sort64), then the particles of denser cells one
// thread each, ranked by counting (grid-stride over the particles: a dense cluster is spread over
// the whole grid)
template <int S>
__global__ __launch_bounds__(kVB) void vr_gather(VArgs A, const int *__restrict__ key, const int *__restrict__ start,
                                                 const int *__restrict__ tmp, const double *__restrict__ state,
                                                 const int *__restrict__ id, const double *__restrict__ nat,
                                                 double *__restrict__ sstate, int *__restrict__ sid,
                                                 double *__restrict__ snat, int *__restrict__ scell) {
  if (gate_off(A)) return;
  const int lane = threadIdx.x & 63;
  auto move = [&](int p, int q, int c) {
#pragma unroll
    for (int k = 0; k < S; ++k) sstate[(long)q * S + k] = state[(long)p * S + k];
    sid[q] = id[p];
    if (nat) snat[q] = nat[p];
    scell[q] = c;
  };
  for (int c = blockIdx.x * (kVB / 64) + (threadIdx.x >> 6); c < A.cells; c += gridDim.x * (kVB / 64)) {
    const int b = start[c], m = start[c + 1] - b;
    if (m > 64) continue;  // the whole wave
    const int v = wave_sort64(lane < m ? tmp[b + lane] : 0x7fffffff, lane);
    if (lane < m) move(v, b + lane, c);
  }
  for (int p = blockIdx.x * kVB + threadIdx.x; p < A.n; p += gridDim.x * kVB) {
    const int c = key[p];
    const int b = start[c], e = start[c + 1];
    if (e - b <= 64) continue;
    int r = 0;
    for (int t = b; t < e; ++t) r += tmp[t] < p ? 1 : 0;
    move(p, b + r, c);
  }
}

// back into the resident arrays, with the snapshot of the new order
template <int S>
__global__ __launch_bounds__(kVB) void vr_settle(DevParams P, VArgs A, const double *__restrict__ sstate,
                                                 const int *__restrict__ sid, const double *__restrict__ snat,
                                                 const int *__restrict__ scell, double *__restrict__ state,
                                                 int *__restrict__ id, double *__restrict__ nat,
                                                 int *__restrict__ cell, float4 *__restrict__ snap,
                                                 float *__restrict__ omega, const int *__restrict__ start,
                                                 int4 *__restrict__ tiles, int wcap, int cmax) {
  if (gate_off(A)) return;
  // the row tiles of the new order (tiled grids), one wave per row in the first workgroups: they
  // need only the final cell starts
  if (tiles && (int)blockIdx.x * 4 < A.ncv) rows_to_tiles([&](int c) { return start[c]; }, A.ncv, wcap, cmax, tiles);
  for (int q = blockIdx.x * kVB + threadIdx.x; q < A.n; q += gridDim.x * kVB) {
    double st[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      st[k] = sstate[(long)q * S + k];
      state[(long)q * S + k] = st[k];
    }
    id[q] = sid[q];
    if (nat) nat[q] = snat[q];
    cell[q] = scell[q];
    snap[q] = vsnap_of(P, st);
    omega[q] = S == 4 ? (float)st[3] : 0.f;
  }
}

// The lists: the candidates of the 3 x 3 cells around the particle's cell in the stencil order of
// ParticleSystemFirstOrderPtr.cpp:83-88, ascending slot inside a cell; j is listed when its
// distance from the wrapped x_i to the stored x_j is within r_max (:694-739).  FP32 test outside a
// guard band, the reference's FP64 ClassC formula inside it.
// The list of slot q read from global memory (lists of a grid too wide for row tiles, and of tiles
// whose stencil cells overflow the LDS stage).
template <int S>
__device__ __forceinline__ void list_of_slot(const DevParams &P, const VArgs &A, const int *__restrict__ start,
                                             const int *__restrict__ cell, const double *__restrict__ state,
                                             const float4 *__restrict__ snap, int *__restrict__ list,
                                             int *__restrict__ len, int q) {
  const int c = cell[q];
  const int cx = c % A.ncv, cy = c / A.ncv;
  const float4 me = snap[q];
  double xi, yi;
  apply_pbc(P, state[(long)q * S], state[(long)q * S + 1], xi, yi);
  int w = 0;
  for (int dyc = -1; dyc <= 1; ++dyc) {
    const int jy = adjust_cell(cy + dyc, A.ncv);
    for (int dxc = -1; dxc <= 1; ++dxc) {
      const int cj = jy * A.ncv + adjust_cell(cx + dxc, A.ncv);
      const int e = start[cj + 1];
      for (int u = start[cj]; u < e; ++u) {
        if (u == q) continue;
        const float dd = dd_of(snap[u], me.x, me.y, A.m2c);
        bool in = dd < 0.f;
        if (__builtin_fabsf(dd) <= A.m2h) {
          double dx, dy;
          class_c_unsigned(P, xi, yi, state[(long)u * S], state[(long)u * S + 1], dx, dy);
          in = dx * dx + dy * dy <= A.rmax2;
        }
        if (in) {
          if (w < A.K) list[lslot(q, w, A.K)] = u;
          ++w;
        }
      }
    }
  }
  // a list longer than the width is not stored: the step walks that particle's 3 x 3 cells instead
  len[q] = w <= A.K ? w : -1;
  if (w > A.K) atomicAdd(&A.ctl->overflow, 1ull);
}

template <int S>
__global__ __launch_bounds__(kVB) void vr_lists(DevParams P, VArgs A, const int *__restrict__ start,
                                                const int *__restrict__ cell, const double *__restrict__ state,
                                                const float4 *__restrict__ snap, int *__restrict__ list,
                                                int *__restrict__ len) {
  if (gate_off(A)) return;
  const int q = blockIdx.x * kVB + threadIdx.x;
  if (q < A.n) list_of_slot<S>(P, A, start, cell, state, snap, list, len, q);
}

// ---------------------------------------------------------------------------
// the step

// Four candidates against particle p: weights, phasor partials (and omega-weighted for SRK).  The
// candidates are list entries t .. t+3 of slot p (CELLS = false) or the slots t .. t+3 of a cell of
// the last rebuild (CELLS = true: a particle whose list outgrew the width; p itself is skipped).
// TAIL: only those below `end` count.
template <int S, bool SRK, bool TAIL, bool CELLS>
__device__ __forceinline__ void walk4(const DevParams &P, const VArgs &A, const int *__restrict__ list,
                                      const float4 *__restrict__ snap, const float *__restrict__ omega,
                                      const double *__restrict__ state, int p, int t, int end, float xi, float yi,
                                      const double *pi, float *ax, float *ay, float &an, float *bx, float *by) {
  int j[4];
  float4 o[4];
  float dd[4], m[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool in = !TAIL || t + k < end;
    j[k] = CELLS ? (in ? t + k : p) : (in ? list[lslot(p, t + k, A.K)] : p);
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    o[k] = snap[j[k]];
    dd[k] = dd_of(o[k], xi, yi, A.r2c);
    m[k] = __builtin_amdgcn_fmed3f(-A.r2k * dd[k], 0.f, 1.f);
  }
  const float dmin = __builtin_fminf(__builtin_fminf(__builtin_fabsf(dd[0]), __builtin_fabsf(dd[1])),
                                     __builtin_fminf(__builtin_fabsf(dd[2]), __builtin_fabsf(dd[3])));
  if (dmin <= A.r2h) {  // rare: the reference's FP64 test (x_i as stored, rhs_verlet's :566-581)
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (__builtin_fabsf(dd[k]) <= A.r2h) {
        double dx, dy;
        class_c_unsigned(P, pi[0], pi[1], state[(long)j[k] * S], state[(long)j[k] * S + 1], dx, dy);
        m[k] = dx * dx + dy * dy <= P.rho2 ? 1.f : 0.f;
      }
  }
  if (TAIL || CELLS) {
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = ((!TAIL || t + k < end) && (!CELLS || j[k] != p)) ? m[k] : 0.f;
  }
  an += (m[0] + m[1]) + (m[2] + m[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    ax[k] = __builtin_fmaf(m[k], o[k].z, ax[k]);
    ay[k] = __builtin_fmaf(m[k], o[k].w, ay[k]);
    if (SRK) {
      const float mo = m[k] * omega[j[k]];
      bx[k] = __builtin_fmaf(mo, o[k].z, bx[k]);
      by[k] = __builtin_fmaf(mo, o[k].w, by[k]);
    }
  }
}

// candidates [a, b) (list entries, or cell slots with CELLS) in groups of four, four FP32 partials per
// sum flushed into FP64 every 16 candidates
template <int S, bool SRK, bool CELLS>
__device__ __forceinline__ void walk_range(const DevParams &P, const VArgs &A, const int *__restrict__ list,
                                           const float4 *__restrict__ snap, const float *__restrict__ omega,
                                           const double *__restrict__ state, int p, int a, int b, float xi, float yi,
                                           const double *pi, double &sx, double &sy, double &tx, double &ty,
                                           int &nnb) {
  for (int t0 = a; t0 < b; t0 += 16) {
    float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, an = 0.f;
    float bx[4] = {0.f, 0.f, 0.f, 0.f}, by[4] = {0.f, 0.f, 0.f, 0.f};
    const int te = min(t0 + 16, b);
    int t = t0;
    for (; t + 4 <= te; t += 4)
      walk4<S, SRK, false, CELLS>(P, A, list, snap, omega, state, p, t, te, xi, yi, pi, ax, ay, an, bx, by);
    if (t < te) walk4<S, SRK, true, CELLS>(P, A, list, snap, omega, state, p, t, te, xi, yi, pi, ax, ay, an, bx, by);
    sx += ((double)ax[0] + (double)ax[1]) + ((double)ax[2] + (double)ax[3]);
    sy += ((double)ay[0] + (double)ay[1]) + ((double)ay[2] + (double)ay[3]);
    nnb += (int)an;
    if (SRK) {
      tx += ((double)bx[0] + (double)bx[1]) + ((double)bx[2] + (double)bx[3]);
      ty += ((double)by[0] + (double)by[1]) + ((double)by[2] + (double)by[3]);
    }
  }
}

// The list entries [0, L) of slot p, sixteen per pass, in the order and 4-slot grouping of
// walk_range: each half of a pass issues its eight list loads, then its eight gathers of the
// snapshot (and omega), before any arithmetic, so a wave keeps eight dependent load pairs in
// flight instead of four; entries past L are the slot itself with weight zero.
template <int S, bool SRK>
__device__ __forceinline__ void walk_list(const DevParams &P, const VArgs &A, const int *__restrict__ list,
                                          const float4 *__restrict__ snap, const float *__restrict__ omega,
                                          const double *__restrict__ state, int p, int L, float xi, float yi,
                                          const double *pi, double &sx, double &sy, double &tx, double &ty,
                                          int &nnb) {
  for (int t0 = 0; t0 < L; t0 += 16) {
    float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, an = 0.f;
    float bx[4] = {0.f, 0.f, 0.f, 0.f}, by[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int h = 0; h < 16; h += 8) {
      int j[8];
      float4 o[8];
      float om[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) j[k] = t0 + h + k < L ? list[lslot(p, t0 + h + k, A.K)] : p;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        o[k] = snap[j[k]];
        if (SRK) om[k] = omega[j[k]];
      }
#pragma unroll
      for (int g = 0; g < 8; g += 4) {
        float dd[4], m[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          dd[k] = dd_of(o[g + k], xi, yi, A.r2c);
          m[k] = __builtin_amdgcn_fmed3f(-A.r2k * dd[k], 0.f, 1.f);
        }
        const float dmin = __builtin_fminf(__builtin_fminf(__builtin_fabsf(dd[0]), __builtin_fabsf(dd[1])),
                                           __builtin_fminf(__builtin_fabsf(dd[2]), __builtin_fabsf(dd[3])));
        if (dmin <= A.r2h) {  // rare: the reference's FP64 test, as in walk4
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (__builtin_fabsf(dd[k]) <= A.r2h) {
              double dx, dy;
              class_c_unsigned(P, pi[0], pi[1], state[(long)j[g + k] * S], state[(long)j[g + k] * S + 1], dx, dy);
              m[k] = dx * dx + dy * dy <= P.rho2 ? 1.f : 0.f;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) m[k] = t0 + h + g + k < L ? m[k] : 0.f;
        an += (m[0] + m[1]) + (m[2] + m[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          ax[k] = __builtin_fmaf(m[k], o[g + k].z, ax[k]);
          ay[k] = __builtin_fmaf(m[k], o[g + k].w, ay[k]);
          if (SRK) {
            const float mo = m[k] * om[g + k];
            bx[k] = __builtin_fmaf(mo, o[g + k].z, bx[k]);
            by[k] = __builtin_fmaf(mo, o[g + k].w, by[k]);
          }
        }
      }
    }
    sx += ((double)ax[0] + (double)ax[1]) + ((double)ax[2] + (double)ax[3]);
    sy += ((double)ay[0] + (double)ay[1]) + ((double)ay[2] + (double)ay[3]);
    nnb += (int)an;
    if (SRK) {
      tx += ((double)bx[0] + (double)bx[1]) + ((double)bx[2] + (double)bx[3]);
      ty += ((double)by[0] + (double)by[1]) + ((double)by[2] + (double)by[3]);
    }
  }
}

// workgroups in an XCD-aware order: the hardware deals consecutive workgroups to the 8 XCDs in
// turn, so workgroup b takes the particle block (b % 8) * per + b / 8 and every XCD walks one
// contiguous eighth of the slots (neighbours' snapshots then stay in that XCD's L2); the grid is
// a multiple of 8 and blocks past the slots only take part in the displacement maximum
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  const int per = nblocks >> 3;
  return (b & 7) * per + (b >> 3);
}

// The step's largest squared displacement, for vc_flag: wave, workgroup, then one atomic per
// workgroup on the maximum of its group (kGroups words on their own cache lines).  (A last-workgroup
// count in the step kernel needs a release fence per workgroup, which writes the XCD's L2 back:
// 0.45 ms per step at N = 1M with 7,000 workgroups.)
__device__ __forceinline__ void step_max(const VArgs &A, double d2, double *wred) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) d2 = fmax(d2, __shfl_xor(d2, o, 64));
  if ((threadIdx.x & 63) == 0) wred[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = wred[0];
    for (int w = 1; w < kVB / 64; ++w) bm = fmax(bm, wred[w]);
    if (bm > 0.0)
      atomicMax(&A.ctl->gmax[(blockIdx.x & (kGroups - 1)) * kLineU64], (unsigned long long)__double_as_longlong(bm));
  }
}

template <int S, bool SRK>
__global__ __launch_bounds__(kVB) void vs_step(DevParams P, VArgs A, StepArgs U, const int *__restrict__ list,
                                               const int *__restrict__ len, const int *__restrict__ start,
                                               const int *__restrict__ cell, const double *__restrict__ state,
                                               const int *__restrict__ id, const double *__restrict__ nat,
                                               const float4 *__restrict__ snap, const float *__restrict__ omega,
                                               double *__restrict__ nstate, float4 *__restrict__ nsnap,
                                               float *__restrict__ nomega, double *__restrict__ compl_out) {
  __shared__ double wmax[kVB / 64];
  const int p = xcd_block((int)blockIdx.x, (int)gridDim.x) * kVB + threadIdx.x;
  double d2 = 0.0;
  if (p < A.n) {
    const double *pi = state + (long)p * S;
    const float4 me = snap[p];
    const int L = len[p];
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    int nnb = 0;
    if (L >= 0) {
      walk_list<S, SRK>(P, A, list, snap, omega, state, p, L, me.x, me.y, pi, sx, sy, tx, ty, nnb);
    } else {  // (rare) the list outgrew the width: the 3 x 3 cells of the last rebuild hold every candidate
      const int c = cell[p];
      const int cx = c % A.ncv, cy = c / A.ncv;
      for (int dyc = -1; dyc <= 1; ++dyc)
        for (int dxc = -1; dxc <= 1; ++dxc) {
          const int cj = adjust_cell(cy + dyc, A.ncv) * A.ncv + adjust_cell(cx + dxc, A.ncv);
          walk_range<S, SRK, true>(P, A, list, snap, omega, state, p, start[cj], start[cj + 1], me.x, me.y, pi, sx,
                                   sy, tx, ty, nnb);
        }
    }
    Coef K;
    K.xi = P.xi;
    K.sigma = P.sigma;
    K.int_phi = U.int_phi;
    K.sq2d = U.sq2d;
    K.seed = U.seed;
    const uint64_t step = U.step_dev ? (uint64_t)*U.step_dev : U.step;
    const int i = id[p];
    double out[S], C;
    step_particle<S, SRK>(P, K, pi, i, nat ? nat[p] : 0.0, make_double2(sx, sy), make_double2(tx, ty), nnb, U,
                          step, out, C, &d2);
#pragma unroll
    for (int k = 0; k < S; ++k) nstate[(long)p * S + k] = out[k];
    nsnap[p] = vsnap_of(P, out);
    nomega[p] = S == 4 ? (float)out[3] : 0.f;
    if (SRK && compl_out) compl_out[i] = C;
  }
  step_max(A, d2, wmax);
}

// ---------------------------------------------------------------------------
// Row tiles of the rebuild's cell order (grids of at most kRowTileCols columns).  One workgroup
// per tile -- a span of whole cells of one row, as for the cell-list pair kernel -- stages the
// snapshots (and omega) of the 3 x (W + 2) cells around its W cells into LDS.  A list entry is
// then also kept as the candidate's LDS position in its tile's stage (16 bits), and a step reads
// its candidates from LDS instead of gathering them from global memory: the same values in the
// same order, so the sums equal the global walk's bit for bit.  A tile whose stencil cells
// overflow the stage builds and walks its lists through global memory as before.

constexpr int kVStage = 1024;     // staged snapshots per tile
constexpr int kVFar = kVStage;    // one more stage slot, far from everything: the lists' padding
constexpr int kVCols = 48;     // staged columns W + 2

template <bool SRK>
struct VtLds {
  float4 sdat[kVStage + 1];  // [kVFar]: x, y far away, phasor 0
  float somega[SRK ? kVStage + 1 : 1];
  int cglob[3 * kVCols];     // first slot of staged cell (r, q), at r * (W + 2) + q
  int ccnt[3 * kVCols];      // its particle count
  int cpos[3 * kVCols + 1];  // its LDS start; [cells] = the staged total
  double wmax[kVB / 64];
};

// the staging plan of tile (row, ca, cb): rows row-1 .. row+1, columns ca-1 .. cb (each wrapped
// once); returns the staged total (the same for every thread)
__device__ int vt_plan(const VArgs &A, const int *__restrict__ start, int row, int ca, int cb, int *cglob, int *ccnt,
                       int *cpos) {
  const int ncols = cb - ca + 2, ncells = 3 * ncols;
  for (int e = threadIdx.x; e < ncells; e += kVB) {
    const int r = e / ncols, q = e - r * ncols;
    int rr = row + r - 1, cc = ca - 1 + q;
    rr += rr < 0 ? A.ncv : (rr >= A.ncv ? -A.ncv : 0);
    cc += cc < 0 ? A.ncv : (cc >= A.ncv ? -A.ncv : 0);
    const int c = rr * A.ncv + cc;
    const int g0 = start[c];
    cglob[e] = g0;
    ccnt[e] = start[c + 1] - g0;
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // one wave: exclusive prefix of the counts, 64 cells per pass
    const int lane = threadIdx.x;
    int run = 0;
    for (int e0 = 0; e0 < ncells; e0 += 64) {
      const int e = e0 + lane;
      const int v = e < ncells ? ccnt[e] : 0;
      const int incl = wave_incl_scan(v, lane);
      if (e < ncells) cpos[e] = run + incl - v;
      run += __shfl(incl, 63, 64);
    }
    if (lane == 0) cpos[ncells] = run;
  }
  __syncthreads();
  return cpos[ncells];
}

// the staged cells' snapshots (and omega) into LDS, G lanes per cell, each position taken as its
// periodic image nearest the centre of its staged cell (the cells are those of the last rebuild,
// which the particles may have left by up to half the skin, across the box's edge too) and re-based
// to the stage's corner (col0, row0 in cells of side 1/ncv): within a stage the separations are
// then plain differences, with no nearest-image correction per candidate.  Each staged coordinate
// takes at most two more FP32 roundings (<= 2^-24 each, coordinates below 1), within what the
// guard bands allow for.
template <bool SRK>
__device__ void vt_stage(const VArgs &A, int ncols, int col0, int row0, const float4 *__restrict__ snap,
                         const float *__restrict__ omega, VtLds<SRK> &L) {
  const int ncells = 3 * ncols;
  const float h = 1.f / (float)A.ncv;
  const float x0 = (float)col0 * h, y0 = (float)row0 * h;
  const int G = ncells <= kVB / 8 ? 8 : (ncells <= kVB / 4 ? 4 : (ncells <= kVB / 2 ? 2 : 1));
  const int sub = (int)threadIdx.x % G;
  for (int c = (int)threadIdx.x / G; c < ncells; c += kVB / G) {
    const int r = c / ncols, q = c - r * ncols;
    // centre of the staged cell, unwrapped (its column and row may lie one cell past the edges)
    const float xc = ((float)(col0 + q) + 0.5f) * h, yc = ((float)(row0 + r) + 0.5f) * h;
    const int e0 = L.cpos[c], n0 = L.ccnt[c], g0 = L.cglob[c];
    for (int j = sub; j < n0; j += G) {
      const float4 v = snap[g0 + j];
      const float xu = v.x - __builtin_rintf(v.x - xc), yu = v.y - __builtin_rintf(v.y - yc);
      L.sdat[e0 + j] = make_float4(xu - x0, yu - y0, v.z, v.w);
      if (SRK) L.somega[e0 + j] = omega[g0 + j];
    }
  }
  if (threadIdx.x == 0) {  // d^2 of the far slot is ~1e36: out of range, never in a guard band
    L.sdat[kVFar] = make_float4(1e18f, 1e18f, 0.f, 0.f);
    if (SRK) L.somega[kVFar] = 0.f;
  }
  __syncthreads();
}

// d^2 - c2 of two staged entries (plain differences: see vt_stage)
__device__ __forceinline__ float dd_staged(float4 o, float xi, float yi, float c2) {
  const float dx = o.x - xi, dy = o.y - yi;
  return __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, -c2));
}

// global slot of staged entry t: the last staged cell starting at or before t (an empty cell
// starts where the next one does, so the cell found holds t)
__device__ __forceinline__ int vt_slot(const int *cglob, const int *cpos, int ncells, int t) {
  int lo = 0, hi = ncells - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cpos[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return cglob[lo] + (t - cpos[lo]);
}

// The lists of a tile's particles, in list_of_slot's order (stencil cells in order, ascending slot
// inside a cell) and with its decisions, from the staged snapshots; each entry is written twice:
// as a slot (list, slot-major over all particles) and as its LDS position in the tile's stage
// (list16, tile-local chunks of four, l16chunk); both written a chunk at a time.
template <int S>
__global__ __launch_bounds__(kVB) void vt_lists(DevParams P, VArgs A, const int4 *__restrict__ tiles, int ntiles,
                                                const int *__restrict__ start, const int *__restrict__ cell,
                                                const double *__restrict__ state, const float4 *__restrict__ snap,
                                                int *__restrict__ list, uint2 *__restrict__ list16,
                                                int *__restrict__ len) {
  if (gate_off(A)) return;
  __shared__ VtLds<false> L;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {  // (the same for the whole workgroup)
    const int4 td = tiles[tile];
    const int row = td.y, ca = td.z, cb = td.w;
    if (ca >= cb) continue;  // an unused tile slot
    const int ncols = cb - ca + 2, ncells = 3 * ncols;
    __syncthreads();  // the previous tile's stage is consumed
    const int total = vt_plan(A, start, row, ca, cb, L.cglob, L.ccnt, L.cpos);
    const int gtb = L.cglob[ncols + 1];
    const int ntile = start[row * A.ncv + cb] - gtb;
    if (ntile > kVB && threadIdx.x == 0) atomicAdd(&A.ctl->big_tiles, 1);
    if (total > kVStage) {
      for (int u = threadIdx.x; u < ntile; u += kVB) list_of_slot<S>(P, A, start, cell, state, snap, list, len, gtb + u);
      continue;
    }
    vt_stage<false>(A, ncols, ca - 1, row - 1, snap, nullptr, L);
    const int ltb = L.cpos[ncols + 1];
    for (int u = threadIdx.x; u < ntile; u += kVB) {
      const int q = gtb + u, li = ltb + u;
      // entries go out four at a time (a chunk of each list): c32 / c16 collect the current chunk
      int4 c32 = make_int4(0, 0, 0, 0);
      unsigned int c16lo = 0u, c16hi = 0u;
      auto flush = [&](int w) {  // the chunk holding entries 4 * (w >> 2) ..
        *reinterpret_cast<int4 *>(list + lslot(q, w & ~3, A.K)) = c32;
        if (u < kVB) list16[l16chunk(tile, u, w >> 2, A.K)] = make_uint2(c16lo, c16hi);
      };
      const float4 me = L.sdat[li];
      double xi, yi;
      apply_pbc(P, state[(long)q * S], state[(long)q * S + 1], xi, yi);
      const int qi = 1 + cell[q] % A.ncv - ca;  // staged column of the particle's cell
      int w = 0;
      for (int r = 0; r < 3; ++r)
        for (int dq = -1; dq <= 1; ++dq) {
          const int e = r * ncols + qi + dq;
          const int t0 = L.cpos[e], t1 = t0 + L.ccnt[e], g0 = L.cglob[e];
          for (int t = t0; t < t1; ++t) {
            if (t == li) continue;
            const float dd = dd_staged(L.sdat[t], me.x, me.y, A.m2c);
            bool in = dd < 0.f;
            if (__builtin_fabsf(dd) <= A.m2h) {
              const long g = g0 + (t - t0);
              double dx, dy;
              class_c_unsigned(P, xi, yi, state[g * S], state[g * S + 1], dx, dy);
              in = dx * dx + dy * dy <= A.rmax2;
            }
            if (in) {
              if (w < A.K) {
                const int g = g0 + (t - t0), k = w & 3;
                c32.x = k == 0 ? g : c32.x;
                c32.y = k == 1 ? g : c32.y;
                c32.z = k == 2 ? g : c32.z;
                c32.w = k == 3 ? g : c32.w;
                const unsigned int h = (unsigned int)t << ((k & 1) * 16);
                if (k < 2) c16lo = (k == 0 ? 0u : c16lo) | h;
                else c16hi = (k == 2 ? 0u : c16hi) | h;
                if (k == 3) flush(w);
              }
              ++w;
            }
          }
        }
      if (w < A.K) {  // the 16-bit list padded with the far slot to a whole pass of 16 (K is a multiple of 16)
        for (int t = w; t < ((w + 15) & ~15); ++t) {
          const int k = t & 3;
          const unsigned int h = (unsigned int)kVFar << ((k & 1) * 16);
          if (k < 2) c16lo = (k == 0 ? 0u : c16lo) | h;
          else c16hi = (k == 2 ? 0u : c16hi) | h;
          if (k == 3) flush(t);
        }
      }
      len[q] = w <= A.K ? w : -1;
      if (w > A.K) atomicAdd(&A.ctl->overflow, 1ull);
    }
  }
}

// walk_list with the candidates read from the tile's stage (l16: the lane's first chunk of the
// tile-local 16-bit list, chunk c at l16[c * kVB]; the entries past L up to the pass's end are the
// far slot, weight 0, so no entry needs a mask): the pass's sixteen list entries are loaded first
// (global memory, four chunks), the candidates four at a time from LDS, in walk_list's order and
// grouping
template <int S, bool SRK>
__device__ __forceinline__ void walk_list_lds(const DevParams &P, const VArgs &A, const uint2 *__restrict__ l16,
                                              const VtLds<SRK> &Ls, int ncells, const double *__restrict__ state,
                                              int p, int li, int L, float xi, float yi, const double *pi, double &sx,
                                              double &sy, double &tx, double &ty, int &nnb, const uint2 *first) {
  uint2 nxt[4];  // the next pass's four chunks, loaded while the current pass computes
#pragma unroll
  for (int c = 0; c < 4; ++c) nxt[c] = first[c];  // the first pass's, loaded with the list's length
  for (int t0 = 0; t0 < L; t0 += 16) {
    float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, an = 0.f;
    float bx[4] = {0.f, 0.f, 0.f, 0.f}, by[4] = {0.f, 0.f, 0.f, 0.f};
    uint2 cur[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) cur[c] = nxt[c];
    if (t0 + 16 < L) {  // within the list's whole passes (at most K entries)
#pragma unroll
      for (int c = 0; c < 4; ++c) nxt[c] = l16[(long)(((t0 + 16) >> 2) + c) * kVB];
    }
    int j[16];
#pragma unroll
    for (int c = 0; c < 4; ++c) {  // four chunks of four 16-bit entries (K is a multiple of 16)
      const uint2 w = cur[c];
      const int e[4] = {(int)(w.x & 0xffffu), (int)(w.x >> 16), (int)(w.y & 0xffffu), (int)(w.y >> 16)};
#pragma unroll
      for (int k = 0; k < 4; ++k) j[4 * c + k] = e[k];  // past L: the far slot (weight 0)
    }
#pragma unroll
    for (int g8 = 0; g8 < 16; g8 += 8) {  // two groups' candidates read from LDS together
      float4 o8[8];
      float om8[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        o8[k] = Ls.sdat[j[g8 + k]];
        om8[k] = SRK ? Ls.somega[j[g8 + k]] : 0.f;
      }
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        const int g = g8 + h;
        const float4 *o = o8 + h;
        float dd[4], m[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          dd[k] = dd_staged(o[k], xi, yi, A.r2c);
          m[k] = __builtin_amdgcn_fmed3f(-A.r2k * dd[k], 0.f, 1.f);
        }
        const float dmin = __builtin_fminf(__builtin_fminf(__builtin_fabsf(dd[0]), __builtin_fabsf(dd[1])),
                                           __builtin_fminf(__builtin_fabsf(dd[2]), __builtin_fabsf(dd[3])));
        if (dmin <= A.r2h) {  // rare: the reference's FP64 test, as in walk4
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (__builtin_fabsf(dd[k]) <= A.r2h) {
              const long g2 = vt_slot(Ls.cglob, Ls.cpos, ncells, j[g + k]);
              double dx, dy;
              class_c_unsigned(P, pi[0], pi[1], state[g2 * S], state[g2 * S + 1], dx, dy);
              m[k] = dx * dx + dy * dy <= P.rho2 ? 1.f : 0.f;
            }
        }
        an += (m[0] + m[1]) + (m[2] + m[3]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          ax[k] = __builtin_fmaf(m[k], o[k].z, ax[k]);
          ay[k] = __builtin_fmaf(m[k], o[k].w, ay[k]);
          if (SRK) {
            const float mo = m[k] * om8[h + k];
            bx[k] = __builtin_fmaf(mo, o[k].z, bx[k]);
            by[k] = __builtin_fmaf(mo, o[k].w, by[k]);
          }
        }
      }
    }
    sx += ((double)ax[0] + (double)ax[1]) + ((double)ax[2] + (double)ax[3]);
    sy += ((double)ay[0] + (double)ay[1]) + ((double)ay[2] + (double)ay[3]);
    nnb += (int)an;
    if (SRK) {
      tx += ((double)bx[0] + (double)bx[1]) + ((double)bx[2] + (double)bx[3]);
      ty += ((double)by[0] + (double)by[1]) + ((double)by[2] + (double)by[3]);
    }
  }
}

// The step of slot p (vs_step's per-particle work); candidates from the tile's stage when
// `staged` (li: p's LDS position, list16: p's tile-local 16-bit list), else through global
// memory.  Returns |dr|^2.
template <int S, bool SRK>
__device__ __forceinline__ double vt_particle(const DevParams &P, const VArgs &A, const StepArgs &U, const Coef &K,
                                              uint64_t step, const int *__restrict__ list,
                                              const uint2 *__restrict__ list16, const int *__restrict__ len,
                                              const int *__restrict__ start, const int *__restrict__ cell,
                                              const double *__restrict__ state, const int *__restrict__ id,
                                              const double *__restrict__ nat, const float4 *__restrict__ snap,
                                              const float *__restrict__ omega, double *__restrict__ nstate,
                                              float4 *__restrict__ nsnap, float *__restrict__ nomega,
                                              double *__restrict__ compl_out, const VtLds<SRK> *Ls, bool staged,
                                              int ncells, int p, int li) {
  const double *pi = state + (long)p * S;
  uint2 first[4];  // the first pass of the 16-bit list (allocated for every lane of a tile)
  if (staged) {
#pragma unroll
    for (int c = 0; c < 4; ++c) first[c] = list16[(long)c * kVB];
  }
  const float4 me = snap[p];
  const int Lp = len[p];
  double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
  int nnb = 0;
  if (Lp >= 0 && staged) {
    const float4 ms = Ls->sdat[li];  // the particle in the stage's coordinates
    walk_list_lds<S, SRK>(P, A, list16, *Ls, ncells, state, p, li, Lp, ms.x, ms.y, pi, sx, sy, tx, ty, nnb, first);
  } else if (Lp >= 0) {
    walk_list<S, SRK>(P, A, list, snap, omega, state, p, Lp, me.x, me.y, pi, sx, sy, tx, ty, nnb);
  } else {  // (rare) the list outgrew the width: the 3 x 3 cells of the last rebuild hold every candidate
    const int c = cell[p];
    const int cx = c % A.ncv, cy = c / A.ncv;
    for (int dyc = -1; dyc <= 1; ++dyc)
      for (int dxc = -1; dxc <= 1; ++dxc) {
        const int cj = adjust_cell(cy + dyc, A.ncv) * A.ncv + adjust_cell(cx + dxc, A.ncv);
        walk_range<S, SRK, true>(P, A, list, snap, omega, state, p, start[cj], start[cj + 1], me.x, me.y, pi, sx,
                                 sy, tx, ty, nnb);
      }
  }
  const int i = id[p];
  double out[S], C, dp2;
  step_particle<S, SRK>(P, K, pi, i, nat ? nat[p] : 0.0, make_double2(sx, sy), make_double2(tx, ty), nnb, U,
                        step, out, C, &dp2);
#pragma unroll
  for (int k = 0; k < S; ++k) nstate[(long)p * S + k] = out[k];
  nsnap[p] = vsnap_of(P, out);
  nomega[p] = S == 4 ? (float)out[3] : 0.f;
  if (SRK && compl_out) compl_out[i] = C;
  return dp2;
}

__device__ __forceinline__ Coef vt_coef(const DevParams &P, const StepArgs &U) {
  Coef K;
  K.xi = P.xi;
  K.sigma = P.sigma;
  K.int_phi = U.int_phi;
  K.sq2d = U.sq2d;
  K.seed = U.seed;
  return K;
}

// The particles of an oversized tile past the first kVB (a single cell holding more than a tile's
// particles: dense clusters), through global memory; launched before vt_step; their displacements
// join the group maxima too.  (A launch of its own: as extra workgroups of vt_step, its particle
// loop would raise vt_step's registers.)
template <int S, bool SRK>
__global__ __launch_bounds__(kVB) void vt_rest(DevParams P, VArgs A, StepArgs U, const int4 *__restrict__ tiles,
                                               int ntiles, const int *__restrict__ list, const int *__restrict__ len,
                                               const int *__restrict__ start, const int *__restrict__ cell,
                                               const double *__restrict__ state, const int *__restrict__ id,
                                               const double *__restrict__ nat, const float4 *__restrict__ snap,
                                               const float *__restrict__ omega, double *__restrict__ nstate,
                                               float4 *__restrict__ nsnap, float *__restrict__ nomega,
                                               double *__restrict__ compl_out) {
  __shared__ double wmax[kVB / 64];
  if (A.ctl->big_tiles == 0) return;  // the usual case: every tile fits its workgroup
  const Coef K = vt_coef(P, U);
  const uint64_t step = U.step_dev ? (uint64_t)*U.step_dev : U.step;
  double d2 = 0.0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 td = tiles[t];
    if (td.z >= td.w) continue;
    const int g0 = start[td.y * A.ncv + td.z], ntile = start[td.y * A.ncv + td.w] - g0;
    for (int u = kVB + threadIdx.x; u < ntile; u += kVB)
      d2 = fmax(d2, vt_particle<S, SRK>(P, A, U, K, step, list, nullptr, len, start, cell, state, id, nat, snap,
                                        omega, nstate, nsnap, nomega, compl_out, nullptr, false, 0, g0 + u, 0));
  }
  step_max(A, d2, wmax);
}

// The step of a tile's first kVB particles, one per lane (vt_rest has the others), and the step's
// largest displacement as in vs_step.
template <int S, bool SRK>
__global__ __launch_bounds__(kVB) void vt_step(DevParams P, VArgs A, StepArgs U, const int4 *__restrict__ tiles,
                                               int ntiles, const int *__restrict__ list,
                                               const uint2 *__restrict__ list16, const int *__restrict__ len,
                                               const int *__restrict__ start, const int *__restrict__ cell,
                                               const double *__restrict__ state, const int *__restrict__ id,
                                               const double *__restrict__ nat, const float4 *__restrict__ snap,
                                               const float *__restrict__ omega, double *__restrict__ nstate,
                                               float4 *__restrict__ nsnap, float *__restrict__ nomega,
                                               double *__restrict__ compl_out) {
  __shared__ VtLds<SRK> Ls;
  const int tile = xcd_block((int)blockIdx.x, (int)gridDim.x);
  const int4 td = tile < ntiles ? tiles[tile] : make_int4(0, 0, 0, 0);
  double d2 = 0.0;
  if (td.z < td.w) {  // the same for the whole workgroup
    const int row = td.y, ca = td.z, cb = td.w;
    const int ncols = cb - ca + 2, ncells = 3 * ncols;
    const bool staged = vt_plan(A, start, row, ca, cb, Ls.cglob, Ls.ccnt, Ls.cpos) <= kVStage;
    if (staged) vt_stage<SRK>(A, ncols, ca - 1, row - 1, snap, omega, Ls);
    const int gtb = Ls.cglob[ncols + 1];
    const int ntile = start[row * A.ncv + cb] - gtb;
    const int u = threadIdx.x;
    if (u < ntile)
      d2 = vt_particle<S, SRK>(P, A, U, vt_coef(P, U), U.step_dev ? (uint64_t)*U.step_dev : U.step, list,
                               list16 + l16chunk(tile, u, 0, A.K), len,
                               start, cell, state, id, nat, snap, omega, nstate, nsnap, nomega, compl_out, &Ls, staged,
                               ncells, gtb + u, Ls.cpos[ncols + 1] + u);
  }
  step_max(A, d2, Ls.wmax);
}

// the largest particle count of any 3 x 3 block of cells (the candidates of one list)
__global__ __launch_bounds__(kVB) void vc_block_max(VArgs A, const int *__restrict__ count, int *__restrict__ out) {
  const int c = blockIdx.x * kVB + threadIdx.x;
  if (c >= A.cells) return;
  const int cx = c % A.ncv, cy = c / A.ncv;
  int s = 0;
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) s += count[adjust_cell(cy + dy, A.ncv) * A.ncv + adjust_cell(cx + dx, A.ncv)];
  atomicMax(out, s);
}

__global__ void vc_enter(VCtl *c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->flag = 1;  // should_update_lists_ starts true (ParticleSystemFirstOrderPtr.cpp:99)
    c->acc = 0.0;
    c->overflow = 0ull;
  }
  for (int g = threadIdx.x; g < kGroups; g += blockDim.x) c->gmax[g * kLineU64] = 0ull;
}

// After the step: CalculateMaxDisplacement's bookkeeping (ParticleSystemFirstOrderPtr.cpp:851-860)
// in the exact mode (one step of margin) from the group maxima, which are cleared for the next
// step; the graph's Philox step counter advances (the step kernel has read it).  One wave.
__global__ void vc_flag(VCtl *c, double half_skin, unsigned long long *step_inc) {
  const int lane = threadIdx.x;
  double m = __longlong_as_double((long long)c->gmax[lane * kLineU64]);
  c->gmax[lane * kLineU64] = 0ull;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) {
    const double stepd = sqrt(m);
    double acc = c->acc + stepd;
    const int flag = acc + stepd > half_skin ? 1 : 0;
    if (flag) acc = 0.0;
    c->acc = acc;
    c->flag = flag;
    if (step_inc) *step_inc += 1ull;
  }
}

__global__ void vc_set_step(VCtl *c, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c->step = v;
}

__global__ __launch_bounds__(kVB) void vc_iota(int n, int *__restrict__ id) {
  const int p = blockIdx.x * kVB + threadIdx.x;
  if (p < n) id[p] = p;
}

template <int S>
__global__ __launch_bounds__(kVB) void vc_unpermute(int n, const double *__restrict__ vstate, const int *__restrict__ id,
                                                    double *__restrict__ state) {
  const int p = blockIdx.x * kVB + threadIdx.x;
  if (p >= n) return;
  const int i = id[p];
#pragma unroll
  for (int k = 0; k < S; ++k) state[(long)i * S + k] = vstate[(long)p * S + k];
}

// ---------------------------------------------------------------------------
// host side

struct VerletFast {
  int n = 0, s = 4, ncv = 1, cells = 1, K = 0, nsb = 1;
  double rho = 0.0, r_max = 0.0;
  VArgs A{};
  double *vstate[2] = {nullptr, nullptr};
  float4 *vsnap[2] = {nullptr, nullptr};
  float *vomega[2] = {nullptr, nullptr};
  int *vid = nullptr, *vcell = nullptr;
  double *vnat = nullptr;
  int *list = nullptr, *len = nullptr;
  // row tiles of the rebuild's cell order (grids of at most kRowTileCols columns): ncv rows of
  // vt_cmax slots, and the lists as LDS positions in the tiles' stages
  bool tiled = false;
  int vt_wcap = 0, vt_cmax = 0;
  int4 *vtiles = nullptr;
  uint2 *list16 = nullptr;  // chunks of four 16-bit entries
  int *key = nullptr, *slotc = nullptr, *count = nullptr, *start = nullptr, *tmp = nullptr, *blocksum = nullptr;
  double *sstate = nullptr, *snat = nullptr;
  int *sid = nullptr, *scell = nullptr;
  VCtl *ctl = nullptr;
  int *bmax = nullptr;
  size_t list_cap = 0;    // ints allocated for the lists
  size_t list16_cap = 0;  // entries allocated for the tile-local 16-bit lists
  bool resident = false, resident_nat = false;
  int cur = 0;
  struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } sg[2][2];  // [bulk, single step][parity]
  int sg_srk = -1, sg_higher = -1;
  double sg_dt = 0.0, sg_dphi = 0.0, sg_friction = 0.0;
  uint64_t sg_seed = 0;
  const double *sg_nat = nullptr;
  double *sg_compl = nullptr;
  long captures = 0;
};

const char *vf_last_error() { return g_verr.c_str(); }

static int vfail(int code, const std::string &m) {
  g_verr = m;
  return code;
}

#define VCHK(call)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (call);                                                                             \
    if (e_ != hipSuccess) return vfail(SPSS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static unsigned vblocks(long n) { return (unsigned)((n + kVB - 1) / kVB); }

static void drop_graphs(VerletFast *v) {
  for (auto &row : v->sg)
    for (auto &g : row) {
      if (g.exec) hipGraphExecDestroy(g.exec);
      if (g.graph) hipGraphDestroy(g.graph);
      g.exec = nullptr;
      g.graph = nullptr;
    }
}

void vf_destroy(VerletFast *v) {
  if (!v) return;
  drop_graphs(v);
  void *b[] = {v->vstate[0], v->vstate[1], v->vsnap[0], v->vsnap[1], v->vomega[0], v->vomega[1], v->vid,
               v->vcell,     v->vnat,      v->list,     v->len,      v->key,       v->slotc,     v->count,
               v->start,     v->tmp,       v->blocksum, v->sstate,   v->snat,      v->sid,       v->scell,
               v->ctl,       v->bmax,      v->vtiles,   v->list16};
  for (void *p : b) hipFree(p);
  delete v;
}

int vf_create(VerletFast **out, hipStream_t st, const DevParams &P, double rho, double r_max) {
  *out = nullptr;
  if (P.s != 3 && P.s != 4) return vfail(SPSS_ERR_UNSUPPORTED, "FP32C Verlet lists are implemented for models 3 and 4");
  if (P.xs != 1.0 || P.ys != 1.0)
    return vfail(SPSS_ERR_UNSUPPORTED, "FP32C Verlet lists assume the unit box of the reference's cell grid");
  if (!(r_max > rho)) return vfail(SPSS_ERR_INVALID, "the Verlet skin must be positive");
  const int ncv = (int)(1.0 / r_max);
  if (ncv < 3)
    return vfail(SPSS_ERR_UNSUPPORTED, "FP32C Verlet lists need int(1/r_max) >= 3 (use the FP64 precision there)");
  VerletFast *v = new VerletFast();
  const int n = P.n;
  v->n = n;
  v->s = P.s;
  v->ncv = ncv;
  v->cells = ncv * ncv;
  v->nsb = scan_blocks(v->cells);
  v->rho = rho;
  v->r_max = r_max;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  const size_t nn = (size_t)n;
  for (int b = 0; b < 2; ++b) {
    A(hipMalloc(&v->vstate[b], sizeof(double) * nn * P.s));
    A(hipMalloc(&v->vsnap[b], sizeof(float4) * nn));
    A(hipMalloc(&v->vomega[b], sizeof(float) * nn));
  }
  A(hipMalloc(&v->vid, sizeof(int) * nn));
  A(hipMalloc(&v->vcell, sizeof(int) * nn));
  A(hipMalloc(&v->vnat, sizeof(double) * nn));
  A(hipMalloc(&v->len, sizeof(int) * nn));
  A(hipMalloc(&v->key, sizeof(int) * nn));
  A(hipMalloc(&v->slotc, sizeof(int) * nn));
  A(hipMalloc(&v->count, sizeof(int) * (size_t)v->cells));
  A(hipMalloc(&v->start, sizeof(int) * ((size_t)v->cells + 1)));
  A(hipMalloc(&v->tmp, sizeof(int) * nn));
  A(hipMalloc(&v->blocksum, sizeof(int) * ((size_t)v->nsb + 1)));
  A(hipMalloc(&v->sstate, sizeof(double) * nn * P.s));
  A(hipMalloc(&v->snat, sizeof(double) * nn));
  A(hipMalloc(&v->sid, sizeof(int) * nn));
  A(hipMalloc(&v->scell, sizeof(int) * nn));
  A(hipMalloc(&v->ctl, sizeof(VCtl)));
  A(hipMalloc(&v->bmax, sizeof(int)));
  v->tiled = ncv <= kRowTileCols;
  if (v->tiled) {
    // slots per row: the greedy spans of a row 1.5x the mean at about 94% of a tile each, plus 2,
    // and at least enough spans of vt_wcap columns to cover a row
    v->vt_wcap = kVCols - 2;
    const double row_mean = (double)n / (double)ncv;
    v->vt_cmax = std::max((ncv + v->vt_wcap - 1) / v->vt_wcap,
                          std::min(ncv, (int)std::ceil(1.5 * row_mean / (0.94 * kTileCap)) + 2));
    A(hipMalloc(&v->vtiles, sizeof(int4) * (size_t)ncv * v->vt_cmax));
  }
  if (e != hipSuccess) {
    vf_destroy(v);
    return vfail(SPSS_ERR_DEVICE, std::string("Verlet list buffers: ") + hipGetErrorString(e));
  }
  launch_fill_int(st, v->cells, v->count, 0);
  launch_fill_int(st, (long)(sizeof(VCtl) / sizeof(int)), reinterpret_cast<int *>(v->ctl), 0);
  // guard bands: coordinates in [0, 1] rounded once to FP32 (<= 2^-25 each), the difference and
  // the nearest-integer shift exact, so |dx| errs by <= 2^-24; the squared distance then by
  // <= 2 sqrt(2) r 2^-24 plus a few ulps -- times a safety factor of 8
  const double eps = std::ldexp(1.0, -24);
  auto band = [&](double r) { return r * r * std::max(8.0 * (2.0 * std::sqrt(2.0) * eps / r + 4.0 * eps), 64.0 * eps); };
  VArgs &a = v->A;
  a.n = n;
  a.ncv = ncv;
  a.cells = v->cells;
  a.rmax2 = r_max * r_max;
  a.half_skin = 0.5 * (r_max - rho);
  a.r2c = (float)(rho * rho);
  a.r2h = (float)band(rho);
  a.r2k = (float)std::ldexp(1.0, std::ilogb(4.0 / (double)a.r2h) + 1);  // power of two, r2k * r2h > 4
  a.m2c = (float)(r_max * r_max);
  a.m2h = (float)band(r_max);
  a.ctl = v->ctl;
  *out = v;
  return SPSS_OK;
}

int vf_list_width(const VerletFast *v) { return v ? v->K : 0; }
bool vf_is_resident(const VerletFast *v) { return v && v->resident; }
void vf_drop(VerletFast *v) {
  if (v) v->resident = false;
}
long vf_graph_captures(const VerletFast *v) { return v ? v->captures : 0; }

bool vf_graph_ready(const VerletFast *v, const double *nat, int srk) {
  (void)srk;  // omega is kept for every model-4 step, so either stepper may follow the other
  return v && v->resident && (nat != nullptr) == v->resident_nat;
}

// Entry into the resident copy.  The list width K is sized here from the state entered: the
// densest 3 x 3 block of cells holds bmax candidates, i.e. a local density of bmax / (9 Hv^2), so a
// list holds about lam = bmax pi r_max^2 / (9 Hv^2); K = 1.5 lam + 6 sqrt(lam) + 32 (at most bmax
// + 16) leaves room for the density to grow.  A list that outgrows K later is counted
// (vf_status); entering again from a fresh state sizes K anew.
int vf_enter(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat) {
  (void)srk;
  const unsigned nb = vblocks(v->n);
  VCHK(hipMemcpyAsync(v->vstate[0], state, sizeof(double) * (size_t)v->n * v->s, hipMemcpyDeviceToDevice, st));
  vc_iota<<<nb, kVB, 0, st>>>(v->n, v->vid);
  if (nat) VCHK(hipMemcpyAsync(v->vnat, nat, sizeof(double) * (size_t)v->n, hipMemcpyDeviceToDevice, st));
  vc_enter<<<1, 64, 0, st>>>(v->ctl);
  // the cell counts of the state entered (the flag is set: vr_keys runs), their densest 3 x 3
  // block, and the counts cleared again for the first rebuild
  launch_fill_int(st, 1, v->bmax, 0);
  if (v->s == 4)
    vr_keys<4><<<nb, kVB, 0, st>>>(P, v->A, v->vstate[0], v->key, v->slotc, v->count);
  else
    vr_keys<3><<<nb, kVB, 0, st>>>(P, v->A, v->vstate[0], v->key, v->slotc, v->count);
  vc_block_max<<<vblocks(v->cells), kVB, 0, st>>>(v->A, v->count, v->bmax);
  int bmax = 0;
  VCHK(hipMemcpyAsync(&bmax, v->bmax, sizeof(int), hipMemcpyDeviceToHost, st));
  launch_fill_int(st, v->cells, v->count, 0);
  VCHK(hipStreamSynchronize(st));
  const double Hv = 1.0 / (double)v->ncv;
  const double lam = (double)bmax * M_PI * v->r_max * v->r_max / (9.0 * Hv * Hv);
  int K = (int)std::ceil(1.5 * lam + 6.0 * std::sqrt(lam) + 32.0);
  K = std::max(16, std::min(K, bmax + 16));
  K = (K + 15) / 16 * 16;
  // an upper bound on the width from the environment (memory cap; lists beyond it walk their cells)
  if (const char *cap = std::getenv("SPSS_VERLET_WIDTH_CAP")) K = std::max(4, std::min(K, std::atoi(cap)));
  K = (K + 15) / 16 * 16;  // whole chunks of four entries, whole passes of sixteen
  const size_t need = ((size_t)v->n + 255) / 256 * 256 * (size_t)K;  // lslot's blocks of 256 slots
  const size_t need16 = v->tiled ? (size_t)v->ncv * v->vt_cmax * (size_t)K * kVB : 0;  // tile-local
  if (need > v->list_cap || need16 > v->list16_cap) {
    drop_graphs(v);  // captured launches hold the old list pointer
    hipFree(v->list);
    hipFree(v->list16);
    v->list = nullptr;
    v->list16 = nullptr;
    v->list_cap = 0;
    v->list16_cap = 0;
    VCHK(hipMalloc(&v->list, sizeof(int) * need));
    if (v->tiled) VCHK(hipMalloc(&v->list16, sizeof(unsigned short) * need16));  // need16 entries
    v->list_cap = need;
    v->list16_cap = need16;
  }
  if (K != v->K) drop_graphs(v);  // the width is a launch argument of the captured steps
  v->K = K;
  v->A.K = K;
  VCHK(hipGetLastError());
  v->cur = 0;
  v->resident = true;
  v->resident_nat = nat != nullptr;
  return SPSS_OK;
}

// the rebuild launches (gated on the device flag) and the step, from buffer v->cur
template <int S, bool SRK>
static void launch_step(VerletFast *v, hipStream_t st, const DevParams &P, const StepArgs &U, double *compl_out) {
  const int c = v->cur, o = 1 - c;
  const unsigned nb = vblocks(v->n);
  const VArgs &A = v->A;
  const double *nat = v->resident_nat ? v->vnat : nullptr;
  // the rebuild launches return at once on most steps: a fixed grid of at most kGateGrid
  // workgroups (grid-stride loops) keeps what they cost then down to a dispatch
  const unsigned gg = std::min(nb, (unsigned)kGateGrid);
  vr_keys<S><<<gg, kVB, 0, st>>>(P, A, v->vstate[c], v->key, v->slotc, v->count);
  scan_local<<<v->nsb, 256, 0, st>>>(v->count, nullptr, v->cells, v->start, v->blocksum, ctl_flag(A.ctl));
  scan_add<<<v->nsb, 256, 0, st>>>(v->count, nullptr, v->cells, v->start, v->blocksum, v->nsb, nullptr,
                                   ctl_flag(A.ctl));
  vr_scatter<<<gg, kVB, 0, st>>>(A, v->key, v->slotc, v->start, v->tmp);
  vr_gather<S><<<gg, kVB, 0, st>>>(A, v->key, v->start, v->tmp, v->vstate[c], v->vid, nat, v->sstate, v->sid, v->snat,
                                   v->scell);
  // tiled grids: the row tiles of the new order are built by vr_settle's first workgroups, so its
  // grid covers a wave per row as well
  const unsigned sg = v->tiled ? std::max(gg, (unsigned)((v->ncv + 3) / 4)) : gg;
  vr_settle<S><<<sg, kVB, 0, st>>>(P, A, v->sstate, v->sid, v->snat, v->scell, v->vstate[c], v->vid,
                                   nat ? v->vnat : nullptr, v->vcell, v->vsnap[c], v->vomega[c], v->start,
                                   v->tiled ? v->vtiles : nullptr, v->vt_wcap, v->vt_cmax);
  if (v->tiled) {
    const int ntiles = v->ncv * v->vt_cmax;
    const unsigned tg = (unsigned)((ntiles + 7) / 8 * 8);  // a multiple of 8 for xcd_block
    const unsigned gt = std::min(tg, (unsigned)kGateGrid);
    vt_lists<S><<<gt, kVB, 0, st>>>(P, A, v->vtiles, ntiles, v->start, v->vcell, v->vstate[c], v->vsnap[c], v->list,
                                    v->list16, v->len);
    vt_rest<S, SRK><<<gt, kVB, 0, st>>>(P, A, U, v->vtiles, ntiles, v->list, v->len, v->start, v->vcell, v->vstate[c],
                                        v->vid, nat, v->vsnap[c], v->vomega[c], v->vstate[o], v->vsnap[o],
                                        v->vomega[o], SRK ? compl_out : nullptr);
    vt_step<S, SRK><<<tg, kVB, 0, st>>>(P, A, U, v->vtiles, ntiles, v->list, v->list16, v->len, v->start, v->vcell,
                                        v->vstate[c], v->vid, nat, v->vsnap[c], v->vomega[c], v->vstate[o],
                                        v->vsnap[o], v->vomega[o], SRK ? compl_out : nullptr);
  } else {
    vr_lists<S><<<nb, kVB, 0, st>>>(P, A, v->start, v->vcell, v->vstate[c], v->vsnap[c], v->list, v->len);
    const unsigned nb8 = (nb + 7u) & ~7u;  // a multiple of 8 for xcd_block
    vs_step<S, SRK><<<nb8, kVB, 0, st>>>(P, A, U, v->list, v->len, v->start, v->vcell, v->vstate[c], v->vid, nat,
                                         v->vsnap[c], v->vomega[c], v->vstate[o], v->vsnap[o], v->vomega[o],
                                         SRK ? compl_out : nullptr);
  }
  vc_flag<<<1, kGroups, 0, st>>>(A.ctl, A.half_skin, U.step_inc);
}

static StepArgs vstep_args(double dt, double D_phi, double friction, int noise_mode, const double *noise,
                           uint64_t seed, uint64_t step) {
  StepArgs U{};
  U.dt = dt;
  U.sqrt_dt = std::sqrt(dt);
  U.int_phi = std::sqrt(2.0 * D_phi);
  U.c15 = 0.5 * std::pow(dt, 1.5);
  U.is3 = 1.0 / std::sqrt(3.0);
  U.sq2d = std::sqrt(2.0 * D_phi);
  U.friction = friction;
  U.noise_mode = noise_mode;
  U.noise = noise;
  U.seed = seed;
  U.step = step;
  return U;
}

static void step_impl(VerletFast *v, hipStream_t st, const DevParams &P, bool do_srk, const StepArgs &U,
                      double *compl_out) {
  if (do_srk)
    launch_step<4, true>(v, st, P, U, compl_out);
  else if (v->s == 4)
    launch_step<4, false>(v, st, P, U, compl_out);
  else
    launch_step<3, false>(v, st, P, U, compl_out);
  v->cur = 1 - v->cur;
}

int vf_step(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
            double *compl_out, double dt, double D_phi, double friction, int noise_mode, const double *noise,
            uint64_t seed, uint64_t step) {
  const bool do_srk = srk && v->s == 4;
  if (!vf_graph_ready(v, nat, do_srk)) {
    const int rc = vf_enter(v, st, P, do_srk, state, nat);
    if (rc != SPSS_OK) return rc;
  }
  step_impl(v, st, P, do_srk, vstep_args(dt, D_phi, friction, noise_mode, noise, seed, step), compl_out);
  VCHK(hipGetLastError());
  return SPSS_OK;
}

static int capture(VerletFast *v, hipStream_t st, const DevParams &P, bool do_srk, const StepArgs &U,
                   double *compl_out, int steps, VerletFast::StepGraph &g) {
  const int cur0 = v->cur;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return vfail(SPSS_ERR_DEVICE, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  for (int k = 0; k < steps; ++k) step_impl(v, st, P, do_srk, U, compl_out);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(st, &graph);
  v->cur = cur0;  // nothing ran
  if (e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    return vfail(SPSS_ERR_DEVICE, std::string("capture of the Verlet step graph failed: ") + hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    return vfail(SPSS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  g.graph = graph;
  g.exec = exec;
  ++v->captures;
  return SPSS_OK;
}

static StepArgs graph_setup(VerletFast *v, const DevParams &P, bool do_srk, const double *nat, double *compl_out,
                            double dt, double D_phi, double friction, uint64_t seed) {
  const bool same = v->sg_srk == (int)do_srk && v->sg_higher == P.higher_order && v->sg_dt == dt &&
                    v->sg_dphi == D_phi && v->sg_friction == friction && v->sg_seed == seed && v->sg_nat == nat &&
                    v->sg_compl == compl_out;
  if (!same) {
    drop_graphs(v);
    v->sg_srk = do_srk;
    v->sg_higher = P.higher_order;
    v->sg_dt = dt;
    v->sg_dphi = D_phi;
    v->sg_friction = friction;
    v->sg_seed = seed;
    v->sg_nat = nat;
    v->sg_compl = compl_out;
  }
  StepArgs U = vstep_args(dt, D_phi, friction, 3, nullptr, seed, 0);
  U.step_dev = ctl_step(v->ctl);
  U.step_inc = ctl_step(v->ctl);
  return U;
}

static int ensure_graph(VerletFast *v, hipStream_t st, const DevParams &P, bool do_srk, const StepArgs &U,
                        double *compl_out, int size) {
  VerletFast::StepGraph &g = v->sg[size == 1 ? 1 : 0][v->cur];
  if (g.exec) return SPSS_OK;
  return capture(v, st, P, do_srk, U, compl_out, size, g);
}

int vf_prepare_graphs(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state,
                      const double *nat, double *compl_out, double dt, double D_phi, double friction, uint64_t seed) {
  (void)state;
  const bool do_srk = srk && v->s == 4;
  if (!vf_graph_ready(v, nat, do_srk)) return vfail(SPSS_ERR_INVALID, "no resident copy to capture steps of");
  const StepArgs U = graph_setup(v, P, do_srk, nat, compl_out, dt, D_phi, friction, seed);
  for (int side = 0; side < 2; ++side) {
    for (int size : {kVGraphSteps, 1}) {
      const int rc = ensure_graph(v, st, P, do_srk, U, compl_out, size);
      if (rc != SPSS_OK) {
        if (side) v->cur = 1 - v->cur;
        return rc;
      }
    }
    v->cur = 1 - v->cur;
  }
  return SPSS_OK;
}

long vf_run_graph(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
                  double *compl_out, double dt, double D_phi, double friction, uint64_t seed, uint64_t step,
                  long nsteps) {
  (void)state;
  const bool do_srk = srk && v->s == 4;
  if (nsteps < 1 || !vf_graph_ready(v, nat, do_srk)) return 0;
  const StepArgs U = graph_setup(v, P, do_srk, nat, compl_out, dt, D_phi, friction, seed);
  const long reps = nsteps / kVGraphSteps, rest = nsteps % kVGraphSteps;
  int rc = SPSS_OK;
  if (reps > 0) rc = ensure_graph(v, st, P, do_srk, U, compl_out, kVGraphSteps);
  for (int side = 0; side < 2 && side < rest && rc == SPSS_OK; ++side) {
    if (side) v->cur = 1 - v->cur;
    rc = ensure_graph(v, st, P, do_srk, U, compl_out, 1);
    if (side) v->cur = 1 - v->cur;
  }
  if (rc != SPSS_OK) return -1;
  vc_set_step<<<1, 64, 0, st>>>(v->ctl, (unsigned long long)step);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    g_verr = std::string("vc_set_step: ") + hipGetErrorString(e);
    return -1;
  }
  for (long r = 0; r < reps; ++r)
    if ((e = hipGraphLaunch(v->sg[0][v->cur].exec, st)) != hipSuccess) {
      g_verr = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
      return -1;
    }
  for (long r = 0; r < rest; ++r) {
    if ((e = hipGraphLaunch(v->sg[1][v->cur].exec, st)) != hipSuccess) {
      g_verr = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
      return -1;
    }
    v->cur = 1 - v->cur;
  }
  return nsteps;
}

int vf_sync(VerletFast *v, hipStream_t st, double *state) {
  if (!v->resident) return SPSS_OK;
  const unsigned nb = vblocks(v->n);
  if (v->s == 4)
    vc_unpermute<4><<<nb, kVB, 0, st>>>(v->n, v->vstate[v->cur], v->vid, state);
  else
    vc_unpermute<3><<<nb, kVB, 0, st>>>(v->n, v->vstate[v->cur], v->vid, state);
  VCHK(hipGetLastError());
  return SPSS_OK;
}

int vf_status(VerletFast *v, hipStream_t st, int *should_update, double *acc, long *overflow) {
  VCtl c{};
  if (v->resident) {
    VCHK(hipMemcpyAsync(&c, v->ctl, sizeof(VCtl), hipMemcpyDeviceToHost, st));
    VCHK(hipStreamSynchronize(st));
  } else {
    c.flag = 1;  // the next step enters and rebuilds
  }
  if (should_update) *should_update = c.flag;
  if (acc) *acc = c.acc;
  if (overflow) *overflow = (long)c.overflow;
  return SPSS_OK;
}

int vf_lists(VerletFast *v, hipStream_t st, std::vector<int> *counts, std::vector<int> *flat) {
  const int n = v->n;
  counts->assign(n, 0);
  flat->clear();
  if (!v->resident) return SPSS_OK;
  std::vector<int> id(n), len(n), list(((size_t)n + 255) / 256 * 256 * (size_t)v->K);
  VCHK(hipMemcpyAsync(id.data(), v->vid, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  VCHK(hipMemcpyAsync(len.data(), v->len, sizeof(int) * n, hipMemcpyDeviceToHost, st));
  VCHK(hipMemcpyAsync(list.data(), v->list, sizeof(int) * list.size(), hipMemcpyDeviceToHost, st));
  VCHK(hipStreamSynchronize(st));
  for (int p = 0; p < n; ++p)
    if (len[p] < 0)
      return vfail(SPSS_ERR_CAPACITY, "some Verlet lists outgrew the width " + std::to_string(v->K) +
                                          " and are not stored (those particles walk their 3 x 3 cells)");
  std::vector<int> slot_of(n);
  for (int p = 0; p < n; ++p) slot_of[id[p]] = p;
  for (int i = 0; i < n; ++i) {
    const int p = slot_of[i];
    (*counts)[i] = len[p];
    for (int t = 0; t < len[p]; ++t) flat->push_back(id[list[lslot(p, t, v->K)]]);
  }
  return SPSS_OK;
}

}  // namespace spss
