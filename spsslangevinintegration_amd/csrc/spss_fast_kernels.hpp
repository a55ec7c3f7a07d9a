// spss_fast_kernels.hpp -- device code of the FP32 cell-list path (SPSS_PRECISION_FP32C)
// (included by spss_kernels_fast.hip; host logic lives there).
//
// Two ways to feed the pair kernel:
//  * rebuilt (MODE 0 RHS, MODE 1/2 neighbour export): every call sorts the particles
//    of an FP64 stage state (original order) into cells -- fk_keys, a two-level scan,
//    fk_scatter, fk_rank (stable: ascending particle index inside a cell);
//  * resident (MODE 3, the stochastic steps): the FP64 state, the particle ids and the
//    natural frequencies stay in cell order between steps.  The pair kernel writes the
//    phasor sums of every particle in the old order; fr_move decides each particle's
//    move (its new position depends on phase and omega only) and counts the moves per
//    (cell, direction) and the arrivals per cell, a scan of the arrival counts with a decoupled
//    look-back gives the new cell starts (extra workgroups of the pair kernel, fr_pairs), and
//    fr_scatter applies the step and writes every particle to its new slot: inside a cell,
//    particles that stayed come first in their previous order, then arrivals grouped by the
//    direction they came from, each group in its previous order, then the rare particles that
//    jumped farther than the 3x3 block around their cell, by previous position (placed by the scan).
//    No atomic decides an order, so results are deterministic.
#pragma once

#include <stdint.h>

#include "spss_device.hpp"
#include "spss_fast.hpp"
#include "spss_scan.hpp"
#include "spss_step.hpp"
#include "spss_tiles.hpp"

namespace spss {

namespace {
constexpr int kFB = kTileCap;   // threads per pair workgroup (one particle of a row tile each)
constexpr int kStageMax = 1024; // staged entries per workgroup (particles + row padding)
constexpr int kPad = 4;         // padding entries after each staged row (>= the unroll - 1)
constexpr int kMaxCells = 336;  // staged cells (2ky+1) x (W+2kx) of one workgroup
constexpr int kMaxSpan = 25;    // stencil columns 2kx+1 (kx = k * xsplit <= 12)
}  // namespace

// The cell grid: ncx x ncy cells of side Hx x Hy per replica (cell c = cy * ncx + cx).  A particle's
// stencil is the (2ky+1) rows x (2kx+1) columns around its cell, kx * Hx = ky * Hy = 1/int(1/rho):
// square cells (kx = ky = k), or cells split into kx/ky columns in x (a narrower stencil in x).
struct FastArgs {
  int n, ncx, ncy, kx, ky, ntiles, cpr;  // cpr: cells per replica (ncx * ncy)
  const int4 *tiles;           // per tile: first cell of its replica, row, first column, end column
  int4 *ovf_list;              // work of tiles whose staged cells overflow the LDS stage, for fk_pairs_ovf:
                               // {first cell of the replica, row, ca, cb} a column span that fits, or
                               // {.., .., column, -1 - j} particles j*kFB.. of a cell that does not
  int *ovf_count;              // their number, zeroed by scan_add or fr_scatter
  double Hx, Hy;
  // guard band of the FP32 range test: |d^2 - r2c| <= r2h is decided in FP64; r2k is a power of
  // two with r2k * r2h > 4, so clamp(-r2k * (d^2 - r2c), 0, 1) is exactly 1 below the band, 0 above
  float r2c, r2h, r2k;
};

// Replicas: independent systems of the same cell grid stored one after another (particles in
// replica order, cells [r * cpr, (r + 1) * cpr) for replica r).  A single system is one replica.
constexpr int kChunkShift = 11;  // rep_chunk[p >> kChunkShift]: replica holding position p rounded down
struct Layout {
  const int *rep_first;  // [R + 1] first particle of each replica
  const int *rep_chunk;
  int cpr;
  int nrep;  // R (a single system: 1, every particle and cell in replica 0)
};

__device__ __forceinline__ int replica_of(const Layout &L, int p) {
  if (L.nrep == 1) return 0;
  int r = L.rep_chunk[p >> kChunkShift];
  while (L.rep_first[r + 1] <= p) ++r;
  return r;
}

__device__ __forceinline__ int cell_base(int c, int cpr) { return (c / cpr) * cpr; }


// buffers of one pair-kernel launch (unused ones are null)
struct PairBufs {
  const double *stage;   // rebuilt modes: stage state, original order (AoS FP64)
  const int *sorted;     // rebuilt modes: particle id of each sorted position
  const double *rstate;  // resident: FP64 state in cell order
  const int *rid;        // resident: particle id of each sorted position
  const double *rnat;    // resident: natural frequency of each sorted position (or null)
  const int *start;      // [cells + 1] first sorted position of each cell
  const float4 *snap;    // FP32 snapshot in cell order: offsets in the cell, cos phi, sin phi
  const float *omega_f;  // SRK: omega of each sorted position
  const double *nat;     // rebuilt modes: natural frequency by particle id (or null)
  double *k_next;        // MODE 0, by particle id
  double *compl_out;     // complementary force by particle id (SRK)
  int *pcount;           // MODE 1, by particle id
  const int *pstart;     // MODE 2, by particle id
  int *plist;            // MODE 2
  double2 *ssum;         // MODE 3: phasor sum S of each sorted position
  double2 *tsum;         // MODE 3, SRK: omega-weighted phasor sum T
  int *scnt;             // MODE 3: neighbour count
};


// ---------------------------------------------------------------------------
// moves of the resident re-sort: d = (dy+1)*3 + (dx+1) for a move into one of the 3x3 cells
// around the old cell, kFar for a longer jump.  Inside a destination cell the stayers (d = 4)
// come first, then moves 0, 1, 2, 3, 5, 6, 7, 8, then the far arrivals by old position.
constexpr int kFar = 9;

__device__ __forceinline__ int move_order(int d) { return d == 4 ? 0 : (d < 4 ? d + 1 : d); }

// source cell of an arrival into cell c by move d (within c's replica)
__device__ __forceinline__ int move_source(int c, int d, int ncx, int ncy, int cpr) {
  const int base = cell_base(c, cpr), l = c - base;
  const int cx = l % ncx, cy = l / ncx;
  const int sx = (cx - (d % 3 - 1) + ncx) % ncx;
  const int sy = (cy - (d / 3 - 1) + ncy) % ncy;
  return base + sy * ncx + sx;
}

// a cell coordinate at most one cell outside [0, nc) wrapped into it
__device__ __forceinline__ int wrap_cell(int c, int nc) { return c < 0 ? c + nc : (c >= nc ? c - nc : c); }

__device__ __forceinline__ int new_count_of(const int *cnt9, int c, int ncx, int ncy, int cpr) {
  int s = 0;
#pragma unroll
  for (int d = 0; d < 9; ++d) s += cnt9[(long)move_source(c, d, ncx, ncy, cpr) * 9 + d];
  return s;
}

// ---------------------------------------------------------------------------
// rebuilt cell list

template <int S>
__global__ void fk_keys(DevParams P, Layout L, const double *__restrict__ stage, int ncx, int ncy,
                        int *__restrict__ key, int *__restrict__ slot, int *__restrict__ count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  double xw, yw;
  apply_pbc(P, stage[(long)i * S], stage[(long)i * S + 1], xw, yw);
  const int c = replica_of(L, i) * L.cpr + cell_coord(yw, ncy) * ncx + cell_coord(xw, ncx);
  key[i] = c;
  slot[i] = atomicAdd(count + c, 1);
}

__global__ void fk_scatter(int n, const int *__restrict__ key, const int *__restrict__ slot,
                           const int *__restrict__ start, int *__restrict__ tmp) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) tmp[start[key[i]] + slot[i]] = i;
}

// snapshot entry of a particle: offsets inside cell (cx, cy) and the unit phasor
// (cos phi, sin phi), each computed in FP64 and rounded once to FP32
template <int S>
__device__ __forceinline__ float4 snap_entry(const DevParams &P, const double *st, int cx, int cy, double Hx,
                                             double Hy) {
  double xw, yw;
  apply_pbc(P, st[0], st[1], xw, yw);
  double sn, cn;
  sincos(st[2], &sn, &cn);
  return make_float4((float)(xw - (double)cx * Hx), (float)(yw - (double)cy * Hy), (float)cn, (float)sn);
}

// stable rank + snapshot in sorted order
template <int S>
__global__ void fk_rank(DevParams P, const double *__restrict__ stage, int ncx, int cpr, double Hx, double Hy,
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
  const int l = c - cell_base(c, cpr);
  snap[p] = snap_entry<S>(P, st, l % ncx, l / ncx, Hx, Hy);
  if (omega_f) omega_f[p] = (S == 4) ? (float)st[3] : 0.f;
}

// ---------------------------------------------------------------------------
// the pair kernel


// FP64 re-test of a guard-band pair with the reference's formula
// (x_i wrapped, x_j as stored, ClassCEffectiveParticleDistanceUnsigned, dr2 <= rho^2)
__device__ __forceinline__ bool fp64_in_range(const DevParams &P, const double *pj, double x_i, double y_i) {
  double dx, dy;
  class_c_unsigned(P, x_i, y_i, pj[0], pj[1], dx, dy);
  double dr2 = dx * dx + dy * dy;
  return dr2 <= P.rho2;
}

// per particle after its pair sums: the RHS (MODE 0), the neighbour count (MODE 1) or the
// phasor sums of a stochastic step (MODE 3)
template <int S, bool SRK, int MODE>
__device__ __forceinline__ void finish_particle(const DevParams &P, const PairBufs &B, int i, int gp,
                                                const double *pi, double sx, double sy, double tx, double ty,
                                                int nnb) {
  const double card = (double)nnb;
  // e^{-i theta}, theta = phi_i + alpha, from the FP64 phase
  double sn = 0.0, cn = 0.0, A64 = 0.0, C64 = 0.0;
  if (MODE == 0) {
    sincos(pi[2], &sn, &cn);
    const double st = sn * P.cos_a + cn * P.sin_a;
    const double ct = cn * P.cos_a - sn * P.sin_a;
    A64 = ct * sy - st * sx;
    if (SRK) {
      const double wi = pi[3];
      C64 = ct * (tx - wi * sx) + st * (ty - wi * sy);
    }
  }
  if (MODE == 1) {
    B.pcount[i] = nnb;
  } else if (MODE == 0) {
    const long ii = (long)i * S;
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
      if (SRK && B.compl_out) B.compl_out[i] = nnb ? C64 * (P.sigma / card) : C64;
    } else {
      B.k_next[ii + 2] = nnb ? nat_i + P.sigma * A64 / card : nat_i;
    }
  } else if (MODE == 3) {
    B.ssum[gp] = make_double2(sx, sy);
    if (SRK) B.tsum[gp] = make_double2(tx, ty);
    B.scnt[gp] = nnb;
  }
}

template <int S, bool SRK, int MODE>
struct PairLds {
  float4 sdat[kStageMax];                          // staged x, y in its cell, cos phi, sin phi
  float somega[SRK ? kStageMax : 1];
  int sg[(MODE == 1 || MODE == 2) ? kStageMax : 1];  // global sorted position (pair export)
  int cglob[kMaxCells];                            // global sorted start of staged cell (r, q) at r*ncols+q
  int ccnt[kMaxCells];                             // its particle count
  int cpos[kMaxCells + 1];                         // its LDS start; [cells] = staged total with padding
};

// FP64 decision of a guard-band candidate (global sorted position g) for particle pi
template <int S, int MODE>
__device__ __forceinline__ bool band_in(const DevParams &P, const PairBufs &B, const double *pi, int g) {
  const double *pj = (MODE >= 3) ? B.rstate + (long)g * S : B.stage + (long)B.sorted[g] * S;
  double x_i64, y_i64;
  apply_pbc(P, pi[0], pi[1], x_i64, y_i64);
  return fp64_in_range(P, pj, x_i64, y_i64);
}

// The candidates staged at LDS positions [lo, hi) (plus far-away padding up to hi + 3) against
// one particle, four at a time: FP32 distance tests with the FP64 decision in the guard band,
// phasor sums in 4 FP32 partial slots (slot = (t - lo) mod 4), counts and list entries.
// dd = d^2 - r2c is formed as fma(dy, dy, fma(dx, dx, -r2c)); its rounding error near the band is
// at most that of fma(dy, dy, dx * dx) (a few ulps of r2c), which the band width covers.  The
// in-range weight m is clamp(-r2k * dd, 0, 1): exactly 1 or 0 outside the band (one VALU op with
// the output clamp), replaced by the FP64 decision inside it.
// gof(t) gives the global sorted position of LDS entry t for band decisions.
template <int S, bool SRK, int MODE, class GOF>
__device__ __forceinline__ void scan_range(const DevParams &P, const PairBufs &B, const float4 *sdat,
                                           const float *somega, int lo, int hi, float xi_t, float syi,
                                           float r2c, float r2h, float r2k, const double *pi, int self_t, GOF gof,
                                           float *ax, float *ay, float &an, float *bx, float *by, int &nnb,
                                           int &w, int wbase) {
  constexpr bool SUMS = MODE == 0 || MODE >= 3;
  const float4 *p = sdat + lo;
  const float4 *const pe = sdat + hi;
  for (; p < pe; p += 4) {
    float4 o[4];
    float m[4], dd[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o[j] = p[j];
      const float dx = o[j].x - xi_t;
      const float dy = o[j].y - syi;
      dd[j] = __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, -r2c));
      m[j] = __builtin_amdgcn_fmed3f(-r2k * dd[j], 0.f, 1.f);
    }
    // one test for the four: the smallest |dd| inside the band
    const float dmin = __builtin_fminf(__builtin_fminf(__builtin_fabsf(dd[0]), __builtin_fabsf(dd[1])),
                                       __builtin_fminf(__builtin_fabsf(dd[2]), __builtin_fabsf(dd[3])));
    if (dmin <= r2h) {  // rare: FP64 decision inside the stencil
      const int t = (int)(p - sdat);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (__builtin_fabsf(dd[j]) <= r2h) m[j] = ((t + j < hi) && band_in<S, MODE>(P, B, pi, gof(t + j))) ? 1.f : 0.f;
    }
    if (SUMS) {
      an += (m[0] + m[1]) + (m[2] + m[3]);  // a count: exact in any order
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ax[j] = __builtin_fmaf(m[j], o[j].z, ax[j]);
        ay[j] = __builtin_fmaf(m[j], o[j].w, ay[j]);
        if (SRK) {
          const float mo = m[j] * somega[(p - sdat) + j];
          bx[j] = __builtin_fmaf(mo, o[j].z, bx[j]);
          by[j] = __builtin_fmaf(mo, o[j].w, by[j]);
        }
      }
    } else {
      const int t = (int)(p - sdat);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (m[j] != 0.f && t + j < hi && t + j != self_t) {
          ++nnb;
          if (MODE == 2) B.plist[wbase + w++] = B.sorted[gof(t + j)];
        }
      }
    }
  }
}

template <bool SRK>
__device__ __forceinline__ void flush_row(const float *ax, const float *ay, float an, const float *bx,
                                          const float *by, double &sx, double &sy, double &tx, double &ty,
                                          int &nnb) {
  sx += ((double)ax[0] + (double)ax[1]) + ((double)ax[2] + (double)ax[3]);
  sy += ((double)ay[0] + (double)ay[1]) + ((double)ay[2] + (double)ay[3]);
  nnb += (int)an;
  if (SRK) {
    tx += ((double)bx[0] + (double)bx[1]) + ((double)bx[2] + (double)bx[3]);
    ty += ((double)by[0] + (double)by[1]) + ((double)by[2] + (double)by[3]);
  }
}

// The work of a span of cells whose staged cells overflow the LDS stage, appended to the overflow
// list: the widest column spans that fit, and for a column that does not fit by itself its
// particles in chunks of kFB.  ccnt holds the particles of every staged cell (row-major over the
// 2ky+1 rows and the W+2kx staged columns); colsum is LDS scratch for W+2kx column sums.
__device__ __forceinline__ void split_span(const FastArgs &A, int cbase, int row, int ca, int cb, const int *ccnt,
                                           int *colsum) {
  const int kx = A.kx, ky = A.ky, nrows = 2 * ky + 1, ncols = cb - ca + 2 * kx;
  __syncthreads();  // every thread has read the staged total that colsum overwrites
  for (int q = threadIdx.x; q < ncols; q += kFB) {
    int sum = 0;
    for (int r = 0; r < nrows; ++r) sum += ccnt[r * ncols + q];
    colsum[q] = sum;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int a = ca; a < cb;) {
    int tot = nrows * kPad;
    for (int q = a - ca; q <= a - ca + 2 * kx; ++q) tot += colsum[q];
    int b = a + 1;
    if (tot <= kStageMax) {
      while (b < cb && tot + colsum[b - ca + 2 * kx] <= kStageMax) tot += colsum[b++ - ca + 2 * kx];
      A.ovf_list[atomicAdd(A.ovf_count, 1)] = make_int4(cbase, row, a, b);
    } else {
      const int nch = (ccnt[ky * ncols + (a - ca) + kx] + kFB - 1) / kFB;
      const int at = atomicAdd(A.ovf_count, nch);
      for (int j = 0; j < nch; ++j) A.ovf_list[at + j] = make_int4(cbase, row, a, -1 - j);
    }
    a = b;
  }
}

// MODE 0: RHS; 1: count neighbours; 2: write neighbour lists; 3: phasor sums of the resident
// state for the stochastic step
//
// One workgroup per tile of W cells of one row.  The (2ky+1) x (W+2kx) cells around the tile
// are staged into LDS (x re-based to the staged column 0 in FP64, rounded once; y as the
// offset in its cell), each staged row followed by kPad entries far outside the box.  A
// lane owns one particle of the tile and walks the (2ky+1) contiguous LDS ranges of its
// stencil four candidates at a time without per-candidate bounds: reads past the end of a
// range land in farther cells of the same staged row (out of range, see below) or in the
// padding, so the loop body has no branch on the common path.  A tile whose cells do not fit
// the stage is split into smaller work for fk_pairs_ovf (split_span).  pairs_span runs the cells
// [ca, cb) of one row; it returns false when their staged cells do not fit.
//
// Range test: an FP32 squared distance below the guard band r2c +- r2h is in range, above it
// out of range; the band (rare) is decided by the reference's FP64 ClassC formula, and only
// for candidates inside the lane's stencil.  A candidate beyond the stencil is at least
// kx cells = 1/int(1/rho) >= rho away in x, so its FP32 distance is never below the band.
//
// Coupling sums go through phasors:
//   sum_j sin(phi_j - phi_i - alpha) = Im(e^{-i(phi_i + alpha)} S_i),  S_i = sum_j e^{i phi_j}
//   sum_j (omega_j - omega_i) cos(phi_j - phi_i - alpha) = Re(e^{-i(phi_i + alpha)} (T_i - omega_i S_i)),
//   T_i = sum_j omega_j e^{i phi_j},
// with the phasor components rounded once from FP64 to FP32, accumulated in four FP32
// partial sums per stencil row and combined in FP64.  The particle itself is among its
// candidates exactly once (distinct stencil columns since int(1/rho) >= 3) and is removed
// afterwards (distinct stencil columns and rows: ncx >= 2kx+1, ncy >= 2ky+1).  A particle's
// candidates, their order and grouping depend on its own stencil
// cells only (not on the tile), so every tiling and the overflow pass give identical sums.
template <int S, bool SRK, int MODE, bool SPLIT>
__device__ __forceinline__ bool pairs_span(const DevParams &P, const FastArgs &A, const PairBufs &B, int cbase,
                                           int row, int ca, int cb, PairLds<S, SRK, MODE> &L) {
  constexpr bool RES = MODE >= 3;
  constexpr bool SUMS = MODE == 0 || MODE >= 3;
  constexpr bool LIST = MODE == 1 || MODE == 2;
  float4 *sdat = L.sdat;
  int *cglob = L.cglob, *ccnt = L.ccnt, *cpos = L.cpos;
  const int W = cb - ca;
  const int kx = A.kx, ky = A.ky;
  const int nrows = 2 * ky + 1;
  const int ncols = W + 2 * kx;
  const int ncx = A.ncx, ncy = A.ncy;
  const double Hx = A.Hx, Hy = A.Hy;
  const int *start = B.start;
  const int ncells = nrows * ncols;  // <= kMaxCells (the host caps W)

  // ---- staging plan: global extent of every staged cell, then the LDS layout
  for (int e = threadIdx.x; e < ncells; e += kFB) {
    const int r = e / ncols, q = e % ncols;
    // at most ky < ncy rows and kx < ncx columns outside the grid: one wrap each
    int rrow = row + r - ky, col = ca - kx + q;
    rrow += rrow < 0 ? ncy : (rrow >= ncy ? -ncy : 0);
    col += col < 0 ? ncx : (col >= ncx ? -ncx : 0);
    const int c = cbase + rrow * ncx + col;
    const int g0 = start[c];
    cglob[e] = g0;
    ccnt[e] = start[c + 1] - g0;
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // one wave: exclusive prefix of the LDS sizes (count + row padding)
    const int lane = threadIdx.x;
    const int step = 64 % ncols;
    int q = lane % ncols;  // staged column of cell e = e0 + lane
    int run = 0;
    for (int e0 = 0; e0 < ncells; e0 += 64) {  // 64 cells per pass
      const int e = e0 + lane;
      const int v = (e < ncells) ? ccnt[e] + (q == ncols - 1 ? kPad : 0) : 0;
      const int incl = wave_incl_scan(v, lane);
      if (e < ncells) cpos[e] = run + incl - v;
      run += __shfl(incl, 63, 64);
      q += step;
      if (q >= ncols) q -= ncols;
    }
    if (lane == 0) cpos[ncells] = run;
  }
  __syncthreads();
  const int total = cpos[ncells];
  if (total > kStageMax) {  // uniform: every thread reads the same total
    if (SPLIT) split_span(A, cbase, row, ca, cb, ccnt, cpos);
    return false;
  }

  // the staged cells, G lanes per cell (as many as fit the workgroup: no search for the cell of
  // an entry); the last cell of every staged row is followed by its row's padding
  {
    const int G = ncells <= kFB / 8 ? 8 : (ncells <= kFB / 4 ? 4 : (ncells <= kFB / 2 ? 2 : 1));
    const int sub = (int)threadIdx.x % G;
    for (int c = (int)threadIdx.x / G; c < ncells; c += kFB / G) {
      const int q = c % ncols;
      const int e0 = cpos[c], n0 = ccnt[c], g0 = cglob[c];
      const double xq = (double)q * Hx;
      for (int j = sub; j < n0; j += G) {
        const int g = g0 + j;
        const float4 v = B.snap[g];
        sdat[e0 + j] = make_float4((float)(xq + (double)v.x), v.y, v.z, v.w);
        if (SRK) L.somega[e0 + j] = B.omega_f[g];
        if (LIST) L.sg[e0 + j] = g;
      }
      if (q == ncols - 1) {
        for (int j = sub; j < kPad; j += G) {  // row padding: never within range
          sdat[e0 + n0 + j] = make_float4(1e30f, 0.f, 0.f, 0.f);
          if (SRK) L.somega[e0 + n0 + j] = 0.f;
          if (LIST) L.sg[e0 + n0 + j] = -1;
        }
      }
    }
  }
  __syncthreads();

  // particles of the tile: staged row ky, columns kx .. kx+W-1 (contiguous in LDS and globally)
  const int gtb = cglob[ky * ncols + kx];
  const int ntile = start[cbase + row * ncx + cb] - start[cbase + row * ncx + ca];
  const int ltb = cpos[ky * ncols + kx];
  const float r2c = A.r2c, r2h = A.r2h, r2k = A.r2k;

  for (int u = threadIdx.x; u < ntile; u += kFB) {
    const int li = ltb + u;  // LDS index
    const int gp = gtb + u;  // global sorted position
    const int i = RES ? B.rid[gp] : B.sorted[gp];
    const double *pi = RES ? B.rstate + (long)gp * S : B.stage + (long)i * S;
    // own staged column: last tile cell whose global start is <= gp
    int qi = kx;
    {
      int lo = kx, hi = kx + W - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cglob[ky * ncols + mid] <= gp) lo = mid; else hi = mid - 1;
      }
      qi = lo;
    }
    const float4 me = sdat[li];
    const float xi_t = me.x, yi = me.y, ci = me.z, si = me.w;
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    int nnb = 0;
    int w = 0;
    const int wbase = (MODE == 2) ? B.pstart[i] : 0;

    for (int r = 0; r < nrows; ++r) {
      const float syi = (float)((double)yi - (double)(r - ky) * Hy);
      const int e0 = r * ncols + qi - kx, e1 = r * ncols + qi + kx;
      const int lo = cpos[e0];
      const int hi = cpos[e1] + ccnt[e1];  // end of the last stencil cell (before any padding)
      // global position of staged entry t inside this row's stencil (band decisions, lists)
      auto gof = [&](int t) -> int {
        int lo2 = e0, hi2 = e1;
        while (lo2 < hi2) {
          const int mid = (lo2 + hi2 + 1) >> 1;
          if (cpos[mid] <= t) lo2 = mid; else hi2 = mid - 1;
        }
        return LIST ? L.sg[t] : cglob[lo2] + (t - cpos[lo2]);
      };
      float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, an = 0.f;
      float bx[4] = {0.f, 0.f, 0.f, 0.f}, by[4] = {0.f, 0.f, 0.f, 0.f};
      scan_range<S, SRK, MODE>(P, B, sdat, L.somega, lo, hi, xi_t, syi, r2c, r2h, r2k, pi, li, gof, ax, ay, an, bx, by,
                               nnb, w, wbase);
      if (SUMS) flush_row<SRK>(ax, ay, an, bx, by, sx, sy, tx, ty, nnb);
    }
    if (SUMS) {  // remove the self term
      --nnb;
      sx -= (double)ci;
      sy -= (double)si;
      if (SRK) {
        const float omi = L.somega[li];
        tx -= (double)(omi * ci);
        ty -= (double)(omi * si);
      }
    }
    finish_particle<S, SRK, MODE>(P, B, i, gp, pi, sx, sy, tx, ty, nnb);
  }
  return true;
}

template <int S, bool SRK, int MODE>
__global__ __launch_bounds__(kFB) void fk_pairs(DevParams P, FastArgs A, PairBufs B) {
  __shared__ PairLds<S, SRK, MODE> L;
  // XCD-aware tile order: consecutive tiles (neighbouring rows) share an XCD's L2
  const int per = gridDim.x / 8;  // the grid is a multiple of 8
  const int tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= A.ntiles) return;
  const int4 td = A.tiles[tile];
  if (td.z >= td.w) return;  // an unused row-tile slot
  pairs_span<S, SRK, MODE, true>(P, A, B, td.x, td.y, td.z, td.w, L);  // overflow: split for fk_pairs_ovf
}

template <bool SRK>
struct OvfLds {
  float4 sdat[kStageMax];
  float somega[SRK ? kStageMax : 1];
  int sg[kStageMax];
  int cglob[kMaxSpan], ccnt[kMaxSpan];
};

// A cell whose own stencil does not fit the LDS stage (a dense cluster): its particles
// chunk*kFB .. (chunk+1)*kFB - 1 share one stencil; each stencil row is staged by itself in chunks of kStageMax - kPad
// entries that start at multiples of 4 of the row range, which keeps every lane's candidate
// order and 4-slot grouping identical to fk_pairs.
template <int S, bool SRK, int MODE>
__device__ __forceinline__ void pairs_dense_cell(const DevParams &P, const FastArgs &A, const PairBufs &B,
                                                 int cbase, int row, int q, int chunk, OvfLds<SRK> &L) {
  constexpr bool RES = MODE >= 3;
  constexpr bool SUMS = MODE == 0 || MODE >= 3;
  constexpr int kChunk = kStageMax - kPad;
  const int kx = A.kx, ky = A.ky, nrows = 2 * ky + 1, nspan = 2 * kx + 1, ncx = A.ncx, ncy = A.ncy;
  const double Hx = A.Hx, Hy = A.Hy;
  const float r2c = A.r2c, r2h = A.r2h, r2k = A.r2k;
  const int *start = B.start;
  const int cown = cbase + row * ncx + q;
  const int gp1 = start[cown + 1];
  {
    const int pc = start[cown] + chunk * kFB;
    const int gp = pc + threadIdx.x;
    const bool act = gp < gp1;
    const int i = act ? (RES ? B.rid[gp] : B.sorted[gp]) : 0;
    const double *pi = act ? (RES ? B.rstate + (long)gp * S : B.stage + (long)i * S) : nullptr;
    const float4 me = act ? B.snap[gp] : make_float4(0.f, 0.f, 0.f, 0.f);
    // x re-based to the window's column 0 (the particle's own column is kx)
    const float xi_t = (float)((double)kx * Hx + (double)me.x);
    const float yi = me.y, ci = me.z, si = me.w;
    double sx = 0.0, sy = 0.0, tx = 0.0, ty = 0.0;
    int nnb = 0;
    int w = 0;
    const int wbase = (MODE == 2 && act) ? B.pstart[i] : 0;
    for (int r = 0; r < nrows; ++r) {
      __syncthreads();  // the previous row's stage is consumed
      if (threadIdx.x < nspan) {
        const int dc = threadIdx.x;
        const int rrow = ((row + r - ky) % ncy + ncy) % ncy;
        const int col = ((q - kx + dc) % ncx + ncx) % ncx;
        const int c = cbase + rrow * ncx + col;
        L.cglob[dc] = start[c];
        L.ccnt[dc] = start[c + 1] - start[c];
      }
      __syncthreads();
      int len = 0;
      for (int dc = 0; dc < nspan; ++dc) len += L.ccnt[dc];
      const float syi = (float)((double)yi - (double)(r - ky) * Hy);
      float ax[4] = {0.f, 0.f, 0.f, 0.f}, ay[4] = {0.f, 0.f, 0.f, 0.f}, an = 0.f;
      float bx[4] = {0.f, 0.f, 0.f, 0.f}, by[4] = {0.f, 0.f, 0.f, 0.f};
      for (int c0 = 0; c0 < len; c0 += kChunk) {
        const int clen = min(kChunk, len - c0);
        if (c0 > 0) __syncthreads();
        for (int e = threadIdx.x; e < clen + kPad; e += kFB) {
          if (e < clen) {
            const int t = c0 + e;
            int dc = 0, base = 0;
            while (t >= base + L.ccnt[dc]) base += L.ccnt[dc++];
            const int g = L.cglob[dc] + (t - base);
            const float4 v = B.snap[g];
            L.sdat[e] = make_float4((float)((double)dc * Hx + (double)v.x), v.y, v.z, v.w);
            if (SRK) L.somega[e] = B.omega_f[g];
            L.sg[e] = g;
          } else {
            L.sdat[e] = make_float4(1e30f, 0.f, 0.f, 0.f);
            if (SRK) L.somega[e] = 0.f;
            L.sg[e] = -1;
          }
        }
        __syncthreads();
        if (act) {
          int st = -1;  // the particle itself, in its own row
          if (r == ky) {
            const int s0 = L.cglob[kx] <= gp && gp < L.cglob[kx] + L.ccnt[kx] ? gp - L.cglob[kx] : -1;
            int base = 0;
            for (int dc = 0; dc < kx; ++dc) base += L.ccnt[dc];
            if (s0 >= 0 && base + s0 >= c0 && base + s0 < c0 + clen) st = base + s0 - c0;
          }
          auto gof = [&](int t) -> int { return L.sg[t]; };
          scan_range<S, SRK, MODE>(P, B, L.sdat, L.somega, 0, clen, xi_t, syi, r2c, r2h, r2k, pi, st, gof, ax,
                                   ay, an, bx, by, nnb, w, wbase);
        }
      }
      if (act && SUMS) flush_row<SRK>(ax, ay, an, bx, by, sx, sy, tx, ty, nnb);
    }
    if (act) {
      if (SUMS) {  // remove the self term
        --nnb;
        sx -= (double)ci;
        sy -= (double)si;
        if (SRK) {
          const float omi = B.omega_f[gp];
          tx -= (double)(omi * ci);
          ty -= (double)(omi * si);
        }
      }
      finish_particle<S, SRK, MODE>(P, B, i, gp, pi, sx, sy, tx, ty, nnb);
    }
  }
}

// The work fk_pairs split off overflowing tiles (dense clusters), a grid-stride loop over the
// items: column spans that fit the stage run like fk_pairs, chunks of dense cells through
// pairs_dense_cell.
template <int S, bool SRK, int MODE>
__device__ __forceinline__ void overflow_items(const DevParams &P, const FastArgs &A, const PairBufs &B, int first,
                                               int stride) {
  // the span stage and the dense-cell stage are never live together: one LDS buffer holds either
  constexpr size_t kLds = sizeof(PairLds<S, SRK, MODE>) > sizeof(OvfLds<SRK>) ? sizeof(PairLds<S, SRK, MODE>)
                                                                               : sizeof(OvfLds<SRK>);
  __shared__ __attribute__((aligned(16))) unsigned char lds[kLds];
  PairLds<S, SRK, MODE> &Ls = *reinterpret_cast<PairLds<S, SRK, MODE> *>(lds);
  OvfLds<SRK> &Lc = *reinterpret_cast<OvfLds<SRK> *>(lds);
  const int cnt = *A.ovf_count;
  for (int it = first; it < cnt; it += stride) {
    const int4 w = A.ovf_list[it];
    __syncthreads();  // the previous item's LDS is consumed
    if (w.w >= 0)
      pairs_span<S, SRK, MODE, false>(P, A, B, w.x, w.y, w.z, w.w, Ls);
    else
      pairs_dense_cell<S, SRK, MODE>(P, A, B, w.x, w.y, w.z, -1 - w.w, Lc);
  }
}

template <int S, bool SRK, int MODE>
__global__ __launch_bounds__(kFB) void fk_pairs_ovf(DevParams P, FastArgs A, PairBufs B) {
  overflow_items<S, SRK, MODE>(P, A, B, blockIdx.x, gridDim.x);
}

template <int S>
__device__ __forceinline__ int cell_of_position(const DevParams &P, double x, double y, int ncx, int ncy) {
  double xw, yw;
  apply_pbc(P, x, y, xw, yw);
  return cell_coord(yw, ncy) * ncx + cell_coord(xw, ncx);
}

// ---------------------------------------------------------------------------
// resident re-sort

// The move of one resident particle from its state pi (at slot p of the cell order): old cell * 16
// + (dy+1)*3 + (dx+1), or kFar with an unordered arrival slot in farcnt of the destination; then
// the near moves counted per (old cell, direction) into cnt9 and per destination into newcnt.
// Particles lie in cell order, so a wave's equal (cell, direction) keys form runs and one lane per
// run adds the run's length.  Every lane of the wave calls this (valid = false for lanes past
// the end): the run search reads its neighbours.
// (sn, cn: sin and cos of the phase pi[2]; xw, yw: pi[0], pi[1] wrapped into the box)
template <int S, bool SRK>
__device__ __forceinline__ void move_and_count(const DevParams &P, const Layout &L, int ncx, int ncy, const double *pi,
                                               double sn, double cn, double xw, double yw, const StepArgs &U, int p,
                                               bool valid, int *__restrict__ mv, int *__restrict__ farcnt,
                                               int *__restrict__ farslot, int *__restrict__ cnt9,
                                               int *__restrict__ newcnt, int *__restrict__ farlist,
                                               int *__restrict__ farlist_n) {
  int cell = -1, d = kFar, dest = 0;  // invalid lanes: a key of their own that is never counted
  if (valid) {
    double x1, y1;
    position_from<S, SRK>(P, pi, U, sn, cn, x1, y1);
    const int ox = cell_coord(xw, ncx), oy = cell_coord(yw, ncy);
    apply_pbc(P, x1, y1, xw, yw);
    const int nx = cell_coord(xw, ncx), ny = cell_coord(yw, ncy);
    const int base = replica_of(L, p) * L.cpr;
    int dxc = nx - ox, dyc = ny - oy;
    if (dxc == ncx - 1) dxc = -1;
    if (dxc == 1 - ncx) dxc = 1;
    if (dyc == ncy - 1) dyc = -1;
    if (dyc == 1 - ncy) dyc = 1;
    if (dxc >= -1 && dxc <= 1 && dyc >= -1 && dyc <= 1) {
      d = (dyc + 1) * 3 + (dxc + 1);
    } else {
      d = kFar;
      farslot[p] = atomicAdd(farcnt + base + ny * ncx + nx, 1);
      const int k = atomicAdd(farlist_n, 1);  // rare: the arrivals scan places these
      farlist[2 * k] = p;
      farlist[2 * k + 1] = base + ny * ncx + nx;
    }
    cell = base + oy * ncx + ox;
    dest = base + ny * ncx + nx;
    mv[p] = (cell << 4) | d;
  }
  const int lane = threadIdx.x & 63;
  const int pc = __shfl_up(cell, 1, 64), pd = __shfl_up(d, 1, 64);
  const bool head = lane == 0 || pc != cell || pd != d;
  const unsigned long long heads = __ballot(head);
  if (head && valid && d != kFar) {
    const unsigned long long later = lane == 63 ? 0ull : heads >> (lane + 1);
    const int len = later ? __ffsll((long long)later) : 64 - lane;
    atomicAdd(cnt9 + (long)cell * 9 + d, len);
    atomicAdd(newcnt + dest, len);
  }
}

// the moves of a resident order whose moves no fr_scatter has computed yet (entry into the
// resident order, or a step with another dt or stepper); cnt9, farcnt and newcnt start at zero
template <int S, bool SRK>
__global__ __launch_bounds__(256) void fr_move(DevParams P, Layout L, int n, int ncx, int ncy,
                                               const double *__restrict__ rstate, StepArgs U, int *__restrict__ mv,
                                               int *__restrict__ farcnt, int *__restrict__ farslot,
                                               int *__restrict__ cnt9, int *__restrict__ newcnt,
                                               int *__restrict__ farlist, int *__restrict__ farlist_n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = p < n;
  const double *pi = valid ? rstate + (long)p * S : rstate;
  double sn = 0.0, cn = 1.0, xw = 0.0, yw = 0.0;
  if (valid) {
    sincos(pi[2], &sn, &cn);
    apply_pbc(P, pi[0], pi[1], xw, yw);
  }
  move_and_count<S, SRK>(P, L, ncx, ncy, pi, sn, cn, xw, yw, U, p, valid, mv, farcnt, farslot, cnt9, newcnt,
                         farlist, farlist_n);
}

// the row tiles of a cell order with final starts (entry into the resident order)
__global__ __launch_bounds__(256) void fk_row_tiles(const int *__restrict__ start, int ncx, int ncy, int wcap,
                                                    int cmax, int4 *__restrict__ tiles) {
  rows_to_tiles([&](int c) { return start[c]; }, ncx, ncy, wcap, cmax, tiles);
}

// The arrivals of a resident step: the new cell starts, computed by extra workgroups of the pair
// kernel (fr_pairs), and the far arrivals' slots (fr_far, the launch after it).  The move words and
// arrival counts they read were written by the previous step's fr_scatter (or fr_move on entry), so
// the scan needs nothing from the pair sums and runs beside the tiles instead of after them.
//  * Workgroups [0, nsb) scan 2048 cells each (the arrival counts newcnt + farcnt, cleared as they
//    are read) and find the total of the cells before theirs by a decoupled look-back over the
//    earlier workgroups' status words (aggregate, then inclusive prefix; workgroups are dispatched
//    in index order and a workgroup only waits on lower ones, so every wait ends; the waits are
//    also bounded), then write the final starts, and clear the next step's move counters.  The
//    status words carry the values they publish, so relaxed atomics suffice: no fence, no acquire
//    (on this GPU those write back or invalidate the XCD's L2 under the tiles running beside).
//  * fr_far places the step's far movers (a compact list filled where the moves were computed)
//    behind the near arrivals of their destination cells, clears the status words, and runs the
//    pair kernel's overflow work (its sums are needed only by fr_scatter).
struct Arrivals {
  int *newcnt, *farcnt;        // arrival counts per cell (read, then cleared)
  int m, nsb;                  // cells, scan workgroups
  int *nstart;                 // [m + 1] new cell starts
  unsigned long long *status;  // [nsb] look-back words: kLbAgg | total, or kLbPre | inclusive prefix
  const int *farlist;          // this step's far movers {slot, destination cell}
  const int *farlist_n;        // their number
  int *farlist_n_next;         // the next step's count, cleared here for this step's fr_scatter
  const int *cnt9;             // this step's move counts per (cell, direction)
  const int *farslot;          // each far mover's slot among its destination's far arrivals
  int *bucket;                 // far arrivals by new position
  int *cnt9_next;              // the other move-count buffer, cleared for the next step
  long nclear;
  int ncx, ncy, cpr, n;
  int *error;                  // set when a look-back wait ran out (never, when dispatch is ordered)
};

constexpr unsigned long long kLbAgg = 1ull << 62, kLbPre = 2ull << 62, kLbVal = (1ull << 62) - 1;

// one scan workgroup b < Q.nsb of the arrivals; sh: kFB ints of LDS, sprefix an LDS word
__device__ __forceinline__ void arrivals_scan(const Arrivals &Q, const StepArgs &U, int b, int *sh,
                                              long long &sprefix) {
  {
    const long stride = (long)Q.nsb * kFB;
    for (long e = (long)b * kFB + threadIdx.x; e < Q.nclear; e += stride) Q.cnt9_next[e] = 0;
  }
  // graph-replayed steps: the counter read by this step's fr_scatter moves on here, one launch
  // after the previous step's fr_scatter read it
  if (b == 0 && threadIdx.x == 0) {
    if (U.step_inc) *U.step_inc += 1ull;
    *Q.farlist_n_next = 0;  // fr_scatter of this step appends the next step's far movers
  }
  const long base = (long)b * kFB * kScanItems + (long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = 0;
    if (base + q < Q.m) {
      v[q] = Q.newcnt[base + q] + Q.farcnt[base + q];
      Q.newcnt[base + q] = 0;
      Q.farcnt[base + q] = 0;
    }
    s += v[q];
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < kFB; off <<= 1) {
    const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (threadIdx.x < 64) {
    // look-back by the first wave: a window of 64 predecessors read at once (one round trip per
    // window instead of one per predecessor), summed up to the nearest one holding its prefix
    const int lane = threadIdx.x;
    const unsigned long long total = (unsigned long long)sh[kFB - 1];
    long long excl = 0;
    if (b == 0) {
      if (lane == 0) __hip_atomic_store(&Q.status[0], kLbPre | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0) __hip_atomic_store(&Q.status[b], kLbAgg | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      for (int top = b - 1;; top -= 64) {
        const int j = top - lane;
        unsigned long long w = 0;
        if (j >= 0) {
          for (long spin = 0; spin < (1L << 26); ++spin) {
            w = __hip_atomic_load(&Q.status[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (w) break;
            __builtin_amdgcn_s_sleep(1);
          }
          if (!w) atomicOr(Q.error, 1);  // a wait that ran out: flagged (the starts are then wrong)
        }
        // lanes past block 0, holding a prefix, or whose wait ran out end the walk
        const unsigned long long stops = __ballot(j < 0 || (w & kLbPre) || !w);
        const int first = stops ? __ffsll((long long)stops) - 1 : 64;
        long long v = (j >= 0 && lane <= first) ? (long long)(w & kLbVal) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        excl += v;
        if (stops) break;
      }
      if (lane == 0)
        __hip_atomic_store(&Q.status[b], kLbPre | (unsigned long long)(excl + (long long)total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) sprefix = excl;
  }
  __syncthreads();
  int run = (int)sprefix + sh[threadIdx.x] - s;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < Q.m) Q.nstart[base + q] = run;
    run += v[q];
  }
  if (b == Q.nsb - 1 && threadIdx.x == kFB - 1) Q.nstart[Q.m] = run;  // every cell's particles: n
}

// The pair kernel of a resident step with the step's arrivals: workgroups [0, nsb8) are the scan
// workgroups of arrivals_scan (nsb of them; nsb8 = nsb rounded up to a multiple of 8, so that
// workgroup b still runs on XCD b % 8 below), the others the row tiles of fk_pairs in its
// XCD-aware order.  Dispatched first, the scan is done long before the tiles (placed last, it
// ended after them: 58 against 54.5 us per launch at N = 1M).  The scan's LDS words live in the
// tile stage, which a scan workgroup does not use, so the kernel holds no more LDS than fk_pairs.
template <int S, bool SRK>
__global__ __launch_bounds__(kFB) void fr_pairs(DevParams P, FastArgs A, PairBufs B, Arrivals Q, StepArgs U) {
  __shared__ PairLds<S, SRK, 3> L;
  const int nsb8 = (Q.nsb + 7) & ~7;
  if ((int)blockIdx.x < nsb8) {
    if ((int)blockIdx.x < Q.nsb) {
      static_assert(sizeof(L.cglob) >= kFB * sizeof(int), "scan scratch");
      long long *words = reinterpret_cast<long long *>(L.sdat);
      arrivals_scan(Q, U, (int)blockIdx.x, L.cglob, words[0]);
    }
    return;
  }
  const int b = (int)blockIdx.x - nsb8;
  const int per = ((int)gridDim.x - nsb8) / 8;  // the tile part of the grid is a multiple of 8
  const int tile = (b % 8) * per + b / 8;
  if (tile >= A.ntiles) return;
  const int4 td = A.tiles[tile];
  if (td.z >= td.w) return;  // an unused row-tile slot
  pairs_span<S, SRK, 3, true>(P, A, B, td.x, td.y, td.z, td.w, L);  // overflow: split for fk_pairs_ovf
}

// The launch after fr_pairs: the step's far movers behind the near arrivals of their destination
// cells (the starts are final now), in the unordered slots their atomics gave them (fr_scatter
// orders them); the look-back status words cleared; the pair kernel's overflow work.
template <int S, bool SRK>
__global__ __launch_bounds__(kFB) void fr_far(DevParams P, FastArgs A, PairBufs B, Arrivals Q) {
  const int nfar = *Q.farlist_n;
  for (int k = blockIdx.x * kFB + threadIdx.x; k < nfar; k += gridDim.x * kFB) {
    const int pf = Q.farlist[2 * k], cd = Q.farlist[2 * k + 1];
    const int q = Q.nstart[cd] + new_count_of(Q.cnt9, cd, Q.ncx, Q.ncy, Q.cpr) + Q.farslot[pf];
    if (q < Q.n) Q.bucket[q] = pf;  // always, when the counts match the moves
  }
  if (blockIdx.x == 0)
    for (int j = threadIdx.x; j < Q.nsb; j += kFB) Q.status[j] = 0ull;
  overflow_items<S, SRK, 3>(P, A, B, blockIdx.x, gridDim.x);
}

// the step of every particle, written to its new slot with its id, natural frequency and
// snapshot; then (mvn set) the move of the next step from the new state, counted into cnt9n,
// farcnt and newcnt as fr_move would (the new position depends on phase and omega only)
template <int S, bool SRK>
__global__ __launch_bounds__(256) void fr_scatter(
    DevParams P, Layout L, int n, int ncx, int ncy, double Hx, double Hy, const int *__restrict__ start,
    const int *__restrict__ nstart,
    const int *__restrict__ mv, const int *__restrict__ cnt9, const int *__restrict__ bucket,
    const double *__restrict__ rstate, const int *__restrict__ rid, const double *__restrict__ rnat,
    const double2 *__restrict__ ssum, const double2 *__restrict__ tsum, const int *__restrict__ scnt, StepArgs U,
    double *__restrict__ compl_out, double *__restrict__ nstate, int *__restrict__ nrid, double *__restrict__ nnat,
    float4 *__restrict__ snap, float *__restrict__ omega_f, int4 *__restrict__ rtiles, int wcap, int cmax,
    int *__restrict__ mvn, int *__restrict__ farcnt, int *__restrict__ farslot, int *__restrict__ cnt9n,
    int *__restrict__ newcnt, int *__restrict__ farlist, int *__restrict__ farlist_n, int *__restrict__ ovf_count) {
  // the overflow work of this step was read by fk_pairs_ovf: the next pair kernel starts at zero
  if (blockIdx.x == 0 && threadIdx.x == 0) *ovf_count = 0;
  // the next pair kernel's row tiles (single systems) from the new starts, in the first
  // workgroups: their serial span search hides behind the other workgroups' memory traffic
  if (rtiles && (int)blockIdx.x * 4 < ncy)
    rows_to_tiles([&](int c) { return nstart[c]; }, ncx, ncy, wcap, cmax, rtiles);
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = p < n;
  const int lane = threadIdx.x & 63;
  const int m = valid ? mv[p] : -1;  // lanes past the end: cell -1, no move
  const int cs = m >> 4, d = m & 15;
  // rank of p among the earlier particles of its old cell with the same near move: the wave's
  // lanes from the cell's first lane in the wave (one ballot per move), plus, for a cell begun in
  // an earlier wave, the slots before the wave's first (four independent loads per pass)
  const unsigned long long below = (1ull << lane) - 1ull;
  const int pcs = __shfl_up(cs, 1, 64);  // with every lane active: the shuffle reads lane - 1
  const unsigned long long heads = __ballot(lane == 0 || pcs != cs);
  const int seg = 63 - __clzll((long long)(heads & (below | (1ull << lane))));
  unsigned long long same = 0ull;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    const unsigned long long b = __ballot(d == k);
    if (d == k) same = b;
  }
  int rank = __popcll(same & below & ~((1ull << seg) - 1ull));
  if (valid && seg == 0 && d != kFar) {
    const int p0 = p - lane;
    for (int q = start[cs]; q < p0; q += 4) {
      const int a0 = mv[q];
      const int a1 = q + 1 < p0 ? mv[q + 1] : -1;
      const int a2 = q + 2 < p0 ? mv[q + 2] : -1;
      const int a3 = q + 3 < p0 ? mv[q + 3] : -1;
      rank += (((a0 & 15) == d) + ((a1 & 15) == d)) + (((a2 & 15) == d) + ((a3 & 15) == d));
    }
  }
  double out[S];
#pragma unroll
  for (int q = 0; q < S; ++q) out[q] = 0.0;
  int pos = 0;
  double sn2 = 0.0, cn2 = 1.0, xw2 = 0.0, yw2 = 0.0;
  if (valid) {
    const int cpr = L.cpr;
    const int r = L.nrep == 1 ? 0 : cs / cpr, base = r * cpr;
    Coef K;
    if (U.coef) {
      K = U.coef[r];
    } else {
      K.xi = P.xi;
      K.sigma = P.sigma;
      K.int_phi = U.int_phi;
      K.sq2d = U.sq2d;
      K.seed = U.seed;
    }
    const uint64_t step = U.step_dev ? (uint64_t)*U.step_dev : U.step;
    const int id = rid[p];
    const double nat_i = rnat ? rnat[p] : 0.0;
    double C;
    // noise and injected draws are indexed by the particle's index inside its replica
    step_particle<S, SRK>(P, K, rstate + (long)p * S, id - L.rep_first[r], nat_i, ssum[p],
                          SRK ? tsum[p] : make_double2(0.0, 0.0), scnt[p], U, step, out, C);
    if (SRK && compl_out) compl_out[id] = C;
    int cd;
    if (d == kFar) {
      cd = base + cell_of_position<S>(P, out[0], out[1], ncx, ncy);
      const int b = nstart[cd] + new_count_of(cnt9, cd, ncx, ncy, cpr);
      const int e = nstart[cd + 1];
      int rank = 0;
      for (int q = b; q < e; ++q) rank += (bucket[q] < p) ? 1 : 0;
      pos = b + rank;
    } else {
      // destination cell, then the arrivals ordered before move d: the counts of the moves e
      // with move_order(e) < move_order(d) out of their source cells (ncx, ncy >= 3: one wrap each)
      const int ls = cs - base;
      const int cx = ls % ncx, cy = ls / ncx;
      const int tx = wrap_cell(cx + d % 3 - 1, ncx), ty = wrap_cell(cy + d / 3 - 1, ncy);
      cd = base + ty * ncx + tx;
      int b0 = 0;
      const int od = move_order(d);
      if (od > 0) {  // stayers (od = 0) come first: no counts to add (whole waves of them skip the loads)
#pragma unroll
        for (int e = 0; e < 9; ++e) {
          const int sx = wrap_cell(tx - (e % 3 - 1), ncx), sy = wrap_cell(ty - (e / 3 - 1), ncy);
          const int v = cnt9[(long)(base + sy * ncx + sx) * 9 + e];
          b0 += move_order(e) < od ? v : 0;
        }
      }
      pos = nstart[cd] + b0 + rank;
    }
    valid = pos < n;  // always, when the counts match the moves
    if (valid) {
      double *dst = nstate + (long)pos * S;
#pragma unroll
      for (int q = 0; q < S; ++q) dst[q] = out[q];
      nrid[pos] = id;
      if (rnat) nnat[pos] = nat_i;
      // the snapshot (as snap_entry), keeping the phase's sine and cosine and the wrapped
      // position for the next move
      sincos(out[2], &sn2, &cn2);
      apply_pbc(P, out[0], out[1], xw2, yw2);
      const int ld = cd - base;
      snap[pos] = make_float4((float)(xw2 - (double)(ld % ncx) * Hx), (float)(yw2 - (double)(ld / ncx) * Hy),
                              (float)cn2, (float)sn2);
      if (omega_f) omega_f[pos] = (S == 4) ? (float)out[3] : 0.f;
    }
  }
  if (mvn)
    move_and_count<S, SRK>(P, L, ncx, ncy, out, sn2, cn2, xw2, yw2, U, pos, valid, mvn, farcnt, farslot, cnt9n,
                           newcnt, farlist, farlist_n);
}

// advance the device step counter of graph-replayed runs
__global__ void fb_step_inc(unsigned long long *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1ull;
}

// per replica: the phase order parameter Z = (1/N) sum e^{i phi} of the resident snapshot,
// summed in FP64 in a fixed tree; when the step counter has reached `from`, |Z| and |Z|^2 are
// added to the replica's running sums (one block per replica)
__global__ __launch_bounds__(256) void fb_order_stats(Layout L, const float4 *__restrict__ snap,
                                                      const unsigned long long *__restrict__ step,
                                                      unsigned long long from, double *__restrict__ zlast,
                                                      double *__restrict__ racc) {
  __shared__ double sc[256], ss[256];
  const int r = blockIdx.x;
  const int b = L.rep_first[r], e = L.rep_first[r + 1];
  double c = 0.0, s = 0.0;
  for (int p = b + threadIdx.x; p < e; p += 256) {
    const float4 v = snap[p];
    c += (double)v.z;
    s += (double)v.w;
  }
  sc[threadIdx.x] = c;
  ss[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sc[threadIdx.x] += sc[threadIdx.x + off];
      ss[threadIdx.x] += ss[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double inv = 1.0 / (double)(e - b);
    const double zr = sc[0] * inv, zi = ss[0] * inv;
    zlast[2 * r] = zr;
    zlast[2 * r + 1] = zi;
    if (*step >= from) {
      const double R2 = zr * zr + zi * zi;
      racc[3 * r] += sqrt(R2);
      racc[3 * r + 1] += R2;
      racc[3 * r + 2] += 1.0;
    }
  }
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
