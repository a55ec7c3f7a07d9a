// spss_tiles.hpp -- row tiles of a cell order: greedy spans of whole cells of one row of cells,
// each with at most kTileCap particles, for the workgroups of the cell-list pair kernel and of the
// Verlet-list step (one workgroup per tile, one particle per lane).
#pragma once

#include <hip/hip_runtime.h>

namespace spss {

constexpr int kTileCap = 256;       // particles per tile (threads of the tile workgroups)
constexpr int kRowTileCols = 1024;  // rows of at most this many cells get tiles

// inclusive prefix sum over the 64 lanes of a wave
__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// The tiles of one row of cells, built by one wave from the row's cell starts P[0..ncx] (LDS):
// greedy spans of whole cells, each as wide as possible with at most kTileCap particles and at
// most wcap (<= 64) columns, span i in slot i of the row's cmax slots, empty spans in the unused
// slots.  Full spans keep the workgroup's waves full where fixed-width
// tiles leave the last wave of most tiles part-empty.  With cmax * wcap >= ncx every span stays
// within wcap columns: a span is widened past the particle bound (a dense row) only as far as
// the remaining slots need to cover the rest of the row.
__device__ __forceinline__ void row_tiles(const int *P, int cbase, int row, int ncx, int wcap, int cmax,
                                          int4 *__restrict__ slots) {
  const int lane = threadIdx.x & 63;
  int a = 0, i = 0;
  while (a < ncx) {  // invariant: ncx - a <= (cmax - i) * wcap
    const int e = a + 1 + lane;
    // true for a leading run of lanes: P is non-decreasing
    const bool fits = lane < wcap && e <= ncx && P[e] - P[a] <= kTileCap;
    const int b = max(a + max(__popcll(__ballot(fits)), 1), ncx - (cmax - i - 1) * wcap);
    if (lane == 0) slots[i] = make_int4(cbase, row, a, b);
    a = b;
    ++i;
  }
  for (int j = i + lane; j < cmax; j += 64) slots[j] = make_int4(cbase, row, 0, 0);
}

// rows [4 * blockIdx.x, 4 * blockIdx.x + 4) of a single system of ncy rows of ncx cells, one
// wave each; start(c) gives the starts of the cell order the tiles are for (c = 0 .. ncx * ncy).
// Uniform per workgroup, so the barrier is reached by all or none of its threads.
template <class START>
__device__ __forceinline__ void rows_to_tiles(START start, int ncx, int ncy, int wcap, int cmax,
                                              int4 *__restrict__ tiles) {
  __shared__ int P[4][kRowTileCols + 1];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + w;
  if (row < ncy)
    for (int c = lane; c <= ncx; c += 64) P[w][c] = start(row * ncx + c);
  __syncthreads();
  if (row < ncy) row_tiles(P[w], 0, row, ncx, wcap, cmax, tiles + (long)row * cmax);
}

// square grids (ncf x ncf)
template <class START>
__device__ __forceinline__ void rows_to_tiles(START start, int ncf, int wcap, int cmax, int4 *__restrict__ tiles) {
  rows_to_tiles(start, ncf, ncf, wcap, cmax, tiles);
}

}  // namespace spss
