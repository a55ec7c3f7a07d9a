// spss_scan.hpp -- two-level exclusive scan of int counts over many workgroups (cell starts of the
// counting sorts, Verlet list offsets), and a wave-wide sort of 64 values.  Device code shared by the FP64 and FP32C translation
// units: device functions are inline and the kernels are templates, so each unit may include it.
//
//   scan_local<I>   one workgroup of 256 threads per 256*I counts: block-local exclusive starts
//                   and the block's total
//   scan_add<I>     the totals of the earlier blocks added to every start (one wave sums them),
//                   start[m] = the grand total, the counts cleared for their next use
#pragma once

#include <hip/hip_runtime.h>

namespace spss {

// ascending sort of one value per lane across a wave of 64 (unused lanes pass INT_MAX): a bitonic
// network of 21 compare-exchange stages by lane shuffles
__device__ __forceinline__ int wave_sort64(int v, int lane) {
#pragma unroll
  for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int o = __shfl_xor(v, j, 64);
      // the lower lane of a pair keeps the smaller value in an ascending run of size k
      v = (((lane & j) == 0) == ((lane & k) == 0)) ? min(v, o) : max(v, o);
    }
  return v;
}

constexpr int kScanItems = 8;  // counts per thread in the two-level scans (2048 per workgroup)

// block-local exclusive scan of 256 * kScanItems values of block `blk`; a value is
// count[c] + count2[c] (count2 may be null)
__device__ __forceinline__ void scan_local_block(const int *__restrict__ count, const int *__restrict__ count2, int m,
                                                 int *__restrict__ start, int *__restrict__ blocksum, int blk) {
  __shared__ int sh[256];
  const long base = (long)blk * 256 * kScanItems + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    v[q] = 0;
    if (base + q < m) {
      v[q] = count[base + q];
      if (count2) v[q] += count2[base + q];
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
  if (threadIdx.x == 255) blocksum[blk] = sh[255];
}

// the sum of the first b scan-block totals, by one wave (every lane gets it; all lanes call)
__device__ __forceinline__ int wave_prefix_of(const int *__restrict__ blocksum, int b) {
  int s = 0;
  for (int j = threadIdx.x & 63; j < b; j += 64) s += blocksum[j];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  return s;
}

// gate (may be null): when *gate == 0 the launch does nothing (work decided on the device)
template <int I = kScanItems>
__global__ __launch_bounds__(256) void scan_local(const int *__restrict__ count, const int *__restrict__ count2, int m,
                                                  int *__restrict__ start, int *__restrict__ blocksum,
                                                  const int *__restrict__ gate = nullptr) {
  static_assert(I == kScanItems, "one block size");
  if (gate && *gate == 0) return;
  scan_local_block(count, count2, m, start, blocksum, blockIdx.x);
}

// adds the scanned block totals, writes start[m], zeroes count[] and count2[] (either may be
// null) for the next use, and the counter `zero` (may be null)
template <int I = kScanItems>
__global__ __launch_bounds__(256) void scan_add(int *__restrict__ count, int *__restrict__ count2, int m,
                                                int *__restrict__ start, const int *__restrict__ blocksum,
                                                int nblocks, int *__restrict__ zero,
                                                const int *__restrict__ gate = nullptr) {
  static_assert(I == kScanItems, "one block size");
  if (gate && *gate == 0) return;
  __shared__ int off;
  if (blockIdx.x == 0 && threadIdx.x == 0 && zero) *zero = 0;
  if (threadIdx.x < 64) {
    const int s = wave_prefix_of(blocksum, (int)blockIdx.x);
    if (threadIdx.x == 0) {
      off = s;
      if ((int)blockIdx.x == nblocks - 1) start[m] = s + blocksum[blockIdx.x];
    }
  }
  __syncthreads();
  const long base = (long)blockIdx.x * 256 * I;
  for (int q = threadIdx.x; q < 256 * I; q += 256) {
    const long c = base + q;
    if (c < m) {
      start[c] += off;
      if (count) count[c] = 0;
      if (count2) count2[c] = 0;
    }
  }
}

// workgroups of a scan over m counts
inline int scan_blocks(long m) { return (int)((m + 256L * kScanItems - 1) / (256L * kScanItems)); }

}  // namespace spss
