// spss_dopri5.hip -- device side of the Dormand-Prince 5(4) integrator with dense output that the
// reference drives through Boost.Odeint (SimulationEngineFirstOrder.cpp:158-165:
// integrate_const(make_dense_output(1e-6, 1e-6, runge_kutta_dopri5<vector<Real>>()), ...), and
// :273-281: integrate_adaptive(make_dense_output(1e-10, 1e-10, ...), ...)).
//
// Boost.Odeint is not part of the reference tree; these kernels restate the published method:
// the stage and output combinations are Boost's scale_sum operations, out = 1.0 * x + a_1 * k_1
// + a_2 * k_2 + ... with every coefficient already multiplied by dt on the host and the terms
// added left to right; the error norm is Boost's default_error_checker,
// max_i |err_i| / (eps_abs + eps_rel * (|x_old,i| + dt * |dxdt_old,i|)).  All arithmetic is FP64,
// compiled without contraction, so a term is a multiply followed by an add as in the reference
// build.
#include <hip/hip_runtime.h>

#include "spss_launch.hpp"

namespace spss {

namespace {

constexpr int kDpBlock = 256;

__global__ __launch_bounds__(kDpBlock) void k_dp_combine(long m, const double *__restrict__ x, DpTerms T,
                                                         double *__restrict__ out) {
  const long e = blockIdx.x * (long)kDpBlock + threadIdx.x;
  if (e >= m) return;
  double v = 1.0 * x[e];
#pragma unroll
  for (int j = 0; j < kDpMaxTerms; ++j)
    if (j < T.n) v = v + T.c[j] * T.k[j][e];
  out[e] = v;
}

// per block: the largest relative error over the s values of its 256 particles, read as s
// coalesced passes of 256 consecutive values (the largest value does not depend on the order)
__global__ __launch_bounds__(kDpBlock) void k_dp_error(long m, int s, const double *__restrict__ x_old,
                                                       const double *__restrict__ dxdt_old, DpTerms T, double dt,
                                                       double eps_abs, double eps_rel,
                                                       double *__restrict__ partial) {
  __shared__ double red[kDpBlock];
  const long e0 = (long)blockIdx.x * kDpBlock * s + threadIdx.x;
  double worst = 0.0;
  for (int q = 0; q < s; ++q) {
    const long e = e0 + (long)q * kDpBlock;
    if (e < m) {
      double err = T.c[0] * T.k[0][e];
#pragma unroll
      for (int j = 1; j < kDpMaxTerms; ++j)
        if (j < T.n) err = err + T.c[j] * T.k[j][e];
      const double rel = fabs(err) / (eps_abs + eps_rel * (1.0 * fabs(x_old[e]) + (1.0 * dt) * fabs(dxdt_old[e])));
      worst = fmax(worst, rel);
    }
  }
  red[threadIdx.x] = worst;
  __syncthreads();
  for (int off = kDpBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

}  // namespace

void launch_dp_combine(hipStream_t st, long m, const double *x, const DpTerms &T, double *out) {
  const unsigned nb = (unsigned)((m + kDpBlock - 1) / kDpBlock);
  if (nb) k_dp_combine<<<dim3(nb), dim3(kDpBlock), 0, st>>>(m, x, T, out);
}

int launch_dp_error(hipStream_t st, int n, int s, const double *x_old, const double *dxdt_old, const DpTerms &T,
                    double dt, double eps_abs, double eps_rel, double *partial, double *out) {
  const int nb = (n + kDpBlock - 1) / kDpBlock;  // partial holds at least one value per 256 particles
  k_dp_error<<<dim3(nb), dim3(kDpBlock), 0, st>>>((long)n * s, s, x_old, dxdt_old, T, dt, eps_abs, eps_rel,
                                                  partial);
  launch_max_final(st, nb, partial, out);
  return nb;
}

}  // namespace spss
