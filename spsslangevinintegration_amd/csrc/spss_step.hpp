// spss_step.hpp -- the FP32C stochastic step of one particle from its neighbour phasor sums
// (Euler or stochastic Runge-Kutta, FP64), shared by the cell-list and the Verlet-list paths so
// both apply the same update.  Device functions only (inline), included by several units.
#pragma once

#include <stdint.h>

#include "spss_device.hpp"
#include "spss_fast.hpp"

namespace spss {

// the stochastic update of a fused step (Euler or SRK)
struct StepArgs {
  double dt, sqrt_dt, int_phi, c15, is3, sq2d, friction;
  int noise_mode;  // 0 none, 1 device array indexed by particle id, 3 Philox
  const double *noise;
  uint64_t seed, step;
  const unsigned long long *step_dev;  // when set, the Philox step is read here (graph replays)
  unsigned long long *step_inc;        // when set, fr_move advances this counter by one first
  const Coef *coef;                    // when set, per-replica xi, sigma, noise intensities, seed
};

__device__ __forceinline__ double philox_z(uint64_t seed, int i, uint32_t block, uint64_t step) {
  double z0, z1;
  philox_normal2(seed, (uint32_t)i, block, step, z0, z1);
  return z0;
}

// The stochastic step of one particle from its phasor sums (StochasticEulerStepperPtr.hpp:50-67,
// StochasticRungeKuttaStepperPtr.hpp:50-71): the derivative as the zero-filled derivative vector of
// StochasticEulerStepperPtr.hpp:42-47 would hold it, then the update.  The new position depends on
// the phase and omega only, so step_position gives it bit-identically without the sums.
// (ddx, ddy: the position slots of the step's total displacement, the vector CalculateMaxDisplacement
// reads: StochasticEulerStepperPtr.hpp:59-60, StochasticRungeKuttaStepperPtr.hpp:59-60)
template <int S, bool SRK>
__device__ __forceinline__ void displacement_from(const DevParams &P, const double *pi, const StepArgs &U, double sn,
                                                  double cn, double &ddx, double &ddy) {
  const double f0 = P.v0 * cn, f1 = P.v0 * sn;
  if (S == 4 && SRK) {
    const double f2 = pi[3], dt = U.dt;
    ddx = f0 * dt - 0.5 * f2 * f1 * dt * dt;
    ddy = f1 * dt + 0.5 * f2 * f0 * dt * dt;
  } else {
    ddx = f0 * U.dt;
    ddy = f1 * U.dt;
  }
}

template <int S, bool SRK>
__device__ __forceinline__ void position_from(const DevParams &P, const double *pi, const StepArgs &U, double sn,
                                              double cn, double &x1, double &y1) {
  double ddx, ddy;
  displacement_from<S, SRK>(P, pi, U, sn, cn, ddx, ddy);
  x1 = pi[0] + ddx;
  y1 = pi[1] + ddy;
}

// the new position from the state pi; sn, cn = sin, cos of its phase
template <int S, bool SRK>
__device__ __forceinline__ void step_position(const DevParams &P, const double *pi, const StepArgs &U, double &sn,
                                              double &cn, double &x1, double &y1) {
  sincos(pi[2], &sn, &cn);
  position_from<S, SRK>(P, pi, U, sn, cn, x1, y1);
}

template <int S, bool SRK>
__device__ __forceinline__ void step_particle(const DevParams &P, const Coef &K, const double *pi, int i,
                                              double nat_i, double2 sm, double2 tm, int nnb, const StepArgs &U,
                                              uint64_t step, double *out, double &C, double *disp2 = nullptr) {
  double sn, cn, ddx, ddy;
  sincos(pi[2], &sn, &cn);
  displacement_from<S, SRK>(P, pi, U, sn, cn, ddx, ddy);
  out[0] = pi[0] + ddx;
  out[1] = pi[1] + ddy;
  if (disp2) *disp2 = ddx * ddx + ddy * ddy;  // CalculateMaxDisplacement's dist2, ParticleSystemFirstOrderPtr.cpp:844-846
  const double st = sn * P.cos_a + cn * P.sin_a;  // sin(phi_i + alpha)
  const double ct = cn * P.cos_a - sn * P.sin_a;  // cos(phi_i + alpha)
  const double A64 = ct * sm.y - st * sm.x;
  const double card = (double)nnb;
  C = 0.0;
  if (S == 4) {
    const double omega_i = pi[3];
    const double f2 = omega_i;
    double f3 = 0.0;
    if (P.natfreq_local) {
      f3 = nnb ? -K.xi * omega_i + nat_i + K.sigma * A64 / card : -K.xi * omega_i + nat_i;
    } else if (nnb) {
      f3 = -K.xi * omega_i + K.sigma * A64 / card;
    }
    if (!SRK) {
      const double z = (U.noise_mode == 1) ? U.noise[i] : (U.noise_mode == 3 ? philox_z(K.seed, i, 0, step) : 0.0);
      const double dW = U.sqrt_dt * z;
      out[2] = pi[2] + f2 * U.dt;
      out[3] = pi[3] + (f3 * U.dt + K.int_phi * dW);
    } else {
      const double C64 = ct * (tm.x - omega_i * sm.x) + st * (tm.y - omega_i * sm.y);
      C = nnb ? C64 * (K.sigma / card) : C64;
      double U1 = 0.0, U2 = 0.0;
      if (U.noise_mode == 1) {
        U1 = U.noise[2L * i];
        U2 = U.noise[2L * i + 1];
      } else if (U.noise_mode == 3) {
        philox_normal2(K.seed, (uint32_t)i, 0, step, U1, U2);
      }
      const double dt = U.dt;
      const double dW = U1 * U.sqrt_dt;
      const double dZ = U.c15 * (U1 + U.is3 * U2);
      out[2] = pi[2] + (f2 * dt + 0.5 * f3 * dt * dt + K.sq2d * dZ);
      out[3] = pi[3] + (f3 * dt + 0.5 * (C - K.xi * f3) * dt * dt + K.sq2d * (dW - K.xi * dZ));
    }
  } else {  // model 3 Euler
    const double f2 = nnb ? nat_i + K.sigma * A64 / card : nat_i;
    const double z = (U.noise_mode == 1) ? U.noise[i] : (U.noise_mode == 3 ? philox_z(K.seed, i, 0, step) : 0.0);
    out[2] = pi[2] + (f2 * U.dt + K.int_phi * (U.sqrt_dt * z));
  }
}

}  // namespace spss
