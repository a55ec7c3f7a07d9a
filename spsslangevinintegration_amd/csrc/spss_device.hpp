// spss_device.hpp -- device-side building blocks shared by the HIP kernels.
//
// FP64 helpers restate the reference's expressions operation by operation;
// the library is compiled with -ffp-contract=off so each one rounds as the
// separate IEEE operation the reference performs.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spss {

constexpr int kBlock = 256;

// Scalars every RHS kernel needs, computed once on the host in FP64.
struct DevParams {
  int n;
  int s;
  double v0;
  double xi;       // model 4: xi; model 6: xi_phi
  double sigma;
  double rho2;     // rho * rho (ParticleSystemFirstOrderPtr.cpp:26)
  double alpha;
  double cos_a;    // std::cos(alpha) (ParticleSystemFirstOrderPtr.cpp:28)
  double sin_a;    // std::sin(alpha)
  double xi_r;     // model 6
  double xs, ys;   // PeriodicBoundaryConditions x_size_, y_size_
  double xrs, yrs; // 1.0 / x_size (the constructor, PeriodicBoundaryConditions.cpp:7-21)
  int natfreq_local;
  int higher_order;
};

// ---- periodic boundary conditions (BoundaryConfiguration/PeriodicBoundaryConditions.cpp)

// ClassCEffectiveParticleDistanceUnsigned, PeriodicBoundaryConditions.cpp:107-119
__device__ __forceinline__ void class_c_unsigned(const DevParams &P, double xi, double yi, double xj,
                                                 double yj, double &dx, double &dy) {
  dx = fabs(xj - xi);
  dx -= (double)((int)(dx * P.xrs + 0.5)) * P.xs;
  dy = fabs(yj - yi);
  dy -= (double)((int)(dy * P.yrs + 0.5)) * P.ys;
}

// ClassAEffectiveParticleDistance, PeriodicBoundaryConditions.cpp:26-38
__device__ __forceinline__ void class_a(const DevParams &P, double xi, double yi, double xj, double yj,
                                        double &dx, double &dy) {
  dx = xj - xi;
  dx -= (double)((int)(dx * 2.0 * P.xrs)) * P.xs;
  dy = yj - yi;
  dy -= (double)((int)(dy * 2.0 * P.yrs)) * P.ys;
}

// ApplyPeriodicBoundaryConditions(x, y, x_pbc, y_pbc), PeriodicBoundaryConditions.cpp:121-125
__device__ __forceinline__ void apply_pbc(const DevParams &P, double x, double y, double &xw, double &yw) {
  xw = x - floor(x * P.xrs) * P.xs;
  yw = y - floor(y * P.yrs) * P.ys;
}

// The wrap applied before cell keys: unconditional for models 3/4
// (ParticleSystemFirstOrderPtr.cpp:374), only outside the box for model 6
// (ParticleSystemSecondOrderPtr.cpp:292-295).
template <int S>
__device__ __forceinline__ void key_wrap(const DevParams &P, double x, double y, double &xw, double &yw) {
  if (S == 6) {
    if (x < 0.0 || x >= 1.0 || y < 0.0 || y >= 1.0) {
      apply_pbc(P, x, y, xw, yw);
    } else {
      xw = x;
      yw = y;
    }
  } else {
    apply_pbc(P, x, y, xw, yw);
  }
}

// int(x / x_size_ * nc) with x_size_ = 1.0 (ParticleSystemFirstOrderPtr.cpp:377-379).  The
// reference leaves x - floor(x) == 1.0 (tiny negative x) unguarded; clamp to the last cell.
__device__ __forceinline__ int cell_coord(double xw, int nc) {
  int c = (int)(xw / 1.0 * (double)nc);
  c = c < 0 ? 0 : c;
  return c >= nc ? nc - 1 : c;
}

// AdjustNeighboringCellToPeriodicBoundaries, ParticleSystemFirstOrderPtr.cpp:817-834
__device__ __forceinline__ int adjust_cell(int c, int nc) { return c < 0 ? nc - 1 : (c >= nc ? 0 : c); }

template <int S>
struct Slots {
  static constexpr int phi = (S == 6) ? 4 : 2;
  static constexpr int omega = (S == 6) ? 5 : 3;
};

// ---- pair term of every local branch
// model 4/3: ParticleSystemFirstOrderPtr.cpp:285-299; model 6: ParticleSystemSecondOrderPtr.cpp:208-215
struct PairAcc {
  double align;
  double card;
  double compl_f;
};

template <int S>
__device__ __forceinline__ void pair_term(const DevParams &P, double x_i, double y_i, double phi_i,
                                          double omega_i, double x_j, double y_j, double phi_j,
                                          double omega_j, PairAcc &a) {
  double dx, dy;
  if (S == 6)
    class_a(P, x_i, y_i, x_j, y_j, dx, dy);
  else
    class_c_unsigned(P, x_i, y_i, x_j, y_j, dx, dy);
  double dr2 = dx * dx + dy * dy;
  if (dr2 <= P.rho2) {
    if (S == 6) {
      a.align += sin(phi_j - phi_i - P.alpha);
      a.card += 1.0;
    } else {
      a.align += sin(phi_j - phi_i) * P.cos_a - cos(phi_j - phi_i) * P.sin_a;
      a.card += 1.0;
      if (S == 4 && P.higher_order) {
        a.compl_f += (omega_j - omega_i) * (cos(phi_j - phi_i) * P.cos_a + sin(phi_j - phi_i) * P.sin_a);
      }
    }
  }
}

// ---- output row of the local branches
// model 4: ParticleSystemFirstOrderPtr.cpp:303-313; model 6: ParticleSystemSecondOrderPtr.cpp:219-227;
// model 3 (m = 0): k = (v0 cos phi, v0 sin phi, w_i + sigma A / n)
template <int S>
__device__ __forceinline__ void write_k(const DevParams &P, const double *st, int i, const PairAcc &a,
                                        double nat_i, double *k_next, double *compl_out) {
  const long ii = (long)i * S;
  if (S == 4) {
    double phi_i = st[2], omega_i = st[3];
    k_next[ii] = P.v0 * cos(phi_i);
    k_next[ii + 1] = P.v0 * sin(phi_i);
    k_next[ii + 2] = omega_i;
    if (P.natfreq_local) {
      if (a.card != 0.0)
        k_next[ii + 3] = -P.xi * omega_i + nat_i + P.sigma * a.align / a.card;
      else
        k_next[ii + 3] = -P.xi * omega_i + nat_i;
    } else if (a.card != 0.0) {
      k_next[ii + 3] = -P.xi * omega_i + P.sigma * a.align / a.card;
    }
    if (P.higher_order && compl_out) {
      double c = a.compl_f;
      if (a.card != 0.0) c *= P.sigma / a.card;
      compl_out[i] = c;
    }
  } else if (S == 6) {
    double vx = st[2], vy = st[3], phi_i = st[4], omega_i = st[5];
    k_next[ii] = vx;
    k_next[ii + 1] = vy;
    k_next[ii + 2] = -P.xi_r * vx + P.v0 * cos(phi_i);
    k_next[ii + 3] = -P.xi_r * vy + P.v0 * sin(phi_i);
    k_next[ii + 4] = omega_i;
    if (a.card != 0.0) k_next[ii + 5] = -P.xi * omega_i + P.sigma * a.align / a.card;
  } else {
    double phi_i = st[2];
    k_next[ii] = P.v0 * cos(phi_i);
    k_next[ii + 1] = P.v0 * sin(phi_i);
    if (a.card != 0.0)
      k_next[ii + 2] = nat_i + P.sigma * a.align / a.card;
    else
      k_next[ii + 2] = nat_i;
  }
}

// ---- counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) written from the
// published round function: two 32x32->64 multiplies per round, Weyl key schedule.
struct U4 {
  uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c.x;
    uint64_t p1 = (uint64_t)M1 * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// two FP64 standard normals from one Philox block: 53-bit uniforms u1 in (0,1], u2 in [0,1),
// Box-Muller (r = sqrt(-2 ln u1), angle 2 pi u2)
__device__ __forceinline__ void philox_normal2(uint64_t seed, uint32_t particle, uint32_t block,
                                               uint64_t step, double &z0, double &z1) {
  U4 c{particle, block, (uint32_t)step, (uint32_t)(step >> 32)};
  U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  uint64_t a = ((uint64_t)r.x << 21) ^ (uint64_t)(r.y >> 11);
  uint64_t b = ((uint64_t)r.z << 21) ^ (uint64_t)(r.w >> 11);
  const double inv53 = 1.0 / 9007199254740992.0;
  double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * inv53;
  double u2 = (double)(b & ((1ull << 53) - 1)) * inv53;
  double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  z0 = rad * cs;
  z1 = rad * sn;
}

}  // namespace spss
