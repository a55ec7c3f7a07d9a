// spss_kernels_fast.hip -- FP32-compensated cell-list RHS (SPSS_PRECISION_FP32C), models 3 and 4.
//
// Data path of one RHS evaluation (stage state in FP64 AoS, original order):
//   fk_keys      cell of every particle on a grid of side 1/(k*nc), histogram by atomics
//   scan         exclusive scan of the cell counts (two kernels)
//   fk_scatter   unstable bucket fill
//   fk_rank      stable position inside the cell (ascending particle index) -> sorted order,
//                and the snapshot: FP32 offsets of the particle inside its cell, phase in
//                revolutions reduced to [0,1), the cell column; omega for the SRK force
//   fk_pairs     one workgroup per tile of W cells of one row: the (2k+1) rows x (W+2k)
//                cells around it are staged into LDS with x re-based to the tile
//                (FP64 rounding, once per staged particle); per lane, phase 1 tests the
//                candidates of its (2k+1) contiguous row ranges in FP32 and compacts the
//                in-range ones into an LDS list, phase 2 evaluates the coupling term for
//                the list and accumulates it with a two-sum (FP32 value + FP32 error).
//                Pairs whose FP32 distance lies in a guard band around rho^2 are re-tested
//                with the reference's FP64 ClassC formula, so the neighbour sets equal
//                the FP64 ones exactly.
//   self terms (v0 cos phi, v0 sin phi, omega, -xi omega + sigma A / n) in FP64.
#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/spss_hip.h"
#include "spss_fast.hpp"
#include "spss_launch.hpp"

#include "spss_fast_kernels.hpp"

namespace spss {

namespace {
thread_local std::string g_fast_err;
}  // namespace

struct FastPath {
  int n = 0, s = 4;
  double rho = 0.0, rho2 = 0.0;
  int nc = 1, k = 1, ncf = 1;
  double H = 1.0;
  int W = 8, chunks = 1, ntiles = 1;
  float r2lo = 0.f, r2hi = 0.f;
  float alpha_rev = 0.f;
  int *key = nullptr, *slot = nullptr, *count = nullptr, *start = nullptr, *tmp = nullptr, *sorted = nullptr;
  int *blocksum = nullptr;
  float4 *snap = nullptr;
  float *omega_f = nullptr;
  hipStream_t st = nullptr;
  KernelTimer *timer = nullptr;
};

// ---------------------------------------------------------------------------
// host side

const char *fast_last_error() { return g_fast_err.c_str(); }

static int ffail(const std::string &m) {
  g_fast_err = m;
  return SPSS_ERR_DEVICE;
}

#define FCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return ffail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int fast_create(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv) {
  *out = nullptr;
  if (subdiv < 1 || subdiv > 3) {
    g_fast_err = "cell_subdivision must be 1, 2 or 3";
    return SPSS_ERR_INVALID;
  }
  if (P.xs != 1.0 || P.ys != 1.0) {
    g_fast_err = "the FP32C cell list assumes the unit box of the reference's cell grid";
    return SPSS_ERR_UNSUPPORTED;
  }
  if (nc < 3) {
    g_fast_err = "the FP32C cell list needs int(1/rho) >= 3: on narrower grids the reference visits cells "
                 "more than once (use the FP64 precision there)";
    return SPSS_ERR_UNSUPPORTED;
  }
  FastPath *f = new FastPath();
  f->n = P.n;
  f->s = P.s;
  f->rho = rho;
  f->rho2 = rho * rho;
  f->nc = nc;
  f->k = subdiv;
  f->ncf = nc * subdiv;
  f->H = 1.0 / (double)f->ncf;
  // tile width: ~200 particles per tile at the mean density, at most kMaxCols - 2k
  const double per_cell = (double)P.n / ((double)f->ncf * f->ncf);
  int W = (int)std::lround(200.0 / std::max(per_cell, 1e-9));
  const int wstage = (int)(0.75 * kStageMax / std::max(per_cell * (2 * f->k + 1), 1e-9)) - 2 * f->k;
  W = std::min(W, std::max(wstage, 1));
  W = std::max(1, std::min(W, std::min(kMaxCols - 2 * f->k, f->ncf)));
  f->W = W;
  f->chunks = (f->ncf + W - 1) / W;
  f->ntiles = f->chunks * f->ncf;
  // guard band: FP32 distance error bound (coordinates <= (W+2k) cells from the staged
  // origin, rounded once; offsets < one cell) times a safety factor of 8
  const double span = (double)(W + 2 * f->k + 1) * f->H;
  const double eps = std::ldexp(1.0, -24);
  const double dxerr = 2.0 * span * eps + 2.0 * f->H * eps;
  double g = 8.0 * (2.0 * dxerr / rho + 4.0 * eps);
  g = std::max(g, 64.0 * eps);
  f->r2lo = (float)(f->rho2 * (1.0 - g));
  f->r2hi = (float)(f->rho2 * (1.0 + g));
  f->alpha_rev = (float)(P.alpha / (2.0 * M_PI));
  f->st = st;
  const int cells = f->ncf * f->ncf;
  const int nsb = (cells + 256 * kScanItems - 1) / (256 * kScanItems);
  FCHK(hipMalloc(&f->key, sizeof(int) * P.n));
  FCHK(hipMalloc(&f->slot, sizeof(int) * P.n));
  FCHK(hipMalloc(&f->count, sizeof(int) * cells));
  FCHK(hipMalloc(&f->start, sizeof(int) * (cells + 1)));
  FCHK(hipMalloc(&f->tmp, sizeof(int) * P.n));
  FCHK(hipMalloc(&f->sorted, sizeof(int) * P.n));
  FCHK(hipMalloc(&f->blocksum, sizeof(int) * (nsb + 1)));
  FCHK(hipMalloc(&f->snap, sizeof(float4) * P.n));
  FCHK(hipMalloc(&f->omega_f, sizeof(float) * P.n));
  launch_fill_int(st, cells, f->count, 0);
  *out = f;
  return SPSS_OK;
}

void fast_destroy(FastPath *f) {
  if (!f) return;
  void *b[] = {f->key, f->slot, f->count, f->start, f->tmp, f->sorted, f->blocksum, f->snap, f->omega_f};
  for (void *p : b) hipFree(p);
  delete f;
}

void fast_set_stream(FastPath *f, hipStream_t st) {
  if (f) f->st = st;
}

void fast_set_timer(FastPath *f, KernelTimer *t) {
  if (f) f->timer = t;
}

// the XCD-aware tile order needs a grid that is a multiple of the 8 XCDs
static int tile_grid(int ntiles) { return ((ntiles + 7) / 8) * 8; }

static FastArgs args_of(const FastPath *f) {
  FastArgs a;
  a.n = f->n;
  a.ncf = f->ncf;
  a.k = f->k;
  a.W = f->W;
  a.chunks = f->chunks;
  a.ntiles = f->ntiles;
  a.H = f->H;
  a.r2lo = f->r2lo;
  a.r2hi = f->r2hi;
  a.alpha_rev = f->alpha_rev;
  return a;
}

static void build(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, bool omega) {
  const int n = f->n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  const int cells = f->ncf * f->ncf;
  const int nsb = (cells + 256 * kScanItems - 1) / (256 * kScanItems);
  if (P.s == 4)
    fk_keys<4><<<nb, 256, 0, st>>>(P, stage, f->ncf, f->key, f->slot, f->count);
  else
    fk_keys<3><<<nb, 256, 0, st>>>(P, stage, f->ncf, f->key, f->slot, f->count);
  fk_scan_local<<<nsb, 256, 0, st>>>(f->count, cells, f->start, f->blocksum);
  fk_scan_add<<<nsb, 256, 0, st>>>(f->count, cells, f->start, f->blocksum, nsb);
  fk_scatter<<<nb, 256, 0, st>>>(n, f->key, f->slot, f->start, f->tmp);
  if (P.s == 4)
    fk_rank<4><<<nb, 256, 0, st>>>(P, stage, f->ncf, f->H, f->key, f->start, f->tmp, f->sorted, f->snap,
                                   omega ? f->omega_f : nullptr);
  else
    fk_rank<3><<<nb, 256, 0, st>>>(P, stage, f->ncf, f->H, f->key, f->start, f->tmp, f->sorted, f->snap, nullptr);
}

int fast_rhs(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const double *nat,
             double *k_next, double *compl_out) {
  const bool srk = P.higher_order && compl_out && P.s == 4;
  build(f, st, P, stage, srk);
  const FastArgs a = args_of(f);
  const unsigned grid = (unsigned)tile_grid(a.ntiles);
  if (f->timer) f->timer->begin(st);
  if (P.s == 4) {
    if (srk)
      fk_pairs<4, true, 0><<<grid, kFB, 0, st>>>(P, a, stage, f->start, f->snap, f->omega_f, nat, k_next,
                                                 compl_out, nullptr, nullptr, nullptr, StepArgs{});
    else
      fk_pairs<4, false, 0><<<grid, kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nat, k_next,
                                                  nullptr, nullptr, nullptr, nullptr, StepArgs{});
  } else {
    fk_pairs<3, false, 0><<<grid, kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nat, k_next,
                                                nullptr, nullptr, nullptr, nullptr, StepArgs{});
  }
  if (f->timer) f->timer->end(st);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_step(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state_in, double *state_out,
              const double *nat, double *compl_out, double dt, double D_phi, double friction, int noise_mode,
              const double *noise, uint64_t seed, uint64_t step) {
  build(f, st, P, state_in, srk != 0);
  const FastArgs a = args_of(f);
  StepArgs U{};
  U.state_out = state_out;
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
  const unsigned grid = (unsigned)tile_grid(a.ntiles);
  if (f->timer) f->timer->begin(st);
  if (P.s == 4 && srk)
    fk_pairs<4, true, 4><<<grid, kFB, 0, st>>>(P, a, state_in, f->start, f->snap, f->omega_f, nat, nullptr,
                                               compl_out, nullptr, nullptr, nullptr, U);
  else if (P.s == 4)
    fk_pairs<4, false, 3><<<grid, kFB, 0, st>>>(P, a, state_in, f->start, f->snap, nullptr, nat, nullptr, nullptr,
                                                nullptr, nullptr, nullptr, U);
  else
    fk_pairs<3, false, 3><<<grid, kFB, 0, st>>>(P, a, state_in, f->start, f->snap, nullptr, nat, nullptr, nullptr,
                                                nullptr, nullptr, nullptr, U);
  if (f->timer) f->timer->end(st);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_pairs_count(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, int *count) {
  build(f, st, P, stage, false);
  const FastArgs a = args_of(f);
  if (P.s == 4)
    fk_pairs<4, false, 1><<<tile_grid(a.ntiles), kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nullptr,
                                                    nullptr, nullptr, count, nullptr, nullptr, StepArgs{});
  else
    fk_pairs<3, false, 1><<<tile_grid(a.ntiles), kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nullptr,
                                                    nullptr, nullptr, count, nullptr, nullptr, StepArgs{});
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_pairs_fill(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const int *pstart,
                    int *plist) {
  const FastArgs a = args_of(f);
  if (P.s == 4)
    fk_pairs<4, false, 2><<<tile_grid(a.ntiles), kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nullptr,
                                                    nullptr, nullptr, nullptr, pstart, plist, StepArgs{});
  else
    fk_pairs<3, false, 2><<<tile_grid(a.ntiles), kFB, 0, st>>>(P, a, stage, f->start, f->snap, nullptr, nullptr,
                                                    nullptr, nullptr, nullptr, pstart, plist, StepArgs{});
  FCHK(hipGetLastError());
  return SPSS_OK;
}

}  // namespace spss
