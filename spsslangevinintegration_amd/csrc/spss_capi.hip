// spss_capi.hip -- the C ABI of libspss_hip.so (declared in include/spss_hip.h).
//
// A handle owns the device-resident particle system: the FP64 master state,
// the stage state, the stepper k-buffers, cell and Verlet structures, and the
// Verlet bookkeeping of the reference (should_update_lists_, the accumulated
// displacement).  Kernels run on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <string>
#include <vector>

#include "../../include/spss_hip.h"
#include "spss_fast.hpp"
#include "spss_launch.hpp"
#include "spss_scan.hpp"
#include "spss_verlet_fast.hpp"

using namespace spss;

namespace {
thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace

namespace spss {
// the message spss_last_error() returns, for the other translation units of the library
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace spss

#define HIPCHECK(call)                                                                           \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail(SPSS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define TRY(call)                 \
  do {                            \
    int r_ = (call);              \
    if (r_ != SPSS_OK) return r_; \
  } while (0)

struct spss_system {
  spss_params p{};
  int s = 4, n = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  DevParams P{};
  int ncx = 1, ncy = 1;    // linked-list grid, int(1.0/rho) (ParticleSystemFirstOrderPtr.cpp:37-38)
  int ncvx = 1, ncvy = 1;  // grid the Verlet candidates come from
  double r_min = 0.0, r_max = 0.0;
  // device buffers
  double *state = nullptr, *stage = nullptr, *k[4] = {nullptr, nullptr, nullptr, nullptr};
  double *deriv = nullptr, *compl_f = nullptr, *nat = nullptr, *noise = nullptr;
  double *partial = nullptr, *scal = nullptr, *acc = nullptr;
  double *io[3] = {nullptr, nullptr, nullptr};
  CellBuffers cells, vcells;
  int *vcount = nullptr, *vstart = nullptr, *vlist = nullptr;
  long vcap = 0;
  int *pcount = nullptr, *pstart = nullptr, *plist = nullptr;
  int *scan_blk = nullptr;  // scratch of the scans over n counts (Verlet lists, pair export)
  long pcap = 0;
  // host bookkeeping mirroring the reference members
  bool should_update = true;      // should_update_lists_
  double acc_disp = 0.0;          // accumulated_displacement_
  bool uses_verlet = false;       // uses_verlet_list_
  bool higher_order = false;      // is_higher_order_sde_integration_
  bool have_nat = false;
  long step = 0;
  long avg_count = 0;
  // steps run since the last spss_run_stats reset: launched one by one, and replayed from graphs
  long eager_steps = 0, replayed_steps = 0, captures0 = 0;
  FastPath *fast = nullptr;
  VerletFast *vfast = nullptr;  // FP32C Verlet lists (exact mode), created on the first such step
  // the fused FP32C steps advance a resident copy of the state kept in cell order (of the cell
  // list: kind 1, of the Verlet lists' last rebuild: kind 2); `state` is then behind it until
  // the next sync_master
  bool master_stale = false;
  int resident_kind = 0;
  int ap_config = kAllPairsDefault;  // all-pairs workgroup and j-tile sizes (spss_autotune_all_pairs)
  int xsplit = 1;                    // FP32C cell grid: columns per square cell (spss_set_cell_layout)
  KernelTimer *timer = nullptr;  // around the pair kernel of the RHS
  // Dormand-Prince 5(4) buffers (allocated on first use): two states and derivatives, k2..k6,
  // a stage state and a pinned host copy for observers
  double *dp[10] = {};
  double *dp_host = nullptr;
};

namespace {

// brings `state` (original order) up to date with the resident copy of the fused steps
int sync_master(spss_system *h) {
  if (!h->master_stale) return SPSS_OK;
  if (h->resident_kind == 2) {
    int rc = vf_sync(h->vfast, h->stream, h->state);
    if (rc != SPSS_OK) return fail(rc, vf_last_error());
  } else {
    int rc = fast_resident_sync(h->fast, h->stream, h->state);
    if (rc == SPSS_OK) rc = fast_check(h->fast, h->stream);
    if (rc != SPSS_OK) return fail(rc, fast_last_error());
  }
  h->master_stale = false;
  return SPSS_OK;
}

// before `state` changes by any other means than a fused step: the resident copies are dropped
int master_write(spss_system *h, bool keep_values) {
  if (keep_values) TRY(sync_master(h));
  h->master_stale = false;
  h->resident_kind = 0;
  if (h->fast) fast_resident_drop(h->fast);
  if (h->vfast) vf_drop(h->vfast);
  return SPSS_OK;
}

// a fused step of the other kind follows: the resident copy of the current kind goes back into
// the master state first
int switch_resident(spss_system *h, int kind) {
  if (h->resident_kind != 0 && h->resident_kind != kind) TRY(master_write(h, true));
  return SPSS_OK;
}

int dalloc(void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(SPSS_ERR_DEVICE, std::string("hipMalloc: ") + hipGetErrorString(e));
  return SPSS_OK;
}

template <class T>
int dalloc_n(T **p, size_t count) {
  return dalloc((void **)p, count * sizeof(T));
}

int ensure_cells(spss_system *h, CellBuffers &cb, int ncx, int ncy) {
  const int cells = ncx * ncy;
  if (cb.key && cb.cells >= cells) return SPSS_OK;
  hipFree(cb.key);
  hipFree(cb.slot);
  hipFree(cb.count);
  hipFree(cb.start);
  hipFree(cb.tmp);
  hipFree(cb.sorted);
  hipFree(cb.blocksum);
  cb = CellBuffers{};
  TRY(dalloc_n(&cb.key, h->n));
  TRY(dalloc_n(&cb.slot, h->n));
  TRY(dalloc_n(&cb.count, cells));
  TRY(dalloc_n(&cb.start, cells + 1));
  TRY(dalloc_n(&cb.tmp, h->n));
  TRY(dalloc_n(&cb.sorted, h->n));
  TRY(dalloc_n(&cb.blocksum, scan_blocks(cells) + 1));
  HIPCHECK(hipMemsetAsync(cb.count, 0, sizeof(int) * cells, h->stream));
  cb.cells = cells;
  return SPSS_OK;
}

// scratch of the scans over n counts
int ensure_scan(spss_system *h) {
  if (h->scan_blk) return SPSS_OK;
  return dalloc_n(&h->scan_blk, scan_blocks(h->n) + 1);
}

// EvaluateRhs's dispatch on rho (ParticleSystemFirstOrderPtr.cpp:131-153; model 6:
// ParticleSystemSecondOrderPtr.cpp:125-138)
int resolve_method(const spss_system *h, int method) {
  if (method != SPSS_METHOD_AUTO) return method;
  const double rho = h->p.rho;
  if (h->s == 6) return rho <= 0.25 ? SPSS_METHOD_LINKED_LIST : SPSS_METHOD_ALL_PAIRS;
  if (rho >= M_SQRT1_2) return SPSS_METHOD_GLOBAL_FIELD;
  if (rho <= 0.25) return SPSS_METHOD_LINKED_LIST_VERLET;
  if (rho <= 0.5) return SPSS_METHOD_VERLET;
  return SPSS_METHOD_ALL_PAIRS;
}

int check_method(const spss_system *h, int m) {
  if (m < SPSS_METHOD_GLOBAL_FIELD || m > SPSS_METHOD_LINKED_LIST_VERLET)
    return fail(SPSS_ERR_INVALID, "unknown method");
  if (h->s == 6 && (m == SPSS_METHOD_GLOBAL_FIELD || m == SPSS_METHOD_VERLET || m == SPSS_METHOD_LINKED_LIST_VERLET))
    return fail(SPSS_ERR_UNSUPPORTED,
                "model 6 has no global-field or Verlet branch (ParticleSystemSecondOrderPtr.cpp:393-400 is a stub)");
  if (h->p.precision == SPSS_PRECISION_FP32C && m != SPSS_METHOD_LINKED_LIST)
    return fail(SPSS_ERR_UNSUPPORTED, "FP32C precision evaluates a single RHS with the linked-list method only "
                                      "(Verlet lists run as fused stochastic steps)");
  if (h->p.precision == SPSS_PRECISION_FP32C && h->s == 6)
    return fail(SPSS_ERR_UNSUPPORTED, "FP32C precision is implemented for models 3 and 4");
  return SPSS_OK;
}

// Verlet list (re)build into vstart/vlist; two passes with the list size read back in between
int build_verlet(spss_system *h, int m) {
  hipStream_t st = h->stream;
  const double rmax2 = h->r_max * h->r_max;
  if (!h->vcount) {
    TRY(dalloc_n(&h->vcount, h->n + 1));
    TRY(dalloc_n(&h->vstart, h->n + 1));
  }
  if (m == SPSS_METHOD_VERLET) {
    launch_verlet_count_all(st, h->P, h->stage, rmax2, h->vcount);
  } else {
    TRY(ensure_cells(h, h->vcells, h->ncvx, h->ncvy));
    launch_cell_build(st, h->P, h->stage, h->ncvx, h->ncvy, h->vcells);
    launch_verlet_count_cells(st, h->P, h->stage, h->ncvx, h->ncvy, h->vcells, rmax2, h->vcount);
  }
  TRY(ensure_scan(h));
  launch_scan(st, h->vcount, h->vstart, h->n, h->scan_blk);
  int total = 0;
  HIPCHECK(hipMemcpyAsync(&total, h->vstart + h->n, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (total < 0) return fail(SPSS_ERR_CAPACITY, "Verlet list size overflows int32");
  if ((long)total > h->vcap) {
    hipFree(h->vlist);
    h->vlist = nullptr;
    h->vcap = std::max<long>((long)total + total / 4, 1024);
    TRY(dalloc_n(&h->vlist, (size_t)h->vcap));
  }
  if (m == SPSS_METHOD_VERLET)
    launch_verlet_fill_all(st, h->P, h->stage, rmax2, h->vstart, h->vlist);
  else
    launch_verlet_fill_cells(st, h->P, h->stage, h->ncvx, h->ncvy, h->vcells, rmax2, h->vstart, h->vlist);
  return SPSS_OK;
}

int rhs_device(spss_system *h, int method, const double *d_state, const double *d_kprev, double *d_knext,
               double k_coef, double dt) {
  const int m = resolve_method(h, method);
  TRY(check_method(h, m));
  hipStream_t st = h->stream;
  h->P.higher_order = h->higher_order ? 1 : 0;
  const double *nat = h->have_nat ? h->nat : nullptr;
  double *compl_out = h->higher_order ? h->compl_f : nullptr;
  launch_stage(st, (long)h->s * h->n, d_state, d_kprev, k_coef * dt, h->stage);
  switch (m) {
    case SPSS_METHOD_GLOBAL_FIELD:
      launch_global(st, h->P, h->stage, h->partial, h->scal, nat, d_knext, compl_out);
      break;
    case SPSS_METHOD_ALL_PAIRS:
      launch_all_pairs(st, h->P, h->stage, nat, d_knext, compl_out, h->ap_config);
      break;
    case SPSS_METHOD_LINKED_LIST:
      if (h->p.precision == SPSS_PRECISION_FP32C) {
        TRY(fast_rhs(h->fast, st, h->P, h->stage, nat, d_knext, compl_out));
      } else {
        TRY(ensure_cells(h, h->cells, h->ncx, h->ncy));
        launch_cell_build(st, h->P, h->stage, h->ncx, h->ncy, h->cells);
        if (h->timer) h->timer->begin(st);
        launch_linked_list(st, h->P, h->stage, h->ncx, h->ncy, h->cells, nat, d_knext, compl_out);
        if (h->timer) h->timer->end(st);
      }
      break;
    case SPSS_METHOD_VERLET:
    case SPSS_METHOD_LINKED_LIST_VERLET:
      if (h->should_update) {
        TRY(build_verlet(h, m));
        h->should_update = false;
      }
      launch_verlet_pairs(st, h->P, h->stage, h->vstart, h->vlist, nat, d_knext, compl_out);
      h->uses_verlet = true;
      break;
    default:
      return fail(SPSS_ERR_INVALID, "unknown method");
  }
  HIPCHECK(hipGetLastError());
  return SPSS_OK;
}

// CalculateMaxDisplacement's bookkeeping, ParticleSystemFirstOrderPtr.cpp:851-860.  Exact mode
// adds the step's own displacement as a margin for the RK stages of the next step.
void update_verlet_flag(spss_system *h, double max2) {
  const double stepd = std::sqrt(max2);
  h->acc_disp += stepd;
  bool flag = false;
  if (h->p.verlet_mode == SPSS_VERLET_EXACT) {
    if (h->acc_disp + stepd > 0.5 * (h->r_max - h->r_min)) {
      h->acc_disp = 0.0;
      flag = true;
    }
  } else if (h->acc_disp > 0.5 * (h->r_max - h->r_min)) {
    h->acc_disp = 0.0;
    flag = true;
  }
  h->should_update = flag;
}

int finish_step_displacement(spss_system *h) {
  if (!h->uses_verlet || h->s == 6) return SPSS_OK;
  const int nb = (h->n + kBlock - 1) / kBlock;
  launch_max_final(h->stream, nb, h->partial, h->scal + 8);
  double max2 = 0.0;
  HIPCHECK(hipMemcpyAsync(&max2, h->scal + 8, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  update_verlet_flag(h, max2);
  return SPSS_OK;
}

int noise_source(spss_system *h, int mode, const double *noise, size_t count, int *kmode, const double **kptr) {
  *kptr = nullptr;
  switch (mode) {
    case SPSS_NOISE_NONE: *kmode = 0; return SPSS_OK;
    case SPSS_NOISE_HOST:
      if (!noise) return fail(SPSS_ERR_INVALID, "noise array is NULL");
      HIPCHECK(hipMemcpyAsync(h->noise, noise, count * sizeof(double), hipMemcpyHostToDevice, h->stream));
      *kmode = 1;
      *kptr = h->noise;
      return SPSS_OK;
    case SPSS_NOISE_DEVICE:
      if (!noise) return fail(SPSS_ERR_INVALID, "noise array is NULL");
      *kmode = 1;
      *kptr = noise;
      return SPSS_OK;
    case SPSS_NOISE_PHILOX: *kmode = 3; return SPSS_OK;
    default: return fail(SPSS_ERR_INVALID, "unknown noise mode");
  }
}

int rk4_step(spss_system *h, int method, double dt) {
  TRY(master_write(h, true));
  TRY(rhs_device(h, method, h->state, h->k[0], h->k[0], 0.0, dt));
  TRY(rhs_device(h, method, h->state, h->k[0], h->k[1], 0.5, dt));
  TRY(rhs_device(h, method, h->state, h->k[1], h->k[2], 0.5, dt));
  TRY(rhs_device(h, method, h->state, h->k[2], h->k[3], 1.0, dt));
  const bool track = h->uses_verlet && h->s != 6;
  launch_rk4_combine(h->stream, h->s, h->n, h->state, h->k[0], h->k[1], h->k[2], h->k[3], dt,
                     track ? h->partial : nullptr);
  HIPCHECK(hipGetLastError());
  ++h->step;
  ++h->eager_steps;
  return finish_step_displacement(h);
}

// FP32C linked list: the fused step of the fast path on its resident copy of the state
int fast_fused_step(spss_system *h, int srk, double dt, int noise_mode, const double *noise) {
  TRY(switch_resident(h, 1));
  int kmode;
  const double *kptr;
  TRY(noise_source(h, noise_mode, noise, (size_t)(srk ? 2 : 1) * h->n, &kmode, &kptr));
  h->P.higher_order = srk ? 1 : 0;
  int rc = fast_step(h->fast, h->stream, h->P, srk, h->state, h->have_nat ? h->nat : nullptr,
                     srk ? h->compl_f : nullptr, dt, h->p.D_phi, h->p.xi, kmode, kptr, h->p.seed,
                     (uint64_t)h->step);
  if (rc != SPSS_OK) return fail(rc, fast_last_error());
  h->master_stale = true;
  h->resident_kind = 1;
  ++h->step;
  ++h->eager_steps;
  return SPSS_OK;
}

bool use_fused(const spss_system *h, int method) {
  return h->p.precision == SPSS_PRECISION_FP32C && resolve_method(h, method) == SPSS_METHOD_LINKED_LIST &&
         h->s != 6;
}

// FP32C with Verlet lists: fused stochastic steps on the Verlet path's resident copy
bool use_fused_verlet(const spss_system *h, int method) {
  const int m = resolve_method(h, method);
  return h->p.precision == SPSS_PRECISION_FP32C && (m == SPSS_METHOD_VERLET || m == SPSS_METHOD_LINKED_LIST_VERLET) &&
         h->s != 6;
}

int ensure_vfast(spss_system *h) {
  if (h->vfast) return SPSS_OK;
  if (h->p.verlet_mode != SPSS_VERLET_EXACT)
    return fail(SPSS_ERR_UNSUPPORTED, "FP32C Verlet lists are implemented in the exact mode (SPSS_VERLET_EXACT); "
                                      "the reference-compatible mode runs in FP64");
  int rc = vf_create(&h->vfast, h->stream, h->P, h->p.rho, h->r_max);
  if (rc != SPSS_OK) return fail(rc, vf_last_error());
  return SPSS_OK;
}

int verlet_fused_step(spss_system *h, int srk, double dt, int noise_mode, const double *noise) {
  TRY(ensure_vfast(h));
  TRY(switch_resident(h, 2));
  int kmode;
  const double *kptr;
  TRY(noise_source(h, noise_mode, noise, (size_t)(srk ? 2 : 1) * h->n, &kmode, &kptr));
  h->P.higher_order = srk ? 1 : 0;
  int rc = vf_step(h->vfast, h->stream, h->P, srk, h->state, h->have_nat ? h->nat : nullptr,
                   srk ? h->compl_f : nullptr, dt, h->p.D_phi, h->p.xi, kmode, kptr, h->p.seed, (uint64_t)h->step);
  if (rc != SPSS_OK) return fail(rc, vf_last_error());
  h->master_stale = true;
  h->resident_kind = 2;
  h->uses_verlet = true;
  ++h->step;
  ++h->eager_steps;
  return SPSS_OK;
}

int euler_step(spss_system *h, int method, double dt, int noise_mode, const double *noise) {
  if (use_fused(h, method)) return fast_fused_step(h, 0, dt, noise_mode, noise);
  if (use_fused_verlet(h, method)) return verlet_fused_step(h, 0, dt, noise_mode, noise);
  TRY(master_write(h, true));
  const size_t m = (size_t)h->s * h->n;
  launch_fill_double(h->stream, (long)m, h->deriv, 0.0);  // std::fill(derivative, 0), :43
  TRY(rhs_device(h, method, h->state, h->deriv, h->deriv, 0.0, dt));
  int kmode;
  const double *kptr;
  TRY(noise_source(h, noise_mode, noise, (size_t)(h->s == 6 ? 3 : 1) * h->n, &kmode, &kptr));
  const bool track = h->uses_verlet && h->s != 6;
  launch_euler(h->stream, h->s, h->n, h->state, h->deriv, dt, std::sqrt(dt), std::sqrt(2.0 * h->p.D_phi),
               std::sqrt(2.0 * h->p.D_r), kmode, kptr, h->p.seed, (uint64_t)h->step,
               track ? h->partial : nullptr);
  HIPCHECK(hipGetLastError());
  ++h->step;
  ++h->eager_steps;
  return finish_step_displacement(h);
}

int srk_step(spss_system *h, int method, double dt, int noise_mode, const double *noise) {
  if (h->s != 4) return fail(SPSS_ERR_UNSUPPORTED, "the stochastic Runge-Kutta stepper is defined for model 4");
  if (use_fused(h, method)) {
    h->higher_order = true;  // StochasticRungeKuttaStepperPtr.hpp:44
    return fast_fused_step(h, 1, dt, noise_mode, noise);
  }
  if (use_fused_verlet(h, method)) {
    h->higher_order = true;
    return verlet_fused_step(h, 1, dt, noise_mode, noise);
  }
  TRY(master_write(h, true));
  const size_t m = (size_t)h->s * h->n;
  launch_fill_double(h->stream, (long)m, h->deriv, 0.0);
  h->higher_order = true;  // StochasticRungeKuttaStepperPtr.hpp:44
  TRY(rhs_device(h, method, h->state, h->deriv, h->deriv, 0.0, dt));
  int kmode;
  const double *kptr;
  TRY(noise_source(h, noise_mode, noise, (size_t)2 * h->n, &kmode, &kptr));
  const bool track = h->uses_verlet;
  launch_srk(h->stream, h->n, h->state, h->deriv, h->compl_f, dt, std::sqrt(dt), 0.5 * std::pow(dt, 1.5),
             1.0 / std::sqrt(3.0), std::sqrt(2.0 * h->p.D_phi), h->p.xi, kmode, kptr, h->p.seed, (uint64_t)h->step,
             track ? h->partial : nullptr);
  HIPCHECK(hipGetLastError());
  ++h->step;
  ++h->eager_steps;
  return finish_step_displacement(h);
}

}  // namespace

extern "C" {

int spss_abi_version(void) { return SPSS_ABI_VERSION; }

const char *spss_last_error(void) { return g_err.c_str(); }

int spss_params_default(spss_params *p) {
  if (!p) return fail(SPSS_ERR_INVALID, "params is NULL");
  std::memset(p, 0, sizeof(*p));
  p->model = 4;
  p->n = 1000;
  p->v0 = 1.0;  // SimulationEngineFirstOrderPtr.cpp:145-150
  p->xi = 0.1;
  p->sigma = 1.0;
  p->rho = 0.8;
  p->alpha = 0.3;
  p->D_phi = 0.01;
  p->xi_r = 1.0;
  p->x_size = 1.0;
  p->y_size = 1.0;
  p->verlet_skin = 0.05;
  p->verlet_mode = SPSS_VERLET_COMPAT;
  p->precision = SPSS_PRECISION_FP64;
  p->cell_subdivision = 1;
  p->seed = 0x5EED0000ull;
  return SPSS_OK;
}

int spss_create(const spss_params *p, spss_system **out) {
  if (!p || !out) return fail(SPSS_ERR_INVALID, "NULL argument");
  *out = nullptr;
  if (p->model != 3 && p->model != 4 && p->model != 6) return fail(SPSS_ERR_INVALID, "model must be 3, 4 or 6");
  if (p->n <= 0) return fail(SPSS_ERR_INVALID, "n must be positive");
  if (!(p->rho > 0.0)) return fail(SPSS_ERR_INVALID, "rho must be positive");
  if (!(p->x_size > 0.0) || !(p->y_size > 0.0)) return fail(SPSS_ERR_INVALID, "box sizes must be positive");
  if ((long)p->n * p->model > (1L << 31) - 1) return fail(SPSS_ERR_CAPACITY, "s*n exceeds int32 indexing");
  spss_system *h = new spss_system();
  h->p = *p;
  h->s = p->model;
  h->n = p->n;
  if (h->p.cell_subdivision <= 0) h->p.cell_subdivision = 1;
  DevParams &P = h->P;
  P.n = h->n;
  P.s = h->s;
  P.v0 = p->v0;
  P.xi = p->xi;
  P.sigma = p->sigma;
  P.rho2 = p->rho * p->rho;
  P.alpha = p->alpha;
  P.cos_a = std::cos(p->alpha);
  P.sin_a = std::sin(p->alpha);
  P.xi_r = p->xi_r;
  P.xs = p->x_size;
  P.ys = p->y_size;
  P.xrs = 1.0 / p->x_size;
  P.yrs = 1.0 / p->y_size;
  P.natfreq_local = p->natfreq_local;
  h->ncx = h->ncy = std::max(1, (int)(1.0 / p->rho));
  h->r_min = p->rho;
  if (p->verlet_mode == SPSS_VERLET_EXACT) {
    if (!(p->verlet_skin > 0.0)) {
      delete h;
      return fail(SPSS_ERR_INVALID, "exact Verlet mode needs a positive skin");
    }
    h->r_max = p->rho + p->verlet_skin;
    h->ncvx = h->ncvy = std::max(1, (int)(1.0 / h->r_max));
  } else {
    h->r_max = (h->s == 6) ? 2 * p->rho : p->rho + 0.05;  // ParticleSystemFirstOrderPtr.cpp:97-98
    h->ncvx = h->ncx;
    h->ncvy = h->ncy;
  }
  int rc = SPSS_OK;
  auto step = [&](int r) {
    if (rc == SPSS_OK) rc = r;
  };
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete h;
    return fail(SPSS_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  h->own_stream = true;
  const size_t m = (size_t)h->s * h->n;
  const size_t nb = (size_t)(h->n + kBlock - 1) / kBlock;
  step(dalloc_n(&h->state, m));
  step(dalloc_n(&h->stage, m));
  for (int q = 0; q < 4; ++q) step(dalloc_n(&h->k[q], m));
  step(dalloc_n(&h->deriv, m));
  step(dalloc_n(&h->compl_f, h->n));
  step(dalloc_n(&h->nat, h->n));
  step(dalloc_n(&h->noise, (size_t)3 * h->n));
  step(dalloc_n(&h->partial, nb * 4 + 64));
  step(dalloc_n(&h->scal, 64));
  if (rc != SPSS_OK) {
    spss_destroy(h);
    return rc;
  }
  // the stepper k-buffers start at zero (RungeKutta4StepperPtr.cpp:48-51) and are never cleared
  for (int q = 0; q < 4; ++q) launch_fill_double(h->stream, (long)m, h->k[q], 0.0);
  launch_fill_double(h->stream, (long)m, h->state, 0.0);
  launch_fill_double(h->stream, (long)m, h->deriv, 0.0);
  launch_fill_double(h->stream, h->n, h->compl_f, 0.0);
  launch_fill_double(h->stream, h->n, h->nat, 0.0);
  if (h->p.precision == SPSS_PRECISION_FP32C) {
    if (h->s == 6) {
      spss_destroy(h);
      return fail(SPSS_ERR_UNSUPPORTED, "FP32C precision is implemented for models 3 and 4");
    }
    rc = fast_create(&h->fast, h->stream, h->P, h->p.rho, h->ncx, h->p.cell_subdivision, h->xsplit);
    if (rc != SPSS_OK) {
      spss_destroy(h);
      return fail(rc, fast_last_error());
    }
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    spss_destroy(h);
    return fail(SPSS_ERR_DEVICE, "initialisation failed");
  }
  *out = h;
  return SPSS_OK;
}

int spss_destroy(spss_system *h) {
  if (!h) return SPSS_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  fast_destroy(h->fast);
  vf_destroy(h->vfast);
  delete h->timer;
  void *bufs[] = {h->state, h->stage, h->k[0], h->k[1], h->k[2], h->k[3], h->deriv, h->compl_f, h->nat,
                  h->noise, h->partial, h->scal, h->acc, h->io[0], h->io[1], h->io[2], h->vcount, h->vstart,
                  h->vlist, h->pcount, h->pstart, h->plist, h->scan_blk};
  for (void *b : bufs) hipFree(b);
  for (double *b : h->dp) hipFree(b);
  if (h->dp_host) hipHostFree(h->dp_host);
  for (CellBuffers *cb : {&h->cells, &h->vcells}) {
    hipFree(cb->key);
    hipFree(cb->slot);
    hipFree(cb->count);
    hipFree(cb->start);
    hipFree(cb->tmp);
    hipFree(cb->sorted);
    hipFree(cb->blocksum);
  }
  if (h->own_stream && h->stream) hipStreamDestroy(h->stream);
  delete h;
  return SPSS_OK;
}

int spss_set_stream(spss_system *h, void *stream) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  HIPCHECK(hipStreamSynchronize(h->stream));
  if (h->own_stream) hipStreamDestroy(h->stream);
  if (stream) {
    h->stream = (hipStream_t)stream;
    h->own_stream = false;
  } else {
    HIPCHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    h->own_stream = true;
  }
  if (h->fast) fast_set_stream(h->fast, h->stream);
  return SPSS_OK;
}

int spss_synchronize(spss_system *h) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

int spss_set_state(spss_system *h, const double *host_state) {
  if (!h || !host_state) return fail(SPSS_ERR_INVALID, "NULL argument");
  TRY(master_write(h, false));
  HIPCHECK(hipMemcpyAsync(h->state, host_state, sizeof(double) * h->s * h->n, hipMemcpyHostToDevice, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

int spss_get_state(spss_system *h, double *host_state) {
  if (!h || !host_state) return fail(SPSS_ERR_INVALID, "NULL argument");
  TRY(sync_master(h));
  HIPCHECK(hipMemcpyAsync(host_state, h->state, sizeof(double) * h->s * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

int spss_set_state_device(spss_system *h, const double *d) {
  if (!h || !d) return fail(SPSS_ERR_INVALID, "NULL argument");
  TRY(master_write(h, false));
  HIPCHECK(hipMemcpyAsync(h->state, d, sizeof(double) * h->s * h->n, hipMemcpyDeviceToDevice, h->stream));
  return SPSS_OK;
}

// the caller may read and write the returned buffer: the resident copy is synced and dropped
double *spss_state_device(spss_system *h) {
  if (!h || master_write(h, true) != SPSS_OK) return nullptr;
  return h->state;
}

int spss_set_natural_frequencies(spss_system *h, const double *w, int n) {
  if (!h || !w) return fail(SPSS_ERR_INVALID, "NULL argument");
  if (n != h->n) return fail(SPSS_ERR_INVALID, "natural frequencies must have n entries");
  TRY(master_write(h, true));
  HIPCHECK(hipMemcpyAsync(h->nat, w, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  h->have_nat = true;
  return SPSS_OK;
}

int spss_set_higher_order(spss_system *h, int on) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (on && h->s != 4) return fail(SPSS_ERR_UNSUPPORTED, "the complementary force is defined for model 4");
  h->higher_order = on != 0;
  return SPSS_OK;
}

int spss_get_complementary_force(spss_system *h, double *out) {
  if (!h || !out) return fail(SPSS_ERR_INVALID, "NULL argument");
  HIPCHECK(hipMemcpyAsync(out, h->compl_f, sizeof(double) * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

int spss_evaluate_rhs_device(spss_system *h, int method, const double *d_state, const double *d_kprev,
                             double *d_knext, double k_coef, double dt) {
  if (!h || !d_state || !d_kprev || !d_knext) return fail(SPSS_ERR_INVALID, "NULL argument");
  return rhs_device(h, method, d_state, d_kprev, d_knext, k_coef, dt);
}

int spss_evaluate_rhs(spss_system *h, int method, const double *state, const double *k_prev, double *k_next,
                      double k_coef, double dt) {
  if (!h || !state || !k_prev || !k_next) return fail(SPSS_ERR_INVALID, "NULL argument");
  const size_t m = (size_t)h->s * h->n;
  for (int q = 0; q < 3; ++q)
    if (!h->io[q]) TRY(dalloc_n(&h->io[q], m));
  hipStream_t st = h->stream;
  HIPCHECK(hipMemcpyAsync(h->io[0], state, m * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(h->io[1], k_prev, m * sizeof(double), hipMemcpyHostToDevice, st));
  // k_next entries the RHS leaves untouched (k[3] of isolated particles) keep their values
  HIPCHECK(hipMemcpyAsync(h->io[2], k_next, m * sizeof(double), hipMemcpyHostToDevice, st));
  TRY(rhs_device(h, method, h->io[0], h->io[1], h->io[2], k_coef, dt));
  HIPCHECK(hipMemcpyAsync(k_next, h->io[2], m * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  return SPSS_OK;
}

int spss_rk4_step(spss_system *h, int method, double dt) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  return rk4_step(h, method, dt);
}

int spss_euler_step(spss_system *h, int method, double dt, int noise_mode, const double *noise) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  return euler_step(h, method, dt, noise_mode, noise);
}

int spss_srk_step(spss_system *h, int method, double dt, int noise_mode, const double *noise) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  return srk_step(h, method, dt, noise_mode, noise);
}

// the fused FP32C stochastic steps with Philox noise that spss_run can replay from graphs:
// 1 on the cell-list path, 2 on the Verlet-list path, 0 none
int graph_run_kind(const spss_system *h, int stepper, int method, int noise_mode) {
  const bool srk = stepper == SPSS_STEPPER_SRK && h->s == 4;
  if (!(stepper == SPSS_STEPPER_EULER || srk) || noise_mode != SPSS_NOISE_PHILOX) return 0;
  if (use_fused(h, method)) return 1;
  if (use_fused_verlet(h, method) && h->p.verlet_mode == SPSS_VERLET_EXACT) return 2;
  return 0;
}

bool kind_ready(const spss_system *h, int kind, const double *nat, int srk) {
  if (h->resident_kind != kind) return false;
  return kind == 1 ? fast_graph_ready(h->fast, nat, srk) : vf_graph_ready(h->vfast, nat, srk);
}

int spss_run(spss_system *h, int stepper, int method, long nsteps, double dt, int noise_mode) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (noise_mode != SPSS_NOISE_NONE && noise_mode != SPSS_NOISE_PHILOX)
    return fail(SPSS_ERR_INVALID, "spss_run draws noise from NONE or PHILOX only");
  long k = 0;
  // fused FP32C stochastic steps with Philox noise: a step that enters the resident order runs
  // eagerly, every other step is replayed from HIP graphs
  const bool srk = stepper == SPSS_STEPPER_SRK && h->s == 4;
  const double *nat = h->have_nat ? h->nat : nullptr;
  const int kind = graph_run_kind(h, stepper, method, noise_mode);
  if (kind && nsteps > 0) {
    if (!kind_ready(h, kind, nat, srk ? 1 : 0)) {
      TRY(srk ? srk_step(h, method, dt, noise_mode, nullptr) : euler_step(h, method, dt, noise_mode, nullptr));
      k = 1;
    }
    h->P.higher_order = srk ? 1 : 0;
    if (srk) h->higher_order = true;  // StochasticRungeKuttaStepperPtr.hpp:44
    const long g =
        kind == 1 ? fast_run_graph(h->fast, h->stream, h->P, srk ? 1 : 0, h->state, nat, srk ? h->compl_f : nullptr,
                                   dt, h->p.D_phi, h->p.xi, h->p.seed, (uint64_t)h->step, nsteps - k)
                  : vf_run_graph(h->vfast, h->stream, h->P, srk ? 1 : 0, h->state, nat, srk ? h->compl_f : nullptr,
                                 dt, h->p.D_phi, h->p.xi, h->p.seed, (uint64_t)h->step, nsteps - k);
    if (g < 0) return fail(SPSS_ERR_DEVICE, kind == 1 ? fast_last_error() : vf_last_error());
    if (g > 0) h->master_stale = true;
    h->step += g;
    h->replayed_steps += g;
    k += g;
  }
  for (; k < nsteps; ++k) {
    switch (stepper) {
      case SPSS_STEPPER_RK4: TRY(rk4_step(h, method, dt)); break;
      case SPSS_STEPPER_EULER: TRY(euler_step(h, method, dt, noise_mode, nullptr)); break;
      case SPSS_STEPPER_SRK: TRY(srk_step(h, method, dt, noise_mode, nullptr)); break;
      default: return fail(SPSS_ERR_INVALID, "unknown stepper");
    }
  }
  return SPSS_OK;
}

int spss_prepare_run(spss_system *h, int stepper, int method, double dt, int noise_mode) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  const int kind = graph_run_kind(h, stepper, method, noise_mode);
  if (!kind) return SPSS_OK;  // spss_run steps eagerly here
  const bool srk = stepper == SPSS_STEPPER_SRK && h->s == 4;
  const double *nat = h->have_nat ? h->nat : nullptr;
  if (kind == 2) TRY(ensure_vfast(h));
  if (!kind_ready(h, kind, nat, srk ? 1 : 0)) {
    TRY(master_write(h, true));  // the resident order is entered from the master state
    int rc = kind == 1 ? fast_enter(h->fast, h->stream, h->P, srk ? 1 : 0, h->state, nat)
                       : vf_enter(h->vfast, h->stream, h->P, srk ? 1 : 0, h->state, nat);
    if (rc != SPSS_OK) return fail(rc, kind == 1 ? fast_last_error() : vf_last_error());
    h->resident_kind = kind;
    if (kind == 2) h->uses_verlet = true;
  }
  h->P.higher_order = srk ? 1 : 0;
  if (srk) h->higher_order = true;
  int rc = kind == 1 ? fast_prepare_graphs(h->fast, h->stream, h->P, srk ? 1 : 0, h->state, nat,
                                           srk ? h->compl_f : nullptr, dt, h->p.D_phi, h->p.xi, h->p.seed)
                     : vf_prepare_graphs(h->vfast, h->stream, h->P, srk ? 1 : 0, h->state, nat,
                                         srk ? h->compl_f : nullptr, dt, h->p.D_phi, h->p.xi, h->p.seed);
  if (rc != SPSS_OK) return fail(rc, kind == 1 ? fast_last_error() : vf_last_error());
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

int spss_run_stats(spss_system *h, long *eager_steps, long *replayed_steps, long *graphs_captured, int reset) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  const long cap = (h->fast ? fast_graph_captures(h->fast) : 0) + (h->vfast ? vf_graph_captures(h->vfast) : 0);
  if (eager_steps) *eager_steps = h->eager_steps;
  if (replayed_steps) *replayed_steps = h->replayed_steps;
  if (graphs_captured) *graphs_captured = cap - h->captures0;
  if (reset) {
    h->eager_steps = 0;
    h->replayed_steps = 0;
    h->captures0 = cap;
  }
  return SPSS_OK;
}

int spss_get_step(spss_system *h, long *step) {
  if (!h || !step) return fail(SPSS_ERR_INVALID, "NULL argument");
  *step = h->step;
  return SPSS_OK;
}

int spss_set_step(spss_system *h, long step) {
  if (!h || step < 0) return fail(SPSS_ERR_INVALID, "bad argument");
  h->step = step;
  return SPSS_OK;
}

int spss_autotune_cells(spss_system *h, int stepper, double dt, int probe_steps, double *table_ms, int table_cap,
                        int *chosen) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (h->p.precision != SPSS_PRECISION_FP32C || !h->fast)
    return fail(SPSS_ERR_UNSUPPORTED, "the cell autotuner needs the FP32C linked-list path");
  if (stepper != SPSS_STEPPER_EULER && stepper != SPSS_STEPPER_SRK)
    return fail(SPSS_ERR_INVALID, "the cell autotuner times the stochastic Euler or Runge-Kutta step");
  if (probe_steps < 1 || !(dt > 0.0)) return fail(SPSS_ERR_INVALID, "probe_steps >= 1 and dt > 0 required");
  // what the probes change and must be put back: the state, the complementary force, the
  // step counter and the higher-order flags
  TRY(sync_master(h));
  const size_t m = (size_t)h->s * h->n;
  double *keep = nullptr, *keep_c = nullptr;
  TRY(dalloc_n(&keep, m));
  if (dalloc_n(&keep_c, h->n) != SPSS_OK) {
    hipFree(keep);
    return SPSS_ERR_DEVICE;
  }
  HIPCHECK(hipMemcpyAsync(keep, h->state, m * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HIPCHECK(hipMemcpyAsync(keep_c, h->compl_f, h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  const long step0 = h->step;
  const bool ho0 = h->higher_order;
  const int pho0 = h->P.higher_order;
  auto restore = [&]() -> int {
    TRY(master_write(h, false));
    HIPCHECK(hipMemcpyAsync(h->state, keep, m * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHECK(hipMemcpyAsync(h->compl_f, keep_c, h->n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    h->step = step0;
    h->higher_order = ho0;
    h->P.higher_order = pho0;
    return SPSS_OK;
  };
  const int k0 = h->p.cell_subdivision, x0 = h->xsplit;
  int best = k0, best_x = x0;
  double best_ms = 1e300;
  int rc = SPSS_OK;
  for (int c = 0; c < kCellLayouts && rc == SPSS_OK; ++c) {
    const int k = kCellLayout[c][0], xs = kCellLayout[c][1];
    FastPath *f = nullptr;
    if (fast_create(&f, h->stream, h->P, h->p.rho, h->ncx, k, xs) != SPSS_OK) {  // grid too fine: not a candidate
      if (c < table_cap && table_ms) table_ms[c] = -1.0;
      continue;
    }
    if ((rc = restore()) != SPSS_OK) {
      fast_destroy(f);
      break;
    }
    fast_destroy(h->fast);
    h->fast = f;
    h->p.cell_subdivision = k;
    h->xsplit = xs;
    const int warm = 2;
    double ms = 0.0;
    for (int s = 0; s < warm + probe_steps && rc == SPSS_OK; ++s) {
      if (s == warm) {
        if (hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(SPSS_ERR_DEVICE, "probe step failed");
        ms = -std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
      }
      if (rc == SPSS_OK)
        rc = stepper == SPSS_STEPPER_EULER ? euler_step(h, SPSS_METHOD_LINKED_LIST, dt, SPSS_NOISE_PHILOX, nullptr)
                                           : srk_step(h, SPSS_METHOD_LINKED_LIST, dt, SPSS_NOISE_PHILOX, nullptr);
    }
    if (rc == SPSS_OK && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(SPSS_ERR_DEVICE, "probe failed");
    if (rc != SPSS_OK) break;
    ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    ms /= probe_steps;
    if (c < table_cap && table_ms) table_ms[c] = ms;
    if (ms < best_ms) best_ms = ms, best = k, best_x = xs;
  }
  // keep the fastest grid (or the original one after a failure) with the original state
  if (rc != SPSS_OK) best = k0, best_x = x0;
  if (best != h->p.cell_subdivision || best_x != h->xsplit) {
    FastPath *f = nullptr;
    const int rc1 = fast_create(&f, h->stream, h->P, h->p.rho, h->ncx, best, best_x);
    if (rc1 == SPSS_OK) {
      fast_destroy(h->fast);
      h->fast = f;
      if (h->timer) fast_set_timer(h->fast, h->timer);
      h->p.cell_subdivision = best;
      h->xsplit = best_x;
    } else if (rc == SPSS_OK) {
      rc = fail(rc1, fast_last_error());
    }
  }
  const int rc2 = restore();
  hipFree(keep);
  hipFree(keep_c);
  if (rc != SPSS_OK) return rc;
  if (rc2 != SPSS_OK) return rc2;
  if (chosen) *chosen = h->p.cell_subdivision;
  return SPSS_OK;
}

int spss_autotune_all_pairs(spss_system *h, int probe_evals, double *table_ms, int table_cap, int *chosen) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (probe_evals < 1) return fail(SPSS_ERR_INVALID, "probe_evals >= 1 required");
  TRY(sync_master(h));
  const size_t m = (size_t)h->s * h->n;
  double *k = nullptr;
  TRY(dalloc_n(&k, m));
  const double *nat = h->have_nat ? h->nat : nullptr;
  double *compl_out = h->higher_order ? h->compl_f : nullptr;
  h->P.higher_order = h->higher_order ? 1 : 0;
  launch_stage(h->stream, (long)m, h->state, nullptr, 0.0, h->stage);
  double best = 1e300;
  int best_c = h->ap_config;
  int rc = SPSS_OK;
  for (int c = 0; c < kAllPairsConfigCount && rc == SPSS_OK; ++c) {
    launch_all_pairs(h->stream, h->P, h->stage, nat, k, compl_out, c);  // warm-up
    EventPair ev;
    hipError_t e = ev.create();
    if (e == hipSuccess) e = hipEventRecord(ev.start, h->stream);
    for (int r = 0; r < probe_evals && e == hipSuccess; ++r)
      launch_all_pairs(h->stream, h->P, h->stage, nat, k, compl_out, c);
    if (e == hipSuccess) e = hipEventRecord(ev.stop, h->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = ev.elapsed(&ms);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
      rc = fail(SPSS_ERR_DEVICE, std::string("all-pairs autotuner: ") + hipGetErrorString(e));
      break;
    }
    const double per = ms / probe_evals;
    if (table_ms && c < table_cap) table_ms[c] = per;
    if (per < best) {
      best = per;
      best_c = c;
    }
  }
  hipStreamSynchronize(h->stream);
  hipFree(k);
  if (rc != SPSS_OK) return rc;
  h->ap_config = best_c;
  if (chosen) *chosen = best_c;
  return SPSS_OK;
}

int spss_set_all_pairs_config(spss_system *h, int config) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (config < 0 || config >= kAllPairsConfigCount) return fail(SPSS_ERR_INVALID, "all-pairs shape index out of range");
  h->ap_config = config;
  return SPSS_OK;
}

int spss_get_all_pairs_config(spss_system *h, int *block, int *jtile) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (block) *block = kAllPairsConfigs[h->ap_config][0];
  if (jtile) *jtile = kAllPairsConfigs[h->ap_config][1];
  return SPSS_OK;
}

int spss_get_cell_subdivision(spss_system *h, int *k) {
  if (!h || !k) return fail(SPSS_ERR_INVALID, "NULL argument");
  *k = h->p.cell_subdivision;
  return SPSS_OK;
}

int spss_set_cell_layout(spss_system *h, int k, int xsplit) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (h->p.precision != SPSS_PRECISION_FP32C || !h->fast)
    return fail(SPSS_ERR_UNSUPPORTED, "the cell layout belongs to the FP32C linked-list path");
  TRY(master_write(h, true));  // the state in original order; the next fused step re-enters
  FastPath *f = nullptr;
  const int rc = fast_create(&f, h->stream, h->P, h->p.rho, h->ncx, k, xsplit);
  if (rc != SPSS_OK) return fail(rc, fast_last_error());
  fast_destroy(h->fast);
  h->fast = f;
  if (h->timer) fast_set_timer(h->fast, h->timer);
  h->p.cell_subdivision = k;
  h->xsplit = xsplit;
  return SPSS_OK;
}

int spss_get_cell_layout(spss_system *h, int *k, int *xsplit) {
  if (!h || !k || !xsplit) return fail(SPSS_ERR_INVALID, "NULL argument");
  *k = h->p.cell_subdivision;
  *xsplit = h->xsplit;
  return SPSS_OK;
}

int spss_calculate_max_displacement(spss_system *h, const double *tot) {
  if (!h || !tot) return fail(SPSS_ERR_INVALID, "NULL argument");
  const size_t m = (size_t)h->s * h->n;
  if (!h->io[0]) TRY(dalloc_n(&h->io[0], m));
  HIPCHECK(hipMemcpyAsync(h->io[0], tot, m * sizeof(double), hipMemcpyHostToDevice, h->stream));
  launch_disp_max(h->stream, h->n, h->s, h->io[0], h->partial);
  const int nb = (h->n + kBlock - 1) / kBlock;
  launch_max_final(h->stream, nb, h->partial, h->scal + 8);
  double max2 = 0.0;
  HIPCHECK(hipMemcpyAsync(&max2, h->scal + 8, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  update_verlet_flag(h, max2);
  return SPSS_OK;
}

int spss_verlet_status(spss_system *h, int *should_update, double *acc, int *uses) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (h->vfast && h->resident_kind == 2) {  // FP32C Verlet lists: the flag lives on the device
    int rc = vf_status(h->vfast, h->stream, should_update, acc, nullptr);
    if (rc != SPSS_OK) return fail(rc, vf_last_error());
    if (uses) *uses = 1;
    return SPSS_OK;
  }
  if (should_update) *should_update = h->should_update ? 1 : 0;
  if (acc) *acc = h->acc_disp;
  if (uses) *uses = h->uses_verlet ? 1 : 0;
  return SPSS_OK;
}

int spss_get_verlet_lists(spss_system *h, int *counts, int *flat, long cap, long *total) {
  if (!h || !counts || !total) return fail(SPSS_ERR_INVALID, "NULL argument");
  if (h->vfast && h->resident_kind == 2) {  // FP32C: the lists of the last rebuild, by particle id
    std::vector<int> c, f;
    int rc = vf_lists(h->vfast, h->stream, &c, &f);
    if (rc != SPSS_OK) return fail(rc, vf_last_error());
    std::copy(c.begin(), c.end(), counts);
    *total = (long)f.size();
    if (flat && cap > 0) std::copy(f.begin(), f.begin() + std::min<long>(cap, (long)f.size()), flat);
    return SPSS_OK;
  }
  if (!h->vstart) {
    std::fill(counts, counts + h->n, 0);
    *total = 0;
    return SPSS_OK;
  }
  std::vector<int> start(h->n + 1);
  HIPCHECK(hipMemcpyAsync(start.data(), h->vstart, sizeof(int) * (h->n + 1), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  for (int i = 0; i < h->n; ++i) counts[i] = start[i + 1] - start[i];
  *total = start[h->n];
  if (flat && cap > 0) {
    const long cnt = std::min<long>(cap, start[h->n]);
    HIPCHECK(hipMemcpyAsync(flat, h->vlist, sizeof(int) * cnt, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
  }
  return SPSS_OK;
}

int spss_neighbour_pairs(spss_system *h, int method, const double *host_stage, int32_t *pairs, long cap,
                         long *count) {
  if (!h || !host_stage || !count) return fail(SPSS_ERR_INVALID, "NULL argument");
  const int m = resolve_method(h, method);
  if (h->s == 6) return fail(SPSS_ERR_UNSUPPORTED, "pair export is implemented for models 3 and 4");
  const size_t sn = (size_t)h->s * h->n;
  hipStream_t st = h->stream;
  HIPCHECK(hipMemcpyAsync(h->stage, host_stage, sn * sizeof(double), hipMemcpyHostToDevice, st));
  if (!h->pcount) {
    TRY(dalloc_n(&h->pcount, h->n + 1));
    TRY(dalloc_n(&h->pstart, h->n + 1));
  }
  const double rho2 = h->P.rho2;
  int total = 0;
  if (m == SPSS_METHOD_LINKED_LIST && h->p.precision == SPSS_PRECISION_FP32C) {
    TRY(fast_pairs_count(h->fast, st, h->P, h->stage, h->pcount));
  } else if (m == SPSS_METHOD_LINKED_LIST) {
    TRY(ensure_cells(h, h->cells, h->ncx, h->ncy));
    launch_cell_build(st, h->P, h->stage, h->ncx, h->ncy, h->cells);
    launch_verlet_count_cells(st, h->P, h->stage, h->ncx, h->ncy, h->cells, rho2, h->pcount);
  } else if (m == SPSS_METHOD_ALL_PAIRS) {
    launch_verlet_count_all(st, h->P, h->stage, rho2, h->pcount);
  } else {
    return fail(SPSS_ERR_UNSUPPORTED, "pair export is implemented for the all-pairs and linked-list methods");
  }
  TRY(ensure_scan(h));
  launch_scan(st, h->pcount, h->pstart, h->n, h->scan_blk);
  HIPCHECK(hipMemcpyAsync(&total, h->pstart + h->n, sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (!pairs) {  // count only
    *count = total;
    return SPSS_OK;
  }
  if ((long)total > h->pcap) {
    hipFree(h->plist);
    h->plist = nullptr;
    h->pcap = std::max<long>(total, 1024);
    TRY(dalloc_n(&h->plist, (size_t)h->pcap));
  }
  if (m == SPSS_METHOD_LINKED_LIST && h->p.precision == SPSS_PRECISION_FP32C)
    TRY(fast_pairs_fill(h->fast, st, h->P, h->stage, h->pstart, h->plist));
  else if (m == SPSS_METHOD_LINKED_LIST)
    launch_verlet_fill_cells(st, h->P, h->stage, h->ncx, h->ncy, h->cells, rho2, h->pstart, h->plist);
  else
    launch_verlet_fill_all(st, h->P, h->stage, rho2, h->pstart, h->plist);
  std::vector<int> start(h->n + 1), js(total > 0 ? total : 1);
  HIPCHECK(hipMemcpyAsync(start.data(), h->pstart, sizeof(int) * (h->n + 1), hipMemcpyDeviceToHost, st));
  if (total > 0) HIPCHECK(hipMemcpyAsync(js.data(), h->plist, sizeof(int) * total, hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  *count = total;
  if (pairs) {
    long w = 0;
    for (int i = 0; i < h->n && w < cap; ++i) {
      std::sort(js.begin() + start[i], js.begin() + start[i + 1]);
      for (int u = start[i]; u < start[i + 1] && w < cap; ++u, ++w) {
        pairs[2 * w] = i;
        pairs[2 * w + 1] = js[u];
      }
    }
  }
  return SPSS_OK;
}

int spss_order_parameter(spss_system *h, int slot, double *re, double *im) {
  if (!h || !re || !im) return fail(SPSS_ERR_INVALID, "NULL argument");
  if (slot < 0 || slot >= h->s) return fail(SPSS_ERR_INVALID, "slot out of range");
  TRY(sync_master(h));
  launch_order_parameter(h->stream, h->n, h->s, slot, h->state, h->partial, h->scal + 16);
  double v[2];
  HIPCHECK(hipMemcpyAsync(v, h->scal + 16, sizeof(v), hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  *re = v[0];
  *im = v[1];
  return SPSS_OK;
}

}  // extern "C"

namespace spss {
// the master state brought up to date (the resident copy of the fused steps is kept), for the
// other translation units of the library (the observers)
int system_view(spss_system *h, SystemView *v) {
  if (!h || !v) return fail(SPSS_ERR_INVALID, "NULL argument");
  TRY(sync_master(h));
  v->stream = h->stream;
  v->state = h->state;
  v->n = h->n;
  v->s = h->s;
  v->x_size = h->P.xs;
  v->y_size = h->P.ys;
  v->x_rsize = h->P.xrs;
  v->y_rsize = h->P.yrs;
  return SPSS_OK;
}
}  // namespace spss

extern "C" {

int spss_apply_periodic_boundary_conditions(spss_system *h) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  // a wrap can move a particle at x = -0 from the last cell to the first: the resident copy is
  // rebuilt from the wrapped state at the next fused step
  TRY(master_write(h, true));
  launch_apply_pbc(h->stream, h->n, h->s, h->state, h->P.xs, h->P.ys, h->P.xrs, h->P.yrs);
  HIPCHECK(hipGetLastError());
  return SPSS_OK;
}

int spss_device_copy_bandwidth(long bytes, int repeats, double *gbps) {
  if (bytes <= 0 || repeats <= 0 || !gbps) return fail(SPSS_ERR_INVALID, "bad argument");
  bytes = (bytes / 16) * 16;
  struct Probe {  // released on every return
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    ~Probe() {
      if (st) (void)hipStreamSynchronize(st);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (st) (void)hipStreamDestroy(st);
    }
  } P;
  HIPCHECK(hipStreamCreateWithFlags(&P.st, hipStreamNonBlocking));
  HIPCHECK(hipMalloc(&P.a, bytes));
  HIPCHECK(hipMalloc(&P.b, bytes));
  HIPCHECK(hipMemsetAsync(P.a, 1, bytes, P.st));
  EventPair ev;
  HIPCHECK(ev.create());
  launch_copy16(P.st, bytes / 16, P.a, P.b);  // warm-up
  HIPCHECK(hipEventRecord(ev.start, P.st));
  for (int r = 0; r < repeats; ++r) launch_copy16(P.st, bytes / 16, P.a, P.b);
  HIPCHECK(hipEventRecord(ev.stop, P.st));
  float ms = 0.f;
  HIPCHECK(ev.elapsed(&ms));
  *gbps = 2.0 * (double)bytes * repeats / (ms * 1e-3) / 1e9;  // read + write
  return SPSS_OK;
}

int spss_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return fail(SPSS_ERR_INVALID, "NULL argument");
  U4 r = philox4x32_10(U4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
  return SPSS_OK;
}

// the host twin of philox_normal2: identical counter, key and 53-bit uniforms
int spss_philox_normals(uint64_t seed, uint64_t step, uint32_t block, long first, long count, double *z0,
                        double *z1) {
  if (count < 0 || first < 0 || (count > 0 && (!z0 || !z1))) return fail(SPSS_ERR_INVALID, "bad argument");
  const double inv53 = 1.0 / 9007199254740992.0;
  for (long k = 0; k < count; ++k) {
    const uint32_t particle = (uint32_t)(first + k);
    U4 r = philox4x32_10(U4{particle, block, (uint32_t)step, (uint32_t)(step >> 32)}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)r.x << 21) ^ (uint64_t)(r.y >> 11);
    const uint64_t b = ((uint64_t)r.z << 21) ^ (uint64_t)(r.w >> 11);
    const double u1 = ((double)(a & ((1ull << 53) - 1)) + 1.0) * inv53;
    const double u2 = (double)(b & ((1ull << 53) - 1)) * inv53;
    const double rad = std::sqrt(-2.0 * std::log(u1));
    z0[k] = rad * std::cos(2.0 * M_PI * u2);
    z1[k] = rad * std::sin(2.0 * M_PI * u2);
  }
  return SPSS_OK;
}

int spss_kernel_timer_start(spss_system *h) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (!h->timer) h->timer = new KernelTimer();
  h->timer->on = true;
  h->timer->used = 0;
  if (h->fast) fast_set_timer(h->fast, h->timer);
  return SPSS_OK;
}

int spss_kernel_timer_stop(spss_system *h, double *total_ms, long *launches) {
  if (!h || !total_ms || !launches) return fail(SPSS_ERR_INVALID, "NULL argument");
  *total_ms = 0.0;
  *launches = 0;
  if (!h->timer) return SPSS_OK;
  HIPCHECK(hipStreamSynchronize(h->stream));
  *total_ms = h->timer->total_ms();
  *launches = h->timer->used;
  h->timer->on = false;
  h->timer->used = 0;
  if (h->fast) fast_set_timer(h->fast, nullptr);
  return SPSS_OK;
}

int spss_averaging_reset(spss_system *h) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  const size_t m = (size_t)h->s * h->n;
  if (!h->acc) TRY(dalloc_n(&h->acc, m));
  launch_fill_double(h->stream, (long)m, h->acc, 0.0);
  h->avg_count = 0;
  return SPSS_OK;
}

int spss_averaging_accumulate(spss_system *h) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (!h->acc) TRY(spss_averaging_reset(h));
  TRY(sync_master(h));
  launch_accumulate(h->stream, (long)h->s * h->n, h->state, h->acc);
  ++h->avg_count;
  return SPSS_OK;
}

int spss_averaging_get(spss_system *h, double *sum, long *count) {
  if (!h || !sum || !count) return fail(SPSS_ERR_INVALID, "NULL argument");
  if (!h->acc) TRY(spss_averaging_reset(h));
  HIPCHECK(hipMemcpyAsync(sum, h->acc, sizeof(double) * h->s * h->n, hipMemcpyDeviceToHost, h->stream));
  HIPCHECK(hipStreamSynchronize(h->stream));
  *count = h->avg_count;
  return SPSS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Dormand-Prince 5(4) with dense output: the reference's Boost.Odeint integration
// (SimulationEngineFirstOrder.cpp:158-165 integrate_const, :273-281 integrate_adaptive, both with
// make_dense_output(eps_abs, eps_rel, runge_kutta_dopri5<vector<Real>>())).  Boost is not in the
// reference tree; this restates the published algorithm of Boost.Odeint (1.6x): the dopri5 stepper
// with first-same-as-last derivative, controlled_runge_kutta with default_error_checker (a_x =
// a_dxdt = 1) and default_step_adjuster, dense_output_runge_kutta, and the dense-output versions of
// integrate_const and integrate_adaptive.  The system is the reference's functor
// (ParticleSystemFirstOrder.cpp:114-120): EvaluateRhs(state, zeros, derivative, 0, t).

namespace {

enum { DP_X0, DP_X1, DP_D0, DP_D1, DP_K2, DP_K3, DP_K4, DP_K5, DP_K6, DP_TMP, DP_COUNT };

// the coefficients as Boost forms them: quotients of integers in double
struct DpTableau {
  double a21, a31, a32, a41, a42, a43, a51, a52, a53, a54, a61, a62, a63, a64, a65;
  double b1, b3, b4, b5, b6;
  double d1, d3, d4, d5, d6, d7;  // fifth- minus fourth-order weights (the error estimate)
  DpTableau() {
    a21 = 1.0 / 5.0;
    a31 = 3.0 / 40.0, a32 = 9.0 / 40.0;
    a41 = 44.0 / 45.0, a42 = -56.0 / 15.0, a43 = 32.0 / 9.0;
    a51 = 19372.0 / 6561.0, a52 = -25360.0 / 2187.0, a53 = 64448.0 / 6561.0, a54 = -212.0 / 729.0;
    a61 = 9017.0 / 3168.0, a62 = -355.0 / 33.0, a63 = 46732.0 / 5247.0, a64 = 49.0 / 176.0;
    a65 = -5103.0 / 18656.0;
    b1 = 35.0 / 384.0, b3 = 500.0 / 1113.0, b4 = 125.0 / 192.0, b5 = -2187.0 / 6784.0, b6 = 11.0 / 84.0;
    d1 = b1 - 5179.0 / 57600.0;
    d3 = b3 - 7571.0 / 16695.0;
    d4 = b4 - 393.0 / 640.0;
    d5 = b5 - (-92097.0 / 339200.0);
    d6 = b6 - 187.0 / 2100.0;
    d7 = -1.0 / 40.0;
  }
};

struct DpRun {
  spss_system *h;
  int method;
  long m;
  double eps_abs, eps_rel;
  int cur = 0;  // which of the two state/derivative buffers holds the current point
  double t = 0.0, t_old = 0.0, dt = 0.0;
  bool deriv_init = false;
  long steps = 0, rejected = 0;
  double *x(int q) const { return h->dp[q ? DP_X1 : DP_X0]; }
  double *dx(int q) const { return h->dp[q ? DP_D1 : DP_D0]; }
};

DpTerms dp_terms(std::initializer_list<const double *> k, std::initializer_list<double> c) {
  DpTerms T{};
  int j = 0;
  for (const double *p : k) T.k[j++] = p;
  T.n = j;
  j = 0;
  for (double v : c) T.c[j++] = v;
  return T;
}

int dp_rhs(DpRun &R, const double *x, double *out) {
  // the functor's EvaluateRhs(x, zeros, dxdt, 0, t): k_prev all zeros (no buffer read)
  return rhs_device(R.h, R.method, x, nullptr, out, 0.0, 0.0);
}

// controlled_runge_kutta::try_step (FSAL): one trial step from the current point; on success the
// new point is in the other buffers and t has advanced; dt is the next (or the retried) step size
int dp_try_step(DpRun &R, bool *ok) {
  static const DpTableau A;
  spss_system *h = R.h;
  hipStream_t st = h->stream;
  const double *in = R.x(R.cur), *din = R.dx(R.cur);
  double *out = R.x(1 - R.cur), *dout = R.dx(1 - R.cur);
  double *const *k = h->dp;
  double *tmp = h->dp[DP_TMP];
  const double dt = R.dt;
  launch_dp_combine(st, R.m, in, dp_terms({din}, {dt * A.a21}), tmp);
  TRY(dp_rhs(R, tmp, k[DP_K2]));
  launch_dp_combine(st, R.m, in, dp_terms({din, k[DP_K2]}, {dt * A.a31, dt * A.a32}), tmp);
  TRY(dp_rhs(R, tmp, k[DP_K3]));
  launch_dp_combine(st, R.m, in, dp_terms({din, k[DP_K2], k[DP_K3]}, {dt * A.a41, dt * A.a42, dt * A.a43}), tmp);
  TRY(dp_rhs(R, tmp, k[DP_K4]));
  launch_dp_combine(st, R.m, in,
                    dp_terms({din, k[DP_K2], k[DP_K3], k[DP_K4]}, {dt * A.a51, dt * A.a52, dt * A.a53, dt * A.a54}),
                    tmp);
  TRY(dp_rhs(R, tmp, k[DP_K5]));
  launch_dp_combine(st, R.m, in,
                    dp_terms({din, k[DP_K2], k[DP_K3], k[DP_K4], k[DP_K5]},
                             {dt * A.a61, dt * A.a62, dt * A.a63, dt * A.a64, dt * A.a65}),
                    tmp);
  TRY(dp_rhs(R, tmp, k[DP_K6]));
  launch_dp_combine(st, R.m, in,
                    dp_terms({din, k[DP_K3], k[DP_K4], k[DP_K5], k[DP_K6]},
                             {dt * A.b1, dt * A.b3, dt * A.b4, dt * A.b5, dt * A.b6}),
                    out);
  TRY(dp_rhs(R, out, dout));
  launch_dp_error(st, h->n, h->s, in, din,
                  dp_terms({din, k[DP_K3], k[DP_K4], k[DP_K5], k[DP_K6], dout},
                           {dt * A.d1, dt * A.d3, dt * A.d4, dt * A.d5, dt * A.d6, dt * A.d7}),
                  dt, R.eps_abs, R.eps_rel, h->partial, h->scal + 24);
  HIPCHECK(hipGetLastError());
  double err = 0.0;
  HIPCHECK(hipMemcpyAsync(&err, h->scal + 24, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  if (err > 1.0) {  // default_step_adjuster::decrease_step, error order 4
    R.dt = dt * std::max(0.9 * std::pow(err, -1.0 / 3.0), 0.2);
    *ok = false;
    return SPSS_OK;
  }
  R.t += dt;
  if (err < 0.5) {  // default_step_adjuster::increase_step, stepper order 5
    err = std::max(std::pow(5.0, -5.0), err);
    R.dt = dt * (0.9 * std::pow(err, -1.0 / 5.0));
  }
  *ok = true;
  return SPSS_OK;
}

// dense_output_runge_kutta::initialize
int dp_initialize(DpRun &R, const double *x0, double t0, double dt0) {
  if (x0 != R.x(R.cur))
    HIPCHECK(hipMemcpyAsync(R.x(R.cur), x0, sizeof(double) * R.m, hipMemcpyDeviceToDevice, R.h->stream));
  R.t = t0;
  R.dt = dt0;
  R.deriv_init = false;
  return SPSS_OK;
}

// dense_output_runge_kutta::do_step (FSAL): trial steps until one is accepted (at most 500)
int dp_do_step(DpRun &R) {
  if (!R.deriv_init) {
    TRY(dp_rhs(R, R.x(R.cur), R.dx(R.cur)));
    R.deriv_init = true;
  }
  R.t_old = R.t;
  for (int fails = 0;; ++fails) {
    if (fails >= 500)
      return fail(SPSS_ERR_INVALID, "dopri5: 500 failed step size adjustments (Boost's failed_step_checker)");
    bool ok = false;
    TRY(dp_try_step(R, &ok));
    if (ok) break;
    ++R.rejected;
  }
  R.cur = 1 - R.cur;
  ++R.steps;
  return SPSS_OK;
}

// runge_kutta_dopri5::calc_state: the fourth-order continuous extension of the last step (the
// coefficient polynomials of Hairer's DOPRI5 dense output in Boost's form) at time t
void dp_calc_state(DpRun &R, double t, double *out) {
  static const DpTableau A;
  double *const *k = R.h->dp;
  const double dt = R.t - R.t_old;
  const double theta = (t - R.t_old) / dt;
  const double X1 = 5.0 * (2558722523.0 - 31403016.0 * theta) / 11282082432.0;
  const double X3 = 100.0 * (882725551.0 - 15701508.0 * theta) / 32700410799.0;
  const double X4 = 25.0 * (443332067.0 - 31403016.0 * theta) / 1880347072.0;
  const double X5 = 32805.0 * (23143187.0 - 3489224.0 * theta) / 199316789632.0;
  const double X6 = 55.0 * (29972135.0 - 7076736.0 * theta) / 822651844.0;
  const double X7 = 10.0 * (7414447.0 - 829305.0 * theta) / 29380423.0;
  const double theta_m_1 = theta - 1.0;
  const double theta_sq = theta * theta;
  const double Aq = theta_sq * (3.0 - 2.0 * theta);
  const double Bq = theta_sq * theta_m_1;
  const double Cq = theta_sq * theta_m_1 * theta_m_1;
  const double Dq = theta * theta_m_1 * theta_m_1;
  const double b1t = Aq * A.b1 - Cq * X1 + Dq;
  const double b3t = Aq * A.b3 + Cq * X3;
  const double b4t = Aq * A.b4 - Cq * X4;
  const double b5t = Aq * A.b5 + Cq * X5;
  const double b6t = Aq * A.b6 - Cq * X6;
  const double b7t = Bq + Cq * X7;
  launch_dp_combine(R.h->stream, R.m, R.x(1 - R.cur),
                    dp_terms({R.dx(1 - R.cur), k[DP_K3], k[DP_K4], k[DP_K5], k[DP_K6], R.dx(R.cur)},
                             {dt * b1t, dt * b3t, dt * b4t, dt * b5t, dt * b6t, dt * b7t}),
                    out);
}

// odeint's comparisons of times with the sign of the step and a tolerance of one machine epsilon
bool less_with_sign(double t1, double t2, double dt) {
  const double eps = std::numeric_limits<double>::epsilon();
  return dt > 0 ? t2 - t1 > eps : t1 - t2 > eps;
}
bool less_eq_with_sign(double t1, double t2, double dt) {
  const double eps = std::numeric_limits<double>::epsilon();
  return dt > 0 ? t1 - t2 <= eps : t2 - t1 <= eps;
}

struct DpObserver {
  spss_observer_fn fn;
  void *user;
  int on_host;
  int call(spss_system *h, const double *d_state, double t) {
    if (!fn) return SPSS_OK;
    const double *p = d_state;
    if (on_host) {
      HIPCHECK(hipMemcpyAsync(h->dp_host, d_state, sizeof(double) * h->s * h->n, hipMemcpyDeviceToHost, h->stream));
      HIPCHECK(hipStreamSynchronize(h->stream));
      p = h->dp_host;
    } else {
      HIPCHECK(hipStreamSynchronize(h->stream));
    }
    if (fn(user, t, p, h->n, h->s) != 0) return fail(SPSS_ERR_INVALID, "dopri5: the observer stopped the integration");
    return SPSS_OK;
  }
};

// integrate_const with a dense-output stepper: observations at t0 + j dt from the continuous
// extension; the state ends as the last observed state
int dp_integrate_const(DpRun &R, double t0, double t1, double dt, DpObserver &obs) {
  spss_system *h = R.h;
  if (!less_with_sign(t0, t1, dt)) return SPSS_OK;
  TRY(dp_initialize(R, h->state, t0, dt));
  TRY(obs.call(h, h->state, t0));
  long obs_step = 1;
  double time = t0 + (double)obs_step * dt;
  while (less_eq_with_sign(time + dt, t1, dt)) {
    while (less_eq_with_sign(time, R.t, dt)) {
      dp_calc_state(R, time, h->state);
      TRY(obs.call(h, h->state, time));
      ++obs_step;
      time = t0 + (double)obs_step * dt;
    }
    if (less_with_sign(R.t + R.dt, t1, R.dt)) {  // not at the end yet: real steps past the next observation
      while (less_eq_with_sign(R.t, time, dt)) TRY(dp_do_step(R));
    } else if (less_with_sign(R.t, t1, R.dt)) {  // the last step, ending on t1
      TRY(dp_initialize(R, R.x(R.cur), R.t, t1 - R.t));
      TRY(dp_do_step(R));
    }
  }
  if (less_eq_with_sign(time, t1, dt)) {  // the last observation
    dp_calc_state(R, time, h->state);
    TRY(obs.call(h, h->state, time));
  }
  return SPSS_OK;
}

// integrate_adaptive with a dense-output stepper: observed after every step; the state ends at t1
int dp_integrate_adaptive(DpRun &R, double t0, double t1, double dt, DpObserver &obs) {
  spss_system *h = R.h;
  TRY(dp_initialize(R, h->state, t0, dt));
  TRY(obs.call(h, h->state, t0));
  while (less_with_sign(R.t, t1, R.dt)) {
    while (less_eq_with_sign(R.t + R.dt, t1, R.dt)) {
      TRY(dp_do_step(R));
      if (obs.fn) {  // the system's state is the observed state during the call
        HIPCHECK(hipMemcpyAsync(h->state, R.x(R.cur), sizeof(double) * R.m, hipMemcpyDeviceToDevice, h->stream));
        TRY(obs.call(h, h->state, R.t));
      }
    }
    // the step size that ends exactly on t1
    TRY(dp_initialize(R, R.x(R.cur), R.t, t1 - R.t));
  }
  HIPCHECK(hipMemcpyAsync(h->state, R.x(R.cur), sizeof(double) * R.m, hipMemcpyDeviceToDevice, h->stream));
  return SPSS_OK;
}

}  // namespace

extern "C" {

int spss_dopri5_integrate(spss_system *h, int method, int mode, double t0, double t1, double dt, double eps_abs,
                          double eps_rel, spss_observer_fn observer, void *user, int observe_on_host, long *steps,
                          long *rejected) {
  if (!h) return fail(SPSS_ERR_INVALID, "NULL handle");
  if (mode != SPSS_DOPRI5_CONST && mode != SPSS_DOPRI5_ADAPTIVE) return fail(SPSS_ERR_INVALID, "unknown mode");
  if (!(dt > 0.0) || !(eps_abs >= 0.0) || !(eps_rel >= 0.0) || !(eps_abs + eps_rel > 0.0))
    return fail(SPSS_ERR_INVALID, "dopri5: dt must be > 0 and the tolerances >= 0, not both 0");
  TRY(check_method(h, resolve_method(h, method)));
  TRY(master_write(h, true));
  const long m = (long)h->s * h->n;
  for (int q = 0; q < DP_COUNT; ++q)
    if (!h->dp[q]) {
      // zeroed like odeint's value-initialised buffers: the RHS leaves k[3] of a particle without
      // neighbours untouched, so such entries keep what the buffer last held, as in the reference
      TRY(dalloc_n(&h->dp[q], (size_t)m));
      launch_fill_double(h->stream, m, h->dp[q], 0.0);
    }
  if (observer && observe_on_host && !h->dp_host)
    HIPCHECK(hipHostMalloc((void **)&h->dp_host, sizeof(double) * (m > 0 ? m : 1), hipHostMallocDefault));
  DpRun R;
  R.h = h;
  R.method = method;
  R.m = m;
  R.eps_abs = eps_abs;
  R.eps_rel = eps_rel;
  DpObserver obs{observer, user, observe_on_host};
  const int rc = mode == SPSS_DOPRI5_CONST ? dp_integrate_const(R, t0, t1, dt, obs)
                                           : dp_integrate_adaptive(R, t0, t1, dt, obs);
  if (steps) *steps = R.steps;
  if (rejected) *rejected = R.rejected;
  if (rc != SPSS_OK) return rc;
  HIPCHECK(hipStreamSynchronize(h->stream));
  return SPSS_OK;
}

}  // extern "C"

