// spss_batch.hip -- batches of independent replicas (the spss_batch_* part of the C ABI).
//
// The replicas of one batch share the model, v0, rho and the cell grid; each has its own
// particle count, phase friction, coupling, noise intensity and Philox seed.  Their particles
// lie one replica after another in one resident cell-ordered array, so every kernel of the
// FP32 cell-list step covers all replicas in one launch.  A run captures a fixed number of
// steps (and an order-parameter reading) in a HIP graph once and replays it; the Philox step
// counter lives on the device so that every replay is the same graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/spss_hip.h"
#include "spss_fast.hpp"
#include "spss_launch.hpp"

using namespace spss;

struct spss_batch {
  spss_batch_params bp{};
  int R = 0, n = 0, s = 4;
  hipStream_t st = nullptr;
  DevParams P{};
  FastPath *fast = nullptr;
  double *state = nullptr, *nat = nullptr;  // master state and natural frequencies, replica order
  bool have_nat = false;
  bool master_stale = false;  // the resident copy is ahead of `state`
  std::vector<int> rep_n;
  std::vector<Coef> coef;  // per-replica coefficients (to rebuild the cell path at another cell size)
  int nc = 1, k = 1;       // int(1/rho) and the cell subdivision of the current cell path
  int xsplit = 1;          // its columns per square cell
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_steps = 0;
  long graph_every = -1;
  long graph_from = -1;
  double graph_dt = 0.0;
  int graph_parity = -1;  // double-buffer parity of the resident order the graph was captured at
};

namespace {

#define BCHK(call)                                                                                \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define BTRY(call)                \
  do {                            \
    int r_ = (call);              \
    if (r_ != SPSS_OK) return r_; \
  } while (0)

int fast_fail(int rc) { return set_error(rc, fast_last_error()); }

// the stream is drained first: a graph exec is never destroyed while a replay of it may run
void drop_graph(spss_batch *b) {
  if (!b->exec && !b->graph) return;
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_steps = 0;
}

int sync_master(spss_batch *b) {
  if (!b->master_stale) return SPSS_OK;
  int rc = fast_resident_sync(b->fast, b->st, b->state);
  if (rc != SPSS_OK) return fast_fail(rc);
  b->master_stale = false;
  return SPSS_OK;
}

int one_step(spss_batch *b, double dt) {
  int rc = fast_batch_step(b->fast, b->st, b->P, b->state, b->have_nat ? b->nat : nullptr, dt);
  if (rc != SPSS_OK) return fast_fail(rc);
  b->master_stale = true;
  return SPSS_OK;
}

// `count` steps, then an order-parameter reading
int steps_and_observe(spss_batch *b, int count, double dt, bool observe, long from) {
  for (int k = 0; k < count; ++k) BTRY(one_step(b, dt));
  if (observe) {
    int rc = fast_batch_observe(b->fast, b->st, (uint64_t)from);
    if (rc != SPSS_OK) return fast_fail(rc);
  }
  return SPSS_OK;
}

// a graph of `steps` steps (even, so that the double buffers end where they started) with an
// order-parameter reading after every `every` steps when every > 0
int capture(spss_batch *b, int steps, long every, long from, double dt) {
  drop_graph(b);
  int rp = fast_prepare_moves(b->fast, b->st, b->P, dt);
  if (rp != SPSS_OK) return fast_fail(rp);
  const int parity0 = fast_buffer_parity(b->fast);
  // nothing runs during the capture: the host bookkeeping the recorded steps change is put back
  const FastHostMark mark = fast_host_mark(b->fast);
  const bool stale0 = b->master_stale;
  BCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed));
  int rc = SPSS_OK;
  for (int k = 0; k < steps && rc == SPSS_OK; ++k) {
    rc = one_step(b, dt);
    if (rc == SPSS_OK && every > 0 && (k + 1) % every == 0) {
      int r2 = fast_batch_observe(b->fast, b->st, (uint64_t)from);
      if (r2 != SPSS_OK) rc = fast_fail(r2);
    }
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->st, &g);
  fast_restore_host_mark(b->fast, mark);
  b->master_stale = stale0;
  if (rc != SPSS_OK) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  b->graph = g;
  e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return set_error(SPSS_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  b->graph_steps = steps;
  b->graph_every = every;
  b->graph_from = from;
  b->graph_dt = dt;
  b->graph_parity = parity0;
  return SPSS_OK;
}

static int batch_run(spss_batch *b, long nsteps, double dt, long observe_every, long observe_from, int use_graph) {
  long done = 0;
  // the first step after new inputs enters the resident cell order; it runs eagerly
  if (!fast_is_resident(b->fast)) {
    BTRY(steps_and_observe(b, 1, dt, observe_every == 1, observe_from));
    ++done;
  }
  if (use_graph) {
    int gs = observe_every > 0 ? (int)observe_every : 10;
    if (gs % 2) gs *= 2;
    if (b->graph_steps != gs || b->graph_every != observe_every || b->graph_from != observe_from ||
        b->graph_dt != dt || b->graph_parity != fast_buffer_parity(b->fast) || !b->exec) {
      // keep the observation phase: a graph starts right after a reading
      while (observe_every > 0 && done < nsteps && done % observe_every != 0) {
        BTRY(steps_and_observe(b, 1, dt, (done + 1) % observe_every == 0, observe_from));
        ++done;
      }
      if (nsteps - done >= gs) BTRY(capture(b, gs, observe_every, observe_from, dt));
    }
    while (b->exec && nsteps - done >= b->graph_steps && (observe_every == 0 || done % observe_every == 0)) {
      const int rp = fast_prepare_moves(b->fast, b->st, b->P, dt);
      if (rp != SPSS_OK) return fast_fail(rp);
      BCHK(hipGraphLaunch(b->exec, b->st));
      done += b->graph_steps;
    }
  }
  while (done < nsteps) {
    ++done;
    BTRY(steps_and_observe(b, 1, dt, observe_every > 0 && done % observe_every == 0, observe_from));
  }
  return SPSS_OK;
}

}  // namespace

extern "C" {

int spss_batch_create(const spss_batch_params *bp, int nrep, const spss_replica *reps, spss_batch **out) {
  if (!bp || !reps || !out || nrep <= 0) return set_error(SPSS_ERR_INVALID, "bad argument");
  *out = nullptr;
  if (bp->model != 3 && bp->model != 4) return set_error(SPSS_ERR_INVALID, "batch model must be 3 or 4");
  if (!(bp->rho > 0.0)) return set_error(SPSS_ERR_INVALID, "rho must be positive");
  long total = 0;
  for (int r = 0; r < nrep; ++r) {
    if (reps[r].n <= 0) return set_error(SPSS_ERR_INVALID, "every replica needs particles");
    total += reps[r].n;
  }
  if (total * bp->model > (1L << 31) - 1) return set_error(SPSS_ERR_CAPACITY, "s*n exceeds int32 indexing");
  spss_batch *b = new spss_batch();
  b->bp = *bp;
  b->R = nrep;
  b->n = (int)total;
  b->s = bp->model;
  b->rep_n.resize(nrep);
  std::vector<Coef> coef(nrep);
  for (int r = 0; r < nrep; ++r) {
    b->rep_n[r] = reps[r].n;
    coef[r].xi = reps[r].xi;
    coef[r].sigma = reps[r].sigma;
    coef[r].int_phi = std::sqrt(2.0 * reps[r].D_phi);
    coef[r].sq2d = std::sqrt(2.0 * reps[r].D_phi);
    coef[r].seed = reps[r].seed;
  }
  DevParams &P = b->P;
  P.n = b->n;
  P.s = b->s;
  P.v0 = bp->v0;
  P.xi = 0.0;
  P.sigma = 0.0;
  P.rho2 = bp->rho * bp->rho;
  P.alpha = bp->alpha;
  P.cos_a = std::cos(bp->alpha);
  P.sin_a = std::sin(bp->alpha);
  P.xs = P.ys = P.xrs = P.yrs = 1.0;
  P.natfreq_local = bp->natfreq_local;
  hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&b->state, sizeof(double) * (size_t)b->s * b->n);
  if (e == hipSuccess) e = hipMalloc(&b->nat, sizeof(double) * (size_t)b->n);
  if (e != hipSuccess) {
    spss_batch_destroy(b);
    return set_error(SPSS_ERR_DEVICE, std::string("batch allocation: ") + hipGetErrorString(e));
  }
  const int nc = std::max(1, (int)(1.0 / bp->rho));
  const int k = bp->cell_subdivision > 0 ? bp->cell_subdivision : 1;
  b->coef = coef;
  b->nc = nc;
  b->k = k;
  int rc = fast_create_replicas(&b->fast, b->st, P, bp->rho, nc, k, 1, nrep, b->rep_n.data(), coef.data());
  if (rc != SPSS_OK) {
    spss_batch_destroy(b);
    return fast_fail(rc);
  }
  launch_fill_double(b->st, (long)b->s * b->n, b->state, 0.0);
  launch_fill_double(b->st, b->n, b->nat, 0.0);
  rc = fast_batch_reset(b->fast, b->st, 0);
  if (rc != SPSS_OK) {
    spss_batch_destroy(b);
    return fast_fail(rc);
  }
  *out = b;
  return SPSS_OK;
}

int spss_batch_destroy(spss_batch *b) {
  if (!b) return SPSS_OK;
  if (b->st) hipStreamSynchronize(b->st);
  drop_graph(b);
  fast_destroy(b->fast);
  hipFree(b->state);
  hipFree(b->nat);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
  return SPSS_OK;
}

int spss_batch_set_state(spss_batch *b, const double *host_state) {
  if (!b || !host_state) return set_error(SPSS_ERR_INVALID, "NULL argument");
  drop_graph(b);
  b->master_stale = false;
  fast_resident_drop(b->fast);
  BCHK(hipMemcpyAsync(b->state, host_state, sizeof(double) * (size_t)b->s * b->n, hipMemcpyHostToDevice, b->st));
  BCHK(hipStreamSynchronize(b->st));
  return SPSS_OK;
}

int spss_batch_get_state(spss_batch *b, double *host_state) {
  if (!b || !host_state) return set_error(SPSS_ERR_INVALID, "NULL argument");
  BTRY(sync_master(b));
  BCHK(hipMemcpyAsync(host_state, b->state, sizeof(double) * (size_t)b->s * b->n, hipMemcpyDeviceToHost, b->st));
  BCHK(hipStreamSynchronize(b->st));
  return SPSS_OK;
}

int spss_batch_set_natural_frequencies(spss_batch *b, const double *host_w) {
  if (!b || !host_w) return set_error(SPSS_ERR_INVALID, "NULL argument");
  BTRY(sync_master(b));
  drop_graph(b);
  fast_resident_drop(b->fast);
  BCHK(hipMemcpyAsync(b->nat, host_w, sizeof(double) * (size_t)b->n, hipMemcpyHostToDevice, b->st));
  BCHK(hipStreamSynchronize(b->st));
  b->have_nat = true;
  return SPSS_OK;
}

int spss_batch_reset(spss_batch *b, long step) {
  if (!b || step < 0) return set_error(SPSS_ERR_INVALID, "bad argument");
  int rc = fast_batch_reset(b->fast, b->st, (uint64_t)step);
  return rc == SPSS_OK ? SPSS_OK : fast_fail(rc);
}

// the run, then the stream drained on every return (error returns included)
int spss_batch_run(spss_batch *b, long nsteps, double dt, long observe_every, long observe_from, int use_graph) {
  if (!b || nsteps < 0 || observe_every < 0 || !(dt > 0.0)) return set_error(SPSS_ERR_INVALID, "bad argument");
  if (nsteps == 0) return SPSS_OK;
  const int rc = batch_run(b, nsteps, dt, observe_every, observe_from, use_graph);
  const hipError_t e = hipStreamSynchronize(b->st);
  if (rc != SPSS_OK) return rc;
  if (e != hipSuccess) return set_error(SPSS_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return SPSS_OK;
}

int spss_batch_order_stats(spss_batch *b, double *mean_R, double *var_R, long *count, double *last_re,
                           double *last_im, long *step) {
  if (!b) return set_error(SPSS_ERR_INVALID, "NULL handle");
  std::vector<double> acc(3 * b->R), z(2 * b->R);
  uint64_t st = 0;
  int rc = fast_batch_stats(b->fast, b->st, acc.data(), z.data(), &st);
  if (rc != SPSS_OK) return fast_fail(rc);
  for (int r = 0; r < b->R; ++r) {
    const double c = acc[3 * r + 2];
    const double m = c > 0 ? acc[3 * r] / c : 0.0;
    if (mean_R) mean_R[r] = m;
    if (var_R) var_R[r] = c > 0 ? std::max(0.0, acc[3 * r + 1] / c - m * m) : 0.0;
    if (count) count[r] = (long)c;
    if (last_re) last_re[r] = z[2 * r];
    if (last_im) last_im[r] = z[2 * r + 1];
  }
  if (step) *step = (long)st;
  return SPSS_OK;
}

int spss_batch_get_tile_widths(spss_batch *b, int *W) {
  if (!b || !W) return set_error(SPSS_ERR_INVALID, "NULL argument");
  std::vector<int> w;
  fast_tile_widths(b->fast, &w);
  std::copy(w.begin(), w.end(), W);
  return SPSS_OK;
}

int spss_batch_set_tile_widths(spss_batch *b, const int *W) {
  if (!b || !W) return set_error(SPSS_ERR_INVALID, "NULL argument");
  drop_graph(b);
  BCHK(hipStreamSynchronize(b->st));
  int rc = fast_set_tile_widths(b->fast, std::vector<int>(W, W + b->R));
  return rc == SPSS_OK ? SPSS_OK : fast_fail(rc);
}

// the cell path rebuilt at cell subdivision k and x split xs (state, natural frequencies and step
// counter are the caller's to restore)
static int rebuild_cells(spss_batch *b, int k, int xs) {
  drop_graph(b);
  FastPath *f = nullptr;
  int rc = fast_create_replicas(&f, b->st, b->P, b->bp.rho, b->nc, k, xs, b->R, b->rep_n.data(), b->coef.data());
  if (rc != SPSS_OK) return rc;
  fast_destroy(b->fast);
  b->fast = f;
  b->k = k;
  b->xsplit = xs;
  return SPSS_OK;
}

// Cell size of the batch: times probe_steps graph-replayed steps for each cell layout of
// kCellLayout (cells 1/(k xs int(1/rho)) wide and 1/(k int(1/rho)) high, each at its default
// tile widths), keeps the fastest and restores the state and the step counter; the time averages
// start again.  The cell layout changes the FP32 candidate order, so the results of later steps
// depend on the choice (to the FP32C accuracy); a replica stays bit-identical to a single FP32C
// system of the same cell layout.
int spss_batch_autotune_cells(spss_batch *b, double dt, int probe_steps, double *table_ms, int table_cap,
                              int *chosen) {
  if (!b || probe_steps < 2 || !(dt > 0.0)) return set_error(SPSS_ERR_INVALID, "bad argument");
  BTRY(sync_master(b));
  std::vector<double> saved((size_t)b->s * b->n);
  BCHK(hipMemcpyAsync(saved.data(), b->state, sizeof(double) * saved.size(), hipMemcpyDeviceToHost, b->st));
  uint64_t step0 = 0;
  int rc = fast_batch_stats(b->fast, b->st, nullptr, nullptr, &step0);
  if (rc != SPSS_OK) return fast_fail(rc);
  double best = 1e300;
  int best_k = b->k, best_x = b->xsplit;
  for (int c = 0; c < kCellLayouts; ++c) {
    const int k = kCellLayout[c][0], xs = kCellLayout[c][1];
    double ms_step = -1.0;
    if (rebuild_cells(b, k, xs) == SPSS_OK) {
      BTRY(spss_batch_set_state(b, saved.data()));
      BTRY(spss_batch_run(b, 2 + probe_steps % 2, dt, 0, 0, 1));  // enter, warm up
      EventPair ev;
      BCHK(ev.create());
      BCHK(hipEventRecord(ev.start, b->st));
      BTRY(spss_batch_run(b, probe_steps, dt, 0, 0, 1));
      BCHK(hipEventRecord(ev.stop, b->st));
      float ms = 0.f;
      BCHK(ev.elapsed(&ms));
      ms_step = ms / probe_steps;
      if (ms_step < best) {
        best = ms_step;
        best_k = k;
        best_x = xs;
      }
    }
    if (table_ms && c < table_cap) table_ms[c] = ms_step;  // -1: this layout does not build
  }
  rc = rebuild_cells(b, best_k, best_x);
  if (rc != SPSS_OK) return fast_fail(rc);
  BTRY(spss_batch_set_state(b, saved.data()));
  rc = fast_batch_reset(b->fast, b->st, step0);
  if (rc != SPSS_OK) return fast_fail(rc);
  if (chosen) *chosen = best_k;
  return SPSS_OK;
}

int spss_batch_get_cell_subdivision(spss_batch *b, int *k) {
  if (!b || !k) return set_error(SPSS_ERR_INVALID, "NULL argument");
  *k = b->k;
  return SPSS_OK;
}

int spss_batch_set_cell_layout(spss_batch *b, int k, int xsplit) {
  if (!b) return set_error(SPSS_ERR_INVALID, "NULL handle");
  BTRY(sync_master(b));
  std::vector<double> saved((size_t)b->s * b->n);
  BCHK(hipMemcpyAsync(saved.data(), b->state, sizeof(double) * saved.size(), hipMemcpyDeviceToHost, b->st));
  BCHK(hipStreamSynchronize(b->st));
  uint64_t step0 = 0;
  int rc = fast_batch_stats(b->fast, b->st, nullptr, nullptr, &step0);
  if (rc != SPSS_OK) return fast_fail(rc);
  rc = rebuild_cells(b, k, xsplit);
  if (rc != SPSS_OK) return fast_fail(rc);
  BTRY(spss_batch_set_state(b, saved.data()));
  rc = fast_batch_reset(b->fast, b->st, step0);
  return rc == SPSS_OK ? SPSS_OK : fast_fail(rc);
}

int spss_batch_get_cell_layout(spss_batch *b, int *k, int *xsplit) {
  if (!b || !k || !xsplit) return set_error(SPSS_ERR_INVALID, "NULL argument");
  *k = b->k;
  *xsplit = b->xsplit;
  return SPSS_OK;
}

// Tile-width autotuner: times `probe_steps` steps (graph replays, no readings) for tile widths of
// 0.5x ... 2x the default particles-per-workgroup of every replica and keeps the fastest.  The
// state and step counter are restored afterwards and the running averages cleared; the tile
// width does not change the results (a lane's candidates and their order depend on the cells,
// not on the tile).
int spss_batch_autotune(spss_batch *b, double dt, int probe_steps, double *table_ms, int table_cap, int *chosen) {
  if (!b || probe_steps < 2 || !(dt > 0.0)) return set_error(SPSS_ERR_INVALID, "bad argument");
  BTRY(sync_master(b));
  std::vector<double> saved((size_t)b->s * b->n);
  BCHK(hipMemcpyAsync(saved.data(), b->state, sizeof(double) * saved.size(), hipMemcpyDeviceToHost, b->st));
  uint64_t step0 = 0;
  int rc = fast_batch_stats(b->fast, b->st, nullptr, nullptr, &step0);
  if (rc != SPSS_OK) return fast_fail(rc);
  std::vector<int> W0;
  fast_tile_widths(b->fast, &W0);
  const double scales[] = {0.5, 0.75, 1.0, 1.25, 1.5, 2.0};
  const int ncand = 6;
  double best = 1e300;
  int best_k = 2;
  for (int kc = 0; kc < ncand; ++kc) {
    std::vector<int> W(b->R);
    for (int r = 0; r < b->R; ++r) W[r] = std::max(1, (int)std::lround(W0[r] * scales[kc]));
    BTRY(spss_batch_set_tile_widths(b, W.data()));
    BTRY(spss_batch_set_state(b, saved.data()));
    BTRY(spss_batch_run(b, 2 + probe_steps % 2, dt, 0, 0, 1));  // enter, warm up
    EventPair ev;
    BCHK(ev.create());
    BCHK(hipEventRecord(ev.start, b->st));
    BTRY(spss_batch_run(b, probe_steps, dt, 0, 0, 1));
    BCHK(hipEventRecord(ev.stop, b->st));
    float ms = 0.f;
    BCHK(ev.elapsed(&ms));
    if (table_ms && kc < table_cap) table_ms[kc] = ms / probe_steps;
    if (ms < best) {
      best = ms;
      best_k = kc;
    }
  }
  std::vector<int> W(b->R);
  for (int r = 0; r < b->R; ++r) W[r] = std::max(1, (int)std::lround(W0[r] * scales[best_k]));
  BTRY(spss_batch_set_tile_widths(b, W.data()));
  BTRY(spss_batch_set_state(b, saved.data()));
  rc = fast_batch_reset(b->fast, b->st, step0);
  if (rc != SPSS_OK) return fast_fail(rc);
  if (chosen) *chosen = best_k;
  return SPSS_OK;
}

}  // extern "C"
