// spss_fast.hpp -- the FP32 cell-list path (SPSS_PRECISION_FP32C): FP32 pair arithmetic, FP64 sums and state.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/spss_hip.h"
#include "spss_device.hpp"
#include "spss_launch.hpp"

namespace spss {

struct FastPath;

// cells 1/(subdiv*xsplit*nc) wide and 1/(subdiv*nc) high (xsplit 1, 2 or 4: narrower stencil columns)
int fast_create(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv, int xsplit);
int fast_cell_xsplit(const FastPath *f);
// the cell layouts the autotuners time: {subdivision, x split}
constexpr int kCellLayouts = SPSS_CELL_LAYOUT_COUNT;
constexpr int kCellLayout[kCellLayouts][2] = SPSS_CELL_LAYOUTS;

// per-replica coefficients of the stochastic step of a batch (a single system takes them from
// its params): phase friction xi, coupling sigma, sqrt(2 D) for Euler and SRK, Philox key
struct Coef {
  double xi, sigma, int_phi, sq2d;
  uint64_t seed;
};
// a batch of R independent replicas on one cell grid, particles stored replica after replica
// (rep_n[r] each, summing to P.n); coef (R entries) makes it a batch stepped by
// fast_batch_step with per-replica xi, sigma, noise and seed
int fast_create_replicas(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv,
                         int xsplit, int R, const int *rep_n, const Coef *coef);
// tile width (cells per pair workgroup) of every replica: read and set (the autotuner's knob)
int fast_tile_widths(FastPath *f, std::vector<int> *W);
int fast_set_tile_widths(FastPath *f, const std::vector<int> &W);
int fast_replica_count(const FastPath *f);
// one stochastic Euler step of every replica with Philox noise (per-replica seed, particle index
// inside the replica, step from the device counter, which the call advances)
int fast_batch_step(FastPath *f, hipStream_t st, const DevParams &P, const double *state, const double *nat,
                    double dt);
// before capturing or replaying Euler steps of size dt: the resident order's current moves in
// place for dt (each step computes the next step's moves, so a captured step never needs fr_move)
int fast_prepare_moves(FastPath *f, hipStream_t st, const DevParams &P, double dt);
// per-replica order parameter of the current state; |Z|, |Z|^2 and a count are added to the
// running sums once the step counter has reached `from`
int fast_batch_observe(FastPath *f, hipStream_t st, uint64_t from);
// clear the running sums and set the step counter
int fast_batch_reset(FastPath *f, hipStream_t st, uint64_t step);
// running sums (3 per replica: sum |Z|, sum |Z|^2, count), last Z (2 per replica), step counter
int fast_batch_stats(FastPath *f, hipStream_t st, double *racc, double *zlast, uint64_t *step);
void fast_destroy(FastPath *f);
void fast_set_stream(FastPath *f, hipStream_t st);
void fast_set_timer(FastPath *f, KernelTimer *t);
const char *fast_last_error();
// RHS of one stage state (FP64 AoS) with FP32 pair math; writes k_next (and the complementary force)
int fast_rhs(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const double *nat,
             double *k_next, double *compl_out);
// one fused stochastic step (srk = 0: Euler, 1: SRK) on the resident copy of the state: the
// state, the particle ids and the natural frequencies are kept in cell order between steps.
// The first step after fast_resident_drop enters from `state` (original order, read only).
// noise_mode 0 none, 1 device array indexed by particle, 3 Philox
int fast_step(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
              double *compl_out, double dt, double D_phi, double friction, int noise_mode, const double *noise,
              uint64_t seed, uint64_t step);
bool fast_is_resident(const FastPath *f);
// the arrivals scan's failure flag (a look-back wait that ran out): SPSS_OK or an error (waits)
int fast_check(FastPath *f, hipStream_t st);
// which of its double buffers the resident order is in (a captured step graph replays only there)
int fast_buffer_parity(const FastPath *f);
// Fused steps replayed from HIP graphs (Philox noise from the device step counter): runs all
// nsteps from graphs of kGraphSteps steps and of one step, starting at Philox step `step`, and
// returns nsteps (0: not applicable now, see fast_graph_ready; -1: a HIP error, see
// fast_last_error).
long fast_run_graph(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
                    double *compl_out, double dt, double D_phi, double friction, uint64_t seed, uint64_t step,
                    long nsteps);
// whether fast_run_graph can run now: the resident order exists in the form the stepper needs
// (natural frequencies, FP32 omega for SRK) and the kernel timer is off
bool fast_graph_ready(const FastPath *f, const double *nat, int srk);
// enters the resident order from `state` without stepping (no-op when already there)
int fast_enter(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat);
// captures and instantiates the step graphs a run of any length needs, from both buffer parities,
// without running a step (the resident order must exist: fast_enter)
int fast_prepare_graphs(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state,
                        const double *nat, double *compl_out, double dt, double D_phi, double friction,
                        uint64_t seed);
// number of step graphs captured since creation
long fast_graph_captures(const FastPath *f);
// the host bookkeeping a step changes (buffer parity, move and resident flags): a stream capture
// records steps without running them, so a capturer saves it before and puts it back after
struct FastHostMark {
  bool resident, resident_nat, romega_valid, mv_ready, mv_srk;
  int cur, mvc;
  double mv_dt, mv_v0;
};
FastHostMark fast_host_mark(const FastPath *f);
void fast_restore_host_mark(FastPath *f, const FastHostMark &m);
// writes the resident state back into `state` in original order; the copy stays resident
int fast_resident_sync(FastPath *f, hipStream_t st, double *state);
// forgets the resident copy (the next fused step re-enters from the caller's state)
void fast_resident_drop(FastPath *f);
// neighbour set the fast path uses for a stage state: per-particle counts, then lists
int fast_pairs_count(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, int *count);
int fast_pairs_fill(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const int *start,
                    int *list);

}  // namespace spss
