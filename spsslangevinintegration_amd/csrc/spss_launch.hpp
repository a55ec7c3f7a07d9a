// spss_launch.hpp -- host-side launchers of the HIP kernels (one per kernel family).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "spss_device.hpp"

namespace spss {

// records the message spss_last_error() returns and passes the code through
int set_error(int code, const std::string &msg);

// what the observers need of a handle: its stream and its master state, up to date
struct SystemView {
  hipStream_t stream;
  const double *state;  // device, AoS FP64, original particle order
  int n, s;
  double x_size, y_size, x_rsize, y_rsize;
};
}  // namespace spss
struct spss_system;
namespace spss {
int system_view(spss_system *h, SystemView *v);

// two HIP events timing a stretch of one stream, destroyed with the object (also on the early
// returns of an error path)
struct EventPair {
  hipEvent_t start = nullptr, stop = nullptr;
  hipError_t create() {
    const hipError_t e = hipEventCreate(&start);
    return e == hipSuccess ? hipEventCreate(&stop) : e;
  }
  // milliseconds between the two records, after waiting for the second
  hipError_t elapsed(float *ms) {
    const hipError_t e = hipEventSynchronize(stop);
    return e == hipSuccess ? hipEventElapsedTime(ms, start, stop) : e;
  }
  EventPair() = default;
  EventPair(const EventPair &) = delete;
  EventPair &operator=(const EventPair &) = delete;
  ~EventPair() {
    if (start) (void)hipEventDestroy(start);
    if (stop) (void)hipEventDestroy(stop);
  }
};

// HIP-event timer around one kernel family, enabled on request (bench instrumentation):
// records an event pair per launch on the launching stream, sums the durations on read.
struct KernelTimer {
  bool on = false;
  int used = 0;
  hipEvent_t ev[2 * 4096];
  bool created = false;
  void begin(hipStream_t st) {
    if (!on || used >= 4096) return;
    if (!created) {
      for (auto &e : ev) hipEventCreate(&e);
      created = true;
    }
    hipEventRecord(ev[2 * used], st);
  }
  void end(hipStream_t st) {
    if (!on || used >= 4096) return;
    hipEventRecord(ev[2 * used + 1], st);
    ++used;
  }
  double total_ms() {
    double t = 0.0;
    for (int k = 0; k < used; ++k) {
      float ms = 0.f;
      hipEventSynchronize(ev[2 * k + 1]);
      hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]);
      t += ms;
    }
    return t;
  }
  ~KernelTimer() {
    if (created)
      for (auto &e : ev) hipEventDestroy(e);
  }
};

// device buffers of one cell list (stable counting sort over cell keys)
struct CellBuffers {
  int *key = nullptr;     // [n] cell key of each particle
  int *slot = nullptr;    // [n] arrival slot from the histogram atomics
  int *count = nullptr;   // [cells] histogram, left zeroed by the scan
  int *start = nullptr;   // [cells + 1] exclusive scan of count
  int *tmp = nullptr;     // [n] unstable bucket fill
  int *sorted = nullptr;  // [n] particle indices, cell-major, ascending index inside a cell
  int *blocksum = nullptr;  // [scan_blocks(cells) + 1] totals of the scan's workgroups
  int cells = 0;
};

// stage = state + kc_dt * kprev; kprev null: an all-zero k_prev
void launch_stage(hipStream_t st, long m, const double *state, const double *kprev, double kc_dt, double *stage);
// all-pairs RHS with workgroups of kAllPairsConfigs[config][0] particles and j-tiles of
// kAllPairsConfigs[config][1] (identical sums for every config; kAllPairsDefault is {256, 256})
constexpr int kAllPairsConfigCount = 6;
constexpr int kAllPairsDefault = 4;
extern const int kAllPairsConfigs[kAllPairsConfigCount][2];
void launch_all_pairs(hipStream_t st, const DevParams &P, const double *stage, const double *nat, double *k_next,
                      double *compl_out, int config = kAllPairsDefault);
void launch_cell_build(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy, CellBuffers &cb);
void launch_linked_list(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                        const CellBuffers &cb, const double *nat, double *k_next, double *compl_out);
// exclusive scan of count[0..m) into start[0..m]; count is left zeroed; blocksum: scratch of
// scan_blocks(m) + 1 ints (spss_scan.hpp)
void launch_scan(hipStream_t st, int *count, int *start, int m, int *blocksum);
void launch_verlet_count_all(hipStream_t st, const DevParams &P, const double *stage, double rmax2, int *vcount);
void launch_verlet_fill_all(hipStream_t st, const DevParams &P, const double *stage, double rmax2,
                            const int *vstart, int *vlist);
void launch_verlet_count_cells(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                               const CellBuffers &cb, double rmax2, int *vcount);
void launch_verlet_fill_cells(hipStream_t st, const DevParams &P, const double *stage, int ncx, int ncy,
                              const CellBuffers &cb, double rmax2, const int *vstart, int *vlist);
void launch_verlet_pairs(hipStream_t st, const DevParams &P, const double *stage, const int *vstart,
                         const int *vlist, const double *nat, double *k_next, double *compl_out);
void launch_global(hipStream_t st, const DevParams &P, const double *stage, double *partial, double *fields,
                   const double *nat, double *k_next, double *compl_out);
void launch_rk4_combine(hipStream_t st, int s, int n, double *state, const double *k1, const double *k2,
                        const double *k3, const double *k4, double dt, double *max_partial);
void launch_euler(hipStream_t st, int s, int n, double *state, const double *f, double dt, double sqrt_dt,
                  double int_phi, double int_r, int noise_mode, const double *noise, uint64_t seed, uint64_t step,
                  double *max_partial);
void launch_srk(hipStream_t st, int n, double *state, const double *f, const double *C, double dt, double sqrt_dt,
                double c15, double is3, double sq2d, double friction, int noise_mode, const double *noise,
                uint64_t seed, uint64_t step, double *max_partial);
void launch_max_final(hipStream_t st, int nb, const double *partial, double *out);

// Dormand-Prince 5(4): out = 1.0 * x + c[0] * k[0] + ... + c[n-1] * k[n-1] over m values (a stage
// state, the step's output or a dense-output state), and the step's error norm
constexpr int kDpMaxTerms = 6;
struct DpTerms {
  const double *k[kDpMaxTerms];
  double c[kDpMaxTerms];
  int n;
};
void launch_dp_combine(hipStream_t st, long m, const double *x, const DpTerms &T, double *out);
// max over the m = s * n values of |sum_j c_j k_j| / (eps_abs + eps_rel * (|x_old| + dt * |dxdt_old|))
// into out[0] (partial: one double per 256 particles); returns the number of partials
int launch_dp_error(hipStream_t st, int n, int s, const double *x_old, const double *dxdt_old, const DpTerms &T,
                    double dt, double eps_abs, double eps_rel, double *partial, double *out);
void launch_disp_max(hipStream_t st, int n, int s, const double *tot, double *partial);
void launch_order_parameter(hipStream_t st, int n, int s, int slot, const double *state, double *partial,
                            double *out);
void launch_apply_pbc(hipStream_t st, int n, int s, double *state, double xs, double ys, double xrs, double yrs);
void launch_accumulate(hipStream_t st, long m, const double *state, double *acc);
void launch_fill_double(hipStream_t st, long m, double *p, double v);
void launch_fill_int(hipStream_t st, long m, int *p, int v);
void launch_copy16(hipStream_t st, long n16, const void *src, void *dst);

}  // namespace spss
