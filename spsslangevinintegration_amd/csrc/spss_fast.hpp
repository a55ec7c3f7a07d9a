// spss_fast.hpp -- the FP32 cell-list path (SPSS_PRECISION_FP32C): FP32 pair arithmetic, FP64 sums and state.
#pragma once

#include <hip/hip_runtime.h>

#include "spss_device.hpp"
#include "spss_launch.hpp"

namespace spss {

struct FastPath;

int fast_create(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv);
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
// writes the resident state back into `state` in original order; the copy stays resident
int fast_resident_sync(FastPath *f, hipStream_t st, double *state);
// forgets the resident copy (the next fused step re-enters from the caller's state)
void fast_resident_drop(FastPath *f);
// neighbour set the fast path uses for a stage state: per-particle counts, then lists
int fast_pairs_count(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, int *count);
int fast_pairs_fill(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const int *start,
                    int *list);

}  // namespace spss
