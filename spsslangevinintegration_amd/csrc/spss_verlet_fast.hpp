// spss_verlet_fast.hpp -- the Verlet-list path in FP32C precision (exact mode), host interface.
//
// Neighbour lists of radius r_max = rho + skin built from a cell list of side >= r_max, rebuilt
// when the accumulated largest displacement says so (CalculateMaxDisplacement's bookkeeping,
// ParticleSystemFirstOrderPtr.cpp:836-861, with the exact mode's one-step margin) -- decided and
// acted on by the device, so steps need no host synchronisation and replay from HIP graphs.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "spss_device.hpp"

namespace spss {

struct VerletFast;

// models 3 and 4, unit box, int(1/r_max) >= 3 (SPSS_ERR_UNSUPPORTED otherwise)
int vf_create(VerletFast **out, hipStream_t st, const DevParams &P, double rho, double r_max);
void vf_destroy(VerletFast *v);
const char *vf_last_error();
// one stochastic step (srk 0: Euler, 1: SRK) of the resident copy, entered from `state` (original
// order) when there is none; noise_mode 0 none, 1 device array by particle id, 3 Philox
int vf_step(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
            double *compl_out, double dt, double D_phi, double friction, int noise_mode, const double *noise,
            uint64_t seed, uint64_t step);
// as fast_run_graph: all nsteps replayed from graphs of 10 steps and of 1 step per buffer parity,
// Philox noise from the device step counter; returns nsteps, 0 when not applicable now (no
// resident copy in the stepper's form), -1 on a HIP error
long vf_run_graph(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
                  double *compl_out, double dt, double D_phi, double friction, uint64_t seed, uint64_t step,
                  long nsteps);
bool vf_graph_ready(const VerletFast *v, const double *nat, int srk);
int vf_enter(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat);
int vf_prepare_graphs(VerletFast *v, hipStream_t st, const DevParams &P, int srk, const double *state,
                      const double *nat, double *compl_out, double dt, double D_phi, double friction, uint64_t seed);
long vf_graph_captures(const VerletFast *v);
bool vf_is_resident(const VerletFast *v);
// writes the resident state back into `state` (original order); the copy stays resident
int vf_sync(VerletFast *v, hipStream_t st, double *state);
// forgets the resident copy (the next step re-enters and rebuilds)
void vf_drop(VerletFast *v);
// the device's rebuild flag and accumulated displacement, and the number of lists that overflowed
// the fixed list width (waits for the stream)
int vf_status(VerletFast *v, hipStream_t st, int *should_update, double *acc, long *overflow);
// the lists of the last rebuild in original particle ids: counts[n] by id, then each id's
// neighbours in list order (stencil order of the build) into flat
int vf_lists(VerletFast *v, hipStream_t st, std::vector<int> *counts, std::vector<int> *flat);
int vf_list_width(const VerletFast *v);

}  // namespace spss
