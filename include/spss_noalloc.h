/* spss_noalloc.h -- the C ABI entry point of libspss_hip.so that allocates nothing.
 *
 * No counterpart in the reference.  It is for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  It is the
 * allocation-free variant of spss_order_parameter (include/spss_hip.h): the same kernels and the
 * same compensated reduction, with the block partials in a fixed device array of the library
 * (one per GPU, 2 MiB, held until the process ends) instead of a handle's scratch.
 */
#ifndef SPSS_NOALLOC_H
#define SPSS_NOALLOC_H

#include "spss_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* largest n the fixed scratch serves (one block partial of 4 doubles per 256 particles) */
#define SPSS_NOALLOC_MAX_PARTICLES 16777216L

/* The complex order parameter (1/n) sum_k exp(i * d_state[s*k + slot]) of a device state array
 * (AoS FP64, s doubles per particle, the reference layout), as SaveSummaryStatistics forms it
 * (BinaryObserverFirstOrderPtr.cpp:126-146), accumulated in FP64 with compensated sums exactly as
 * spss_order_parameter does: d_out[0] = Re, d_out[1] = Im (device memory, 2 doubles).  Enqueued
 * on hip_stream (a hipStream_t; NULL = the default stream) and not waited for; no allocation, no
 * synchronisation.  n > SPSS_NOALLOC_MAX_PARTICLES is rejected with SPSS_ERR_CAPACITY.  Calls on
 * one GPU share its scratch: calls that overlap on one GPU (other streams or threads) race. */
int spss_order_parameter_noalloc(const double *d_state, int n, int s, int slot, double *d_out, void *hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* SPSS_NOALLOC_H */
