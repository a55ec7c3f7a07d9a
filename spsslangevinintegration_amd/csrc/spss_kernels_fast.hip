// spss_kernels_fast.hip -- placeholder until the FP32C path lands
#include "spss_fast.hpp"
#include "../../include/spss_hip.h"
namespace spss {
struct FastPath { int dummy; };
int fast_create(FastPath **out, hipStream_t, const DevParams &, double, int, int) { *out = nullptr; return SPSS_ERR_UNSUPPORTED; }
void fast_destroy(FastPath *) {}
void fast_set_stream(FastPath *, hipStream_t) {}
const char *fast_last_error() { return "FP32C path not built"; }
int fast_rhs(FastPath *, hipStream_t, const DevParams &, const double *, const double *, double *, double *) { return SPSS_ERR_UNSUPPORTED; }
int fast_pairs_count(FastPath *, hipStream_t, const DevParams &, const double *, int *) { return SPSS_ERR_UNSUPPORTED; }
int fast_pairs_fill(FastPath *, hipStream_t, const DevParams &, const double *, const int *, int *) { return SPSS_ERR_UNSUPPORTED; }
}
