// The exact operator A = K(X_source, X_source) of a uniform grid, applied by FFT convolution
// (gridop.hip).  Internal interface used by the refinement in solver.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hssc.h"

namespace hssc {

// R = B - A X (B given) or R = A X (B null) for m right-hand sides (N x m column-major,
// double or interleaved complex double as the kernel).  Asynchronous on st.
void grid_op_apply(const hssc_grid_op* op, const void* X, const void* B, void* R, int64_t m,
                   hipStream_t st);
// N of the operator and whether it is complex.
int64_t grid_op_n(const hssc_grid_op* op);
bool grid_op_complex(const hssc_grid_op* op);
int grid_op_device(const hssc_grid_op* op);
// Distinct for every operator the process creates (no two live or past operators share one).
uint64_t grid_op_uid(const hssc_grid_op* op);

}  // namespace hssc
