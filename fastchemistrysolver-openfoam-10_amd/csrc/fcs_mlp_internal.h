// fcs_mlp_internal.h -- library-internal entry of the learned seed's
// network used by the integrator (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "fcs.h"

// Forward pass on device arrays, enqueued on the caller's stream: X [rows][w0]
// raw inputs -> Y [rows][wL] raw outputs.  FCS_E_ARG when the network's
// input / output widths are not (w0, wL).
int fcs_mlp_forward_device(fcs_mlp* m, int rows, int w0, int wL, const float* X, float* Y, hipStream_t st);
