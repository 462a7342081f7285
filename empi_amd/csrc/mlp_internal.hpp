// Entry points of the region-scoring network used by the engine inside the
// same library (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include "empi_hip.h"

// y[r] for r < min(*rows_dev, max_rows), on stream s; no host synchronisation
// (the row count stays on the device, so the call can be graph-captured)
empi_status empi_mlp_forward_counted(empi_mlp* m, const float* X, const int* rows_dev, int max_rows, float* y,
                                     hipStream_t s);
int empi_mlp_input_dim(const empi_mlp* m);
