// Device-side hooks of the sizing network used by the graph build.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/genomix_sizer.h"

// Forward pass on device arrays, ordered on the caller's stream.
int gx_mlp_forward_device(gx_mlp* m, const float* dX, int B, float* dY, hipStream_t st);
int gx_mlp_device_of(const gx_mlp* m);
int gx_mlp_dims(const gx_mlp* m, int* din, int* dout);
