// Launch wrappers of the realignment-screen kernels (rv_screen.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "rv_screen.h"

namespace rv {

hipError_t launch_screen_features(const ScreenIn& S, hipStream_t s);
hipError_t launch_mlp_forward(const float* P, const float* X, int64_t n, float* out, hipStream_t s);
hipError_t launch_mlp_grad(const float* P, const float* X, const uint8_t* labels, const int32_t* idx, int64_t nb,
                           float inv_batch, float pos_weight, float* grad, float* loss, hipStream_t s);
hipError_t launch_adam(float* P, float* grad, float* m, float* v, int n, float lr, float beta1, float beta2, float eps,
                       float c1, float c2, hipStream_t s);

}  // namespace rv
