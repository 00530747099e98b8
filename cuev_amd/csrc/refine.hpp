// refine.hpp — mixed-precision solve: FP32 two-stage reduction and back-transform,
// host FP64 divide and conquer, FP64 Ogita-Aishima refinement on the GPU.
#pragma once

#include "common.hpp"

namespace cuev {

// Free the idle cached FP64 refinement buffers of every device.
void mixed_release_all();

int solve_mixed(double *A, int n, double *eval, double *evec, hipStream_t s, cuev_solve_timer *timer,
                const cuev_options &opt, int b, int k);

} // namespace cuev
