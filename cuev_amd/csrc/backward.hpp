// backward.hpp — eigendecomposition backward pass.
#pragma once

#include "common.hpp"

namespace cuev {

template <typename T>
int eig_backward(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar,
                 double gap_tol, int *n_masked, hipStream_t s);
// Free the cached work buffers of eig_backward (all devices).
void backward_release_all();
// Same without any allocation (n <= 1024): work buffers in a fixed per-GPU device array.
template <typename T>
int eig_backward_noalloc(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar,
                         double gap_tol, int *n_masked, hipStream_t s);

} // namespace cuev
