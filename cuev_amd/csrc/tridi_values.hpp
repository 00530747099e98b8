// tridi_values.hpp — eigenvalues il..iu of a symmetric tridiagonal matrix on the GPU.
#pragma once

#include "common.hpp"

namespace cuev {

// w[0 .. iu-il] = eigenvalues il..iu (0-based, ascending) of tridiag(e, d, e);
// d (n) and e (n-1 used) in device memory. Asynchronous on s; no scratch memory.
template <typename T> int tridi_values(hipStream_t s, const T *d, const T *e, int n, int il, int iu, T *w);

} // namespace cuev
