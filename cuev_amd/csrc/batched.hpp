// batched.hpp — eigendecomposition of a batch of small symmetric matrices.
#pragma once

#include "common.hpp"

namespace cuev {

constexpr int kBatchedMaxN = 256;

// A: batch matrices of order n (column-major, each n x n contiguous, lower triangle
// read, overwritten); W: batch x n eigenvalues (ascending per matrix); V: batch
// eigenvector matrices (columns). Blocking. NOT_CONVERGED if any matrix did not
// converge (the count goes to *n_not_converged when given).
template <typename T> int eig_batched(hipStream_t s, T *A, int n, int batch, T *W, T *V, int *n_not_converged);

} // namespace cuev
