// util.hpp — small element-wise kernels and error bookkeeping.
#pragma once

#include "common.hpp"

namespace cuev {

namespace detail {
const char *last_error();
void clear_error();
} // namespace detail

// A[i,j] = (i == j) ? diag : v for the m x n block at A (ld lda).
template <typename T> int fill2d(hipStream_t s, T *A, i64 lda, i64 m, i64 n, T v, T diag);
// D[0:m, 0:n] = S[0:m, 0:n]
template <typename T> int copy2d(hipStream_t s, const T *S, i64 lds, T *D, i64 ldd, i64 m, i64 n);
// dst = (D) src, element-wise over an m x n block
template <typename S, typename D>
int convert2d(hipStream_t s, const S *src, i64 lds, D *dst, i64 ldd, i64 m, i64 n);
// p[0:n] = v
int fill_int(hipStream_t s, int *p, i64 n, int v);

} // namespace cuev
