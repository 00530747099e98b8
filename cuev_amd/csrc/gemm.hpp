// gemm.hpp — host interface of the gfx950 MFMA matrix-multiply engine.
//
// One tile kernel serves every dense product of the solver: the cuBLAS gemm /
// symm / syr2k call sites of the reference (src/cuda/cublas.cu:15-56 used at
// src/cuda/dbbr.cu:140-224 and src/cuda/backtransform.cu:225-228), the
// refinement products and the backward pass. FP64 runs on
// v_mfma_f64_16x16x4_f64, FP32 on v_mfma_f32_16x16x4_f32 with FP32 accumulation.
#pragma once

#include "common.hpp"

namespace cuev {

// One GEMM operand. A is viewed as M x K, B as K x N (column-major storage).
//   A, trans=0: element (m,k) at p[m + k*ld];  trans=1: at p[k + m*ld]
//   B, trans=0: element (k,n) at p[k + n*ld];  trans=1: at p[n + k*ld]
// K-concatenation: for k >= split the element comes from (p1, ld1) at k - split
// (split must be a multiple of 16). sym (A only): A is square symmetric and
// only its lower triangle p0[m + k*ld0], m >= k, is read.
template <typename T> struct Operand {
    const T *p0 = nullptr;
    i64 ld0 = 0;
    const T *p1 = nullptr;
    i64 ld1 = 0;
    i64 split = -1;
    int trans = 0;
    int sym = 0;
};

template <typename T> inline Operand<T> opnd(const T *p, i64 ld, int trans = 0) {
    Operand<T> o;
    o.p0 = p;
    o.ld0 = ld;
    o.trans = trans;
    return o;
}
template <typename T>
inline Operand<T> opnd_cat(const T *p0, i64 ld0, const T *p1, i64 ld1, i64 split, int trans) {
    Operand<T> o;
    o.p0 = p0;
    o.ld0 = ld0;
    o.p1 = p1;
    o.ld1 = ld1;
    o.split = split;
    o.trans = trans;
    return o;
}

// Device scratch the engine may use for split-K partial sums. small_tiles: use
// only 64 x 64 block tiles (fewer registers per wave), for products meant to run
// beside a kernel that holds most of each SIMD's registers.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    int small_tiles = 0;
};

// C <- alpha * op(A) op(B) + beta * C  (lower != 0: only entries row >= col of C
// are written; requires M == N). beta == 0 never reads C.
template <typename T>
int gemm_ex(hipStream_t s, i64 M, i64 N, i64 K, T alpha, const Operand<T> &A,
            const Operand<T> &B, T beta, T *C, i64 ldc, int lower, Scratch scratch);

template <typename T>
inline int gemm(hipStream_t s, int ta, int tb, i64 M, i64 N, i64 K, T alpha, const T *A, i64 lda,
                const T *B, i64 ldb, T beta, T *C, i64 ldc, Scratch scratch) {
    return gemm_ex<T>(s, M, N, K, alpha, opnd(A, lda, ta), opnd(B, ldb, tb), beta, C, ldc, 0,
                      scratch);
}

// C <- alpha * Sym(A) * B + beta * C with A m x m, lower triangle stored.
template <typename T>
inline int symm_lower(hipStream_t s, i64 m, i64 n, T alpha, const T *A, i64 lda, const T *B,
                      i64 ldb, T beta, T *C, i64 ldc, Scratch scratch) {
    Operand<T> a = opnd(A, lda);
    a.sym = 1;
    return gemm_ex<T>(s, m, n, m, alpha, a, opnd(B, ldb, 0), beta, C, ldc, 0, scratch);
}

// lower(C) <- alpha (X Y^T + Y X^T) + beta lower(C); X, Y are n x k.
template <typename T>
int syr2k_lower(hipStream_t s, i64 n, i64 k, T alpha, const T *X, i64 ldx, const T *Y, i64 ldy,
                T beta, T *C, i64 ldc, Scratch scratch);

// Bytes of split-K scratch a product of this shape may ask for (upper bound used
// to size workspaces).
size_t gemm_scratch_bytes(i64 M, i64 N, size_t elem);

} // namespace cuev
