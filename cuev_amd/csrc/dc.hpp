// dc.hpp — divide and conquer for the symmetric tridiagonal eigenproblem with
// the eigenvector matrix resident in device memory.
//
// Replaces the body of the reference's tridi_dc (src/cuda/tridi.cu:425-452, a
// MAGMA-style hybrid: host leaf solves and secular roots, one GPU GEMM pair per
// merge at tridi.cu:187-234). Here the host keeps the O(n) bookkeeping of every
// merge (deflation, column typing) and the leaves; the secular roots, the
// Gu-Eisenstat vector, the merge eigenvector matrix, the GEMMs and the sorted
// placement all run on the GPU, so Q never travels over PCIe.
#pragma once

#include "common.hpp"
#include "gemm.hpp"

namespace cuev {

template <typename T> struct DcState; // device/pinned staging, sized for n

template <typename T> int dc_state_create(DcState<T> *&st, int n);
template <typename T> void dc_state_destroy(DcState<T> *st);

// Eigen-decomposition of the symmetric tridiagonal (d[0:n], e[0:n-1]) given on
// the host. On return Q (n x n, ld ldq, device) holds the orthonormal
// eigenvectors as columns and eval (device, n) the eigenvalues ascending.
// Qc, S and Qn are device work buffers of n*n elements each; they may alias
// nothing else used by the call. All work goes on stream s, which the call
// synchronises before returning. gs is split-K scratch owned by s (may be empty).
// leaf <= 0 selects the default leaf size. gate (may be null) is an event the
// stream waits for before the first tree level whose largest merge has gate_from
// rows or more: a caller running other GPU-wide work beside the D&C lets the
// small, latency-bound merges overlap it and keeps the large merges, whose GEMMs
// want the whole GPU, until it is done.
template <typename T>
int dc_solve(DcState<T> &st, hipStream_t s, int n, const T *d, const T *e, T *Q, i64 ldq, T *eval,
             T *Qc, T *S, T *Qn, Scratch gs, int leaf, hipEvent_t gate = nullptr, int gate_from = 0);

} // namespace cuev
