// panel_qr.hpp — persistent Householder panel factorization (see panel_qr.hip).
#pragma once

#include "common.hpp"

namespace cuev {

// Factor the rows x b panel at A (ld lda) in place (R upper, reflector tails
// lower), write the explicit unit-lower reflectors to Y (ld ldy; may be null),
// W = Y T to W (ld ldw; may be null), T (b x b, ld b) to Tm and tau to tau.
// scratch: device buffer of >= panel_qr_scratch_bytes(rows, b) bytes;
// counter: one device unsigned (reset by this call); err: device int flag
// (set to DEV_WATCHDOG if a bounded wait expired).
template <typename T>
int panel_qr(hipStream_t s, T *A, i64 lda, T *Y, i64 ldy, T *W, i64 ldw, T *Tm, T *tau, int rows,
             int b, void *scratch, size_t scratch_bytes, unsigned *counter, int *err);

size_t panel_qr_scratch_bytes(int rows, int b);

// Rows of the panel each workgroup keeps in LDS.
template <typename T> int panel_qr_chunk(int b);

} // namespace cuev
