// context.hpp — a workspace plus the host staging of the tridiagonal stage.
#pragma once

#include "dc.hpp"
#include "workspace.hpp"

namespace cuev {

template <typename T> struct Context {
    Workspace<T> ws;
    // pinned host staging for the tridiagonal stage
    T *h_d = nullptr;     // n
    T *h_e = nullptr;     // n
    T *h_z = nullptr;     // n x n (Q_d)
    // host LAPACK workspace (pageable)
    T *h_work = nullptr;
    int64_t *h_iwork = nullptr;
    i64 lwork = 0;
    i64 liwork = 0;
    hipEvent_t ev[12] = {};
    int dtype = 0;
    // device-resident D&C, allocated on first use: state, Q and Qn (n x n each),
    // and a stream of its own so it can run beside the back-transform
    DcState<T> *dc = nullptr;
    T *dcQ = nullptr;
    T *dcQn = nullptr;
    void *dc_pool = nullptr;
    hipStream_t dc_stream = nullptr;
    // reordered back-transform: X^T = Q_b^T Q_s^T built on a stream of its own,
    // Q_s^T while the bulge chasing runs (events: start/end of each half)
    hipStream_t bt_stream = nullptr;
    hipEvent_t xev[4] = {};
    int dc_mode = 1; // stage API default: host ?stedc
    int dc_leaf = 0;
};

// Allocate the device D&C state and buffers (need_q: also the separate Q).
template <typename T> int dc_ensure(Context<T> &ctx, bool need_q);
// Create the back-transform stream and its events.
template <typename T> int bt_ensure(Context<T> &ctx);

template <typename T> int context_create(Context<T> *&ctx, int n, int b, int k, hipStream_t s);
template <typename T> void context_destroy(Context<T> *ctx);

// Process-wide cache used by the drop-in entry points: at most one idle context
// per (device, precision); a busy or mismatched entry is bypassed.
template <typename T> int context_acquire(Context<T> *&ctx, int n, int b, int k, hipStream_t s);
template <typename T> void context_release(Context<T> *ctx);
// Free every idle cached context of this precision (all devices).
template <typename T> void context_release_all();

} // namespace cuev
