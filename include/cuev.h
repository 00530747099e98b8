/*
 * cuev.h — C++ drop-in for cuEV's public entry point on MI355X.
 *
 *   cuev::symm_eig_solve<T>(A, n, eval, evec, stream, timer)
 *
 * keeps the reference's signature (include/cuev.h:31-33 of ynkrue/cuEV) with
 * hipStream_t in place of cudaStream_t, and its behaviour: float/double only,
 * column-major device pointers, A overwritten, eval ascending, evec columns,
 * host-blocking, and a failing call prints a message and exits (the reference's
 * CUDA_CHECK/CUSOLVER_CHECK behaviour, include/common.h:9-34). Callers that want
 * status codes instead call the C ABI in cuev_c.h directly.
 */
#pragma once

#include <cstdio>
#include <cstdlib>
#include <type_traits>

#include "cuev_c.h"

namespace cuev {

/// Per-stage timings; field order and meaning follow include/cuda/handle.h:25-37,
/// extra fields are appended. Layout-identical to cuev_solve_timer.
struct SolveTimer {
    float dbbr_ms = 0;
    float bc_ms = 0;
    float dc_ms = 0;
    float bt_ms = 0;
    float total_ms = 0;
    float bt_copyin_ms = 0;
    float bt_qb_ms = 0;
    float bt_qs_ms = 0;
    float bt_copyout_ms = 0;
    float refine_ms = 0;
    float overlap_ms = 0;
    int refine_sweeps = 0;
    int n_cluster_cols = 0;
    float refine_resid = 0;
    float refine_orth = 0;
};
static_assert(sizeof(SolveTimer) == sizeof(cuev_solve_timer), "SolveTimer layout");
static_assert(std::is_standard_layout<SolveTimer>::value, "SolveTimer layout");

inline void solve_timer_print(const SolveTimer &t) {
    std::printf("  -- stage timings (ms) --\n");
    std::printf("    DBBR  full->band              %9.2f\n", t.dbbr_ms);
    std::printf("    BC    band->tridi             %9.2f\n", t.bc_ms);
    std::printf("    D&C   tridiagonal (host)      %9.2f\n", t.dc_ms);
    std::printf("    BT    after D&C               %9.2f\n", t.bt_ms);
    std::printf("      BT.copyin   Q_d upload      %9.2f\n", t.bt_copyin_ms);
    std::printf("      BT.Qb       bc_back         %9.2f\n", t.bt_qb_ms);
    std::printf("      BT.Qs       sbr_back        %9.2f\n", t.bt_qs_ms);
    std::printf("      BT.copyout  final GEMM      %9.2f\n", t.bt_copyout_ms);
    if (t.refine_sweeps)
        std::printf("    refine %d sweeps             %9.2f\n", t.refine_sweeps, t.refine_ms);
    std::printf("    total (wall)                  %9.2f\n", t.total_ms);
}

namespace detail {
[[noreturn]] inline void fail(int status) {
    std::fprintf(stderr, "cuev error %d (%s): %s\n", status, cuev_status_string(status),
                 cuev_last_error_message());
    std::exit(1);
}
} // namespace detail

/// Compute all eigenvalues and eigenvectors of a real symmetric matrix
/// (lower triangle of A read; A destroyed).
template <typename T>
inline void symm_eig_solve(T *A, int n, T *eval, T *evec, hipStream_t stream,
                           SolveTimer *timer = nullptr) {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "symm_eig_solve is instantiated for float and double only");
    auto *t = reinterpret_cast<cuev_solve_timer *>(timer);
    int st;
    if constexpr (std::is_same<T, float>::value)
        st = cuev_ssymm_eig_solve(A, n, eval, evec, stream, t);
    else
        st = cuev_dsymm_eig_solve(A, n, eval, evec, stream, t);
    if (st != CUEV_STATUS_SUCCESS) detail::fail(st);
}

} // namespace cuev
