// backward.hip — gradient of the symmetric eigendecomposition (no reference
// counterpart).
//
// For A = V diag(lambda) V^T with cotangents (Vbar, lambar):
//   Abar = V (diag(lambar) + F o Skw) V^T,   Skw = (V^T Vbar - Vbar^T V) / 2,
//   F_ij = 1 / (lambda_j - lambda_i)  (i != j),  F_ii = 0.
// F o Skw is symmetric (both factors are antisymmetric), so Abar is symmetric;
// it equals the symmetric part of V (diag(lambar) + F o (V^T Vbar)) V^T, the
// form usually quoted. Pairs with |lambda_j - lambda_i| <= gap_tol get F_ij = 0
// and are counted (each unordered pair once).
//
// Products: C = V^T Vbar, then K = diag(lambar) + F o skew(C) in one fused
// element-wise pass over pairs (i < j) that writes K[i,j] and K[j,i] in place,
// then T = V K and Abar = T V^T (rocBLAS for large n, the MFMA tile engine
// below that). Only V and lambda are needed from the forward pass.
#include "backward.hpp"

#include "gemm.hpp"
#include "util.hpp"

namespace cuev {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void eig_bwd_scale_kernel(T *C, i64 ldc, const T *lam,
                                                            const T *lambar, int n, T gap_tol,
                                                            int *n_masked) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (i64)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    if (i > j) return;
    if (i == j) {
        C[i + (i64)i * ldc] = lambar ? lambar[i] : T(0);
        return;
    }
    const T cij = C[i + (i64)j * ldc], cji = C[j + (i64)i * ldc];
    const T gap = lam[j] - lam[i];
    T v = T(0);
    if (fabs((double)gap) > (double)gap_tol) {
        v = T(0.5) * (cij - cji) / gap;
    } else if (n_masked) {
        atomicAdd(n_masked, 1);
    }
    C[i + (i64)j * ldc] = v;
    C[j + (i64)i * ldc] = v;
}

} // namespace

template <typename T>
int eig_backward(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar,
                 double gap_tol, int *n_masked_host, hipStream_t s) {
    if (n < 1 || !V || !lam || !Abar)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "eig_backward: n=%d V=%p lambda=%p Abar=%p", n,
                  (const void *)V, (const void *)lam, (void *)Abar);
    const size_t nn = (size_t)n * n;
    T *C = nullptr, *Tm = nullptr;
    int *dmask = nullptr;
    void *scr = nullptr;
    const size_t scr_bytes = gemm_scratch_bytes(n, n, sizeof(T));
    auto cleanup = [&]() {
        if (C) (void)hipFree(C);
        if (Tm) (void)hipFree(Tm);
        if (dmask) (void)hipFree(dmask);
        if (scr) (void)hipFree(scr);
    };
    if (hipMalloc(&C, nn * sizeof(T)) != hipSuccess || hipMalloc(&Tm, nn * sizeof(T)) != hipSuccess ||
        hipMalloc(&dmask, sizeof(int)) != hipSuccess || hipMalloc(&scr, scr_bytes) != hipSuccess) {
        cleanup();
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "eig_backward: device allocation failed (n=%d)", n);
    }
    Scratch gs{scr, scr_bytes};
    gs.vendor = 1; // three plain n^3 products with the GPU to themselves: rocBLAS when large
    int st = CUEV_STATUS_SUCCESS;
    do {
        if ((st = fill_int(s, dmask, 1, 0))) break;
        if (Vbar) {
            if ((st = gemm<T>(s, 1, 0, n, n, n, T(1), V, n, Vbar, n, T(0), C, n, gs))) break;
        } else {
            if ((st = fill2d<T>(s, C, n, n, n, T(0), T(0)))) break;
        }
        hipLaunchKernelGGL(eig_bwd_scale_kernel<T>, dim3((unsigned)ceil_div((i64)nn, 256)),
                           dim3(256), 0, s, C, (i64)n, lam, lambar, n, (T)gap_tol, dmask);
        if (hipGetLastError() != hipSuccess) {
            detail::set_error("eig_backward: scale kernel launch failed");
            st = CUEV_STATUS_HIP_ERROR;
            break;
        }
        if ((st = gemm<T>(s, 0, 0, n, n, n, T(1), V, n, C, n, T(0), Tm, n, gs))) break;
        if ((st = gemm<T>(s, 0, 1, n, n, n, T(1), Tm, n, V, n, T(0), Abar, n, gs))) break;
        int hm = 0;
        if (hipMemcpyAsync(&hm, dmask, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            detail::set_error("eig_backward: synchronisation failed");
            st = CUEV_STATUS_HIP_ERROR;
            break;
        }
        if (n_masked_host) *n_masked_host = hm;
    } while (0);
    (void)hipStreamSynchronize(s);
    cleanup();
    return st;
}

template int eig_backward<float>(const float *, const float *, const float *, const float *, int,
                                 float *, double, int *, hipStream_t);
template int eig_backward<double>(const double *, const double *, const double *, const double *,
                                  int, double *, double, int *, hipStream_t);

} // namespace cuev
