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
// then T = V K and Abar = T V^T, all on the MFMA tile engine. Only V and lambda
// are needed from the forward pass.
//
// Work buffers (C and T, n x n each, and the masked-pair counter):
//  * cuev_{s,d}symm_eig_backward keeps them per device between calls (allocated
//    on the first call of a size, grown when a larger n comes; a call that finds
//    the cached buffer in use by another call allocates its own);
//  * cuev_{s,d}symm_eig_backward_noalloc allocates nothing: it uses the file-scope
//    device array `scratch` below (n <= kNoAllocMaxN = 1024, 16 MiB per GPU for the
//    life of the process), whose address is looked up once per device. Every call
//    on one GPU shares that GPU's array, so calls that overlap on one GPU race; the
//    counter is cleared (by a fill kernel) at the start of every call. With
//    n_masked == NULL it only queues the work on the stream and returns.
#include "backward.hpp"

#include <mutex>

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

// The backward pass on caller-provided work buffers (C, Tm: n x n; dmask: one int).
template <typename T>
int backward_core(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar, double gap_tol,
                  int *n_masked_host, hipStream_t s, T *C, T *Tm, int *dmask, Scratch gs, bool sync) {
    const size_t nn = (size_t)n * n;
    CUEV_TRY(fill_int(s, dmask, 1, 0));
    if (Vbar) {
        CUEV_TRY(gemm<T>(s, 1, 0, n, n, n, T(1), V, n, Vbar, n, T(0), C, n, gs));
    } else {
        CUEV_TRY(fill2d<T>(s, C, n, n, n, T(0), T(0)));
    }
    hipLaunchKernelGGL(eig_bwd_scale_kernel<T>, dim3((unsigned)ceil_div((i64)nn, 256)), dim3(256), 0, s, C,
                       (i64)n, lam, lambar, n, (T)gap_tol, dmask);
    CUEV_LAUNCH_CHECK();
    CUEV_TRY(gemm<T>(s, 0, 0, n, n, n, T(1), V, n, C, n, T(0), Tm, n, gs));
    CUEV_TRY(gemm<T>(s, 0, 1, n, n, n, T(1), Tm, n, V, n, T(0), Abar, n, gs));
    if (!n_masked_host && !sync) return CUEV_STATUS_SUCCESS;
    int hm = 0;
    CUEV_HIP(hipMemcpyAsync(&hm, dmask, sizeof(int), hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipStreamSynchronize(s));
    if (n_masked_host) *n_masked_host = hm;
    return CUEV_STATUS_SUCCESS;
}

int check_args(const void *V, const void *lam, const void *Abar, int n) {
    if (n < 1 || !V || !lam || !Abar)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "eig_backward: n=%d V=%p lambda=%p Abar=%p", n, V, lam, Abar);
    return CUEV_STATUS_SUCCESS;
}

// ---- cached work buffer of the allocating entry point (one idle buffer per device)
struct BwdBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int device = -1;
};
std::mutex g_bwd_mu;
BwdBuf g_bwd_idle[64];

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// ---- the allocation-free entry point's device array
constexpr int kNoAllocMaxN = 1024;
constexpr size_t kNoAllocBytes = 2 * (size_t)kNoAllocMaxN * kNoAllocMaxN * sizeof(double) + 256;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scr_mu;
void *g_scr_ptr[64] = {};

// Device address of `scratch` on the current device, looked up on the first call there.
int scratch_ptr(void **out) {
    int dev = 0;
    CUEV_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "eig_backward_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lk(g_scr_mu);
    if (!g_scr_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "eig_backward_noalloc: hipGetSymbolAddress(scratch) -> %s",
                      hipGetErrorString(e));
        g_scr_ptr[dev] = p;
    }
    *out = g_scr_ptr[dev];
    return CUEV_STATUS_SUCCESS;
}

} // namespace

template <typename T>
int eig_backward(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar,
                 double gap_tol, int *n_masked_host, hipStream_t s) {
    CUEV_TRY(check_args(V, lam, Abar, n));
    const size_t mat = align256((size_t)n * n * sizeof(T));
    const size_t gsb = gemm_scratch_bytes(n, n, sizeof(T));
    const size_t need = 2 * mat + 256 + gsb;
    int dev = 0;
    CUEV_HIP(hipGetDevice(&dev));
    BwdBuf buf;
    {
        std::lock_guard<std::mutex> lk(g_bwd_mu);
        BwdBuf &slot = g_bwd_idle[dev & 63];
        if (slot.p && slot.device == dev && slot.bytes >= need) {
            buf = slot;
            slot = BwdBuf{};
        }
    }
    if (!buf.p) {
        buf.device = dev;
        buf.bytes = need;
        if (hipMalloc(&buf.p, need) != hipSuccess)
            CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "eig_backward: device allocation of %zu bytes failed (n=%d)", need, n);
    }
    auto *base = static_cast<unsigned char *>(buf.p);
    T *C = reinterpret_cast<T *>(base);
    T *Tm = reinterpret_cast<T *>(base + mat);
    int *dmask = reinterpret_cast<int *>(base + 2 * mat);
    Scratch gs{base + 2 * mat + 256, gsb};
    const int st =
        backward_core<T>(V, lam, Vbar, lambar, n, Abar, gap_tol, n_masked_host, s, C, Tm, dmask, gs, true);
    (void)hipStreamSynchronize(s);
    BwdBuf victim;
    {
        std::lock_guard<std::mutex> lk(g_bwd_mu);
        BwdBuf &slot = g_bwd_idle[dev & 63];
        if (slot.p && slot.bytes >= buf.bytes) {
            victim = buf; // keep the larger of the two
        } else {
            victim = slot;
            slot = buf;
        }
    }
    if (victim.p) (void)hipFree(victim.p);
    return st;
}

void backward_release_all() {
    BwdBuf victims[64];
    {
        std::lock_guard<std::mutex> lk(g_bwd_mu);
        for (int i = 0; i < 64; ++i) {
            victims[i] = g_bwd_idle[i];
            g_bwd_idle[i] = BwdBuf{};
        }
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (BwdBuf &v : victims)
        if (v.p) {
            (void)hipSetDevice(v.device);
            (void)hipDeviceSynchronize();
            (void)hipFree(v.p);
        }
    (void)hipSetDevice(cur);
}

template <typename T>
int eig_backward_noalloc(const T *V, const T *lam, const T *Vbar, const T *lambar, int n, T *Abar,
                         double gap_tol, int *n_masked_host, hipStream_t s) {
    CUEV_TRY(check_args(V, lam, Abar, n));
    if (n > kNoAllocMaxN)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT,
                  "eig_backward_noalloc: n=%d exceeds the fixed work array's limit n <= %d", n, kNoAllocMaxN);
    void *p = nullptr;
    CUEV_TRY(scratch_ptr(&p));
    auto *base = static_cast<unsigned char *>(p);
    const size_t mat = (size_t)kNoAllocMaxN * kNoAllocMaxN * sizeof(double);
    T *C = reinterpret_cast<T *>(base);
    T *Tm = reinterpret_cast<T *>(base + mat);
    int *dmask = reinterpret_cast<int *>(base + 2 * mat);
    // no split-K partials: the products run as single passes over K. Without
    // n_masked the call returns once the work is queued on s.
    return backward_core<T>(V, lam, Vbar, lambar, n, Abar, gap_tol, n_masked_host, s, C, Tm, dmask, Scratch{},
                            false);
}

template int eig_backward<float>(const float *, const float *, const float *, const float *, int,
                                 float *, double, int *, hipStream_t);
template int eig_backward<double>(const double *, const double *, const double *, const double *,
                                  int, double *, double, int *, hipStream_t);
template int eig_backward_noalloc<float>(const float *, const float *, const float *, const float *, int,
                                         float *, double, int *, hipStream_t);
template int eig_backward_noalloc<double>(const double *, const double *, const double *, const double *,
                                          int, double *, double, int *, hipStream_t);

} // namespace cuev
