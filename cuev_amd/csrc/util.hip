// util.hip — error bookkeeping, device queries and small element-wise kernels.
#include "util.hpp"

#include <cstdarg>
#include <cstring>
#include <mutex>

namespace cuev {

namespace detail {
static thread_local char g_last_error[1024] = {0};
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}
const char *last_error() { return g_last_error; }
void clear_error() { g_last_error[0] = 0; }
} // namespace detail

int device_cu_count() {
    static std::mutex mu;
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    std::lock_guard<std::mutex> lk(mu);
    if (cache[dev] == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        cache[dev] = cus;
    }
    return cache[dev];
}

namespace {

template <typename T>
__global__ void fill2d_kernel(T *A, i64 lda, i64 m, i64 n, T v, T diag) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const i64 i = idx % m, j = idx / m;
    A[i + j * lda] = (i == j) ? diag : v;
}

template <typename T>
__global__ void copy2d_kernel(const T *S, i64 lds, T *D, i64 ldd, i64 m, i64 n) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const i64 i = idx % m, j = idx / m;
    D[i + j * ldd] = S[i + j * lds];
}

template <typename S, typename D>
__global__ void convert2d_kernel(const S *src, i64 lds, D *dst, i64 ldd, i64 m, i64 n) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const i64 i = idx % m, j = idx / m;
    dst[i + j * ldd] = (D)src[i + j * lds];
}

__global__ void fill_int_kernel(int *p, i64 n, int v) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

} // namespace

template <typename T> int fill2d(hipStream_t s, T *A, i64 lda, i64 m, i64 n, T v, T diag) {
    if (m <= 0 || n <= 0) return CUEV_STATUS_SUCCESS;
    hipLaunchKernelGGL(fill2d_kernel<T>, dim3((unsigned)ceil_div(m * n, 256)), dim3(256), 0, s, A,
                       lda, m, n, v, diag);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int copy2d(hipStream_t s, const T *S, i64 lds, T *D, i64 ldd, i64 m, i64 n) {
    if (m <= 0 || n <= 0) return CUEV_STATUS_SUCCESS;
    hipLaunchKernelGGL(copy2d_kernel<T>, dim3((unsigned)ceil_div(m * n, 256)), dim3(256), 0, s, S,
                       lds, D, ldd, m, n);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename S, typename D>
int convert2d(hipStream_t s, const S *src, i64 lds, D *dst, i64 ldd, i64 m, i64 n) {
    if (m <= 0 || n <= 0) return CUEV_STATUS_SUCCESS;
    hipLaunchKernelGGL((convert2d_kernel<S, D>), dim3((unsigned)ceil_div(m * n, 256)), dim3(256),
                       0, s, src, lds, dst, ldd, m, n);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

int fill_int(hipStream_t s, int *p, i64 n, int v) {
    if (n <= 0) return CUEV_STATUS_SUCCESS;
    hipLaunchKernelGGL(fill_int_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template int fill2d<float>(hipStream_t, float *, i64, i64, i64, float, float);
template int fill2d<double>(hipStream_t, double *, i64, i64, i64, double, double);
template int copy2d<float>(hipStream_t, const float *, i64, float *, i64, i64, i64);
template int copy2d<double>(hipStream_t, const double *, i64, double *, i64, i64, i64);
template int convert2d<float, double>(hipStream_t, const float *, i64, double *, i64, i64, i64);
template int convert2d<double, float>(hipStream_t, const double *, i64, float *, i64, i64, i64);

} // namespace cuev
