// tridi_values.hip — eigenvalues of a symmetric tridiagonal matrix on the GPU by
// Sturm-sequence multisection (no reference counterpart: the reference always
// computes every eigenpair with its D&C, src/cuda/tridi.cu:12-13).
//
// For T = tridiag(e, d, e) the number of eigenvalues below x equals the number of
// negative pivots of the LDL^T factorisation of T - xI:
//   q_0 = d_0 - x,  q_i = (d_i - x) - e_{i-1}^2 / q_{i-1},
// with a pivot of magnitude below pivmin replaced by -pivmin (the safeguard of
// LAPACK's dstebz / dlaneg, which keeps the count monotone in x and the
// computed count exact for a matrix within a few ulps of T). Eigenvalue j
// (0-based, ascending) is the point where the count passes j.
//
// L consecutive lanes of a wave share one eigenvalue: each iteration they count
// at L interior points of the current interval [lo, hi], split into L + 1 equal
// parts, and keep the part in which the count passes j (a ballot over the
// group: the counts are monotone in x, so the lanes with count <= j form a
// prefix). The interval shrinks by L + 1 per iteration, until its width is below
// max(2 eps max(|lo|, |hi|), eps ||T||, pivmin) — the stopping rule of dstebz. The
// Gershgorin interval that starts it, widened by 2.1 eps ||T|| + 2 pivmin as in
// dstebz, is formed by every wave from (d, e) itself, so the kernel needs no scratch
// memory. Everything runs in FP64 (the float entry converts on load).
#include "tridi_values.hpp"

#include <cfloat>

namespace cuev {
namespace {

constexpr int kLanesPerEig = 8;  // L: lanes sharing one eigenvalue
constexpr int kTvThreads = 256;

template <typename T>
__global__ __launch_bounds__(kTvThreads) void sturm_multisect_kernel(const T *__restrict__ d,
                                                                     const T *__restrict__ e, int n,
                                                                     int il, int iu, T *__restrict__ w) {
    constexpr int L = kLanesPerEig;
    const int lane = threadIdx.x & 63;
    // --- Gershgorin bounds, ||T||_inf and max e^2, over the whole matrix (every wave)
    double gl = DBL_MAX, gu = -DBL_MAX, emax2 = 0.0;
    for (int i = lane; i < n; i += 64) {
        const double di = (double)d[i];
        const double el = (i > 0) ? fabs((double)e[i - 1]) : 0.0;
        const double er = (i < n - 1) ? fabs((double)e[i]) : 0.0;
        gl = fmin(gl, di - el - er);
        gu = fmax(gu, di + el + er);
        emax2 = fmax(emax2, er * er);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        gl = fmin(gl, __shfl_xor(gl, o, 64));
        gu = fmax(gu, __shfl_xor(gu, o, 64));
        emax2 = fmax(emax2, __shfl_xor(emax2, o, 64));
    }
    const double eps = DBL_EPSILON * 0.5; // unit roundoff, as LAPACK's dlamch('P') / 2
    const double safmin = DBL_MIN;
    const double pivmin = safmin * fmax(1.0, emax2);
    const double tnorm = fmax(fabs(gl), fabs(gu));
    const double fudge = 2.1 * eps * tnorm + 2.0 * pivmin;
    gl -= fudge;
    gu += fudge;
    const double atol = 2.0 * eps * tnorm; // absolute part of the stopping width

    // --- one eigenvalue per group of L lanes
    const int grp = (int)((blockIdx.x * (i64)blockDim.x + threadIdx.x) / L);
    const int j = il + grp;
    const bool active = j <= iu;
    const int sub = lane % L;
    const int gbase = lane - sub;
    double lo = gl, hi = gu;
    for (int it = 0; it < 64; ++it) {
        const double width = hi - lo;
        const double tol = fmax(fmax(2.0 * eps * fmax(fabs(lo), fabs(hi)), atol), pivmin);
        // groups stop independently; inactive lanes still join the ballot
        const bool go = active && width > tol;
        if (!__any(go)) break;
        const double x = lo + width * (double)(sub + 1) / (double)(L + 1);
        int cnt = 0;
        if (go) {
            double q = (double)d[0] - x;
            if (fabs(q) < pivmin) q = -pivmin;
            cnt = q < 0.0;
            for (int i = 1; i < n; ++i) {
                const double ei = (double)e[i - 1];
                q = ((double)d[i] - x) - (ei * ei) / q;
                if (fabs(q) < pivmin) q = -pivmin;
                cnt += q < 0.0;
            }
        }
        // lanes of this group whose point has count <= j form a prefix of length p
        const unsigned long long below = __ballot(go && cnt <= j);
        const int p = __popcll((below >> gbase) & ((1ull << L) - 1ull));
        const double xl = __shfl(x, gbase + (p > 0 ? p - 1 : 0), 64);
        const double xh = __shfl(x, gbase + (p < L ? p : L - 1), 64);
        if (go) {
            if (p > 0) lo = xl;
            if (p < L) hi = xh;
        }
    }
    if (active && sub == 0) w[j - il] = (T)(0.5 * (lo + hi));
}

} // namespace

template <typename T> int tridi_values(hipStream_t s, const T *d, const T *e, int n, int il, int iu, T *w) {
    if (n < 1 || !d || !w || (n > 1 && !e))
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "tridi_values: n=%d d=%p e=%p w=%p", n, (const void *)d,
                  (const void *)e, (void *)w);
    if (il < 0 || iu >= n || il > iu)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "tridi_values: index range [%d, %d] outside [0, %d)", il, iu, n);
    const i64 threads = (i64)(iu - il + 1) * kLanesPerEig;
    hipLaunchKernelGGL(sturm_multisect_kernel<T>, dim3((unsigned)ceil_div(threads, kTvThreads)), dim3(kTvThreads),
                       0, s, d, e, n, il, iu, w);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template int tridi_values<float>(hipStream_t, const float *, const float *, int, int, int, float *);
template int tridi_values<double>(hipStream_t, const double *, const double *, int, int, int, double *);

} // namespace cuev
