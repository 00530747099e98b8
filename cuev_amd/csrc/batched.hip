// batched.hip — eigendecomposition of many small symmetric matrices in one launch
// (no reference counterpart: the reference solves one matrix per call,
// include/cuev.h:31-33). For the autograd harness and similar callers with many
// problems of order n <= 256.
//
// One workgroup per matrix runs the cyclic two-sided Jacobi method with the
// round-robin (tournament) ordering: for m = n rounded up to even, each of the
// m - 1 steps of a sweep pairs the indices into m/2 disjoint pairs (the circle
// method: pair 0 = (m-1, r), pair k = ((r+k) mod (m-1), (r-k) mod (m-1)) in round
// r), so all m/2 rotations of a step are independent. For pair (p, q) with
// a = A_pp, d = A_qq, b = A_pq the rotation J = [[c, s], [-s, c]] on (p, q) uses
// Rutishauser's stable form theta = (d - a) / (2b), t = sign(theta) /
// (|theta| + sqrt(theta^2 + 1)), c = 1 / sqrt(1 + t^2), s = t c, which zeroes
// (J^T A J)_pq and gives A_pp - t b, A_qq + t b. A step updates every 2 x 2 block
// (P, R) of A as J_P^T A_PR J_R (blocks are disjoint, so the threads never share
// an element) and every row of V as V J. For odd n a virtual index n with zero
// row and column joins the pairing and never rotates.
//
// A sweep ends with the off-diagonal norm; the iteration stops when
// ||offdiag(A)||_F <= eps ||A||_F (the eigenvalue error is then below a few
// eps ||A||, and the last sweep's rotations are at rounding level) or after
// kMaxSweeps sweeps (reported as not converged). The matrix lives in LDS when it
// fits in 128 KiB (n <= 128 in FP64, n <= 180 in FP32) and in the caller's A
// buffer (which the call overwrites, as the single-matrix solve does) otherwise;
// V is accumulated in evec. Only the lower triangle of A is read. Eigenvalues
// are returned ascending with the columns of V in the same order (the caller's A
// buffer serves as the staging copy for that permutation).
#include "batched.hpp"

#include <cfloat>
#include <cmath>

namespace cuev {
namespace {

constexpr int kJThreads = 256;
constexpr int kMaxSweeps = 40;
constexpr size_t kLdsMatBytes = 128 * 1024;

template <typename T> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr double v = FLT_EPSILON; };

__device__ __forceinline__ void round_pair(int k, int r, int m, int &p, int &q) {
    if (k == 0) {
        p = m - 1;
        q = r;
    } else {
        const int mm = m - 1;
        p = (r + k) % mm;
        q = (r - k + mm) % mm;
    }
}

// block-wide sum of one double per thread (256 threads); every thread gets it
__device__ double block_sum(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < kJThreads / 64; ++i) t += red[i];
    return t;
}

template <typename T, bool IN_LDS>
__global__ __launch_bounds__(kJThreads) void jacobi_batched_kernel(T *Ag, int n, T *W, T *V, int *nconv) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int m = n + (n & 1);
    const int np = m / 2;
    const i64 mat = (i64)n * n;
    T *Aglob = Ag + (i64)blockIdx.x * mat;
    T *Vg = V + (i64)blockIdx.x * mat;
    // LDS: [A (m x m, only when IN_LDS)] [c, s, t: np each (double)] [p, q: np each] [red: 4]
    const size_t abytes = IN_LDS ? (size_t)m * m * sizeof(T) : 0;
    double *cs = reinterpret_cast<double *>(smem + abytes);
    double *ss = cs + np;
    double *ts = ss + np;
    double *red = ts + np;
    int *pp = reinterpret_cast<int *>(red + 4);
    int *qq = pp + np;
    T *A;
    i64 lda;
    if constexpr (IN_LDS) {
        A = reinterpret_cast<T *>(smem);
        lda = m;
    } else {
        A = Aglob;
        lda = n;
    }
    const int tid = threadIdx.x;
    auto at = [&](int i, int j) -> T & { return A[i + (i64)j * lda]; };

    // symmetric copy from the lower triangle, V = I, ||A||_F^2
    double fro2 = 0;
    for (i64 idx = tid; idx < mat; idx += kJThreads) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const T v = (i >= j) ? Aglob[i + (i64)j * n] : Aglob[j + (i64)i * n];
        if constexpr (IN_LDS) {
            A[i + (i64)j * lda] = v;
        }
        fro2 += (double)v * (double)v;
        Vg[idx] = (i == j) ? T(1) : T(0);
    }
    if constexpr (!IN_LDS) {
        // mirror the lower triangle into the upper one in place (each thread writes
        // only upper entries, read from the lower triangle nobody writes)
        __syncthreads();
        for (i64 idx = tid; idx < mat; idx += kJThreads) {
            const int i = (int)(idx % n), j = (int)(idx / n);
            if (i < j) Aglob[i + (i64)j * n] = Aglob[j + (i64)i * n];
        }
        __threadfence_block();
    }
    fro2 = block_sum(fro2, red);
    const double tol2 = (Eps<T>::v * Eps<T>::v) * fro2;
    const int nblk = np * (np + 1) / 2;

    int sweep = 0;
    bool converged = false;
    double prev = -1.0;
    for (; sweep < kMaxSweeps; ++sweep) {
        // off-diagonal norm of the current iterate
        double off = 0;
        for (i64 idx = tid; idx < mat; idx += kJThreads) {
            const int i = (int)(idx % n), j = (int)(idx / n);
            if (i != j) {
                const double v = (double)at(i, j);
                off += v * v;
            }
        }
        off = block_sum(off, red);
        // done at the rounding level: below eps ||A||_F, or stalled within 4 eps ||A||_F
        if (off <= tol2 || (off <= 16.0 * tol2 && prev >= 0.0 && off >= 0.25 * prev)) {
            converged = true;
            break;
        }
        prev = off;
        for (int r = 0; r < m - 1; ++r) {
            // rotation parameters of the step's pairs
            if (tid < np) {
                int p, q;
                round_pair(tid, r, m, p, q);
                double c = 1, s = 0, t = 0;
                if (p < n && q < n) {
                    const double b = (double)at(p, q);
                    if (b != 0.0) {
                        const double a = (double)at(p, p), d = (double)at(q, q);
                        const double theta = (d - a) / (2.0 * b);
                        t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(1.0 + t * t);
                        s = t * c;
                    }
                }
                cs[tid] = c;
                ss[tid] = s;
                ts[tid] = t;
                pp[tid] = p;
                qq[tid] = q;
            }
            __syncthreads();
            // A <- J^T A J over the 2x2 blocks (P, R), R <= P (block (R, P) is its transpose)
            for (int blk = tid; blk < nblk; blk += kJThreads) {
                // enumerate the block pairs R <= P of the np x np block grid
                int P = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
                while ((P + 1) * (P + 2) / 2 <= blk) ++P;
                while (P * (P + 1) / 2 > blk) --P;
                const int R = blk - P * (P + 1) / 2; // R <= P
                const int p = pp[P], q = qq[P], rr = pp[R], sq = qq[R];
                const double cP = cs[P], sP = ss[P];
                if (P == R) {
                    if (p < n && q < n && sP != 0.0) {
                        const double b = (double)at(p, q), t = ts[P];
                        at(p, p) = (T)((double)at(p, p) - t * b);
                        at(q, q) = (T)((double)at(q, q) + t * b);
                        at(p, q) = T(0);
                        at(q, p) = T(0);
                    }
                    continue;
                }
                const double cR = cs[R], sR = ss[R];
                if (sP == 0.0 && sR == 0.0) continue;
                // X = A[(p,q), (r,s)] with virtual index n reading as zero
                const bool vp = p < n, vq = q < n, vr = rr < n, vs = sq < n;
                const double x00 = (vp && vr) ? (double)at(p, rr) : 0.0;
                const double x01 = (vp && vs) ? (double)at(p, sq) : 0.0;
                const double x10 = (vq && vr) ? (double)at(q, rr) : 0.0;
                const double x11 = (vq && vs) ? (double)at(q, sq) : 0.0;
                // Y = X J_R, Z = J_P^T Y
                const double y00 = cR * x00 - sR * x01, y01 = sR * x00 + cR * x01;
                const double y10 = cR * x10 - sR * x11, y11 = sR * x10 + cR * x11;
                const double z00 = cP * y00 - sP * y10, z01 = cP * y01 - sP * y11;
                const double z10 = sP * y00 + cP * y10, z11 = sP * y01 + cP * y11;
                if (vp && vr) { at(p, rr) = (T)z00; at(rr, p) = (T)z00; }
                if (vp && vs) { at(p, sq) = (T)z01; at(sq, p) = (T)z01; }
                if (vq && vr) { at(q, rr) = (T)z10; at(rr, q) = (T)z10; }
                if (vq && vs) { at(q, sq) = (T)z11; at(sq, q) = (T)z11; }
            }
            // V <- V J: rows i, pairs P
            for (int idx = tid; idx < n * np; idx += kJThreads) {
                const int i = idx % n, P = idx / n;
                const double sP = ss[P];
                if (sP == 0.0) continue;
                const int p = pp[P], q = qq[P];
                if (p >= n || q >= n) continue;
                const double cP = cs[P];
                T *vp = Vg + i + (i64)p * n;
                T *vq = Vg + i + (i64)q * n;
                const double a = (double)*vp, b = (double)*vq;
                *vp = (T)(cP * a - sP * b);
                *vq = (T)(sP * a + cP * b);
            }
            if constexpr (!IN_LDS) __threadfence_block();
            __syncthreads();
        }
    }
    if (!converged && tid == 0) atomicAdd(nconv, 1);
    __syncthreads();
    // eigenvalues ascending (rank by value, ties by index); V columns follow,
    // staged through the caller's A buffer
    T *Wg = W + (i64)blockIdx.x * n;
    double *lamv = reinterpret_cast<double *>(smem + abytes); // reuse the parameter area (n <= 2 np)
    if constexpr (!IN_LDS) {
        // the diagonal is read before the staging copy overwrites A
        for (int i = tid; i < n; i += kJThreads) cs[i] = (double)at(i, i);
    } else {
        for (int i = tid; i < n; i += kJThreads) cs[i] = (double)at(i, i);
    }
    __syncthreads();
    for (i64 idx = tid; idx < mat; idx += kJThreads) Aglob[idx] = Vg[idx];
    __threadfence_block();
    __syncthreads();
    for (int i = tid; i < n; i += kJThreads) {
        const double li = lamv[i];
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const double lj = lamv[j];
            rank += (lj < li) || (lj == li && j < i);
        }
        Wg[rank] = (T)li;
        for (int k = 0; k < n; ++k) Vg[k + (i64)rank * n] = Aglob[k + (i64)i * n];
    }
}

} // namespace

template <typename T>
int eig_batched(hipStream_t s, T *A, int n, int batch, T *W, T *V, int *n_not_converged) {
    if (n < 1 || batch < 0 || (batch > 0 && (!A || !W || !V)))
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "eig_batched: n=%d batch=%d A=%p W=%p V=%p", n, batch, (void *)A,
                  (void *)W, (void *)V);
    if (n > kBatchedMaxN)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "eig_batched: n=%d above the batched limit %d", n, kBatchedMaxN);
    if (batch == 0) {
        if (n_not_converged) *n_not_converged = 0;
        return CUEV_STATUS_SUCCESS;
    }
    const int m = n + (n & 1), np = m / 2;
    const size_t params = (size_t)3 * np * 8 + 4 * 8 + (size_t)2 * np * 4;
    const size_t abytes = (size_t)m * m * sizeof(T);
    const bool in_lds = abytes <= kLdsMatBytes;
    const size_t lds = (in_lds ? abytes : 0) + params + 16;
    int *dcount = nullptr;
    CUEV_HIP(hipMallocAsync(reinterpret_cast<void **>(&dcount), sizeof(int), s));
    CUEV_HIP(hipMemsetAsync(dcount, 0, sizeof(int), s));
    if (in_lds) {
        CUEV_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(jacobi_batched_kernel<T, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL((jacobi_batched_kernel<T, true>), dim3((unsigned)batch), dim3(kJThreads), lds, s, A, n,
                           W, V, dcount);
    } else {
        hipLaunchKernelGGL((jacobi_batched_kernel<T, false>), dim3((unsigned)batch), dim3(kJThreads), lds, s, A, n,
                           W, V, dcount);
    }
    const hipError_t le = hipGetLastError();
    int h = 0;
    const hipError_t ce = hipMemcpyAsync(&h, dcount, sizeof(int), hipMemcpyDeviceToHost, s);
    const hipError_t fe = hipFreeAsync(dcount, s);
    const hipError_t se = hipStreamSynchronize(s);
    if (le != hipSuccess || ce != hipSuccess || fe != hipSuccess || se != hipSuccess)
        CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "eig_batched: %s", hipGetErrorString(le != hipSuccess ? le : ce != hipSuccess ? ce : fe != hipSuccess ? fe : se));
    if (n_not_converged) *n_not_converged = h;
    if (h > 0)
        CUEV_FAIL(CUEV_STATUS_NOT_CONVERGED, "eig_batched: %d of %d matrices did not converge in %d sweeps", h,
                  batch, kMaxSweeps);
    return CUEV_STATUS_SUCCESS;
}

template int eig_batched<float>(hipStream_t, float *, int, int, float *, float *, int *);
template int eig_batched<double>(hipStream_t, double *, int, int, double *, double *, int *);

} // namespace cuev
