// panel_qr.hip — Householder QR of one DBBR panel as a single persistent kernel.
//
// Replaces the reference's panel step dbbr_panel_qr (src/cuda/dbbr.cu:124-143):
// cuSOLVER geqrf + dbbr_extract_reflectors (:55-71) + syrk G = Y^T Y (:140) +
// dbbr_larft (:86-109), and additionally forms W = Y T (the reference's GEMM at
// src/cuda/dbbr.cu:171) for the rows each workgroup owns.
//
// Convention (LAPACK dgeqr2/dlarfg, which the reference obtains from geqrf):
// for column c with pivot alpha = A[c,c] and tail x = A[c+1:, c],
//   xnorm == 0  ->  tau = 0, H = I, the column is left as is;
//   otherwise   ->  beta = -sign(alpha) * sqrt(alpha^2 + xnorm^2),
//                   tau  = (beta - alpha) / beta,  v = [1; x / (alpha - beta)],
//                   A[c,c] = beta, A[c+1:, c] = v[1:].
// T is the upper-triangular factor with H_0 ... H_{b-1} = I - Y T Y^T, built by
// the recurrence T[j,j] = tau_j, T[0:j, j] = -tau_j T[0:j,0:j] (Y^T Y)[0:j, j].
//
// Parallel structure: G workgroups, each owns CH consecutive panel rows and
// keeps them in LDS for the whole factorization. Per column one grid-wide
// reduction carries ||x||^2, the b-c-1 dot products x^T a_j and the Gram
// column of the previous reflector; every workgroup sums the G partial vectors
// in the same order, so all of them hold bit-identical scalars. Norms, dot
// products, tau and T are FP64 for both float and double panels.
// The grid is sized to be co-resident (one workgroup per CU at this LDS size);
// every wait is bounded and reports through *err.
#include "panel_qr.hpp"

#include <algorithm>

namespace cuev {
namespace {

constexpr int kPQThreads = 256;

template <typename T> struct PQArgs {
    T *A;
    i64 lda;
    T *Y;
    i64 ldy;
    T *W;
    i64 ldw;
    T *Tm;
    T *tau;
    int rows, b, CH, G;
    double *part;   // [2][G][2b]
    double *pivot;  // [2][b]
    unsigned *counter;
    int *err;
};

// Grid barrier: all G workgroups arrive (counter reaches `target`), then all leave.
__device__ __forceinline__ bool grid_barrier(unsigned *counter, unsigned target, int *err) {
    wg_publish_fence();
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        counter_add_release(counter, 1u);
        long long it = 0;
        int ok = 1;
        while (counter_load_relaxed(counter) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++it > kSpinLimit) {
                __hip_atomic_store(err, (int)DEV_WATCHDOG, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <typename T> __global__ __launch_bounds__(kPQThreads) void panel_qr_kernel(PQArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int b = a.b, CH = a.CH, CHp = CH + 1;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = blockIdx.x;

    // LDS layout
    double *GT = reinterpret_cast<double *>(smem_raw);   // [b][b+1]: Gram, then T (column-major)
    double *sums = GT + (size_t)b * (b + 1);             // [2b] reduced payload
    double *tauv = sums + 2 * b;                         // [b]
    double *prow = tauv + b;                             // [b] pivot row
    double *red4 = prow + b;                             // [256] partial-sum combine buffer
    double *red4b = red4 + 256;                          // [256] second combine buffer
    double *sx = red4b + 256;                            // [CH] x = column c below the diagonal
    double *sv = sx + CH;                                // [CH] reflector c-1 (unit at row c-1)
    double *sjv = sv + CH;                               // [b] tau * v^T a_j of the trailing columns
    T *P = reinterpret_cast<T *>(sjv + b);               // [b][CHp] panel rows of this workgroup

    const int r0g = g * CH;                              // first global panel row owned
    const int chl = max(0, min(CH, a.rows - r0g));       // rows owned
    const int PS = 2 * b;                                // payload stride

    // load the owned rows
    for (int idx = tid; idx < b * CH; idx += kPQThreads) {
        const int j = idx / CH, rl = idx - j * CH;
        P[j * CHp + rl] = (rl < chl) ? a.A[(i64)(r0g + rl) + (i64)j * a.lda] : T(0);
    }
    for (int idx = tid; idx < b * (b + 1); idx += kPQThreads) GT[idx] = 0.0;
    __syncthreads();

    bool alive = true;
    for (int c = 0; c <= b; ++c) {
        const int buf = c & 1;
        const bool has_col = c < b && c < a.rows;
        const int ndot = has_col ? b - c : 0;           // q = 0: ||x||^2, q >= 1: x^T a_{c+q}
        const int ngram = (c >= 2 && c - 1 < a.rows) ? c - 1 : 0; // v_i^T v_{c-1}, i < c-1

        // (1) local partial sums: x^T a_j for j >= c (j = c: ||x||^2) and the Gram
        // column v_i^T v_{c-1}, i < c-1, from masked copies of x and v_{c-1}. Each
        // thread owns one column q and a quarter of the rows; no wave reductions.
        double *mypart = a.part + ((size_t)buf * a.G + g) * PS;
        const int nq = ndot + ngram;
        for (int rl = tid; rl < CH; rl += kPQThreads) {
            const int r = r0g + rl;
            const bool own = rl < chl;
            sx[rl] = (own && has_col && r > c) ? (double)P[c * CHp + rl] : 0.0;
            double v = 0.0;
            if (own && ngram) v = (r > c - 1) ? (double)P[(c - 1) * CHp + rl] : (r == c - 1 ? 1.0 : 0.0);
            sv[rl] = v;
        }
        __syncthreads();
        {
            const int qc = tid & 63, quarter = tid >> 6;
            const int QR = (CH + 3) >> 2;
            const int rb = quarter * QR, re = min(chl, rb + QR);
            double s1 = 0.0, s2 = 0.0;
            if (qc < b) {
                const T *pc = P + (size_t)qc * CHp;
                for (int rl = rb; rl < re; ++rl) {
                    const double pv = (double)pc[rl];
                    s1 += pv * sx[rl];
                    s2 += pv * sv[rl];
                }
            }
            red4[quarter * 64 + qc] = s1;
            red4b[quarter * 64 + qc] = s2;
        }
        __syncthreads();
        if (tid < b) {
            const int qc = tid;
            if (qc >= c && ndot)
                mypart[qc - c] = ((red4[qc] + red4[64 + qc]) + red4[128 + qc]) + red4[192 + qc];
            if (qc < ngram)
                mypart[b + qc] = ((red4b[qc] + red4b[64 + qc]) + red4b[128 + qc]) + red4b[192 + qc];
        }
        // pivot row A[c, c:b] from its owner
        if (has_col && c >= r0g && c < r0g + chl) {
            const int cl = c - r0g;
            for (int j = c + tid; j < b; j += kPQThreads)
                a.pivot[buf * b + j] = (double)P[j * CHp + cl];
        }

        // (2) grid-wide reduction
        if (!grid_barrier(a.counter, (unsigned)a.G * (unsigned)(c + 1), a.err)) alive = false;

        // Sum the G partial vectors: 4 threads per payload entry, each taking every
        // 4th workgroup with its loads issued back to back, then a fixed-order
        // combine, so every workgroup obtains bit-identical sums.
        {
            const int q = tid & 63, part = tid >> 6;
            double s = 0.0;
            if (q < nq) {
                const int slot = (q < ndot) ? q : b + (q - ndot);
                const double *src = a.part + (size_t)buf * a.G * PS + slot;
                // batches of 32 independent loads (one memory round trip per batch)
                for (int g0 = part; g0 < a.G; g0 += 4 * 32) {
                    double v[32];
#pragma unroll
                    for (int u = 0; u < 32; ++u) {
                        const int gg = g0 + 4 * u;
                        v[u] = (gg < a.G) ? src[(size_t)gg * PS] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 32; ++u) s += v[u];
                }
            }
            red4[part * 64 + q] = s;
        }
        __syncthreads();
        if (tid < nq) {
            const int slot = (tid < ndot) ? tid : b + (tid - ndot);
            sums[slot] = ((red4[tid] + red4[64 + tid]) + red4[128 + tid]) + red4[192 + tid];
        }
        if (has_col)
            for (int j = c + tid; j < b; j += kPQThreads) prow[j] = a.pivot[buf * b + j];
        __syncthreads();

        // Gram column c-1
        if (ngram)
            for (int i = tid; i < ngram; i += kPQThreads) GT[i + (size_t)(c - 1) * (b + 1)] = sums[b + i];
        if (!has_col) {
            if (c < b && tid == 0) tauv[c] = 0.0;
            __syncthreads();
            continue;
        }

        // (3) reflector scalars (identical in every workgroup)
        const double alpha = prow[c];
        const double xnorm2 = sums[0];
        double beta, tau, scal;
        if (xnorm2 == 0.0) {
            beta = alpha;
            tau = 0.0;
            scal = 0.0;
        } else {
            const double nrm = sqrt(alpha * alpha + xnorm2);
            beta = (alpha >= 0.0) ? -nrm : nrm;
            tau = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        if (tid == 0) tauv[c] = tau;

        // (4) apply H = I - tau v v^T to columns c+1..b-1 of the owned rows:
        // a_j -= (tau v^T a_j) v with v^T a_j = a_j[c] + scal * (x^T a_j); each wave
        // takes every 4th column, each lane fixed rows (no index division)
        const int ntrail = b - c - 1;
        if (tau != 0.0) {
            for (int jj = tid; jj < ntrail; jj += kPQThreads)
                sjv[jj] = tau * (prow[c + 1 + jj] + scal * sums[1 + jj]);
            __syncthreads();
            for (int rl = lane; rl < chl; rl += 64) {
                const int r = r0g + rl;
                if (r < c) continue;
                const double vr = (r == c) ? 1.0 : (double)P[c * CHp + rl] * scal;
                for (int jj = wave; jj < ntrail; jj += kPQThreads / 64) {
                    T &e = P[(c + 1 + jj) * CHp + rl];
                    e = (T)((double)e - sjv[jj] * vr);
                }
            }
        }
        __syncthreads();
        for (int rl = tid; rl < chl; rl += kPQThreads) {
            const int r = r0g + rl;
            if (r == c) P[c * CHp + rl] = (T)beta;
            else if (r > c) P[c * CHp + rl] = (T)((double)P[c * CHp + rl] * scal);
        }
        __syncthreads();
    }
    (void)alive;

    // (5) T from the Gram matrix: column j overwrites Gram column j
    for (int j = 0; j < b; ++j) {
        double t = 0.0;
        if (tid < j) {
            for (int q = tid; q < j; ++q) t += GT[tid + (size_t)q * (b + 1)] * GT[q + (size_t)j * (b + 1)];
            t *= -tauv[j];
        }
        __syncthreads();
        if (tid < j) GT[tid + (size_t)j * (b + 1)] = t;
        if (tid == j) GT[j + (size_t)j * (b + 1)] = tauv[j];
        __syncthreads();
    }
    for (int idx = tid; idx < b * b; idx += kPQThreads) {
        const int i = idx % b, j = idx / b;
        if (i > j) GT[i + (size_t)j * (b + 1)] = 0.0;
    }
    __syncthreads();

    // (6) outputs for the owned rows: A (R + reflector tails), Y (explicit), W = Y T
    for (int idx = tid; idx < b * chl; idx += kPQThreads) {
        const int j = idx / chl, rl = idx - j * chl;
        const int r = r0g + rl;
        const T pv = P[j * CHp + rl];
        a.A[(i64)r + (i64)j * a.lda] = pv;
        if (a.Y) a.Y[(i64)r + (i64)j * a.ldy] = (r > j) ? pv : (r == j ? T(1) : T(0));
    }
    if (a.W) {
        for (int idx = tid; idx < b * chl; idx += kPQThreads) {
            const int j = idx / chl, rl = idx - j * chl;
            const int r = r0g + rl;
            double s = 0.0;
            const int imax = min(j, r); // Y[r,i] = 0 for i > r
            for (int i = 0; i <= imax; ++i) {
                const double y = (r > i) ? (double)P[i * CHp + rl] : 1.0;
                s += y * GT[i + (size_t)j * (b + 1)];
            }
            a.W[(i64)r + (i64)j * a.ldw] = (T)s;
        }
    }
    if (g == 0) {
        for (int idx = tid; idx < b * b; idx += kPQThreads) {
            const int i = idx % b, j = idx / b;
            a.Tm[idx] = (T)GT[i + (size_t)j * (b + 1)];
        }
        for (int j = tid; j < b; j += kPQThreads) a.tau[j] = (T)tauv[j];
    }
}

__global__ void zero_u32_kernel(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template <typename T> size_t pq_lds_bytes(int b, int CH) {
    return sizeof(double) * ((size_t)b * (b + 1) + 2 * b + b + b + 512 + 2 * (size_t)CH + b) +
           sizeof(T) * (size_t)b * (CH + 1);
}

} // namespace

template <typename T> int panel_qr_chunk(int b) {
    // largest multiple of 64 rows whose panel slice fits in 160 KB of LDS
    const size_t budget = 160 * 1024;
    int ch = 64;
    while (ch + 64 <= 1024 && pq_lds_bytes<T>(b, ch + 64) <= budget) ch += 64;
    return ch;
}

size_t panel_qr_scratch_bytes(int n, int b) {
    // partials [2][G][2b] doubles + pivot [2][b] doubles, G <= n / 64 + 1
    const size_t G = (size_t)n / 64 + 2;
    return sizeof(double) * (2 * G * 2 * b + 2 * b) + 256;
}

template <typename T>
int panel_qr(hipStream_t s, T *A, i64 lda, T *Y, i64 ldy, T *W, i64 ldw, T *Tm, T *tau, int rows,
             int b, void *scratch, size_t scratch_bytes, unsigned *counter, int *err) {
    if (rows < 1 || b < 1 || b > 64)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "panel_qr: rows=%d b=%d", rows, b);
    const int CH = panel_qr_chunk<T>(b);
    const int G = (int)ceil_div(rows, CH);
    const size_t lds = pq_lds_bytes<T>(b, CH);
    static_assert(kPQThreads == 256, "");
    CUEV_HIP(hipFuncSetAttribute((const void *)panel_qr_kernel<T>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, panel_qr_kernel<T>, kPQThreads,
                                                          lds));
    const int resident = per_cu * device_cu_count();
    if (G > resident)
        CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED,
                  "panel_qr: %d rows need %d co-resident workgroups, device holds %d", rows, G,
                  resident);
    const size_t need = sizeof(double) * ((size_t)2 * G * 2 * b + 2 * b);
    if (need > scratch_bytes)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "panel_qr: scratch too small (%zu < %zu)",
                  scratch_bytes, need);
    PQArgs<T> a;
    a.A = A;
    a.lda = lda;
    a.Y = Y;
    a.ldy = ldy;
    a.W = W;
    a.ldw = ldw;
    a.Tm = Tm;
    a.tau = tau;
    a.rows = rows;
    a.b = b;
    a.CH = CH;
    a.G = G;
    a.part = static_cast<double *>(scratch);
    a.pivot = a.part + (size_t)2 * G * 2 * b;
    a.counter = counter;
    a.err = err;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, counter, 1);
    CUEV_LAUNCH_CHECK();
    hipLaunchKernelGGL(panel_qr_kernel<T>, dim3(G), dim3(kPQThreads), lds, s, a);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template int panel_qr_chunk<float>(int);
template int panel_qr_chunk<double>(int);
template int panel_qr<float>(hipStream_t, float *, i64, float *, i64, float *, i64, float *,
                             float *, int, int, void *, size_t, unsigned *, int *);
template int panel_qr<double>(hipStream_t, double *, i64, double *, i64, double *, i64, double *,
                              double *, int, int, void *, size_t, unsigned *, int *);

} // namespace cuev
