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
constexpr size_t kPQGateBytes = 256; // scratch head: the CholQR -> Householder gate word

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
    const int *gate; // when set: run only if *gate != 0 (the CholQR kernel declined the panel)
    unsigned *counter;
    unsigned *reset; // the CholQR kernel's barrier counter, zeroed for the next panel
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
        fence_acquire_agent();
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// Write-through exchange for the CholQR kernel: every global access to data that
// another workgroup of the same launch reads (Gram partials, their sums, U and Mw)
// is an agent-scope relaxed atomic access. On gfx950 these carry the sc1 bit: the
// store is written through past the XCD-private L2 and the load reads from the
// device-coherent level. A grid barrier then needs neither an L2 write-back before
// the arrival nor an L2 invalidate after the wait: each wave waits until its own
// stores have completed (vmcnt(0)), the workgroup meets, one thread increments the
// counter and polls it, and the workgroup meets again.
__device__ __forceinline__ double ld_wt(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_wt(double *p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool grid_barrier_wt(unsigned *counter, unsigned target, int *err) {
    vm_drain();
    __syncthreads();
    __shared__ int s_ok;
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <typename T> __global__ __launch_bounds__(kPQThreads) void panel_qr_kernel(PQArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the CholQR kernel before this one has finished: its counter is free
    if (blockIdx.x == 0 && threadIdx.x == 0)
        __hip_atomic_store(a.reset, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a.gate && *a.gate == 0) return;
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

// ---------------------------------------------------------------------------
// CholQR2 panel factorization with Householder reconstruction (the fast path).
//
// Two passes of Cholesky QR make an orthonormal Q with A = Q R (each pass: the
// Gram matrix P^T P of the panel, its Cholesky factor R_p and P <- P R_p^-1, so
// R = R_2 R_1). The Householder form of the same factorization follows from an
// LU factorization without pivoting of the top b x b block of Q with signs
// S = diag(s_k), s_k = -sign of the k-th pivot before the shift:
//   Q_1 - S = L U,  Y = [L; Q_2 U^-1],  T = -U S L^-T,  R_house = S R,
// so that (I - Y T Y^T)[:, :b] = Q S (Ballard, Demmel, Grigori, Jacquelin, Knight,
// Nguyen, "Reconstructing Householder vectors from tall-skinny QR", J. Parallel
// Distrib. Comput. 85, 2015). In exact arithmetic these are the reflectors,
// T and R of Householder QR on the panel itself, the LAPACK convention above.
// CholQR2 keeps Q orthonormal to working precision when cond(panel) is well
// below eps^-1/2; a panel whose first Gram pivots reveal a (nearly) dependent
// column (d_k <= 1e-10 G_kk, or not finite) is left untouched and flagged for
// the Householder kernel, which then runs in full (*gate = 1).
//
// Parallel structure: G workgroups (one per CU), each with CH rows in LDS.
// Per pass: partial Gram -> global, grid barrier, each workgroup sums a slice of
// the entries over all partials in a fixed order, grid barrier, then every
// workgroup factors the same b x b Gram matrix redundantly (bit-identical
// results everywhere), with X = L'^-1 alongside for P <- P R_p^-1. Workgroup 0 then
// factors Q_1 and publishes U and Mw = -S L^-T (fifth barrier): every workgroup forms
// its rows of W = Y T = (Q - [S; 0]) Mw and of Y = Q_2 U^-1 (by substitution along
// its rows, four lanes per row), and workgroup 0 forms T = U Mw. The b x b factorizations keep the matrix in registers as 4x4 blocks
// per thread and broadcast the pivot column/row through LDS: one LDS barrier per
// elimination step. All arithmetic is FP64.
constexpr int kCQThreads = 256;
constexpr int kCQMaxCH = 96;

template <typename T> struct CQArgs {
    T *A;
    i64 lda;
    T *Y;
    i64 ldy;
    T *W;
    i64 ldw;
    T *Tm;
    T *tau;
    int rows, CH, G;
    double *part; // [G][b*b] partial Gram matrices
    double *gsum; // [b*b]
    double *bc;   // [2*b*b]: U, then Mw = -S L^-T (column-major)
    int *gate;
    unsigned *counter;
    unsigned *reset; // the Householder kernel's barrier counter, zeroed for the launch after this one
    int *err;
};

// Packed lower-triangle index of (i, j), i >= j, of a B x B symmetric matrix.
template <int B> __device__ __forceinline__ int cq_lo(int i, int j) { return i + j * B - (j * (j + 1)) / 2; }

// out (packed lower) = P^T P over the first nr rows (P column-major, ld CHp).
// Thread = strided 4x4 block, rows ti + NB u and columns tj + NB v, so the 16
// lanes with different ti read 16 different banks.
template <int B>
__device__ __forceinline__ void cq_gram(const double *P, int CHp, int nr, double *out, int tid) {
    constexpr int NB = B / 4;
    if (tid >= NB * NB) return;
    const int ti = tid % NB, tj = tid / NB;
    double acc[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int r = 0; r < nr; ++r) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = P[(size_t)(ti + NB * u) * CHp + r];
            y[u] = P[(size_t)(tj + NB * u) * CHp + r];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * y[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = ti + NB * u, j = tj + NB * v;
            if (i >= j) st_wt(out + cq_lo<B>(i, j), acc[u][v]);
        }
}

// acc = X Y for B x B LDS matrices (column-major, ld B + 1; XT: X transposed),
// thread = strided 4x4 block (ti + NB u, tj + NB v).
template <int B, bool XT>
__device__ __forceinline__ void cq_mm(double (&acc)[4][4], const double *X, const double *Y, int ti, int tj) {
    constexpr int NB = B / 4, LD = B + 1;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
    for (int k = 0; k < B; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = ti + NB * u;
            x[u] = XT ? X[k + i * LD] : X[i + k * LD];
            y[u] = Y[k + (tj + NB * u) * LD];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * y[v];
    }
}

// P[rows r0..r1) <- P U for an upper-triangular U (LDS, column-major, ld B + 1), in
// place: rounds of whole 4-row blocks, all column blocks of a row block in the
// same round (a row's new values depend on that row only). Columns are strided
// over threads (tj + NB v) so the U reads of a wave spread over the banks.
template <int B>
__device__ __forceinline__ void cq_trmm(double *P, int CHp, int r0, int r1, const double *U, int tid) {
    constexpr int NB = B / 4, LD = B + 1;
    constexpr int RB_PER_ROUND = kCQThreads / NB;
    const int nrb = (r1 - r0 + 3) / 4;
    for (int rb0 = 0; rb0 < nrb; rb0 += RB_PER_ROUND) {
        const int rb = rb0 + tid / NB, tj = tid % NB;
        const bool act = rb < nrb;
        double acc[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
        const int rr = r0 + 4 * rb;
        if (act) {
            const int imax = tj + NB * 3; // U(i, j) = 0 for i > j
            for (int i = 0; i <= imax; ++i) {
                double x[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = P[(size_t)i * CHp + rr + u];
#pragma unroll
                for (int v = 0; v < 4; ++v) w[v] = U[i + (tj + NB * v) * LD];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * w[v];
            }
        }
        __syncthreads();
        if (act)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (rr + u < r1) P[(size_t)(tj + NB * v) * CHp + rr + u] = acc[u][v];
        __syncthreads();
    }
}

// P[r, :] <- P[r, :] U^-1 for the rows r0 <= r < r1, in place: forward substitution along
// each row, four lanes per row (a quad of one wave) and 64 rows per round. The columns go
// in blocks of 16, lane q of a quad holding columns 16J + 4q .. 16J + 4q + 3 of block J:
// the block's columns first take the contributions of the finished blocks (read back from
// P), then are finished one by one, each broadcast within the quad by a lane shuffle. U is
// upper triangular in LDS (column-major, ld B + 1) and rd[j] = 1 / U(j, j). A row's values
// are written and read back by its own quad only, so no workgroup barrier is needed until
// the end.
template <int B>
__device__ __forceinline__ void cq_rsolve(double *P, int CHp, int r0, int r1, const double *U, const double *rd,
                                          int tid) {
    static_assert(B % 16 == 0 && kCQThreads == 256, "row solve layout");
    constexpr int LD = B + 1;
    const int q = tid & 3, lane = tid & 63;
    for (int rb = r0; rb < r1; rb += kCQThreads / 4) {
        const int r = rb + (tid >> 2);
        const bool act = r < r1;
#pragma unroll 1
        for (int J = 0; J < B / 16; ++J) {
            const int jb = 16 * J + 4 * q; // this lane's first column of the block
            double acc[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] = act ? P[(size_t)(jb + c) * CHp + r] : 0.0;
            for (int k = 0; k < 16 * J; ++k) {
                const double y = act ? P[(size_t)k * CHp + r] : 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[c] -= y * U[k + (jb + c) * LD];
            }
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int qt = t >> 2, ct = t & 3, kt = 16 * J + t;
                const double mine = (q == qt) ? acc[ct] * rd[kt] : 0.0;
                const double y = __shfl(mine, (lane & ~3) + qt, 64);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (4 * q + c > t) acc[c] -= y * U[kt + (jb + c) * LD];
                if (q == qt && act) P[(size_t)kt * CHp + r] = y;
            }
        }
    }
    __syncthreads();
}

template <typename T, int B> __global__ __launch_bounds__(kCQThreads) void panel_cholqr_kernel(CQArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int NB = B / 4, LD = B + 1, BB = B * B, BL = B * (B + 1) / 2;
    const int CH = a.CH, CHp = CH + 1;
    const int tid = threadIdx.x, g = blockIdx.x;
    const bool own = tid < NB * NB;
    // contiguous 4x4 blocks (4 ti + u, 4 tj + v) for the factorizations, strided
    // (si + NB u, sj + NB v) for the products
    const int ti = own ? tid % NB : 0, tj = own ? tid / NB : 0;
    const int si = ti, sj = tj;
    double *P = reinterpret_cast<double *>(smem_raw); // [B][CHp]: panel(r, c) at P[c*CHp + r]
    double *M1 = P + (size_t)B * CHp;                 // [B][LD] column-major
    double *M2 = M1 + B * LD;
    double *M3 = M2 + B * LD;
    double *cb = M3 + B * LD; // [2][B] broadcast: pivot column
    double *rb = cb + 2 * B;  // [2][B] broadcast: pivot row
    double *xb = rb + 2 * B;  // [2][B] broadcast: row of the inverse
    double *dv = xb + 2 * B;  // [B] pivots
    double *gd = dv + B;      // [B] Gram diagonal before the factorization
    double *sv = gd + B;      // [B] signs S
    double *red = sv + B;     // [kCQThreads]

    if (g == 0 && tid == 0) __hip_atomic_store(a.reset, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int r0g = g * CH;
    const int chl = max(0, min(CH, a.rows - r0g));
    {
        // the panel slice: rounds of LB loads per thread, all issued before their LDS stores
        constexpr int LB = 8;
        for (int base = 0; base < B * CH; base += LB * kCQThreads) {
            double v[LB];
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const int idx = base + q * kCQThreads + tid;
                const int j = idx / CH, rl = idx - j * CH;
                v[q] = (idx < B * CH && rl < chl) ? (double)a.A[(i64)(r0g + rl) + (i64)j * a.lda] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const int idx = base + q * kCQThreads + tid;
                const int j = idx / CH, rl = idx - j * CH;
                if (idx < B * CH) P[(size_t)j * CHp + rl] = v[q];
            }
        }
    }
    __syncthreads();

    unsigned nbar = 0;
    for (int pass = 0; pass < 2; ++pass) {
        // ---- Gram matrix of the panel: partials, then fixed-order slice sums
        cq_gram<B>(P, CHp, chl, a.part + (size_t)g * BL, tid);
        if (!grid_barrier_wt(a.counter, (unsigned)a.G * ++nbar, a.err)) return;
        {
            const int S = (BL + a.G - 1) / a.G;
            const int e0 = g * S, ne = max(0, min(BL, e0 + S) - e0);
            int tpe = 1; // threads per entry (power of two)
            while (tpe < 64 && tpe * 2 * max(ne, 1) <= kCQThreads) tpe *= 2;
            const int per_round = kCQThreads / tpe;
            for (int base = 0; base < ne; base += per_round) {
                const int el = base + tid / tpe, pt = tid % tpe;
                double s = 0.0;
                if (el < ne) {
                    const double *src = a.part + e0 + el;
                    for (int g0 = pt; g0 < a.G; g0 += tpe * 16) {
                        double v[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u) {
                            const int gg = g0 + u * tpe;
                            v[u] = (gg < a.G) ? ld_wt(src + (size_t)gg * BL) : 0.0;
                        }
#pragma unroll
                        for (int u = 0; u < 16; ++u) s += v[u];
                    }
                }
                red[tid] = s;
                __syncthreads();
                if (el < ne && pt == 0) {
                    double t = 0.0;
                    for (int q = 0; q < tpe; ++q) t += red[tid + q];
                    st_wt(a.gsum + e0 + el, t);
                }
                __syncthreads();
            }
        }
        if (!grid_barrier_wt(a.counter, (unsigned)a.G * ++nbar, a.err)) return;

        // ---- second pass, G = I + E with E tiny (the usual case): R_2 = I + F,
        // F = up(E - up(E)^T up(E)) with up(M) = strict upper of M + diag(M)/2,
        // leaves R_2^T R_2 - G = O(|E|^3); R_2^-1 = I - F + F^2 + O(|F|^3)
        if (pass == 1) {
            double e[4][4];
            double emax = 0.0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = si + NB * u, j = sj + NB * v;
                    const double gij = own ? ld_wt(a.gsum + ((i >= j) ? cq_lo<B>(i, j) : cq_lo<B>(j, i))) : 0.0;
                    e[u][v] = own ? gij - (i == j ? 1.0 : 0.0) : 0.0;
                    emax = fmax(emax, fabs(e[u][v]));
                    if (own) M2[i + j * LD] = (i < j) ? e[u][v] : (i == j ? 0.5 * e[u][v] : 0.0);
                }
            red[tid] = (emax == emax) ? emax : 1.0; // NaN -> not small
            __syncthreads();
            for (int s2 = kCQThreads / 2; s2 > 0; s2 >>= 1) {
                if (tid < s2) red[tid] = fmax(red[tid], red[tid + s2]);
                __syncthreads();
            }
            // |E| <= 1e-10 entrywise (the usual case): up(E)^T up(E) and F^2 have entries
            // below B 1e-20, under half an ulp of R_2's unit diagonal, so F = up(E) and
            // R_2^-1 = I - F, and only F R_1 is formed
            const bool small = red[0] <= 1e-6, tiny = red[0] <= 1e-10;
            __syncthreads();
            if (small) {
                if (!tiny) {
                    double c[4][4];
                    cq_mm<B, true>(c, M2, M2, si, sj); // up(E)^T up(E)
                    __syncthreads();
                    if (own)
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int i = si + NB * u, j = sj + NB * v;
                                const double f = e[u][v] - c[u][v];
                                M2[i + j * LD] = (i < j) ? f : (i == j ? 0.5 * f : 0.0);
                            }
                    __syncthreads();
                }
                double f2[4][4], fr[4][4];
                if (!tiny) {
                    cq_mm<B, false>(f2, M2, M2, si, sj); // F^2
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) f2[u][v] = 0.0;
                }
                cq_mm<B, false>(fr, M2, M1, si, sj); // F R_1
                double fo[4][4], r1[4][4];
                if (own)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int i = si + NB * u, j = sj + NB * v;
                            fo[u][v] = M2[i + j * LD];
                            r1[u][v] = M1[i + j * LD];
                        }
                __syncthreads();
                if (own)
#pragma unroll
                    for (int u = 0; u < 4; ++u)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int i = si + NB * u, j = sj + NB * v;
                            M2[i + j * LD] = (i == j ? 1.0 : 0.0) - fo[u][v] + f2[u][v]; // R_2^-1
                            M1[i + j * LD] = r1[u][v] + fr[u][v];                         // R_2 R_1
                        }
                __syncthreads();
                cq_trmm<B>(P, CHp, 0, CH, M2, tid);
                continue;
            }
        }

        // ---- G = L' D L'^T with X = L'^-1 (registers), R_p = D^1/2 L'^T. Each step
        // applies masked multiplier vectors to all 16 entries of a thread's block.
        double gr[4][4], xr[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = 4 * ti + u, j = 4 * tj + v;
                gr[u][v] = own ? ld_wt(a.gsum + ((i >= j) ? cq_lo<B>(i, j) : cq_lo<B>(j, i))) : 0.0;
                xr[u][v] = (i == j) ? 1.0 : 0.0;
            }
        double *Rout = (pass == 0) ? M1 : M3;
        if (tid < B) gd[tid] = ld_wt(a.gsum + cq_lo<B>(tid, tid));
        if (own && tj == 0)
#pragma unroll
            for (int u = 0; u < 4; ++u) cb[4 * ti + u] = gr[u][0];
        if (own && ti == 0)
#pragma unroll
            for (int v = 0; v < 4; ++v) xb[4 * tj + v] = xr[0][v];
        const double tol = (pass == 0) ? 1e-10 : 1e-3;
        bool ok = true;
        for (int kb = 0; kb < NB && ok; ++kb) {
#pragma unroll
            for (int v0 = 0; v0 < 4; ++v0) {
                const int k = 4 * kb + v0, p = k & 1;
                __syncthreads();
                const double dk = cb[p * B + k];
                if (!(dk > tol * gd[k]) || !(dk < 1e300)) {
                    ok = false;
                    break;
                }
                const double rk = 1.0 / dk, sq = sqrt(dk), rs = 1.0 / sq;
                double ci[4], cim[4], cjm[4], xk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    ci[u] = cb[p * B + 4 * ti + u];
                    cim[u] = (4 * ti + u > k) ? ci[u] : 0.0;
                    const double cj = cb[p * B + 4 * tj + u];
                    cjm[u] = (4 * tj + u > k) ? cj * rk : 0.0;
                    xk[u] = xb[p * B + 4 * tj + u];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double lx = cim[u] * rk;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        gr[u][v] -= cim[u] * cjm[v];
                        xr[u][v] -= lx * xk[v];
                    }
                }
                if (tid == 0) dv[k] = dk;
                if (own && tj == 0)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = 4 * ti + u;
                        Rout[k + i * LD] = (i > k) ? ci[u] * rs : (i == k ? sq : 0.0);
                    }
                const int kn = k + 1, vn = (v0 + 1) & 3, bn = kn >> 2;
                if (kn < B) {
                    if (own && tj == bn)
#pragma unroll
                        for (int u = 0; u < 4; ++u) cb[(kn & 1) * B + 4 * ti + u] = gr[u][vn];
                    if (own && ti == bn)
#pragma unroll
                        for (int v = 0; v < 4; ++v) xb[(kn & 1) * B + 4 * tj + v] = xr[vn][v];
                }
            }
        }
        if (!ok) {
            if (g == 0 && tid == 0) *a.gate = 1;
            return; // every workgroup takes the same decision
        }
        // R_p^-1 (i, j) = X(j, i) / sqrt(d_j) into M2
        __syncthreads();
        if (own)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    M2[(4 * tj + v) + (4 * ti + u) * LD] = xr[u][v] / sqrt(dv[4 * ti + u]);
        __syncthreads();
        cq_trmm<B>(P, CHp, 0, CH, M2, tid);
        if (pass == 1) {
            // R = R_2 R_1 into M1
            double acc[4][4];
            cq_mm<B, false>(acc, M3, M1, si, sj);
            __syncthreads();
            if (own)
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) M1[(si + NB * u) + (sj + NB * v) * LD] = acc[u][v];
            __syncthreads();
        }
    }

    // ---- workgroup 0: Q_1 - S = L U (signs chosen on the way), L^-1, T, U^-1
    if (g == 0) {
        double mr[4][4], xr[4][4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                mr[u][v] = own ? P[(size_t)(4 * tj + v) * CHp + 4 * ti + u] : 0.0;
                xr[u][v] = (4 * ti + u == 4 * tj + v) ? 1.0 : 0.0;
            }
        if (own && tj == 0)
#pragma unroll
            for (int u = 0; u < 4; ++u) cb[4 * ti + u] = mr[u][0];
        if (own && ti == 0)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                rb[4 * tj + v] = mr[0][v];
                xb[4 * tj + v] = xr[0][v];
            }
        double *Lm = M2, *Um = M3; // unit-lower L, upper U (column-major)
        for (int kb = 0; kb < NB; ++kb) {
#pragma unroll
            for (int v0 = 0; v0 < 4; ++v0) {
                const int k = 4 * kb + v0, p = k & 1;
                __syncthreads();
                const double mkk = cb[p * B + k];
                const double s = (mkk >= 0.0) ? -1.0 : 1.0;
                const double piv = mkk - s, rp = 1.0 / piv;
                double li[4], lim[4], rj[4], rjm[4], xk[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    li[u] = cb[p * B + 4 * ti + u] * rp;
                    lim[u] = (4 * ti + u > k) ? li[u] : 0.0;
                    rj[u] = rb[p * B + 4 * tj + u];
                    rjm[u] = (4 * tj + u > k) ? rj[u] : 0.0;
                    xk[u] = xb[p * B + 4 * tj + u]; // X(k, j) = 0 for j > k
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        mr[u][v] -= lim[u] * rjm[v];
                        xr[u][v] -= lim[u] * xk[v];
                    }
                if (tid == 0) {
                    sv[k] = s;
                    dv[k] = piv;
                }
                if (own && tj == 0)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int i = 4 * ti + u;
                        Lm[i + k * LD] = (i > k) ? li[u] : (i == k ? 1.0 : 0.0);
                    }
                if (own && ti == 0)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int j = 4 * tj + v;
                        Um[k + j * LD] = (j > k) ? rj[v] : (j == k ? piv : 0.0);
                    }
                const int kn = k + 1, vn = (v0 + 1) & 3, bn = kn >> 2;
                if (kn < B) {
                    if (own && tj == bn)
#pragma unroll
                        for (int u = 0; u < 4; ++u) cb[(kn & 1) * B + 4 * ti + u] = mr[u][vn];
                    if (own && ti == bn)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            rb[(kn & 1) * B + 4 * tj + v] = mr[vn][v];
                            xb[(kn & 1) * B + 4 * tj + v] = xr[vn][v];
                        }
                }
            }
        }
        __syncthreads();
        // panel rows 0..b-1 of A: R_house = S R on and above the diagonal, L below
        for (int idx = tid; idx < BB; idx += kCQThreads) {
            const int r = idx % B, j = idx / B;
            const double v = (j >= r) ? sv[r] * M1[r + j * LD] : Lm[r + j * LD];
            a.A[(i64)r + (i64)j * a.lda] = (T)v;
            if (a.Y) a.Y[(i64)r + (i64)j * a.ldy] = (T)((j < r) ? Lm[r + j * LD] : (j == r ? 1.0 : 0.0));
        }
        // Mw = -S L^-T (upper triangular: Mw(k, j) = -s_k L^-1(j, k)) into M1 and, with U,
        // into bc for the other workgroups. Every workgroup then forms its rows of
        // W = Y T = (Q - [S; 0]) Mw from its rows of Q (T = -U S L^-T gives
        // Q_2 U^-1 T = -Q_2 S L^-T below the first b rows, and L T = -(Q_1 - S) S L^-T on
        // them) and Y = Q_2 U^-1 by substitution along its rows; workgroup 0 also forms
        // T = U Mw for the output, after the barrier.
        __syncthreads(); // R in M1 and L in M2 read above
        if (own)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = 4 * ti + u, k = 4 * tj + v; // xr[u][v] = L^-1(j, k)
                    const double m = -sv[k] * xr[u][v];
                    M1[k + j * LD] = m;
                    st_wt(a.bc + BB + k + (size_t)j * B, m);
                }
        for (int idx = tid; idx < BB; idx += kCQThreads) st_wt(a.bc + idx, Um[(idx % B) + (idx / B) * LD]);
        if (tid < B) P[(size_t)tid * CHp + tid] -= sv[tid]; // Q - [S; 0] on the top block
        if (tid == 0) *a.gate = 0;
    }
    if (!grid_barrier_wt(a.counter, (unsigned)a.G * ++nbar, a.err)) return;
    if (g != 0) {
        // U and Mw: rounds of 2 x LB loads per thread, all issued before their LDS stores
        constexpr int LB = 8;
        for (int base = 0; base < BB; base += LB * kCQThreads) {
            double u[LB], t[LB];
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const int idx = base + q * kCQThreads + tid;
                u[q] = (idx < BB) ? ld_wt(a.bc + idx) : 0.0;
                t[q] = (idx < BB) ? ld_wt(a.bc + BB + idx) : 0.0;
            }
#pragma unroll
            for (int q = 0; q < LB; ++q) {
                const int idx = base + q * kCQThreads + tid;
                const int i = idx % B, j = idx / B;
                if (idx < BB) {
                    M3[i + j * LD] = u[q];
                    M1[i + j * LD] = t[q];
                }
            }
        }
        __syncthreads();
    }
    if (tid < B) gd[tid] = 1.0 / M3[tid + tid * LD]; // 1 / U(j, j)
    __syncthreads();
    // W = (Q - [S; 0]) Mw for every row of this workgroup, before Y replaces Q
    if (a.W) {
        const int nrb = (chl + 3) / 4;
        for (int blk = tid; blk < nrb * NB; blk += kCQThreads) {
            const int rbk = blk / NB, cj = blk % NB;
            const int rr = 4 * rbk;
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = 0.0;
            for (int i = 0; i <= cj + NB * 3; ++i) { // Mw(i, j) = 0 for i > j
                double x[4], w[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) x[u] = P[(size_t)i * CHp + rr + u];
#pragma unroll
                for (int v = 0; v < 4; ++v) w[v] = M1[i + (cj + NB * v) * LD];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[u][v] += x[u] * w[v];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (rr + u < chl)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        a.W[(i64)(r0g + rr + u) + (i64)(cj + NB * v) * a.ldw] = (T)acc[u][v];
        }
        __syncthreads();
    }
    // Y = Q U^-1 for the rows below the first b (workgroup 0 wrote its top rows, L)
    const int ylo = (g == 0) ? B : 0;
    cq_rsolve<B>(P, CHp, ylo, chl, M3, gd, tid);
    // outputs: Y and the reflector tails in A
    for (int idx = tid; idx < B * chl; idx += kCQThreads) {
        const int j = idx / chl, rl = idx - j * chl;
        if (rl < ylo) continue;
        const T y = (T)P[(size_t)j * CHp + rl];
        const i64 r = r0g + rl;
        a.A[r + (i64)j * a.lda] = y;
        if (a.Y) a.Y[r + (i64)j * a.ldy] = y;
    }
    if (g == 0) {
        // T = U Mw (upper triangular), and tau = diag(T)
        double acc[4][4];
        cq_mm<B, false>(acc, M3, M1, si, sj);
        if (own)
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = si + NB * u, j = sj + NB * v;
                    a.Tm[i + (size_t)j * B] = (T)acc[u][v];
                    if (i == j) a.tau[j] = (T)acc[u][v];
                }
    }
}

template <int B> size_t cq_lds_bytes(int CH) {
    return sizeof(double) * ((size_t)B * (CH + 1) + 3 * (size_t)B * (B + 1) + 9 * (size_t)B + kCQThreads);
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
    // gate word; Householder: partials [2][G][2b] + pivot [2][b] doubles, G <= n / 64 + 2;
    // CholQR: partial Gram matrices [G][b*b] + sums [b*b] + broadcast [2*b*b], G <= n / 64 + 1
    const size_t G = (size_t)n / 64 + 2;
    const size_t hh = sizeof(double) * (2 * G * 2 * b + 2 * b);
    const size_t cq = sizeof(double) * ((G + 3) * (size_t)b * b);
    return kPQGateBytes + std::max(hh, cq) + 256;
}

namespace {

// CholQR2 launch if the panel qualifies: b in {16, 32, 64}, at least 2b rows, and
// rows spread one workgroup per CU with at most kCQMaxCH rows each.
template <typename T, int B>
int launch_cholqr(hipStream_t s, T *A, i64 lda, T *Y, i64 ldy, T *W, i64 ldw, T *Tm, T *tau, int rows,
                  void *scratch, size_t scratch_bytes, unsigned *counter, int *err, bool *launched) {
    *launched = false;
    const int ncu = device_cu_count();
    const int CH = std::max(64, (int)round_up(ceil_div(rows, ncu), 16));
    if (CH > kCQMaxCH) return CUEV_STATUS_SUCCESS;
    const int G = (int)ceil_div(rows, CH);
    const size_t lds = cq_lds_bytes<B>(CH);
    CUEV_HIP(hipFuncSetAttribute((const void *)panel_cholqr_kernel<T, B>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, panel_cholqr_kernel<T, B>, kCQThreads, lds));
    if (G > per_cu * ncu) return CUEV_STATUS_SUCCESS;
    const size_t need = kPQGateBytes + sizeof(double) * ((size_t)(G + 3) * B * B);
    if (need > scratch_bytes) return CUEV_STATUS_SUCCESS;
    CQArgs<T> a;
    a.A = A;
    a.lda = lda;
    a.Y = Y;
    a.ldy = ldy;
    a.W = W;
    a.ldw = ldw;
    a.Tm = Tm;
    a.tau = tau;
    a.rows = rows;
    a.CH = CH;
    a.G = G;
    a.gate = static_cast<int *>(scratch);
    a.part = reinterpret_cast<double *>(static_cast<unsigned char *>(scratch) + kPQGateBytes);
    a.gsum = a.part + (size_t)G * B * B;
    a.bc = a.gsum + B * B;
    a.counter = counter;
    a.reset = counter + 1;
    a.err = err;
    hipLaunchKernelGGL((panel_cholqr_kernel<T, B>), dim3(G), dim3(kCQThreads), lds, s, a);
    CUEV_LAUNCH_CHECK();
    *launched = true;
    return CUEV_STATUS_SUCCESS;
}

} // namespace

template <typename T>
int panel_qr(hipStream_t s, T *A, i64 lda, T *Y, i64 ldy, T *W, i64 ldw, T *Tm, T *tau, int rows,
             int b, void *scratch, size_t scratch_bytes, unsigned *counter, int *err) {
    if (rows < 1 || b < 1 || b > 64)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "panel_qr: rows=%d b=%d", rows, b);
    if (scratch_bytes < kPQGateBytes)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "panel_qr: scratch too small");
    // counter[0]: CholQR barriers, counter[1]: Householder barriers. Both are zero
    // when a panel starts: the workspace starts them at zero, each kernel zeroes the
    // other's counter when it starts (the CholQR kernel the Householder kernel's that
    // follows it, the Householder kernel, always launched, the CholQR counter for the
    // next panel), and without a CholQR launch counter[1] is zeroed here.
    bool cq = false;
    if (rows >= 2 * b) {
        if (b == 64)
            CUEV_TRY((launch_cholqr<T, 64>(s, A, lda, Y, ldy, W, ldw, Tm, tau, rows, scratch, scratch_bytes,
                                           counter, err, &cq)));
        else if (b == 32)
            CUEV_TRY((launch_cholqr<T, 32>(s, A, lda, Y, ldy, W, ldw, Tm, tau, rows, scratch, scratch_bytes,
                                           counter, err, &cq)));
        else if (b == 16)
            CUEV_TRY((launch_cholqr<T, 16>(s, A, lda, Y, ldy, W, ldw, Tm, tau, rows, scratch, scratch_bytes,
                                           counter, err, &cq)));
    }
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
    const size_t need = kPQGateBytes + sizeof(double) * ((size_t)2 * G * 2 * b + 2 * b);
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
    a.part = reinterpret_cast<double *>(static_cast<unsigned char *>(scratch) + kPQGateBytes);
    a.pivot = a.part + (size_t)2 * G * 2 * b;
    a.gate = cq ? static_cast<const int *>(scratch) : nullptr;
    a.counter = counter + 1;
    a.reset = counter;
    a.err = err;
    if (!cq) {
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, counter + 1, 1);
        CUEV_LAUNCH_CHECK();
    }
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
