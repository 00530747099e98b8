// gemm.hip — MFMA tile engine for gfx950 (FP64 and FP32).
//
// Block tile BM x BN x 16, 256 threads = 4 wave64s arranged 2 x 2, each wave
// owning a (BM/2) x (BN/2) sub-tile of 16 x 16 MFMA fragments. Operands are
// staged global -> registers -> LDS (two LDS buffers, one barrier per k-step,
// the next k-tile's global loads in flight while the current one computes).
//
// The MFMA is issued with the B fragment in the instruction's A slot and the A
// fragment in its B slot, so the accumulator's 16 consecutive lanes hold 16
// consecutive rows of one C column: epilogue stores are coalesced for
// column-major C. Output lane/register layout (measured on gfx950):
//   f64 16x16x4: lane l, reg r -> (row l/16 + 4r, col l%16)
//   f32 16x16x4: lane l, reg r -> (row 4(l/16) + r, col l%16)
// (row/col of the instruction's D; with the swapped operands D's row is C's
// column index and D's column is C's row index.)
//
// Work distribution: blocks are dispatched round-robin over the 8 XCDs, so the
// block index is first remapped so that each XCD receives a contiguous range of
// tiles, then grouped 8 tile-rows at a time, keeping operand panels shared by
// neighbouring tiles inside one XCD's L2.
//
// Tall-skinny products (the DBBR panel products with K = n - j - b, the
// sbr_back Y^T M) split K over blockIdx.y into a scratch buffer, summed in a
// fixed order by a second kernel, so results are deterministic.
#include "gemm.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <utility>
#include <vector>

#include <rocblas/rocblas.h>

namespace cuev {
namespace {

constexpr int kThreads = 256;
constexpr int BK = 16;
constexpr int kGroupM = 8;

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T> struct KParams {
    const T *a0, *a1;
    i64 lda0, lda1, asplit;
    int atrans, asym;
    const T *b0, *b1;
    i64 ldb0, ldb1, bsplit;
    int btrans;
    T *C;
    i64 ldc;
    T *part; // split-K partials: part[s*M*N + m + n*M]
    i64 M, N, K, kchunk;
    T alpha, beta;
    int lower;
    int tiles_m, tiles_n;
};

// Map the (remapped) linear tile id to (tile_m, tile_n).
__device__ __forceinline__ void tile_coords(int t, int tiles_m, int tiles_n, int lower, int &tm,
                                            int &tn) {
    if (lower) {
        // enumerate pairs tm >= tn: t = tm (tm + 1) / 2 + tn
        int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while ((i64)(r + 1) * (r + 2) / 2 <= t) ++r;
        while ((i64)r * (r + 1) / 2 > t) --r;
        tm = r;
        tn = t - r * (r + 1) / 2;
    } else {
        const int group = kGroupM * tiles_n;
        const int gid = t / group;
        const int first_m = gid * kGroupM;
        const int gsz = min(tiles_m - first_m, kGroupM);
        const int w = t - gid * group;
        tm = first_m + w % gsz;
        tn = w / gsz;
    }
}

// Operand modes, fixed at compile time so that each kernel carries only the
// loader it needs (register pressure decides how many waves share a SIMD).
enum AMode { A_N = 0, A_T = 1, A_SYM = 2, A_CAT = 3 }; // A_CAT: two m-contiguous sources
enum BMode { B_N = 0, B_T = 1, B_CATT = 2 };           // B_CATT: two n-contiguous sources

// A tile (BM x BK at rows m0, k in [k0, k0+BK)) into registers, in the order
// store_a expects. Pattern 0: m-contiguous (element (m,k) at p[m + k*ld]);
// pattern 1: k-contiguous (at p[k + m*ld]); pattern 2: symmetric straddling the
// diagonal (lower triangle stored).
template <typename T, int BM, int NA>
__device__ __forceinline__ void load_a(T (&ra)[NA], int pat, const T *pa, i64 lda, i64 m0, i64 k0,
                                       i64 kend, i64 M, int tid) {
    if (pat == 1) {
        const int kk = tid % BK;
        const i64 k = k0 + kk;
        const bool kok = k < kend;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const i64 m = m0 + tid / BK + (kThreads / BK) * i;
            ra[i] = (kok && m < M) ? pa[k + m * lda] : T(0);
        }
    } else {
        const int ml = tid % BM;
        const i64 m = m0 + ml;
        const bool mok = m < M;
        const int kk0 = tid / BM;
        const T *base = pa + m + (k0 + kk0) * lda;
        const i64 step = (i64)(kThreads / BM) * lda;
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const i64 k = k0 + kk0 + (kThreads / BM) * i;
            if (pat == 0) {
                ra[i] = (mok && k < kend) ? base[i * step] : T(0);
            } else {
                const i64 idx = (m >= k) ? m + k * lda : k + m * lda;
                ra[i] = (mok && k < kend) ? pa[idx] : T(0);
            }
        }
    }
}

template <typename T, int BM, int NA>
__device__ __forceinline__ void store_a(T (*sA)[BM + 1], const T (&ra)[NA], int pat, int tid) {
    if (pat == 1) {
#pragma unroll
        for (int i = 0; i < NA; ++i) sA[tid % BK][tid / BK + (kThreads / BK) * i] = ra[i];
    } else {
#pragma unroll
        for (int i = 0; i < NA; ++i) sA[tid / BM + (kThreads / BM) * i][tid % BM] = ra[i];
    }
}

// B tile (BK x BN at columns n0): pattern 0 k-contiguous (p[k + n*ld]), 1 n-contiguous.
template <typename T, int BN, int NB>
__device__ __forceinline__ void load_b(T (&rb)[NB], int pat, const T *pb, i64 ldb, i64 n0, i64 k0,
                                       i64 kend, i64 N, int tid) {
    if (pat == 0) {
        const int kk = tid % BK;
        const i64 k = k0 + kk;
        const bool kok = k < kend;
        const T *base = pb + k + (n0 + tid / BK) * ldb;
        const i64 step = (i64)(kThreads / BK) * ldb;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const i64 n = n0 + tid / BK + (kThreads / BK) * i;
            rb[i] = (kok && n < N) ? base[i * step] : T(0);
        }
    } else {
        const int nl = tid % BN;
        const i64 n = n0 + nl;
        const bool nok = n < N;
        const int kk0 = tid / BN;
        const T *base = pb + n + (k0 + kk0) * ldb;
        const i64 step = (i64)(kThreads / BN) * ldb;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const i64 k = k0 + kk0 + (kThreads / BN) * i;
            rb[i] = (nok && k < kend) ? base[i * step] : T(0);
        }
    }
}

template <typename T, int BN, int NB>
__device__ __forceinline__ void store_b(T (*sB)[BN + 1], const T (&rb)[NB], int pat, int tid) {
    if (pat == 0) {
#pragma unroll
        for (int i = 0; i < NB; ++i) sB[tid % BK][tid / BK + (kThreads / BK) * i] = rb[i];
    } else {
#pragma unroll
        for (int i = 0; i < NB; ++i) sB[tid / BN + (kThreads / BN) * i][tid % BN] = rb[i];
    }
}

// Block tile BM x BN, K step BK, 4 waves (2 x 2) each owning (BM/2) x (BN/2).
// __launch_bounds__(256, 2): at most 256 VGPRs, so two waves share each SIMD and
// one hides the other's LDS and barrier latency.
template <typename T, int BM, int BN, int AMODE, int BMODE>
__global__ __launch_bounds__(kThreads, 2) void gemm_tile_kernel(const KParams<T> p) {
    constexpr int TM = BM / 2, TN = BN / 2;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int NA = BM * BK / kThreads; // A elements per thread per k-tile
    constexpr int NB = BN * BK / kThreads;
    using V4 = typename Mfma<T>::v4;

    __shared__ T sA[2][BK][BM + 1];
    __shared__ T sB[2][BK][BN + 1];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // XCD-contiguous remap of the block index
    const int nbk = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q = nbk >> 3, rmd = nbk & 7;
    const int t = xcd * q + min(xcd, rmd) + (bid >> 3);
    int tm, tn;
    tile_coords(t, p.tiles_m, p.tiles_n, p.lower, tm, tn);
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;

    const i64 kbeg = (i64)blockIdx.y * p.kchunk;
    const i64 kend = min(p.K, kbeg + p.kchunk);
    const int nkt = (int)((kend - kbeg + BK - 1) / BK);

    V4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};

    T ra[NA], rb[NB];
    int patA = 0, patB = 0;
    auto load_tiles = [&](i64 k0) {
        // A operand source and pattern for this k-tile
        const T *pa = p.a0;
        i64 lda = p.lda0, ko = 0;
        if constexpr (AMODE == A_N) {
            patA = 0;
        } else if constexpr (AMODE == A_T) {
            patA = 1;
        } else if constexpr (AMODE == A_SYM) {
            patA = (m0 >= k0 + BK - 1) ? 0 : (m0 + BM - 1 <= k0) ? 1 : 2;
        } else {
            patA = 0;
            if (k0 >= p.asplit) {
                pa = p.a1;
                lda = p.lda1;
                ko = p.asplit;
            }
        }
        load_a<T, BM, NA>(ra, patA, pa, lda, m0, k0 - ko, kend - ko, p.M, tid);
        const T *pb = p.b0;
        i64 ldb = p.ldb0, kob = 0;
        if constexpr (BMODE == B_N) {
            patB = 0;
        } else if constexpr (BMODE == B_T) {
            patB = 1;
        } else {
            patB = 1;
            if (k0 >= p.bsplit) {
                pb = p.b1;
                ldb = p.ldb1;
                kob = p.bsplit;
            }
        }
        load_b<T, BN, NB>(rb, patB, pb, ldb, n0, k0 - kob, kend - kob, p.N, tid);
    };

    if (nkt > 0) {
        load_tiles(kbeg);
        store_a<T, BM, NA>(sA[0], ra, patA, tid);
        store_b<T, BN, NB>(sB[0], rb, patB, tid);
    }
    __syncthreads();

    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nkt;
        if (more) load_tiles(kbeg + (i64)(kt + 1) * BK);
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kr = ks * 4 + (lane >> 4);
            T af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = sA[cur][kr][wm * TM + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = sB[cur][kr][wn * TN + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(bf[j], af[i], acc[i][j]);
        }
        if (more) {
            store_a<T, BM, NA>(sA[cur ^ 1], ra, patA, tid);
            store_b<T, BN, NB>(sB[cur ^ 1], rb, patB, tid);
        }
        __syncthreads();
    }

    // ---- epilogue
    const bool partial = p.part != nullptr;
    T *part = partial ? p.part + (i64)blockIdx.y * p.M * p.N : nullptr;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const i64 m = m0 + wm * TM + i * 16 + (lane & 15);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn * TN + j * 16 + Mfma<T>::out_row(lane, r);
                if (m >= p.M || n >= p.N) continue;
                if (p.lower && m < n) continue;
                const T v = acc[i][j][r];
                if (partial) {
                    part[m + n * p.M] = v;
                } else {
                    T *cp = p.C + m + n * p.ldc;
                    *cp = (p.beta == T(0)) ? p.alpha * v : p.alpha * v + p.beta * (*cp);
                }
            }
        }
    }
}

// C = alpha * sum_s part[s] + beta * C (fixed summation order).
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const T *part, int splits, i64 M,
                                                            i64 N, T alpha, T beta, T *C, i64 ldc,
                                                            int lower) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    const i64 m = idx % M, n = idx / M;
    if (lower && m < n) return;
    T s = T(0);
    for (int k = 0; k < splits; ++k) s += part[(i64)k * M * N + idx];
    T *c = C + m + n * ldc;
    *c = (beta == T(0)) ? alpha * s : alpha * s + beta * (*c);
}

template <typename T, int BM, int BN, int AM, int BMD>
int launch_modes(hipStream_t s, KParams<T> &p, int splits, int ntiles) {
    dim3 grid((unsigned)ntiles, (unsigned)splits);
    hipLaunchKernelGGL((gemm_tile_kernel<T, BM, BN, AM, BMD>), grid, dim3(kThreads), 0, s, p);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename T, int BM, int BN>
int launch_cfg(hipStream_t s, KParams<T> &p, int splits, int ntiles, int am, int bm) {
#define CUEV_GEMM_B(AMV)                                                                           \
    switch (bm) {                                                                                  \
    case B_N: return launch_modes<T, BM, BN, AMV, B_N>(s, p, splits, ntiles);                      \
    case B_T: return launch_modes<T, BM, BN, AMV, B_T>(s, p, splits, ntiles);                      \
    default: return launch_modes<T, BM, BN, AMV, B_CATT>(s, p, splits, ntiles);                    \
    }
    switch (am) {
    case A_N: CUEV_GEMM_B(A_N)
    case A_T: CUEV_GEMM_B(A_T)
    case A_SYM: CUEV_GEMM_B(A_SYM)
    default: CUEV_GEMM_B(A_CAT)
    }
#undef CUEV_GEMM_B
}

} // namespace

size_t gemm_scratch_bytes(i64 M, i64 N, size_t elem) {
    // split-K is only used when the tile count is below ~one wave of blocks, so
    // splits * M * N stays below ~3 * (2 * CUs) * 128 * 128 elements.
    (void)M;
    (void)N;
    return (size_t)3 * 512 * 128 * 128 * elem;
}

namespace {

// rocBLAS handles, one per stream (a handle's workspace must not be shared by
// work queued on two streams).
std::mutex g_vendor_mu;
std::vector<std::pair<hipStream_t, rocblas_handle>> g_vendor;

rocblas_handle vendor_handle(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_vendor_mu);
    for (auto &e : g_vendor)
        if (e.first == s) return e.second;
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return nullptr;
    if (rocblas_set_stream(h, s) != rocblas_status_success) {
        rocblas_destroy_handle(h);
        return nullptr;
    }
    g_vendor.emplace_back(s, h);
    return h;
}

inline rocblas_status vendor_gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n,
                                  int k, const double *alpha, const double *A, int lda, const double *B, int ldb,
                                  const double *beta, double *C, int ldc) {
    return rocblas_dgemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
inline rocblas_status vendor_gemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m, int n,
                                  int k, const float *alpha, const float *A, int lda, const float *B, int ldb,
                                  const float *beta, float *C, int ldc) {
    return rocblas_sgemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

} // namespace

int gemm_vendor_attach(hipStream_t s) {
    if (!vendor_handle(s)) CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "gemm: rocblas handle creation failed");
    return CUEV_STATUS_SUCCESS;
}

void gemm_vendor_release(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_vendor_mu);
    for (size_t i = 0; i < g_vendor.size(); ++i)
        if (g_vendor[i].first == s) {
            rocblas_destroy_handle(g_vendor[i].second);
            g_vendor.erase(g_vendor.begin() + (long)i);
            return;
        }
}

namespace {

template <typename T>
int gemm_vendor(hipStream_t s, i64 M, i64 N, i64 K, T alpha, const Operand<T> &A, const Operand<T> &B,
                T beta, T *C, i64 ldc) {
    rocblas_handle h = vendor_handle(s);
    if (!h) CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "gemm: rocblas handle creation failed");
    const rocblas_operation ta = A.trans ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation tb = B.trans ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_status st = vendor_gemm(h, ta, tb, (int)M, (int)N, (int)K, &alpha, A.p0, (int)A.ld0, B.p0,
                                          (int)B.ld0, &beta, C, (int)ldc);
    if (st != rocblas_status_success) CUEV_FAIL(CUEV_STATUS_HIP_ERROR, "gemm: rocblas_?gemm returned status %d", (int)st);
    return CUEV_STATUS_SUCCESS;
}

} // namespace

template <typename T>
int gemm_ex(hipStream_t s, i64 M, i64 N, i64 K, T alpha, const Operand<T> &A,
            const Operand<T> &B, T beta, T *C, i64 ldc, int lower, Scratch scratch) {
    if (M < 0 || N < 0 || K < 0)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: negative size M=%lld N=%lld K=%lld",
                  (long long)M, (long long)N, (long long)K);
    if (M == 0 || N == 0) return CUEV_STATUS_SUCCESS;
    if (lower && M != N) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: lower mode needs M == N");
    if ((A.split >= 0 && A.split % BK) || (B.split >= 0 && B.split % BK))
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: K split must be a multiple of %d", BK);
    if (ldc < M) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: ldc < M");
    const i64 imax = 0x7fffffffLL;
    if (scratch.vendor && !lower && A.split < 0 && B.split < 0 && !A.sym && std::min(std::min(M, N), K) >= 512 &&
        (double)M * N * K >= 8.6e9 && M <= imax && N <= imax && K <= imax && A.ld0 <= imax && B.ld0 <= imax && ldc <= imax)
        return gemm_vendor<T>(s, M, N, K, alpha, A, B, beta, C, ldc);

    KParams<T> p{};
    p.a0 = A.p0;
    p.a1 = A.p1;
    p.lda0 = A.ld0;
    p.lda1 = A.ld1;
    p.asplit = A.split;
    p.atrans = A.trans;
    p.asym = A.sym;
    p.b0 = B.p0;
    p.b1 = B.p1;
    p.ldb0 = B.ld0;
    p.ldb1 = B.ld1;
    p.bsplit = B.split;
    p.btrans = B.trans;
    p.C = C;
    p.ldc = ldc;
    p.M = M;
    p.N = N;
    p.K = K;
    p.alpha = alpha;
    p.beta = beta;
    p.lower = lower;

    // tile shape
    int BMs = (M > 64) ? 128 : 64;
    int BNs = (N > 64) ? 128 : 64;
    if (lower) BMs = BNs = (N >= 1024) ? 128 : 64;
    if (scratch.small_tiles) BMs = BNs = 64;
    // symmetric A times a panel-wide B (the band reduction's A22 W): bandwidth-bound on
    // A, so small tiles and deep split-K keep more loads in flight (measured at
    // b = 64: 0.87 -> 0.70 ms for 16320 rows, 0.10 -> 0.045 ms for 3000)
    const bool sym_panel = A.sym && N <= 64;
    if (sym_panel) BMs = 64;
    const int cus = device_cu_count();
    auto count_tiles = [&](int bm, int bn) -> i64 {
        const i64 tm = ceil_div(M, bm), tn = ceil_div(N, bn);
        return lower ? tm * (tm + 1) / 2 : tm * tn;
    };
    if (!lower && BMs == 128 && BNs == 128 && count_tiles(128, 128) < cus) {
        // few large tiles: trade tile size for parallelism
        if (count_tiles(64, 128) >= cus || M < N) BMs = 64;
        else BNs = 64;
    }
    const i64 ntiles = count_tiles(BMs, BNs);
    if (ntiles > 0x7fffffffLL) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: too many tiles");
    p.tiles_m = (int)ceil_div(M, BMs);
    p.tiles_n = (int)ceil_div(N, BNs);

    // split-K for products with few output tiles and a long K
    int splits = 1;
    const i64 target = (sym_panel ? 8LL : 2LL) * cus;
    const i64 kmin = sym_panel ? 256 : 512; // shortest K chunk per split
    if (K == 0) {
        splits = 0;
    } else if (ntiles < target / 2 && K >= 1024) {
        i64 want = std::min<i64>(ceil_div(target, ntiles), K / kmin);
        const i64 cap = (i64)(scratch.bytes / ((size_t)M * N * sizeof(T)));
        want = std::min(want, cap);
        if (want > 1) splits = (int)want;
    }
    if (splits <= 1 && K > 0) {
        p.kchunk = round_up(K, BK);
        p.part = nullptr;
        splits = 1;
    } else if (splits > 1) {
        p.kchunk = round_up(ceil_div(K, splits), BK);
        splits = (int)ceil_div(K, p.kchunk);
        p.part = static_cast<T *>(scratch.ptr);
    }

    if (splits >= 1) {
        int am, bm;
        if (A.sym) am = A_SYM;
        else if (A.split >= 0) am = A_CAT;
        else am = A.trans ? A_T : A_N;
        if (B.split >= 0) bm = B_CATT;
        else bm = B.trans ? B_T : B_N;
        if ((A.sym && (A.split >= 0 || A.trans)) || (A.split >= 0 && A.trans) ||
            (B.split >= 0 && !B.trans) || B.sym)
            CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "gemm: operand mode combination not implemented");
        int st;
        if (BMs == 128 && BNs == 128) st = launch_cfg<T, 128, 128>(s, p, splits, (int)ntiles, am, bm);
        else if (BMs == 128) st = launch_cfg<T, 128, 64>(s, p, splits, (int)ntiles, am, bm);
        else if (BNs == 128) st = launch_cfg<T, 64, 128>(s, p, splits, (int)ntiles, am, bm);
        else st = launch_cfg<T, 64, 64>(s, p, splits, (int)ntiles, am, bm);
        if (st) return st;
    }
    if (p.part != nullptr || splits == 0) {
        const i64 total = M * N;
        const int nsplit = (splits == 0) ? 0 : splits;
        hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)ceil_div(total, 256)),
                           dim3(256), 0, s, static_cast<const T *>(p.part), nsplit, M, N, alpha,
                           beta, C, ldc, lower);
        CUEV_LAUNCH_CHECK();
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int syr2k_lower(hipStream_t s, i64 n, i64 k, T alpha, const T *X, i64 ldx, const T *Y, i64 ldy,
                T beta, T *C, i64 ldc, Scratch scratch) {
    if (k % BK == 0) {
        // [X | Y] * [Y | X]^T as one product with K = 2k
        Operand<T> a = opnd_cat(X, ldx, Y, ldy, k, 0);
        Operand<T> b = opnd_cat(Y, ldy, X, ldx, k, 1);
        return gemm_ex<T>(s, n, n, 2 * k, alpha, a, b, beta, C, ldc, 1, scratch);
    }
    CUEV_TRY(gemm_ex<T>(s, n, n, k, alpha, opnd(X, ldx, 0), opnd(Y, ldy, 1), beta, C, ldc, 1,
                        scratch));
    return gemm_ex<T>(s, n, n, k, alpha, opnd(Y, ldy, 0), opnd(X, ldx, 1), T(1), C, ldc, 1,
                      scratch);
}

template int gemm_ex<float>(hipStream_t, i64, i64, i64, float, const Operand<float> &,
                            const Operand<float> &, float, float *, i64, int, Scratch);
template int gemm_ex<double>(hipStream_t, i64, i64, i64, double, const Operand<double> &,
                             const Operand<double> &, double, double *, i64, int, Scratch);
template int syr2k_lower<float>(hipStream_t, i64, i64, float, const float *, i64, const float *,
                                i64, float, float *, i64, Scratch);
template int syr2k_lower<double>(hipStream_t, i64, i64, double, const double *, i64,
                                 const double *, i64, double, double *, i64, Scratch);

} // namespace cuev
