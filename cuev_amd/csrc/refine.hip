// refine.hip — mixed-precision solve: FP32 two-stage reduction and back-transform,
// host FP64 divide and conquer, FP64 Ogita-Aishima refinement on the GPU.
// No counterpart in the reference (the reference runs every stage in T).
//
// Pipeline (double API, precision = CUEV_PRECISION_MIXED):
//   1. A32 = (float) A; DBBR and bulge chasing in FP32 (FP32 MFMA, FP32
//      accumulation; panel norms, dot products, tau and T stay FP64).
//   2. FP64 divide and conquer on the FP32-derived (d, e) (device-resident, or host
//      dstedc), overlapped with the FP32 build of M = Q_b^T Q_s^T on the GPU (Q_s^T
//      beside the bulge chasing, as in the FP64 solve's reordered order);
//      X = (double)(M^T (float)Q_d).
//   3. Ogita-Aishima sweeps in FP64 (T. Ogita, K. Aishima, "Iterative
//      refinement for symmetric eigenvalue decomposition", JJIAM 35, 2018,
//      Algorithm 1):
//        G = X^T X,  R = I - G,  S = X^T A X,  lam_i = s_ii / (1 - r_ii),
//        E_ij = (s_ij + lam_j r_ij) / (lam_j - lam_i)  for a separated pair,
//             = r_ij / 2                              for a clustered one,
//        X <- X + X E.
//      The paper separates pairs by one global bound,
//      |lam_i - lam_j| > delta = 2 (||S - diag(lam)||_2 + ||A||_2 ||R||_2).
//      Here each pair is tested on its own: it is separated when both first-order
//      corrections |s_ij + lam_j r_ij| and |s_ij + lam_i r_ij|, plus the rounding
//      level of S (n u ||A||_F), are at most tau = 1e-4 of |lam_j - lam_i|, so that
//      the step's second-order error (below tau^2 of the pair's coupling) is gone
//      after the next step. On a geometric spectrum the global bound (with its
//      2-norms bounded by Frobenius norms) marks nearly every pair of an FP32 start
//      as clustered; the pair test marks those whose coupling is comparable to
//      their gap. Each evaluation measures the residual
//      ||A X - X diag(lam)||_F / (n ||A||_F) and ||X^T X - I||_F / n from the
//      products it needs anyway; after a Newton step an estimate of both from 16
//      Gaussian probes comes first and ends the refinement when both are well
//      below the tolerance. The columns of the pairs that the first Newton step
//      leaves clustered (orthogonalised, not rotated) split into sets that no
//      clustered pair joins; right after that step one Rayleigh-Ritz step per set
//      resolves them (the cluster treatment of Ogita and Aishima's refinement for
//      clustered eigenvalues, JJIAM 36, 2019), and the next Newton step finishes
//      the separated pairs.
//   A is first filled in full from its lower triangle (the solve overwrites A),
//   so every product with it is a plain one. All products run on the FP64 MFMA
//   engine; the fused element-wise kernels below build lam, E and the norms.
#include "refine.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "backtransform.hpp"
#include "bulge.hpp"
#include "context.hpp"
#include "dbbr.hpp"
#include "dc.hpp"
#include "gemm.hpp"
#include "lapack.hpp"
#include "profile.hpp"
#include "solver.hpp"
#include "util.hpp"

namespace cuev {
namespace {

constexpr int kRedBlocks = 1024; // fixed partial count: deterministic reductions
// clustered columns from which the Rayleigh-Ritz step runs on the GPU (one
// FP64 solve of the projected union) instead of one host dsygvd
constexpr int kRRDeviceMin = 512;
// a pair is separated when its first-order corrections are at most kPairTau of its gap
constexpr double kPairTau = 1e-4;

// the pair test: clustered unless both first-order corrections of the pair (i, j),
// plus nu (the rounding level of S), are at most kPairTau |lam_j - lam_i|; a NaN
// tests as clustered
__device__ inline bool oa_clustered(double sij, double r, double li, double lj, double nu) {
    const double c = fmax(fabs(sij + lj * r), fabs(sij + li * r)) + nu;
    return !(c <= kPairTau * fabs(lj - li));
}

// lam_i = S_ii / G_ii
// with S_ii = x_i . (A X)_i: one workgroup per column, so the convergence check
// needs no product for S (formed only when the refinement goes on)
__global__ __launch_bounds__(256) void oa_lambda_kernel(const double *X, const double *AX, const double *G,
                                                        i64 ld, int n, double *lam) {
    __shared__ double red[256];
    const int i = blockIdx.x;
    const double *x = X + (i64)i * ld, *ax = AX + (i64)i * ld;
    double a = 0;
    for (int r = threadIdx.x; r < n; r += 256) a += x[r] * ax[r];
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lam[i] = red[0] / G[i + (i64)i * ld];
}

// block partial sums of squares (G symmetric, lower triangle stored):
//  part[1*kRedBlocks + b] : ||G - I||_F^2            (= ||R||_F^2)
//  part[2*kRedBlocks + b] : ||AX - X diag(lam)||_F^2
// (part[0*kRedBlocks + b] is written as zero)
__global__ __launch_bounds__(256) void oa_norms_kernel(const double *G, const double *AX, const double *X,
                                                       i64 ld, int n, const double *lam, double *part) {
    __shared__ double red[3][256];
    double a = 0, g = 0, r = 0;
    const i64 total = (i64)n * n;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (i64)gridDim.x * blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        const i64 o = i + (i64)j * ld;
        const double t = AX[o] - X[o] * lam[j];
        r += t * t;
        if (i < j) continue;
        const double w = (i == j) ? 1.0 : 2.0;
        const double q = G[o] - (i == j ? 1.0 : 0.0);
        g += w * q * q;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = g;
    red[2][threadIdx.x] = r;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) part[threadIdx.x * kRedBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// The strict upper triangle of A from its lower one (64 x 64 tiles through LDS):
// block (ti, tj), tj <= ti, reads lower tile (ti, tj) and writes the upper tile
// (tj, ti); a diagonal tile writes only its entries above the diagonal. A is then
// stored in full, so A X runs as a plain product.
__global__ __launch_bounds__(256) void symmetrize_lower_kernel(double *A, i64 ld, int n) {
    __shared__ double t[64][65];
    const int ti = blockIdx.x, tj = blockIdx.y;
    if (tj > ti) return;
    const int r0 = ti * 64, c0 = tj * 64;
    const int lr = threadIdx.x & 63, lc = threadIdx.x >> 6;
    for (int c = lc; c < 64; c += 4) {
        const int r = r0 + lr, col = c0 + c;
        if (r < n && col < n) t[lr][c] = A[r + (i64)col * ld];
    }
    __syncthreads();
    // upper element (row, col) = A(col, row) = t[col - r0][row - c0]
    for (int c = lc; c < 64; c += 4) {
        const int row = c0 + lr, col = r0 + c;
        if (row < n && col < n && col > row) A[row + (i64)col * ld] = t[c][lr];
    }
}

// Gaussian probes for the norm check: R (n x kProbes, ld n) with standard normal
// entries by the Box-Muller transform of two uniforms taken from the splitmix64
// finalizer (Steele, Lea and Flood, OOPSLA 2014) of the entry index, and
// LR = diag(lam) R.
constexpr int kProbes = 16;
__global__ void probe_kernel(double *R, double *LR, const double *lam, int n) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (i64)n * kProbes) return;
    unsigned long long x = (unsigned long long)k + 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    const double u1 = ((double)(x >> 32) + 1.0) * 0x1p-32;       // (0, 1]
    const double u2 = (double)(x & 0xffffffffull) * 0x1p-32;      // [0, 1)
    const double r = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
    R[k] = r;
    LR[k] = lam[k % n] * r;
}

// block partial sums over n x kProbes entries:
//  part[0*kRedBlocks + b] : ||A X R - X diag(lam) R||_F^2
//  part[1*kRedBlocks + b] : ||X^T X R - R||_F^2
__global__ __launch_bounds__(256) void probe_norms_kernel(const double *AY, const double *XLR, const double *XtY,
                                                          const double *R, i64 total, double *part) {
    __shared__ double red[2][256];
    double a = 0, g = 0;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (i64)gridDim.x * blockDim.x) {
        const double t = AY[idx] - XLR[idx];
        const double q = XtY[idx] - R[idx];
        a += t * t;
        g += q * q;
    }
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = g;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) part[threadIdx.x * kRedBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// ||A||_F^2 of the symmetric matrix stored in the lower triangle
__global__ __launch_bounds__(256) void sym_frob_kernel(const double *A, i64 ld, int n, double *part) {
    __shared__ double red[256];
    double a = 0;
    const i64 total = (i64)n * n;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (i64)gridDim.x * blockDim.x) {
        const int i = (int)(idx % n), j = (int)(idx / n);
        if (i < j) continue;
        const double v = A[i + (i64)j * ld];
        a += (i == j) ? v * v : 2.0 * v * v;
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// E (written over G, both triangles) from the lower triangles of S = X^T A X and
// G = X^T X, lam and the pair test. The thread of a pair i > j reads S_ij and G_ij
// and writes E_ij and E_ji, so the update in place is race-free; clustered pairs
// are counted once. One value s_ij serves both entries: E + E^T then equals R to
// rounding, which the orthogonality's convergence needs.
__global__ __launch_bounds__(256) void oa_e_kernel(const double *S, double *G, i64 ld, int n,
                                                   const double *lam, double nu,
                                                   unsigned *nclustered) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (i64)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    if (i < j) return;
    const i64 o = i + (i64)j * ld;
    if (i == j) {
        G[o] = 0.5 * (1.0 - G[o]);
        return;
    }
    const double r = -G[o], sij = S[o];
    const double gap = lam[j] - lam[i]; // E_ij divides by lam_j - lam_i, E_ji by its negative
    double eij, eji;
    if (!oa_clustered(sij, r, lam[i], lam[j], nu)) {
        eij = (sij + lam[j] * r) / gap;
        eji = -(sij + lam[i] * r) / gap;
    } else {
        eij = eji = 0.5 * r;
        atomicAdd(nclustered, 1u);
    }
    G[o] = eij;
    G[j + (i64)i * ld] = eji;
}

// For every clustered pair (i, j): the largest rank (position in the ascending order
// of lam) among each column's clustered partners, hi (set beforehand to -1; hi[i] >= 0
// marks a clustered column).
__global__ __launch_bounds__(256) void oa_flag_kernel(const double *S, const double *G, i64 ld, int n,
                                                      const double *lam, double nu, const int *rank, int *hi) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (i64)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    if (i <= j) return;
    const i64 o = i + (i64)j * ld;
    if (oa_clustered(S[o], -G[o], lam[i], lam[j], nu)) {
        atomicMax(hi + i, rank[j]);
        atomicMax(hi + j, rank[i]);
    }
}

// dst[:, t] = X[:, idx[t]] for t < m (n rows)
__global__ void gather_cols_kernel(const double *X, double *dst, i64 ld, int n, const int *idx, int m) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (i64)n * m) return;
    const int i = (int)(k % n), t = (int)(k / n);
    dst[i + (i64)t * n] = X[i + (i64)idx[t] * ld];
}
// X[:, idx[t]] = src[:, t]
__global__ void scatter_cols_kernel(const double *src, double *X, i64 ld, int n, const int *idx, int m) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (i64)n * m) return;
    const int i = (int)(k % n), t = (int)(k / n);
    X[i + (i64)idx[t] * ld] = src[i + (i64)t * n];
}
// Xp[:, j] = X[:, perm[j]]
__global__ void permute_cols_kernel(const double *X, double *Xp, i64 ld, int n, const int *perm) {
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (i64)n * n) return;
    const int i = (int)(idx % n), j = (int)(idx / n);
    Xp[i + (i64)j * ld] = X[i + (i64)perm[j] * ld];
}

// C = S - (ES + (ES)^T) / 2 over the lower triangle of the m x m block, in place
// over S (the upper triangle is left as it is: the solver reads the lower one)
__global__ void rr_congruence_kernel(double *S, const double *ES, i64 m) {
    const i64 k = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m * m) return;
    const i64 i = k % m, j = k / m;
    if (i < j) return;
    S[k] -= 0.5 * (ES[k] + ES[j + i * m]);
}

// G - I on the diagonal of the m x m block (in place)
__global__ void sub_identity_kernel(double *G, i64 m) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) G[i + i * m] -= 1.0;
}

struct MixedBuffers {
    int n = 0, device = -1;
    double *T1 = nullptr, *S = nullptr, *G = nullptr; // n x n FP64
    double *lam = nullptr;                            // n
    double *part = nullptr;                           // 3 * kRedBlocks
    unsigned *count = nullptr;
    int *perm = nullptr;                              // n
    int *rank = nullptr, *hi = nullptr;               // n each (the cluster sets)
    double *h_d = nullptr, *h_e = nullptr, *h_z = nullptr; // pinned
    double *h_work = nullptr;
    int64_t *h_iwork = nullptr;
    i64 lwork = 0, liwork = 0;
    void *pool = nullptr;
    DcState<double> *dc = nullptr; // device D&C (FP64), created on first use
    hipStream_t dc_stream = nullptr;
};

void mixed_free(MixedBuffers *m) {
    if (!m) return;
    if (m->dc_stream) (void)hipStreamSynchronize(m->dc_stream);
    dc_state_destroy(m->dc);
    if (m->dc_stream) (void)hipStreamDestroy(m->dc_stream);
    if (m->pool) (void)hipFree(m->pool);
    if (m->h_d) (void)hipHostFree(m->h_d);
    if (m->h_e) (void)hipHostFree(m->h_e);
    if (m->h_z) (void)hipHostFree(m->h_z);
    std::free(m->h_work);
    std::free(m->h_iwork);
    delete m;
}

int mixed_alloc(MixedBuffers *&out, int n) {
    auto *m = new MixedBuffers();
    m->n = n;
    (void)hipGetDevice(&m->device);
    const size_t nn = (size_t)n * n;
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t bytes = 3 * al(nn * 8) + al(n * 8) + al(3 * kRedBlocks * 8) + al(64) + 3 * al(n * 4);
    if (hipMalloc(&m->pool, bytes) != hipSuccess) {
        m->pool = nullptr;
        mixed_free(m);
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "mixed: hipMalloc(%zu) failed", bytes);
    }
    auto *p = static_cast<unsigned char *>(m->pool);
    m->T1 = reinterpret_cast<double *>(p);
    p += al(nn * 8);
    m->S = reinterpret_cast<double *>(p);
    p += al(nn * 8);
    m->G = reinterpret_cast<double *>(p);
    p += al(nn * 8);
    m->lam = reinterpret_cast<double *>(p);
    p += al(n * 8);
    m->part = reinterpret_cast<double *>(p);
    p += al(3 * kRedBlocks * 8);
    m->count = reinterpret_cast<unsigned *>(p);
    p += al(64);
    m->perm = reinterpret_cast<int *>(p);
    p += al(n * 4);
    m->rank = reinterpret_cast<int *>(p);
    p += al(n * 4);
    m->hi = reinterpret_cast<int *>(p);
    if (hipHostMalloc(&m->h_d, n * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&m->h_e, n * 8, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(&m->h_z, nn * 8, hipHostMallocDefault) != hipSuccess) {
        mixed_free(m);
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "mixed: pinned host allocation failed");
    }
    m->lwork = stedc_lwork(n);
    m->liwork = stedc_liwork(n);
    m->h_work = static_cast<double *>(std::malloc(8 * (size_t)m->lwork));
    m->h_iwork = static_cast<int64_t *>(std::malloc(8 * (size_t)m->liwork));
    if (!m->h_work || !m->h_iwork) {
        mixed_free(m);
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "mixed: host allocation failed");
    }
    out = m;
    return CUEV_STATUS_SUCCESS;
}

std::mutex g_mix_mu;
MixedBuffers *g_mix_idle[64] = {};

int mixed_acquire(MixedBuffers *&m, int n) {
    int dev = 0;
    CUEV_HIP(hipGetDevice(&dev));
    MixedBuffers *victim = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mix_mu);
        MixedBuffers *&slot = g_mix_idle[dev & 63];
        if (slot && slot->n == n && slot->device == dev) {
            m = slot;
            slot = nullptr;
            return CUEV_STATUS_SUCCESS;
        }
        victim = slot;
        slot = nullptr;
    }
    if (victim) {
        (void)hipDeviceSynchronize();
        mixed_free(victim);
    }
    return mixed_alloc(m, n);
}

} // namespace

void mixed_release_all() {
    MixedBuffers *victims[64] = {};
    {
        std::lock_guard<std::mutex> lk(g_mix_mu);
        for (int i = 0; i < 64; ++i) {
            victims[i] = g_mix_idle[i];
            g_mix_idle[i] = nullptr;
        }
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (MixedBuffers *v : victims)
        if (v) {
            (void)hipSetDevice(v->device);
            (void)hipDeviceSynchronize();
            mixed_free(v);
        }
    (void)hipSetDevice(cur);
}

namespace {

void mixed_release(MixedBuffers *m) {
    MixedBuffers *victim = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mix_mu);
        MixedBuffers *&slot = g_mix_idle[m->device & 63];
        victim = slot;
        slot = m;
    }
    if (victim) mixed_free(victim);
}

using clk = std::chrono::steady_clock;
double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

double sum_parts(const std::vector<double> &h, int which) {
    double s = 0;
    for (int b = 0; b < kRedBlocks; ++b) s += h[(size_t)which * kRedBlocks + b];
    return s;
}

// Rayleigh-Ritz steps on cluster sets of columns of X: idx holds the sets one after
// another, set c at [start[c], start[c+1]). From X_J (gathered into mb.S), A X_J (into
// F) gives the lower triangles of S_JJ = X_J^T A X_J (mb.G) and G_JJ = X_J^T X_J (F,
// over A X_J); each set's step uses its diagonal blocks. (A Rayleigh-Ritz step on a
// union of clusters is one on a larger subspace, so it resolves every cluster in it.)
//  * on_gpu false: host dsygvd solves S_JJ w = theta G_JJ w on each set's block, so
//    the rotated columns also come out G-orthonormal, and X_J <- X_J W with W block
//    diagonal.
//  * on_gpu true (one set): the step follows a Newton step, so G_JJ = I + E with E at
//    that step's orthogonality error (~1e-9): Z = I - E/2 makes X_J Z orthonormal up
//    to 3/4 E^2, which the next Newton step removes. C = Z S_JJ Z is formed to the
//    same order as S_JJ - (E S_JJ + (E S_JJ)^T)/2 and solved by the FP64 two-stage
//    solver of this library, C = W~ Theta W~^T, and X_J <- X_J (Z W~) with
//    Z W~ = W~ - E W~ / 2.
// A is stored in full; F and mb.S, mb.G (n x n each) are scratch, mb.lam takes Theta.
int rayleigh_ritz(MixedBuffers &mb, hipStream_t s, const double *A, double *X, double *F, int n,
                  const std::vector<int> &idx, const std::vector<int> &start, bool on_gpu, Scratch gs) {
    const int m = (int)idx.size();
    const i64 mm = (i64)m * m;
    CUEV_HIP(hipMemcpyAsync(mb.perm, idx.data(), 4 * (size_t)m, hipMemcpyHostToDevice, s));
    const unsigned gcol = (unsigned)ceil_div((i64)n * m, 256);
    double *XJ = mb.S, *Sjj = mb.G, *Gjj = F;
    hipLaunchKernelGGL(gather_cols_kernel, dim3(gcol), dim3(256), 0, s, X, XJ, (i64)n, n, mb.perm, m);
    CUEV_LAUNCH_CHECK();
    CUEV_TRY(gemm<double>(s, 0, 0, n, m, n, 1.0, A, n, XJ, n, 0.0, F, n, gs));
    CUEV_TRY(gemm_ex<double>(s, m, m, n, 1.0, opnd(XJ, n, 1), opnd(static_cast<const double *>(F), n, 0), 0.0,
                             Sjj, m, 1, gs));
    CUEV_TRY(gemm_ex<double>(s, m, m, n, 1.0, opnd(XJ, n, 1), opnd(XJ, n, 0), 0.0, Gjj, m, 1, gs));
    if (!on_gpu) {
        const LapackFns *L = lapack();
        if (!L || !L->dsygvd)
            CUEV_FAIL(CUEV_STATUS_LAPACK_UNAVAILABLE, "mixed: host dsygvd needed for clustered eigenvalues");
        std::vector<double> sh((size_t)mm), gh((size_t)mm), W((size_t)mm, 0.0), sb, gb, wv, work;
        std::vector<int> iwork;
        CUEV_HIP(hipMemcpyAsync(sh.data(), Sjj, 8 * (size_t)mm, hipMemcpyDeviceToHost, s));
        CUEV_HIP(hipMemcpyAsync(gh.data(), Gjj, 8 * (size_t)mm, hipMemcpyDeviceToHost, s));
        CUEV_HIP(hipStreamSynchronize(s));
        for (size_t c = 0; c + 1 < start.size(); ++c) {
            const int a = start[c], sz = start[c + 1] - start[c];
            sb.assign((size_t)sz * sz, 0.0);
            gb.assign((size_t)sz * sz, 0.0);
            for (int q = 0; q < sz; ++q)
                for (int r = q; r < sz; ++r) {
                    sb[r + (size_t)q * sz] = sh[(a + r) + (size_t)(a + q) * m];
                    gb[r + (size_t)q * sz] = gh[(a + r) + (size_t)(a + q) * m];
                }
            wv.assign(sz, 0.0);
            const int itype = 1;
            const char jobz = 'V', uplo = 'L';
            int nn = sz, lda = sz, info = 0;
            int lwork = 1 + 6 * sz + 2 * sz * sz, liwork = 3 + 5 * sz;
            work.resize(lwork);
            iwork.resize(liwork);
            L->dsygvd(&itype, &jobz, &uplo, &nn, sb.data(), &lda, gb.data(), &lda, wv.data(), work.data(), &lwork,
                      iwork.data(), &liwork, &info);
            if (info != 0) CUEV_FAIL(CUEV_STATUS_LAPACK_ERROR, "mixed: dsygvd on a cluster set returned %d", info);
            for (int q = 0; q < sz; ++q)
                for (int r = 0; r < sz; ++r) W[(a + r) + (size_t)(a + q) * m] = sb[r + (size_t)q * sz];
        }
        // W into mb.G, X_J W into F, scattered back into X
        CUEV_HIP(hipMemcpyAsync(mb.G, W.data(), 8 * (size_t)mm, hipMemcpyHostToDevice, s));
    } else {
        const unsigned gsub = (unsigned)ceil_div(mm, 256), mt = (unsigned)ceil_div(m, 64);
        // S_JJ and E = G_JJ - I in full; E S_JJ into mb.S (X_J is gathered again below);
        // C over the lower triangle of S_JJ
        hipLaunchKernelGGL(symmetrize_lower_kernel, dim3(mt, mt), dim3(256), 0, s, Sjj, (i64)m, m);
        CUEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(symmetrize_lower_kernel, dim3(mt, mt), dim3(256), 0, s, Gjj, (i64)m, m);
        CUEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(sub_identity_kernel, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, s, Gjj, (i64)m);
        CUEV_LAUNCH_CHECK();
        const double *E = Gjj;
        CUEV_TRY(gemm<double>(s, 0, 0, m, m, m, 1.0, E, m, Sjj, m, 0.0, mb.S, m, gs));
        hipLaunchKernelGGL(rr_congruence_kernel, dim3(gsub), dim3(256), 0, s, Sjj, mb.S, (i64)m);
        CUEV_LAUNCH_CHECK();
        // C = W~ Theta W~^T: W~ into mb.S, Theta into mb.lam (C is consumed)
        {
            int b = 0, k = 0;
            CUEV_TRY(resolve_blocking(m, b, k));
            if (k > m) k = (int)round_up(m, b);
            Context<double> *ctx = nullptr;
            CUEV_TRY(context_acquire<double>(ctx, m, b, k, s));
            const int st = solve_native<double>(*ctx, Sjj, mb.lam, mb.S, nullptr, CUEV_BT_AUTO, CUEV_DC_AUTO, 0);
            if (st == CUEV_STATUS_SUCCESS) {
                context_release<double>(ctx);
            } else {
                (void)hipStreamSynchronize(s);
                context_destroy<double>(ctx);
                return st;
            }
        }
        // W = W~ - E W~ / 2 into mb.G
        CUEV_TRY(copy2d<double>(s, mb.S, m, mb.G, m, m, m));
        CUEV_TRY(gemm<double>(s, 0, 0, m, m, m, -0.5, E, m, mb.S, m, 1.0, mb.G, m, gs));
        hipLaunchKernelGGL(gather_cols_kernel, dim3(gcol), dim3(256), 0, s, X, XJ, (i64)n, n, mb.perm, m);
        CUEV_LAUNCH_CHECK();
    }
    // X_J W into F, scattered back into X
    CUEV_TRY(gemm<double>(s, 0, 0, n, m, m, 1.0, XJ, n, mb.G, m, 0.0, F, n, gs));
    hipLaunchKernelGGL(scatter_cols_kernel, dim3(gcol), dim3(256), 0, s, F, X, (i64)n, n, mb.perm, m);
    CUEV_LAUNCH_CHECK();
    CUEV_HIP(hipStreamSynchronize(s));
    return CUEV_STATUS_SUCCESS;
}

// The Rayleigh-Ritz steps of a list of cluster sets: each set of kRRDeviceMin columns
// or more on the GPU on its own, the smaller ones together (one pass of products over
// their union, a host solve per set); cols receives the number of columns treated.
int rayleigh_ritz_sets(MixedBuffers &mb, hipStream_t s, const double *A, double *X, double *F, int n,
                       const std::vector<std::vector<int>> &sets, Scratch gs, int &cols) {
    std::vector<int> small, start{0};
    cols = 0;
    for (const auto &J : sets) {
        cols += (int)J.size();
        if ((int)J.size() >= kRRDeviceMin) {
            const std::vector<int> one{0, (int)J.size()};
            CUEV_TRY(rayleigh_ritz(mb, s, A, X, F, n, J, one, true, gs));
        } else {
            small.insert(small.end(), J.begin(), J.end());
            start.push_back((int)small.size());
        }
    }
    if (!small.empty()) CUEV_TRY(rayleigh_ritz(mb, s, A, X, F, n, small, start, false, gs));
    return CUEV_STATUS_SUCCESS;
}

int run_mixed(Context<float> &cf, MixedBuffers &mb, double *A, int n, double *eval, double *evec,
              cuev_solve_timer *timer, const cuev_options &opt) {
    Workspace<float> &ws = cf.ws;
    hipStream_t s = ws.stream;
    Scratch gs32{ws.gscratch, ws.gscratch_bytes};
    const auto wall0 = clk::now();
    hipEvent_t *ev = cf.ev;

    // 1. FP32 reduction (A32 lives in ws.M until the band is packed)
    CUEV_TRY(bt_ensure(cf));
    hipStream_t bs = cf.bt_stream;
    CUEV_HIP(hipEventRecord(ev[0], s));
    CUEV_TRY((convert2d<double, float>(s, A, n, ws.M, ws.ld, n, n)));
    CUEV_TRY(dbbr_reduce<float>(ws, ws.M, ws.B));
    CUEV_HIP(hipEventRecord(ev[1], s));
    // 2. M = Q_b^T Q_s^T in FP32 on the back-transform stream: Q_s^T beside the
    // bulge chasing, Q_b^T after it, beside the FP64 D&C
    CUEV_HIP(hipStreamWaitEvent(bs, ev[1], 0));
    CUEV_HIP(hipEventRecord(cf.xev[0], bs));
    CUEV_TRY(sbr_back_t<float>(ws, bs, ws.Y, ws.W, ws.M, ws.ld));
    CUEV_HIP(hipEventRecord(cf.xev[1], bs));
    CUEV_TRY(bc_chase<float>(ws, ws.B, ws.d, ws.e));
    CUEV_HIP(hipEventRecord(ev[2], s));
    CUEV_HIP(hipMemcpyAsync(cf.h_d, ws.d, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipMemcpyAsync(cf.h_e, ws.e, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipEventRecord(ev[3], s));
    CUEV_HIP(hipStreamWaitEvent(bs, ev[2], 0));
    CUEV_HIP(hipEventRecord(cf.xev[2], bs));
    CUEV_TRY(bc_back_t<float>(ws, bs, ws.U, ws.M, ws.ld));
    CUEV_HIP(hipEventRecord(cf.xev[3], bs));
    CUEV_HIP(hipEventSynchronize(ev[3]));
    for (int i = 0; i < n; ++i) {
        mb.h_d[i] = cf.h_d[i];
        mb.h_e[i] = (i < n - 1) ? cf.h_e[i] : 0.0;
        if (!std::isfinite(mb.h_d[i]) || !std::isfinite(mb.h_e[i]))
            CUEV_FAIL(CUEV_STATUS_NOT_FINITE, "mixed: non-finite tridiagonal entry at %d", i);
    }
    const int dc_mode = opt.dc_mode == CUEV_DC_AUTO ? (n >= 1024 ? CUEV_DC_DEVICE : CUEV_DC_HOST)
                                                   : opt.dc_mode;
    if (dc_mode == CUEV_DC_DEVICE) {
        if (!mb.dc) CUEV_TRY(dc_state_create<double>(mb.dc, n));
        if (!mb.dc_stream) CUEV_HIP(hipStreamCreateWithFlags(&mb.dc_stream, hipStreamNonBlocking));
    }
    const auto dc0 = clk::now();
    if (dc_mode == CUEV_DC_DEVICE) {
        // Q_d into S and the eigenvalues into lam, on a stream of its own beside the
        // FP32 back-transform (small tiles, no split-K; the large merges wait for
        // the back-transform); T1, G and evec (written last) are its work buffers
        Scratch dcs;
        dcs.small_tiles = 1;
        CUEV_TRY(dc_solve<double>(*mb.dc, mb.dc_stream, n, mb.h_d, mb.h_e, mb.S, n, mb.lam, mb.T1, mb.G,
                                  evec, dcs, opt.dc_leaf, cf.xev[3], kDcGateFrom(n)));
    } else {
        CUEV_TRY(stedc<double>(n, mb.h_d, mb.h_e, mb.h_z, n, mb.h_work, mb.lwork, mb.h_iwork, mb.liwork));
    }
    const auto dc1 = clk::now();
    // X = (double)(M^T (float)Q_d): Q_d through S (FP64) into W (FP32), and the
    // product into Y, both free once the back-transform stream is done
    if (dc_mode != CUEV_DC_DEVICE) {
        CUEV_HIP(hipMemcpyAsync(mb.S, mb.h_z, 8 * (size_t)n * n, hipMemcpyHostToDevice, s));
        CUEV_HIP(hipMemcpyAsync(mb.lam, mb.h_d, 8 * (size_t)n, hipMemcpyHostToDevice, s));
    }
    CUEV_HIP(hipStreamWaitEvent(s, cf.xev[3], 0));
    CUEV_HIP(hipEventRecord(ev[4], s));
    CUEV_TRY((convert2d<double, float>(s, mb.S, n, ws.W, ws.ld, n, n)));
    {
        prof::Scope sc(s, "mixed.bt_gemm32", 2.0 * (double)n * n * n);
        CUEV_TRY(gemm<float>(s, 1, 0, n, n, n, 1.f, ws.M, ws.ld, ws.W, ws.ld, 0.f, ws.Y, ws.ld, gs32));
    }
    CUEV_TRY((convert2d<float, double>(s, ws.Y, ws.ld, evec, n, n, n)));
    CUEV_HIP(hipEventRecord(ev[6], s));

    // 3. FP64 refinement (X in evec, A original in A)
    // FP64 products share the FP32 workspace's split-K scratch (the engine limits
    // the split count to what fits)
    Scratch gs64{ws.gscratch, ws.gscratch_bytes};
    const int max_sweeps = opt.refine_sweeps > 0 ? opt.refine_sweeps : 3;
    const double tol = opt.refine_tol > 0 ? opt.refine_tol : 2e-14;
    std::vector<double> hp(3 * (size_t)kRedBlocks);
    hipLaunchKernelGGL(sym_frob_kernel, dim3(kRedBlocks), dim3(256), 0, s, A, (i64)n, n, mb.part);
    CUEV_LAUNCH_CHECK();
    CUEV_HIP(hipMemcpyAsync(hp.data(), mb.part, 8 * (size_t)kRedBlocks, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipStreamSynchronize(s));
    const double afro = std::sqrt(sum_parts(hp, 0));
    // A in full (the solve overwrites A; the refinement reads it only as a whole)
    {
        const unsigned nt = (unsigned)ceil_div(n, 64);
        hipLaunchKernelGGL(symmetrize_lower_kernel, dim3(nt, nt), dim3(256), 0, s, A, (i64)n, n);
        CUEV_LAUNCH_CHECK();
    }
    const double anorm = afro > 0 ? afro : 1.0;
    // the rounding level of an entry of S = X^T A X: n u ||A||_F
    const double nu = (double)n * 0x1p-53 * anorm;
    // X alternates between evec and T1; the other buffer holds A X, then the next iterate
    double *X = evec;
    // sweeps counts every evaluation-and-update (Newton and Rayleigh-Ritz); the
    // limit applies to the Newton steps
    int sweeps = 0, oa_sweeps = 0, nclus_cols = 0;
    bool did_rr = false, probe = false;
    // cluster sets of the Rayleigh-Ritz steps due right after the next Newton step
    std::vector<std::vector<int>> rr_sets;
    double res = 0, orth = 0;
    unsigned nclus = 0;
    const auto ref0 = clk::now();
    for (;;) {
        double *other = (X == evec) ? mb.T1 : evec;
        if (!rr_sets.empty()) {
            int cols = 0;
            CUEV_TRY(rayleigh_ritz_sets(mb, s, A, X, other, n, rr_sets, gs64, cols));
            nclus_cols = std::max(nclus_cols, cols);
            rr_sets.clear();
            did_rr = true;
            ++sweeps;
            continue;
        }
        // After a Newton step, first a randomized check of both norms: with r of
        // independent standard normal entries, E ||M r||^2 = ||M||_F^2, and the sum
        // over kProbes probes is at worst (M of rank one) ||M||_F^2 times a chi-square
        // with kProbes degrees of freedom, so an estimate a factor 4 below the truth
        // has a probability below 1e-7 at 16 probes. The eigenvalue estimates are those of
        // the last evaluation: the step changes them only in second order. Both
        // estimates below a quarter of the tolerance end the refinement without the
        // full evaluation; otherwise it follows. S and G are free here (6 n kProbes
        // doubles of S are used).
        if (probe && n >= 256) {
            prof::Scope sc(s, "mixed.refine_check", 8.0 * (double)n * n * kProbes);
            const i64 nk = (i64)n * kProbes;
            double *R = mb.S, *LR = R + nk, *Y = LR + nk, *AY = Y + nk, *XLR = AY + nk, *XtY = XLR + nk;
            hipLaunchKernelGGL(probe_kernel, dim3((unsigned)ceil_div(nk, 256)), dim3(256), 0, s, R, LR, mb.lam, n);
            CUEV_LAUNCH_CHECK();
            CUEV_TRY(gemm<double>(s, 0, 0, n, kProbes, n, 1.0, X, n, R, n, 0.0, Y, n, gs64));
            CUEV_TRY(gemm<double>(s, 0, 0, n, kProbes, n, 1.0, A, n, Y, n, 0.0, AY, n, gs64));
            CUEV_TRY(gemm<double>(s, 0, 0, n, kProbes, n, 1.0, X, n, LR, n, 0.0, XLR, n, gs64));
            CUEV_TRY(gemm<double>(s, 1, 0, n, kProbes, n, 1.0, X, n, Y, n, 0.0, XtY, n, gs64));
            hipLaunchKernelGGL(probe_norms_kernel, dim3(kRedBlocks), dim3(256), 0, s, AY, XLR, XtY, R, nk, mb.part);
            CUEV_LAUNCH_CHECK();
            CUEV_HIP(hipMemcpyAsync(hp.data(), mb.part, 8 * 2 * (size_t)kRedBlocks, hipMemcpyDeviceToHost, s));
            CUEV_HIP(hipStreamSynchronize(s));
            const double est_res = std::sqrt(sum_parts(hp, 0) / kProbes) / ((double)n * anorm);
            const double est_orth = std::sqrt(sum_parts(hp, 1) / kProbes) / n;
            if (est_res <= 0.25 * tol && est_orth <= 0.25 * tol) {
                res = est_res;
                orth = est_orth;
                break;
            }
        }
        probe = false;
        // A X (2n^3), the lower triangles of G and S (n^3 each; S only when the
        // refinement goes on), X E (2n^3)
        prof::Scope sc(s, "mixed.refine_sweep", 6.0 * (double)n * n * n);
        // AX = A X into the buffer not holding X, G = X^T X (lower triangle),
        // lam from diag(X^T AX) and diag(G)
        CUEV_TRY(gemm<double>(s, 0, 0, n, n, n, 1.0, A, n, X, n, 0.0, other, n, gs64));
        CUEV_TRY(gemm_ex<double>(s, n, n, n, 1.0, opnd(X, n, 1), opnd(X, n, 0), 0.0, mb.G, n, 1, gs64));
        hipLaunchKernelGGL(oa_lambda_kernel, dim3((unsigned)n), dim3(256), 0, s, X, other, mb.G, (i64)n, n,
                           mb.lam);
        CUEV_LAUNCH_CHECK();
        hipLaunchKernelGGL(oa_norms_kernel, dim3(kRedBlocks), dim3(256), 0, s, mb.G, other, X, (i64)n, n,
                           mb.lam, mb.part);
        CUEV_LAUNCH_CHECK();
        CUEV_HIP(hipMemcpyAsync(hp.data(), mb.part, 8 * hp.size(), hipMemcpyDeviceToHost, s));
        CUEV_HIP(hipStreamSynchronize(s));
        const double r_fro = std::sqrt(sum_parts(hp, 1));
        res = std::sqrt(sum_parts(hp, 2)) / ((double)n * anorm);
        orth = r_fro / n;
        if ((res <= tol && orth <= tol) || oa_sweeps >= max_sweeps) break;
        // S = X^T AX (lower triangle: the element-wise kernels read it so)
        CUEV_TRY(gemm_ex<double>(s, n, n, n, 1.0, opnd(X, n, 1), opnd(other, n, 0), 0.0, mb.S, n, 1, gs64));

        // The columns of the clustered pairs. The Newton step only orthogonalises
        // them against each other, so they stay at the FP32 start's accuracy within
        // their clusters, and one Rayleigh-Ritz step on their union, after that step,
        // resolves them in a basis it has refined everywhere else. Before the first
        // Newton step they are taken from this evaluation (the step leaves the same
        // pairs clustered), and the Rayleigh-Ritz step follows the step directly;
        // later, after a Newton step, from the evaluation that follows it. S and G
        // are recomputed for the rotated basis.
        if (!did_rr) {
            // ranks of lam; each clustered column's partner ranks
            std::vector<double> hl(n);
            CUEV_HIP(hipMemcpyAsync(hl.data(), mb.lam, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
            CUEV_HIP(hipStreamSynchronize(s));
            std::vector<int> ord(n), rank(n), hi(n);
            std::iota(ord.begin(), ord.end(), 0);
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return hl[a] < hl[b]; });
            for (int p = 0; p < n; ++p) rank[ord[p]] = p;
            CUEV_HIP(hipMemcpyAsync(mb.rank, rank.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
            CUEV_TRY(fill_int(s, mb.hi, n, -1));
            hipLaunchKernelGGL(oa_flag_kernel, dim3((unsigned)ceil_div((i64)n * n, 256)), dim3(256), 0, s,
                               mb.S, mb.G, (i64)n, n, mb.lam, nu, mb.rank, mb.hi);
            CUEV_LAUNCH_CHECK();
            CUEV_HIP(hipMemcpyAsync(hi.data(), mb.hi, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
            CUEV_HIP(hipStreamSynchronize(s));
            // Cluster sets: the rank intervals [rank, max(rank, hi)] of the clustered
            // columns, merged where they overlap (a pair is seen from both ends, so an
            // earlier partner's interval already reaches a column). No clustered pair
            // joins two sets, so each takes a Rayleigh-Ritz step of its own.
            std::vector<std::vector<int>> sets;
            int end = -1;
            for (int p = 0; p < n; ++p) {
                const int i = ord[p];
                if (hi[i] < 0) continue;
                if (p > end) sets.emplace_back();
                sets.back().push_back(i);
                end = std::max(end, std::max(p, hi[i]));
            }
            if (oa_sweeps == 0) {
                rr_sets = std::move(sets);
            } else if (!sets.empty()) {
                int cols = 0;
                CUEV_TRY(rayleigh_ritz_sets(mb, s, A, X, other, n, sets, gs64, cols));
                nclus_cols = std::max(nclus_cols, cols);
                did_rr = true;
                ++sweeps;
                continue;
            }
        }
        did_rr = false;
        // E over G, then X <- X + X E
        CUEV_TRY(fill_int(s, reinterpret_cast<int *>(mb.count), 1, 0));
        hipLaunchKernelGGL(oa_e_kernel, dim3((unsigned)ceil_div((i64)n * n, 256)), dim3(256), 0, s,
                           mb.S, mb.G, (i64)n, n, mb.lam, nu, mb.count);
        CUEV_LAUNCH_CHECK();
        // next iterate X + X E into `other` (A X is no longer needed)
        CUEV_TRY(copy2d<double>(s, X, n, other, n, n, n));
        CUEV_TRY(gemm<double>(s, 0, 0, n, n, n, 1.0, X, n, mb.G, n, 1.0, other, n, gs64));
        CUEV_HIP(hipMemcpyAsync(&nclus, mb.count, 4, hipMemcpyDeviceToHost, s));
        X = other;
        ++sweeps;
        ++oa_sweeps;
        probe = rr_sets.empty(); // no check when the Rayleigh-Ritz step follows
    }
    const auto ref1 = clk::now();
    // eigenvalues ascending; columns follow
    std::vector<double> hl(n);
    CUEV_HIP(hipMemcpyAsync(hl.data(), mb.lam, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipStreamSynchronize(s));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return hl[a] < hl[b]; });
    bool sorted = true;
    for (int i = 0; i < n; ++i) sorted &= perm[i] == i;
    std::vector<double> hs(n);
    for (int i = 0; i < n; ++i) hs[i] = hl[perm[i]];
    CUEV_HIP(hipMemcpyAsync(eval, hs.data(), 8 * (size_t)n, hipMemcpyHostToDevice, s));
    if (!sorted || X != evec) {
        double *dst = (X != evec) ? evec : mb.T1;
        if (!sorted) {
            CUEV_HIP(hipMemcpyAsync(mb.perm, perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(permute_cols_kernel, dim3((unsigned)ceil_div((i64)n * n, 256)), dim3(256),
                               0, s, X, dst, (i64)n, n, mb.perm);
            CUEV_LAUNCH_CHECK();
        } else {
            CUEV_TRY(copy2d<double>(s, X, n, dst, n, n, n));
        }
        if (dst != evec) CUEV_TRY(copy2d<double>(s, dst, n, evec, n, n, n));
    }
    CUEV_HIP(hipStreamSynchronize(s));
    CUEV_TRY(workspace_check_error(ws));
    const auto wall1 = clk::now();
    if (timer) {
        std::memset(timer, 0, sizeof(*timer));
        auto em = [&](int a, int b) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
            return ms;
        };
        timer->dbbr_ms = em(0, 1);
        timer->bc_ms = em(1, 2);
        timer->dc_ms = (float)ms_between(dc0, dc1);
        auto emx = [&](int a, int b) {
            float ms = 0;
            (void)hipEventElapsedTime(&ms, cf.xev[a], cf.xev[b]);
            return ms;
        };
        timer->bt_qs_ms = emx(0, 1);
        timer->bt_qb_ms = emx(2, 3);
        timer->overlap_ms = timer->bt_qs_ms + timer->bt_qb_ms;
        timer->bt_copyout_ms = em(4, 6);
        timer->bt_ms = (float)ms_between(dc1, ref0);
        timer->refine_ms = (float)ms_between(ref0, ref1);
        timer->total_ms = (float)ms_between(wall0, wall1);
        timer->refine_sweeps = sweeps;
        timer->n_cluster_cols = nclus_cols > 0 ? nclus_cols : (int)nclus;
        timer->refine_resid = (float)res;
        timer->refine_orth = (float)orth;
    }
    // the sweep limit was reached above the tolerance: the outputs hold the last
    // iterate, and the caller is told it is not FP64-accurate
    if (!(res <= tol && orth <= tol))
        CUEV_FAIL(CUEV_STATUS_NOT_CONVERGED,
                  "mixed: %d refinement sweeps left residual %.3e and orthogonality %.3e above "
                  "refine_tol %.3e",
                  oa_sweeps, res, orth, tol);
    return CUEV_STATUS_SUCCESS;
}

} // namespace

int solve_mixed(double *A, int n, double *eval, double *evec, hipStream_t s, cuev_solve_timer *timer,
                const cuev_options &opt, int b, int k) {
    Context<float> *cf = nullptr;
    CUEV_TRY(context_acquire<float>(cf, n, b, k, s));
    MixedBuffers *mb = nullptr;
    int st = mixed_acquire(mb, n);
    if (st) {
        context_release<float>(cf);
        return st;
    }
    st = run_mixed(*cf, *mb, A, n, eval, evec, timer, opt);
    if (st == CUEV_STATUS_SUCCESS || st == CUEV_STATUS_NOT_CONVERGED) {
        context_release<float>(cf);
        mixed_release(mb);
    } else {
        (void)hipStreamSynchronize(s);
        context_destroy<float>(cf);
        mixed_free(mb);
    }
    return st;
}

} // namespace cuev
