// gemm.hip — MFMA tile engine for gfx950 (FP64 and FP32).
//
// Every dense product of the solver runs here: plain products (N/T operands),
// symmetric-from-lower A (symm), K-concatenation of two sources (the band
// reduction's corrections, green update and syr2k as one product each), lower-only
// output, split-K.
//
// Block tile BM x BN x BK with WM x WN wave64s, each owning a (BM/WM) x (BN/WN)
// sub-tile of 16 x 16 MFMA fragments (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32,
// FP32 accumulation for FP32). BK is 16 elements (a K-concatenation split must be a
// multiple of it).
//
// Operand staging: global -> registers -> LDS, double-buffered in both, one barrier
// per k-tile; the next k-tile's global loads are in flight while the current one
// computes. Global loads are 16 bytes wide (2 doubles / 4 floats) wherever the tile
// lies inside the operand and the operand is 16-byte aligned with ld a multiple of
// the vector width; edge tiles, symmetric tiles straddling the diagonal and
// unaligned operands take a per-element path with bounds checks into the same LDS
// layout. In LDS both operands are stored row-major with K contiguous ([m][k] and
// [n][k]), rows padded by 16 bytes so that the fragment reads of 16 consecutive
// rows hit distinct banks. A fragment read is one 16-byte ds_read per lane: lane
// (i = l%16, j = l/16) reads row i at k = E*j .. E*j+E-1 (E = 2 doubles / 4
// floats), which feeds E MFMAs at once. Each MFMA then contracts lane group j's
// k = E*j + e; A and B use the same k assignment, so the sum over the E MFMAs and
// the 4 lane groups covers every k of the group exactly once.
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
// Tall-skinny products (the band reduction's panel products with K = n - j - b,
// sbr_back's Y^T M) split K over blockIdx.y into a scratch buffer, summed in a
// fixed order by a second kernel, so results are deterministic.
#include "gemm.hpp"

#include <algorithm>
#include <cmath>

namespace cuev {
namespace {

constexpr int kGroupM = 8;

typedef __attribute__((address_space(1))) void gvoid_t; // global
typedef __attribute__((address_space(3))) void lvoid_t; // LDS

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    typedef double vec __attribute__((ext_vector_type(2)));
    static constexpr int E = 2;   // elements per 16-byte vector
    static constexpr int BK = 16; // k per tile
    __device__ static __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    typedef float vec __attribute__((ext_vector_type(4)));
    static constexpr int E = 4;
    static constexpr int BK = 16;
    __device__ static __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T> struct KParams {
    const T *a0, *a1;
    i64 lda0, lda1, asplit;
    const T *b0, *b1;
    i64 ldb0, ldb1, bsplit;
    T *C;
    i64 ldc;
    T *part; // split-K partials: part[s*M*N + m + n*M]
    i64 M, N, K, kchunk;
    T alpha, beta;
    int lower;
    int tiles_m, tiles_n;
    int avec, bvec; // 16-byte loads allowed (alignment of both sources and their ld)
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
// loaders it needs (register pressure decides how many waves share a SIMD).
// A_CAT: two m-contiguous sources split along k; A_TCAT: transposed, two k-contiguous
// sources split along m (rows [0, split) of op(A) from a0, the rest from a1), with the
// split a multiple of the tile height so that no tile straddles it
enum AMode { A_N = 0, A_T = 1, A_SYM = 2, A_CAT = 3, A_TCAT = 4 };
enum BMode { B_N = 0, B_T = 1, B_CATT = 2 };           // B_CATT: two n-contiguous sources

// Source layouts of one R x BK operand tile (rows r = m or n of the product, k).
//   KC: k contiguous, element (r, k) at p[k + r*ld]
//   RC: r contiguous, element (r, k) at p[r + k*ld]
//   SYM: symmetric, lower triangle stored: (r, k) at p[r + k*ld] if r >= k, else p[k + r*ld]
enum Layout { L_KC = 0, L_RC = 1, L_SYM = 2 };

// Registers holding one operand tile between its global load and its LDS store.
// Chunk c of the tile is E consecutive elements along the contiguous direction of
// the layout it was loaded with (KC: along k; RC and SYM: along r).
template <typename T, int R, int BK, int NT> struct TileRegs {
    using vec = typename Mfma<T>::vec;
    static constexpr int E = Mfma<T>::E;
    static constexpr int CH = R * BK / E / NT;
    static_assert(CH * E * NT == R * BK, "tile / thread count mismatch");
    static constexpr int KCH = BK / E; // KC: chunks per row
    static constexpr int RCH = R / E;  // RC: chunks per k column
    vec v[CH];

    // 16-byte loads of an interior tile (rows r0.., k0.. valid, source aligned)
    __device__ __forceinline__ void load_vec(int lay, const T *p, i64 ld, i64 r0, i64 k0, int tid) {
        if (lay == L_KC) {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int c = tid + q * NT;
                const int row = c / KCH, kc = c % KCH;
                v[q] = *reinterpret_cast<const vec *>(p + (r0 + row) * ld + k0 + kc * E);
            }
        } else {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int c = tid + q * NT;
                const int kk = c / RCH, rc = c % RCH;
                v[q] = *reinterpret_cast<const vec *>(p + (k0 + kk) * ld + r0 + rc * E);
            }
        }
    }
    // per-element loads with bounds (rows < rlim, k < klim read, the rest zero)
    __device__ __forceinline__ void load_scalar(int lay, const T *p, i64 ld, i64 r0, i64 k0, i64 rlim,
                                                i64 klim, int tid) {
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int c = tid + q * NT;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                i64 r, k;
                if (lay == L_KC) {
                    r = r0 + c / KCH;
                    k = k0 + (c % KCH) * E + e;
                } else {
                    r = r0 + (c % RCH) * E + e;
                    k = k0 + c / RCH;
                }
                T x = T(0);
                if (r < rlim && k < klim) {
                    const i64 idx = (lay == L_KC) ? k + r * ld
                                    : (lay == L_RC || r >= k) ? r + k * ld
                                                              : k + r * ld;
                    x = p[idx];
                }
                v[q][e] = x;
            }
        }
    }
    // into LDS s[row][k] (row stride RS elements)
    template <int RS> __device__ __forceinline__ void store(int lay, T *s, int tid) const {
        if (lay == L_KC) {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int c = tid + q * NT;
                *reinterpret_cast<vec *>(s + (c / KCH) * RS + (c % KCH) * E) = v[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                const int c = tid + q * NT;
                const int kk = c / RCH, r = (c % RCH) * E;
#pragma unroll
                for (int e = 0; e < E; ++e) s[(r + e) * RS + kk] = v[q][e];
            }
        }
    }
};

// Register-staged path: C tile BM x BN of C = alpha op(A) op(B) + beta C over the
// split's K range.
template <typename T, int BM, int BN, int WM, int WN, int AMODE, int BMODE>
__device__ __forceinline__ void gemm_body_regs(const KParams<T> &p) {
    constexpr int NT = 64 * WM * WN;
    constexpr int E = Mfma<T>::E;
    constexpr int BK = Mfma<T>::BK;
    constexpr int RS = BK + E; // LDS row stride in elements: 16 bytes of padding
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr int KG = 4 * E; // k covered by one fragment read
    using V4 = typename Mfma<T>::v4;
    using vec = typename Mfma<T>::vec;

    __shared__ __attribute__((aligned(16))) T sA[2][BM * RS];
    __shared__ __attribute__((aligned(16))) T sB[2][BN * RS];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;

    // XCD-contiguous remap of the block index
    const int nbk = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q8 = nbk >> 3, rmd = nbk & 7;
    const int t = xcd * q8 + min(xcd, rmd) + (bid >> 3);
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

    TileRegs<T, BM, BK, NT> ra;
    TileRegs<T, BN, BK, NT> rb;
    const bool m_in = m0 + BM <= p.M, n_in = n0 + BN <= p.N;
    int layA = L_RC, layB = L_KC;
    auto load_tiles = [&](i64 k0) {
        const bool k_in = k0 + BK <= kend;
        // ---- A
        {
            const T *pa = p.a0;
            i64 lda = p.lda0, ko = 0, mo = 0;
            bool fast = m_in && k_in && p.avec;
            if constexpr (AMODE == A_N) {
                layA = L_RC;
            } else if constexpr (AMODE == A_T) {
                layA = L_KC;
            } else if constexpr (AMODE == A_TCAT) {
                layA = L_KC;
                if (m0 >= p.asplit) {
                    pa = p.a1;
                    lda = p.lda1;
                    mo = p.asplit;
                }
            } else if constexpr (AMODE == A_SYM) {
                if (m0 >= k0 + BK - 1) {
                    layA = L_RC; // below the diagonal
                } else if (m0 + BM - 1 <= k0) {
                    layA = L_KC; // above: the transpose of stored lower entries
                } else {
                    layA = L_SYM;
                    fast = false;
                }
            } else {
                layA = L_RC;
                if (k0 >= p.asplit) {
                    pa = p.a1;
                    lda = p.lda1;
                    ko = p.asplit;
                }
            }
            if (fast) ra.load_vec(layA, pa, lda, m0 - mo, k0 - ko, tid);
            else ra.load_scalar(layA, pa, lda, m0 - mo, k0 - ko, p.M - mo, kend - ko, tid);
            if (layA == L_SYM) layA = L_RC; // stored with the RC chunk mapping
        }
        // ---- B (rows of the tile = n)
        {
            const T *pb = p.b0;
            i64 ldb = p.ldb0, ko = 0;
            const bool fast = n_in && k_in && p.bvec;
            if constexpr (BMODE == B_N) {
                layB = L_KC;
            } else if constexpr (BMODE == B_T) {
                layB = L_RC;
            } else {
                layB = L_RC;
                if (k0 >= p.bsplit) {
                    pb = p.b1;
                    ldb = p.ldb1;
                    ko = p.bsplit;
                }
            }
            if (fast) rb.load_vec(layB, pb, ldb, n0, k0 - ko, tid);
            else rb.load_scalar(layB, pb, ldb, n0, k0 - ko, p.N, kend - ko, tid);
        }
    };

    if (nkt > 0) {
        load_tiles(kbeg);
        ra.template store<RS>(layA, sA[0], tid);
        rb.template store<RS>(layB, sB[0], tid);
    }
    __syncthreads();

    const int li = lane & 15, lj = lane >> 4;
    for (int kt = 0; kt < nkt; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nkt;
        if (more) load_tiles(kbeg + (i64)(kt + 1) * BK);
        const T *a_s = sA[cur] + (wm * TM + li) * RS + lj * E;
        const T *b_s = sB[cur] + (wn * TN + li) * RS + lj * E;
#pragma unroll
        for (int kg = 0; kg < BK / KG; ++kg) {
            vec af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = *reinterpret_cast<const vec *>(a_s + i * 16 * RS + kg * KG);
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = *reinterpret_cast<const vec *>(b_s + j * 16 * RS + kg * KG);
#pragma unroll
            for (int e = 0; e < E; ++e)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<T>::mma(bf[j][e], af[i][e], acc[i][j]);
        }
        if (more) {
            ra.template store<RS>(layA, sA[cur ^ 1], tid);
            rb.template store<RS>(layB, sB[cur ^ 1], tid);
        }
        __syncthreads();
    }

    // ---- epilogue
    const bool partial = p.part != nullptr;
    T *part = partial ? p.part + (i64)blockIdx.y * p.M * p.N : nullptr;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const i64 m = m0 + wm * TM + i * 16 + li;
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


// ---------------------------------------------------------------------------
// LDS-DMA path (FP64, block tiles of at least 128 x 128): interior tiles of aligned
// operands go straight from global memory into LDS with global_load_lds_dwordx4 (no
// register staging, no LDS store instructions), the next k-tile's transfer in flight
// while the current one computes. Layouts:
//   KC operand (k contiguous): [row][k], 8 chunks of 16 bytes per row of BK = 16
//     doubles, chunk c of row r at slot c ^ f(r), f(r) = (r ^ (r >> 3)) & 7, so the
//     fragment reads of 16 consecutive rows are spread over all banks (no padding, so
//     one DMA instruction fills 1 KB of consecutive LDS);
//   RC operand (rows contiguous): [k][row] with a row stride of R + 8 doubles, which
//     puts the k and k+2 reads of one fragment read in different bank halves.
// Edge tiles, K tails and unaligned operands are loaded per element (bounds-checked)
// into the same layouts.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int kc_swz(int r) { return (r ^ (r >> 3)) & 7; }

template <int R> struct DmaTile {
    static constexpr int BK = 16;
    static constexpr int RSTR = R + 8; // RC layout: doubles per k row
    template <bool KC> static constexpr int bytes() { return KC ? R * BK * 8 : BK * RSTR * 8; }
};

// DMA of one interior R x 16 tile; the LDS destination of each instruction is wave-uniform.
template <int R, int NT>
__device__ __forceinline__ void dma_tile(bool kc, const double *p, i64 ld, i64 r0, i64 k0, char *s, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = NT / 64;
    if (kc) {
        constexpr int NI = R * 8 / 64 / NW; // R rows x 8 chunks, 64 chunks per instruction
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int inst = wave * NI + i;
            const int pos = inst * 64 + lane;
            const int row = pos >> 3, c = (pos & 7) ^ kc_swz(row);
            const double *g = p + (r0 + row) * ld + k0 + 2 * c;
            __builtin_amdgcn_global_load_lds((gvoid_t *)g, (lvoid_t *)(s + inst * 1024), 16, 0, 0);
        }
    } else if constexpr (R >= 128) {
        // (row-contiguous tiles of fewer than 128 rows never take this path: the
        // 64-row DMA tiles are k-contiguous B operands only)
        constexpr int PER_K = R / 128; // 1 KB instructions per k row
        constexpr int NI = 16 * PER_K / NW;
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            const int inst = wave * NI + i;
            const int kk = inst / PER_K, part = inst % PER_K;
            const double *g = p + (k0 + kk) * ld + r0 + part * 128 + 2 * lane;
            __builtin_amdgcn_global_load_lds((gvoid_t *)g,
                                             (lvoid_t *)(s + ((size_t)kk * DmaTile<R>::RSTR + part * 128) * 8), 16,
                                             0, 0);
        }
    }
}

// Per-element load of an R x 16 tile (rows < rlim, k < klim; zero elsewhere) into the
// DMA layouts (synchronous: used for edge tiles, K tails and unaligned operands).
// sym: p holds the lower triangle of a symmetric matrix, (r, k) read at p[r + k*ld]
// if r >= k, else p[k + r*ld], into either layout.
template <int R, int NT>
__device__ __forceinline__ void scalar_tile(bool kc, const double *p, i64 ld, i64 r0, i64 k0, i64 rlim, i64 klim,
                                            char *s, int tid, bool sym = false) {
    double *d = reinterpret_cast<double *>(s);
    constexpr int PER = R * 16 / NT;
    // one element at a time (edge tiles only): few live registers beside the accumulators
#pragma unroll 1
    for (int q = 0; q < PER; ++q) {
        const int idx = tid + q * NT;
        int row, kk;
        if (kc) {
            row = idx >> 4;
            kk = idx & 15;
        } else {
            row = idx % R;
            kk = idx / R;
        }
        const i64 r = r0 + row, k = k0 + kk;
        const i64 off = sym ? ((r >= k) ? r + k * ld : k + r * ld) : (kc ? k + r * ld : r + k * ld);
        const double v = (r < rlim && k < klim) ? p[off] : 0.0;
        if (kc) d[(row * 8 + ((kk >> 1) ^ kc_swz(row))) * 2 + (kk & 1)] = v;
        else d[kk * DmaTile<R>::RSTR + row] = v;
    }
}

// fragment read: lane (i, j) gets rows rb + i at k = 8 kg + 2j + {0, 1}
template <int R>
__device__ __forceinline__ Mfma<double>::vec dma_frag(bool kc, const char *s, int rb, int kg, int li, int lj) {
    if (kc) {
        const int row = rb + li, c = kg * 4 + lj;
        return *reinterpret_cast<const Mfma<double>::vec *>(s + (row * 8 + (c ^ kc_swz(row))) * 16);
    }
    const double *d = reinterpret_cast<const double *>(s);
    const int k = kg * 8 + 2 * lj;
    Mfma<double>::vec v;
    v[0] = d[k * DmaTile<R>::RSTR + rb + li];
    v[1] = d[(k + 1) * DmaTile<R>::RSTR + rb + li];
    return v;
}

template <int BM, int BN, int WM, int WN, int AMODE, int BMODE>
__device__ __forceinline__ void gemm_body_dma(const KParams<double> &p) {
    constexpr int NT = 64 * WM * WN;
    constexpr int BK = 16;
    constexpr int TM = BM / WM, TN = BN / WN;
    constexpr int FM = TM / 16, FN = TN / 16;
    constexpr bool AKC = AMODE == A_T, BKC = BMODE == B_N;
    // A_SYM: a k-tile below the diagonal is read row-contiguous, one above it
    // k-contiguous (the transpose of stored lower entries), one across it per element
    constexpr bool ASYM = AMODE == A_SYM;
    constexpr int ABYTES = ASYM ? (DmaTile<BM>::template bytes<true>() > DmaTile<BM>::template bytes<false>()
                                       ? DmaTile<BM>::template bytes<true>()
                                       : DmaTile<BM>::template bytes<false>())
                                : DmaTile<BM>::template bytes<AKC>();
    constexpr int BBYTES = DmaTile<BN>::template bytes<BKC>();
    constexpr int STAGE = ABYTES + BBYTES;
    using V4 = Mfma<double>::v4;
    using vec = Mfma<double>::vec;
    __shared__ __attribute__((aligned(1024))) char smem[2 * STAGE];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int nbk = gridDim.x, bid = blockIdx.x;
    const int xcd = bid & 7, q8 = nbk >> 3, rmd = nbk & 7;
    const int t = xcd * q8 + min(xcd, rmd) + (bid >> 3);
    int tm, tn;
    tile_coords(t, p.tiles_m, p.tiles_n, p.lower, tm, tn);
    const i64 m0 = (i64)tm * BM, n0 = (i64)tn * BN;
    const i64 kbeg = (i64)blockIdx.y * p.kchunk;
    const i64 kend = min(p.K, kbeg + p.kchunk);
    const int nkt = (int)((kend - kbeg + BK - 1) / BK);
    const bool m_in = m0 + BM <= p.M, n_in = n0 + BN <= p.N;

    V4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = V4{0, 0, 0, 0};

    // issue (or, per element, complete) the transfer of k-tile kt into its stage
    auto issue = [&](int kt) {
        char *st = smem + (kt & 1) * STAGE;
        const i64 k0 = kbeg + (i64)kt * BK;
        const bool k_in = k0 + BK <= kend;
        {
            const double *pa = p.a0;
            i64 lda = p.lda0, ko = 0;
            if constexpr (AMODE == A_CAT) {
                if (k0 >= p.asplit) {
                    pa = p.a1;
                    lda = p.lda1;
                    ko = p.asplit;
                }
            }
            if constexpr (ASYM) {
                const bool above = m0 + BM - 1 <= k0, below = m0 >= k0 + BK - 1;
                if ((above || below) && m_in && k_in && p.avec) dma_tile<BM, NT>(above, pa, lda, m0, k0, st, tid);
                else scalar_tile<BM, NT>(above, pa, lda, m0, k0, p.M, kend, st, tid, true);
            } else {
                if (m_in && k_in && p.avec) dma_tile<BM, NT>(AKC, pa, lda, m0, k0 - ko, st, tid);
                else scalar_tile<BM, NT>(AKC, pa, lda, m0, k0 - ko, p.M, kend - ko, st, tid);
            }
        }
        {
            const double *pb = p.b0;
            i64 ldb = p.ldb0, ko = 0;
            if constexpr (BMODE == B_CATT) {
                if (k0 >= p.bsplit) {
                    pb = p.b1;
                    ldb = p.ldb1;
                    ko = p.bsplit;
                }
            }
            if (n_in && k_in && p.bvec) dma_tile<BN, NT>(BKC, pb, ldb, n0, k0 - ko, st + ABYTES, tid);
            else scalar_tile<BN, NT>(BKC, pb, ldb, n0, k0 - ko, p.N, kend - ko, st + ABYTES, tid);
        }
    };

    const int li = lane & 15, lj = lane >> 4;
    if (nkt > 0) issue(0);
    for (int kt = 0; kt < nkt; ++kt) {
        // this wave's transfers of tile kt are complete; the barrier makes every wave's
        // visible and ends every wave's reads of tile kt - 1, whose stage kt + 1 reuses
        __builtin_amdgcn_s_waitcnt(0x0070); // vmcnt(0) lgkmcnt(0)
        __syncthreads();
        if (kt + 1 < nkt) issue(kt + 1);
        const char *st = smem + (kt & 1) * STAGE;
        const bool akc = ASYM ? (m0 + BM - 1 <= kbeg + (i64)kt * BK) : AKC;
#pragma unroll
        for (int kg = 0; kg < BK / 8; ++kg) {
            vec af[FM], bf[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) af[i] = dma_frag<BM>(akc, st, wm * TM + i * 16, kg, li, lj);
#pragma unroll
            for (int j = 0; j < FN; ++j) bf[j] = dma_frag<BN>(BKC, st + ABYTES, wn * TN + j * 16, kg, li, lj);
#pragma unroll
            for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j) acc[i][j] = Mfma<double>::mma(bf[j][e], af[i][e], acc[i][j]);
            // spread this group's fragment reads between its MFMAs
#pragma unroll
            for (int u = 0; u < FM + FN; ++u) {
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, 2 * FM * FN - 2 * (FM + FN), 0);
        }
    }

    // ---- epilogue
    const bool partial = p.part != nullptr;
    double *part = partial ? p.part + (i64)blockIdx.y * p.M * p.N : nullptr;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const i64 m = m0 + wm * TM + i * 16 + li;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const i64 n = n0 + wn * TN + j * 16 + Mfma<double>::out_row(lane, r);
                if (m >= p.M || n >= p.N) continue;
                if (p.lower && m < n) continue;
                const double v = acc[i][j][r];
                if (partial) {
                    part[m + n * p.M] = v;
                } else {
                    double *cp = p.C + m + n * p.ldc;
                    *cp = (p.beta == 0.0) ? p.alpha * v : p.alpha * v + p.beta * (*cp);
                }
            }
        }
    }
}

// The DMA path serves FP64 tiles of at least 128 x 128 with plain or K-concatenated
// operands, and symmetric A in 128 x 64 tiles (A22 W of the band reduction, W's tile
// k-contiguous), when both operands allow 16-byte transfers; unaligned operands and
// everything else take the register-staged path, which overlaps its per-element
// loads with the computation.
template <typename T, int BM, int BN, int AMODE, int BMODE>
constexpr bool dma_capable() {
    if constexpr (AMODE == A_SYM) return sizeof(T) == 8 && BM == 128 && BN == 64 && BMODE == B_N;
    else return sizeof(T) == 8 && BM >= 128 && BN >= 128;
}

template <typename T, int BM, int BN, int WM, int WN, int AMODE, int BMODE, int MINB, bool DMA>
__global__ __launch_bounds__(64 * WM * WN, MINB) void gemm_tile_kernel(const KParams<T> p) {
    if constexpr (DMA && dma_capable<T, BM, BN, AMODE, BMODE>()) {
        gemm_body_dma<BM, BN, WM, WN, AMODE, BMODE>(p);
    } else {
        gemm_body_regs<T, BM, BN, WM, WN, AMODE, BMODE>(p);
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
    // four interleaved partial sums keep eight loads in flight per thread; the
    // order of the additions is fixed, so every run rounds alike
    const i64 MN = M * N;
    const T *q = part + idx;
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
    int k = 0;
#pragma unroll 2
    for (; k + 4 <= splits; k += 4) {
        s0 += q[(i64)k * MN];
        s1 += q[(i64)(k + 1) * MN];
        s2 += q[(i64)(k + 2) * MN];
        s3 += q[(i64)(k + 3) * MN];
    }
    for (; k < splits; ++k) s0 += q[(i64)k * MN];
    const T s = (s0 + s1) + (s2 + s3);
    T *c = C + m + n * ldc;
    *c = (beta == T(0)) ? alpha * s : alpha * s + beta * (*c);
}

// The same sum for few outputs and many splits (the band reduction's w x b and b x b
// products over the whole panel height): each 64 outputs get 16 waves, wave g summing
// the splits g, g + 16, ... (lanes over consecutive outputs, so every load is one
// coalesced row), and the 16 wave sums are added in a fixed order.
constexpr int kWideWaves = 16;
template <typename T>
__global__ __launch_bounds__(64 * kWideWaves) void splitk_reduce_wide_kernel(const T *part, int splits, i64 M,
                                                                             i64 N, T alpha, T beta, T *C,
                                                                             i64 ldc, int lower) {
    __shared__ T red[kWideWaves][64];
    const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
    const i64 MN = M * N;
    const i64 idx = (i64)blockIdx.x * 64 + lane;
    T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
    if (idx < MN) {
        const T *q = part + idx;
        int k = g;
        for (; k + 3 * kWideWaves < splits; k += 4 * kWideWaves) {
            s0 += q[(i64)k * MN];
            s1 += q[(i64)(k + kWideWaves) * MN];
            s2 += q[(i64)(k + 2 * kWideWaves) * MN];
            s3 += q[(i64)(k + 3 * kWideWaves) * MN];
        }
        for (; k < splits; k += kWideWaves) s0 += q[(i64)k * MN];
    }
    red[g][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (g != 0 || idx >= MN) return;
    const i64 m = idx % M, n = idx / M;
    if (lower && m < n) return;
    T s = T(0);
#pragma unroll
    for (int w = 0; w < kWideWaves; ++w) s += red[w][lane];
    T *c = C + m + n * ldc;
    *c = (beta == T(0)) ? alpha * s : alpha * s + beta * (*c);
}

// Tile configurations: (BM, BN, WM, WN, blocks per CU the launch bounds ask for).
// C128x64W8: 128 x 64 on eight waves, the FP64 symmetric panel product's LDS-DMA tiles.
enum Cfg { C128x128 = 0, C128x64 = 1, C64x128 = 2, C64x64 = 3, C256x128 = 4, C128x64W8 = 5 };

template <typename T, int BM, int BN, int WM, int WN, int MINB, int AM, int BMD>
int launch_one(hipStream_t s, const KParams<T> &p, int splits, int ntiles) {
    dim3 grid((unsigned)ntiles, (unsigned)splits);
    if constexpr (dma_capable<T, BM, BN, AM, BMD>()) {
        if (p.avec && p.bvec) {
            hipLaunchKernelGGL((gemm_tile_kernel<T, BM, BN, WM, WN, AM, BMD, MINB, true>), grid, dim3(64 * WM * WN),
                               0, s, p);
            CUEV_LAUNCH_CHECK();
            return CUEV_STATUS_SUCCESS;
        }
    }
    hipLaunchKernelGGL((gemm_tile_kernel<T, BM, BN, WM, WN, AM, BMD, MINB, false>), grid, dim3(64 * WM * WN), 0, s,
                       p);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename T, int AM, int BMD>
int launch_cfg(hipStream_t s, const KParams<T> &p, int splits, int ntiles, int cfg) {
    switch (cfg) {
    case C128x128: return launch_one<T, 128, 128, 2, 2, 2, AM, BMD>(s, p, splits, ntiles);
    case C128x64: return launch_one<T, 128, 64, 2, 2, 2, AM, BMD>(s, p, splits, ntiles);
    case C64x128: return launch_one<T, 64, 128, 2, 2, 2, AM, BMD>(s, p, splits, ntiles);
    case C256x128: return launch_one<T, 256, 128, 4, 2, 1, AM, BMD>(s, p, splits, ntiles);
    case C128x64W8:
        if constexpr (sizeof(T) == 8) return launch_one<T, 128, 64, 4, 2, 1, AM, BMD>(s, p, splits, ntiles);
        else return launch_one<T, 64, 64, 2, 2, 2, AM, BMD>(s, p, splits, ntiles); // not selected for FP32
    default: return launch_one<T, 64, 64, 2, 2, 2, AM, BMD>(s, p, splits, ntiles);
    }
}

template <typename T>
int launch_modes(hipStream_t s, const KParams<T> &p, int splits, int ntiles, int cfg, int am, int bm) {
    if (am == A_N && bm == B_N) return launch_cfg<T, A_N, B_N>(s, p, splits, ntiles, cfg);
    if (am == A_N && bm == B_T) return launch_cfg<T, A_N, B_T>(s, p, splits, ntiles, cfg);
    if (am == A_T && bm == B_N) return launch_cfg<T, A_T, B_N>(s, p, splits, ntiles, cfg);
    if (am == A_T && bm == B_T) return launch_cfg<T, A_T, B_T>(s, p, splits, ntiles, cfg);
    if (am == A_SYM && bm == B_N) return launch_cfg<T, A_SYM, B_N>(s, p, splits, ntiles, cfg);
    if (am == A_CAT && bm == B_N) return launch_cfg<T, A_CAT, B_N>(s, p, splits, ntiles, cfg);
    if (am == A_TCAT && bm == B_N) return launch_one<T, 64, 64, 2, 2, 2, A_TCAT, B_N>(s, p, splits, ntiles);
    if (am == A_CAT && bm == B_CATT) return launch_cfg<T, A_CAT, B_CATT>(s, p, splits, ntiles, cfg);
    CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "gemm: operand mode combination (%d, %d) not implemented", am, bm);
}

int cfg_bm(int cfg) { return (cfg == C128x128 || cfg == C128x64 || cfg == C128x64W8) ? 128 : (cfg == C256x128) ? 256 : 64; }
int cfg_bn(int cfg) { return (cfg == C128x128 || cfg == C64x128 || cfg == C256x128) ? 128 : 64; }

template <typename T> bool vec_ok(const T *p, i64 ld) {
    constexpr int E = 16 / (int)sizeof(T);
    return p == nullptr || ((reinterpret_cast<uintptr_t>(p) & 15) == 0 && ld % E == 0);
}

} // namespace

size_t gemm_scratch_bytes(i64 M, i64 N, size_t elem) {
    // split-K is only used when the tile count is below ~one wave of blocks, so
    // splits * M * N stays below ~3 * (2 * CUs) * 128 * 128 elements.
    (void)M;
    (void)N;
    return (size_t)3 * 512 * 128 * 128 * elem;
}

template <typename T>
int gemm_ex(hipStream_t s, i64 M, i64 N, i64 K, T alpha, const Operand<T> &A,
            const Operand<T> &B, T beta, T *C, i64 ldc, int lower, Scratch scratch) {
    constexpr int BK = Mfma<T>::BK;
    if (M < 0 || N < 0 || K < 0)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: negative size M=%lld N=%lld K=%lld",
                  (long long)M, (long long)N, (long long)K);
    if (M == 0 || N == 0) return CUEV_STATUS_SUCCESS;
    if (lower && M != N) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: lower mode needs M == N");
    if ((A.split >= 0 && A.split % BK) || (B.split >= 0 && B.split % BK))
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: K split must be a multiple of %d", BK);
    if (ldc < M) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: ldc < M");

    KParams<T> p{};
    p.a0 = A.p0;
    p.a1 = A.p1;
    p.lda0 = A.ld0;
    p.lda1 = A.ld1;
    p.asplit = A.split;
    p.b0 = B.p0;
    p.b1 = B.p1;
    p.ldb0 = B.ld0;
    p.ldb1 = B.ld1;
    p.bsplit = B.split;
    p.C = C;
    p.ldc = ldc;
    p.M = M;
    p.N = N;
    p.K = K;
    p.alpha = alpha;
    p.beta = beta;
    p.lower = lower;
    p.avec = vec_ok(A.p0, A.ld0) && vec_ok(A.p1, A.ld1);
    p.bvec = vec_ok(B.p0, B.ld0) && vec_ok(B.p1, B.ld1);

    // tile shape
    const int cus = device_cu_count();
    auto count_tiles = [&](int bm, int bn) -> i64 {
        const i64 tm = ceil_div(M, bm), tn = ceil_div(N, bn);
        return lower ? tm * (tm + 1) / 2 : tm * tn;
    };
    int cfg;
    const bool sym_panel = A.sym && N <= 64;
    // a tall-skinny update (the band reduction's panel-wide products): 64 x 64 tiles
    // for the row parallelism
    const bool tall_skinny = !lower && !A.sym && N <= 64 && M >= 2048 && K < 1024;
    const bool tcat = A.split >= 0 && A.trans;
    if (tcat && (A.split % 64 != 0 || B.split >= 0 || B.trans || A.sym))
        CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED,
                  "gemm: a transposed A split along m needs the split a multiple of 64 and B plain");
    if (tall_skinny || scratch.small_tiles || tcat) {
        cfg = C64x64;
    } else if (lower) {
        cfg = (N >= 1024) ? C128x128 : C64x64;
    } else if (sym_panel) {
        // symmetric A times a panel-wide B (the band reduction's A22 W): FP64 on the
        // LDS-DMA 128 x 64 tiles with eight waves (four waves reach two per SIMD only),
        // FP32 on small register-staged tiles; deep split-K
        cfg = (sizeof(T) == 8 && !B.trans && B.split < 0) ? C128x64W8 : C64x64;
    } else {
        const int bm = (M > 64) ? 128 : 64, bn = (N > 64) ? 128 : 64;
        cfg = (bm == 128 && bn == 128) ? C128x128 : (bm == 128) ? C128x64 : (bn == 128) ? C64x128 : C64x64;
        if (cfg == C128x128 && count_tiles(128, 128) < cus) {
            // few large tiles: trade tile size for parallelism
            cfg = (count_tiles(64, 128) >= cus || M < N) ? C64x128 : C128x64;
        }
    }
    const int BMs = cfg_bm(cfg), BNs = cfg_bn(cfg);
    const i64 ntiles = count_tiles(BMs, BNs);
    if (ntiles > 0x7fffffffLL) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: too many tiles");
    p.tiles_m = (int)ceil_div(M, BMs);
    p.tiles_n = (int)ceil_div(N, BNs);

    // split-K for products with few output tiles and a long K (from 1024, or from 256
    // for at most four tiles). The K chunks are short (64): a workgroup of these
    // products waits on memory latency at every K step, so many short chunks keep more
    // loads in flight than a few long ones. Tall-skinny updates (N <= 64, short K) are
    // split four ways.
    int splits = 1;
    const i64 target = (sym_panel ? (cfg == C128x64W8 ? 8LL : 16LL) : 2LL) * cus;
    const i64 kmin = sym_panel ? 256 : 64; // shortest K chunk per split
    if (K == 0) {
        splits = 0;
    } else if (tall_skinny && K >= 128) {
        const i64 cap = (i64)(scratch.bytes / ((size_t)M * N * sizeof(T)));
        const i64 want = std::min<i64>(std::min<i64>(4, K / 64), cap);
        if (want > 1) splits = (int)want;
    } else if (ntiles < target / 2 && (K >= 1024 || (ntiles <= 4 && K >= 256))) {
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
        else if (A.split >= 0) am = A.trans ? A_TCAT : A_CAT;
        else am = A.trans ? A_T : A_N;
        if (B.split >= 0) bm = B_CATT;
        else bm = B.trans ? B_T : B_N;
        if ((A.sym && (A.split >= 0 || A.trans)) || (B.split >= 0 && !B.trans) || B.sym)
            CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "gemm: operand mode combination not implemented");
        CUEV_TRY(launch_modes<T>(s, p, splits, (int)ntiles, cfg, am, bm));
    }
    if (p.part != nullptr || splits == 0) {
        const i64 total = M * N;
        const int nsplit = (splits == 0) ? 0 : splits;
        if (nsplit >= 2 * kWideWaves && ceil_div(total, 256) < 2 * cus) {
            // fewer than two blocks per CU of the plain sum, each thread summing many
            // splits: give every output 16 waves instead
            hipLaunchKernelGGL(splitk_reduce_wide_kernel<T>, dim3((unsigned)ceil_div(total, 64)),
                               dim3(64 * kWideWaves), 0, s, static_cast<const T *>(p.part), nsplit, M, N,
                               alpha, beta, C, ldc, lower);
        } else {
            hipLaunchKernelGGL(splitk_reduce_kernel<T>, dim3((unsigned)ceil_div(total, 256)),
                               dim3(256), 0, s, static_cast<const T *>(p.part), nsplit, M, N, alpha,
                               beta, C, ldc, lower);
        }
        CUEV_LAUNCH_CHECK();
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int syr2k_lower(hipStream_t s, i64 n, i64 k, T alpha, const T *X, i64 ldx, const T *Y, i64 ldy,
                T beta, T *C, i64 ldc, Scratch scratch) {
    if (k % Mfma<T>::BK == 0) {
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
