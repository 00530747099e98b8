// dbbr.hip — double-blocked band reduction (full symmetric -> band) and bc_pack.
//
// Follows the reference's schedule (src/cuda/dbbr.cu:145-229, Algorithm 1 of
// Wang et al., PPoPP'25): for each block of k columns, panels of b columns are
// factored; each panel's two-sided update is deferred into the companion Z,
//   P = A22 W - Y_blk (Z_blk^T W) - Z_blk (Y_blk^T W),   Z = P - 1/2 Y (W^T P),
// only the next panel's b columns are brought up to date ("green" update), and
// the trailing matrix receives one rank-2w update A -= Y_blk Z_blk^T + Z_blk Y_blk^T
// per block. The algebra is the reference's; the kernels are our own: the panel
// QR is one persistent kernel that also forms W = Y T (panel_qr.hip), and every
// product runs on the MFMA engine (gemm.hip). The two correction products that
// the reference issues as four GEMMs become two small K-split products plus one
// K-concatenated update, and the green update's two GEMMs become one
// K-concatenated product.
#include "dbbr.hpp"

#include <algorithm>

#include "gemm.hpp"
#include "panel_qr.hpp"
#include "profile.hpp"

namespace cuev {
namespace {

// Bp[r + j*2b] = A[j+r, j] for r <= b (and j+r < n), 0 for b < r < 2b
// (reference: bc_pack_kernel, src/cuda/dbbr.cu:31-39).
template <typename T>
__global__ __launch_bounds__(256) void bc_pack_kernel(const T *A, i64 lda, T *Bp, int n, int b) {
    const i64 ldb = 2 * b;
    const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ldb * n) return;
    const int r = (int)(idx % ldb);
    const i64 j = idx / ldb;
    const i64 i = j + r;
    Bp[idx] = (r <= b && i < n) ? A[i + j * lda] : T(0);
}

// The b x b block G of D (rows [w, w + b), ld ldd) holds G = P^T W. With E = D_y^T D_z
// (b x b, ld b), C = W^T (P - Y_blk D_z - Z_blk D_y) = G^T - E - E^T, and G becomes C / 2.
// One thread per unordered pair (r, c), r <= c, which reads and writes both (r, c) and
// (c, r), so no entry is read after another thread has written it.
template <typename T>
__global__ __launch_bounds__(256) void dbbr_c_half_kernel(T *G, i64 ldd, const T *E, int b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= b * (b + 1) / 2) return;
    int c = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5); // t = c (c + 1) / 2 + r, r <= c
    while ((c + 1) * (c + 2) / 2 <= t) ++c;
    while (c * (c + 1) / 2 > t) --c;
    const int r = t - c * (c + 1) / 2;
    const T grc = G[r + (i64)c * ldd], gcr = G[c + (i64)r * ldd];
    const T erc = E[r + (i64)c * b], ecr = E[c + (i64)r * b];
    // C(r, c) = G(c, r) - E(r, c) - E(c, r); C(c, r) = G(r, c) - E(c, r) - E(r, c)
    G[r + (i64)c * ldd] = T(0.5) * (gcr - erc - ecr);
    G[c + (i64)r * ldd] = T(0.5) * (grc - ecr - erc);
}

} // namespace

template <typename T> int dbbr_pack(Workspace<T> &ws, const T *A, T *Bp, int n, int b) {
    if (n < 1 || b < 1) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dbbr_pack: n=%d b=%d", n, b);
    const i64 total = (i64)2 * b * n;
    prof::Scope sc(ws.stream, "dbbr.pack", 0.0, (double)sizeof(T) * ((double)(b + 1) * n + 2.0 * b * n));
    hipLaunchKernelGGL(bc_pack_kernel<T>, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0,
                       ws.stream, A, ws.ld, Bp, n, b);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int dbbr_panel_qr(Workspace<T> &ws, T *A, T *Y, T *W, int rows, int b) {
    return panel_qr<T>(ws.stream, A, ws.ld, Y, ws.ld, W, ws.ld, ws.Tri, ws.tau, rows, b,
                       ws.pqscratch, ws.pq_bytes, ws.counter, ws.err);
}

template <typename T> int dbbr_reduce(Workspace<T> &ws, T *A, T *Bp) {
    const int n = ws.n, b = ws.b, k = ws.k;
    const i64 lda = ws.ld;
    hipStream_t s = ws.stream;
    Scratch gs{ws.gscratch, ws.gscratch_bytes};
    const T one = T(1), zero = T(0), neg1 = T(-1), neg_half = T(-0.5);

    for (int i = 0; i < n; i += k) {
        const int kc = std::min(k, n - i);
        int w = 0; // reflector columns accumulated in this block
        const T *Yblk_base = ws.Y + (i64)i * lda; // block reflector columns i..i+w-1

        for (int j = i; j < i + kc && j + b < n; j += b) {
            const int rows = n - j - b;
            const i64 off = (i64)j * lda + (j + b);       // panel rows j+b.., columns j..j+b-1
            const i64 tr = (i64)(j + b) * lda + (j + b);  // trailing A22
            T *P = ws.Z + (i64)(j - i) * lda + (j + b);   // this panel's P / Z columns
            const T *Yb = Yblk_base + (j + b);            // Y_blk restricted to rows j+b..
            const T *Zb = ws.Z + (j + b);                 // Z_blk restricted to rows j+b..

            const double R = rows, Bw = b;
            // panel QR, explicit Y, T and W = Y T
            {
                prof::Scope sc(s, "dbbr.panel_qr", 2.0 * R * Bw * Bw + 2.0 * R * Bw * Bw / 2.0);
                CUEV_TRY(dbbr_panel_qr(ws, A + off, ws.Y + off, ws.W + off, rows, b));
            }
            // P = A22 W
            {
                prof::Scope sc(s, "dbbr.symm", 2.0 * R * R * Bw);
                CUEV_TRY(symm_lower<T>(s, rows, b, one, A + tr, lda, ws.W + off, lda, zero, P, lda, gs));
            }

            if (w > 0 && w % 64 == 0 && b % 64 == 0) {
                // Fused form. One reduction over the rows gives D = [Z_blk P Y_blk]^T W =
                // [D_z; G; D_y] ((2w + b) x b, ld 2w + b; [Z_blk | P] are adjacent columns of Z),
                // E = D_y^T D_z, and G becomes C / 2 with
                // C = W^T (P - Y_blk D_z - Z_blk D_y) = G^T - E - E^T; one update then gives
                // Z = P - [Y_blk Y_cur Z_blk] [D_z; C/2; D_y] ([Y_blk | Y_cur] are adjacent
                // columns of Y). Same algebra as the two steps below, with two products fewer.
                prof::Scope sc(s, "dbbr.corr", 4.0 * R * Bw * (2.0 * w + Bw) + 2.0 * Bw * Bw * w);
                const i64 ldd = 2 * w + b;
                T *E = ws.Dwk + (i64)(2 * k - b) * b; // past the (2w + b) x b block: 2w + b <= 2k - b
                CUEV_TRY(gemm_ex<T>(s, 2 * w + b, b, rows, one, opnd_cat(Zb, lda, Yb, lda, w + b, 1),
                                    opnd(static_cast<const T *>(ws.W + off), lda, 0), zero, ws.Dwk, ldd,
                                    0, gs));
                CUEV_TRY(gemm<T>(s, 1, 0, b, b, w, one, ws.Dwk + w + b, ldd, ws.Dwk, ldd, zero, E, b, gs));
                hipLaunchKernelGGL(dbbr_c_half_kernel<T>, dim3((unsigned)ceil_div((i64)b * (b + 1) / 2, 256)),
                                   dim3(256), 0, s, ws.Dwk + w, ldd, static_cast<const T *>(E), b);
                CUEV_LAUNCH_CHECK();
                CUEV_TRY(gemm_ex<T>(s, rows, b, 2 * w + b, neg1, opnd_cat(Yb, lda, Zb, lda, w + b, 0),
                                    opnd(static_cast<const T *>(ws.Dwk), ldd, 0), one, P, lda, 0, gs));
            } else {
                if (w > 0) {
                    prof::Scope sc(s, "dbbr.corr", 8.0 * R * Bw * w);
                    // D = [Z_blk Y_blk]^T W = [Z_blk^T W ; Y_blk^T W]   (2w x b, ld 2w): one product
                    // whose transposed A is split along its rows between the two blocks (the split
                    // must be a multiple of 64), else two
                    const i64 ldd = 2 * w;
                    if (w % 64 == 0) {
                        CUEV_TRY(gemm_ex<T>(s, 2 * w, b, rows, one, opnd_cat(Zb, lda, Yb, lda, w, 1),
                                            opnd(static_cast<const T *>(ws.W + off), lda, 0), zero, ws.Dwk, ldd,
                                            0, gs));
                    } else {
                        CUEV_TRY(gemm<T>(s, 1, 0, w, b, rows, one, Zb, lda, ws.W + off, lda, zero, ws.Dwk,
                                         ldd, gs));
                        CUEV_TRY(gemm<T>(s, 1, 0, w, b, rows, one, Yb, lda, ws.W + off, lda, zero,
                                         ws.Dwk + w, ldd, gs));
                    }
                    // P -= [Y_blk | Z_blk] D
                    CUEV_TRY(gemm_ex<T>(s, rows, b, 2 * w, neg1, opnd_cat(Yb, lda, Zb, lda, w, 0),
                                        opnd(static_cast<const T *>(ws.Dwk), ldd, 0), one, P, lda, 0,
                                        gs));
                }

                // C = W^T P  (b x b) ; Z = P - 1/2 Y C
                {
                    prof::Scope sc(s, "dbbr.wtp_z", 4.0 * R * Bw * Bw);
                    CUEV_TRY(gemm<T>(s, 1, 0, b, b, rows, one, ws.W + off, lda, P, lda, zero, ws.Dwk, b, gs));
                    CUEV_TRY(gemm<T>(s, 0, 0, rows, b, b, neg_half, ws.Y + off, lda, ws.Dwk, b, one, P, lda,
                                     gs));
                }
            }

            w += b;

            // bring the next panel's columns up to date with the whole block so far:
            // A[j+b:n, j+b:j+b+nc] -= Z_blk Y_blk[j+b:j+b+nc,:]^T + Y_blk Z_blk[j+b:j+b+nc,:]^T
            // The next panel has nc = min(b, n-j-b) columns: the reference updates b
            // columns unconditionally (dbbr.cu:210-216), which writes past the end of
            // A whenever n - j - b < b in the last block.
            if (j + b < i + kc) {
                const int nc = std::min(b, n - j - b);
                prof::Scope sc(s, "dbbr.green", 2.0 * R * nc * 2.0 * w);
                CUEV_TRY(gemm_ex<T>(s, rows, nc, 2 * w, neg1, opnd_cat(Zb, lda, Yb, lda, w, 0),
                                    opnd_cat(Yb, lda, Zb, lda, w, 1), one, A + tr, lda, 0, gs));
            }
        }

        // deferred trailing update: A[t:n, t:n] -= Y Z^T + Z Y^T (lower)
        const int t = i + kc;
        if (t < n && w > 0) {
            const double nt = n - t;
            prof::Scope sc(s, "dbbr.syr2k", 2.0 * nt * (nt + 1) * w);
            CUEV_TRY(syr2k_lower<T>(s, n - t, w, neg1, Yblk_base + t, lda, ws.Z + t, lda, one,
                                    A + (i64)t * lda + t, lda, gs));
        }
    }
    return dbbr_pack(ws, A, Bp, n, b);
}

template int dbbr_pack<float>(Workspace<float> &, const float *, float *, int, int);
template int dbbr_pack<double>(Workspace<double> &, const double *, double *, int, int);
template int dbbr_panel_qr<float>(Workspace<float> &, float *, float *, float *, int, int);
template int dbbr_panel_qr<double>(Workspace<double> &, double *, double *, double *, int, int);
template int dbbr_reduce<float>(Workspace<float> &, float *, float *);
template int dbbr_reduce<double>(Workspace<double> &, double *, double *);

} // namespace cuev
