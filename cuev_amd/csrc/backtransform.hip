// backtransform.hip — eigenvector back-transform evec <- Q_s Q_b Q_d.
//
// Reference: src/cuda/backtransform.cu:76-275 (bc_back_kernel, sbr_back,
// back_transform).
//
// bc_back (M <- Q_b M). The bulge-chasing reflectors compose to Q_b in forward
// (sweep-major) order, so Q_b M applies them last-first. Reflector (s, h) covers
// rows [s+1+hb, s+(h+1)b]; two reflectors need a fixed relative order only when
// their row ranges overlap, and then the one of the later sweep goes first.
// Overlaps occur only between (s, h) and (s', h) with |s - s'| < b, and between
// (s, h+1) and (s', h) with s' > s. Hence this order is valid:
//   passes of G = 128/b hop indices, in increasing order;
//   inside a pass, groups of 16 consecutive sweeps, in decreasing order;
//   inside a group, its hops in increasing order; for one hop the 16
//   reflectors (sweeps s0..s0+15) are one block H_s0 ... H_s0+15 = I - V T V^T
//   (V: (b+15) x 16 parallelogram read straight from the shifted U layout,
//   T = (I/2 + striu(V^T V))^-1 for reflectors H = I - 2 w w^T).
// Each block is three small MFMA products on 16 columns of M:
//   X = V^T M_J, Y = T X, M_J -= V Y.
// One wave64 owns 16 columns of M and walks every block; waves never share
// data, so the kernel has no barriers. With v_mfma_*_16x16x4, the
// accumulator register r of lane l holds row kidx(r, l/16) of its 16x16 tile;
// the K index of every product is ordered by the same kidx, so the registers of
// X and Y are directly the B operand of the next product and the rows of M_J
// kept in registers are at once the B operand of X and the accumulator of the
// update.
//
// sbr_back (M <- Q_s M): DBBR panels last-first, M_b -= W_j (Y_j^T M_b), as the
// reference (backtransform.cu:210-230), on the MFMA engine.
#include "backtransform.hpp"

#include <algorithm>

#include "gemm.hpp"
#include "profile.hpp"

namespace cuev {
namespace {

constexpr int kBTThreads = 256; // T-factor kernel: one wave per block reflector
constexpr int kBBThreads = 128; // apply kernel: independent waves, 2 per workgroup (LDS <= 64 KB)

template <typename T> struct MfmaBt;
template <> struct MfmaBt<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // f64 16x16x4: accumulator register r of lane l is row l/16 + 4r of the tile
    __device__ static __forceinline__ int kidx(int r, int q) { return 4 * r + q; }
};
template <> struct MfmaBt<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // f32 16x16x4: accumulator register r of lane l is row 4(l/16) + r of the tile
    __device__ static __forceinline__ int kidx(int r, int q) { return 4 * q + r; }
};

// T factor of every block (g, h): one wave per block. V's columns are the
// hop-h reflectors of sweeps 16g..16g+15 (zero for sweeps past n-3 and for rows
// past a sweep's last hop, which U holds as zeros).
template <typename T, int B>
__global__ __launch_bounds__(kBTThreads) void bt_tfactor_kernel(const T *__restrict__ U, i64 ldu,
                                                                const i64 *__restrict__ toff,
                                                                int ngroups, i64 nblocks, int n,
                                                                T *__restrict__ Tf) {
    constexpr int NG = kBtGroup;
    __shared__ T sv[4][NG][B];
    __shared__ T sn[4][NG][NG + 1];
    __shared__ T st[4][NG][NG + 1];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const i64 id = (i64)blockIdx.x * 4 + wave;
    const bool act = id < nblocks;
    int g = 0, h = 0;
    if (act) {
        int lo = 0, hi = ngroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (toff[mid] <= id) lo = mid;
            else hi = mid - 1;
        }
        g = lo;
        h = (int)(id - toff[g]);
    }
    const int s0 = g * NG;
    for (int e = lane; e < NG * B; e += 64) {
        const int j = e / B, k = e - (e / B) * B;
        const int sw = s0 + j;
        sv[wave][j][k] = (act && sw <= n - 3) ? U[(i64)h * B + k + (i64)sw * ldu] : T(0);
    }
    __syncthreads();
    // strictly upper Gram: N[i][j] = v_i . v_j with v_j shifted down by j - i rows
    for (int pidx = lane; pidx < NG * NG; pidx += 64) {
        const int i = pidx % NG, j = pidx / NG;
        T acc = T(0);
        if (i < j) {
            const int dl = j - i;
            for (int t = 0; t + dl < B; ++t) acc += sv[wave][i][t + dl] * sv[wave][j][t];
        }
        sn[wave][i][j] = acc;
    }
    __syncthreads();
    // T = (I/2 + N)^-1, column j by lane j (back substitution)
    if (lane < NG) {
        const int j = lane;
        st[wave][j][j] = T(2);
        for (int i = j - 1; i >= 0; --i) {
            T acc = T(0);
            for (int k = i + 1; k <= j; ++k) acc += sn[wave][i][k] * st[wave][k][j];
            st[wave][i][j] = T(-2) * acc;
        }
        for (int i = j + 1; i < NG; ++i) st[wave][i][j] = T(0);
    }
    __syncthreads();
    if (act)
        for (int pidx = lane; pidx < NG * NG; pidx += 64)
            Tf[id * (NG * NG) + pidx] = st[wave][pidx % NG][pidx / NG];
}

// Register set of one block's operands, in the lane layouts of the three products.
template <typename T, int B> struct BtOperands {
    static constexpr int NKB = B / 16 + 1;
    T vx[4 * NKB]; // X = V^T M_J: A operand of chunk c, V[16(c/4) + kidx(c%4, q)][li]
    T vu[NKB][4];  // M_J -= V Y: A operand of tile t, chunk r, -V[16t + li][kidx(r, q)]
    T t[4];        // Y = T X: A operand of chunk r, T[li][kidx(r, q)]
};

// M <- Q_b M for columns [0, ncols): one wave per 16 columns (see the header).
// Each wave keeps the rows of the current pass's window, WS = G*b + 16 rows of
// its 16 columns, in a private circular LDS buffer (row r in slot r mod WS):
// the window slides up 16 rows per sweep group, so a group costs 16 rows in and
// 16 rows out. The next block's V and T and the next group's entering rows are
// loaded into registers while the current block computes.
template <typename T, int B>
__global__ __launch_bounds__(kBBThreads) void bc_back_kernel(const T *__restrict__ U, i64 ldu,
                                                             const T *__restrict__ Tf,
                                                             const i64 *__restrict__ toff, T *M,
                                                             i64 ldm, int n, int ncols) {
    using MB = MfmaBt<T>;
    using V4 = typename MB::v4;
    using Ops = BtOperands<T, B>;
    constexpr int NG = kBtGroup;
    constexpr int NKB = B / 16 + 1; // 16-row tiles covering a block's b + 15 rows
    constexpr int NC = 4 * NKB;     // 4-row K chunks
    constexpr int G = 128 / B;      // hops per pass
    constexpr int WS = G * B + 16;  // window rows
    constexpr int LDW = 17;         // LDS row stride: rows and 16-column reads both conflict-free
    __shared__ T win_all[kBBThreads / 64][WS * LDW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, li = lane & 15;
    const int wcol = (blockIdx.x * (kBBThreads / 64) + wave) * 16;
    if (wcol >= ncols) return; // whole wave
    T *win = win_all[wave];
    const int last = n - 3;    // last sweep
    for (int e = lane; e < WS * LDW; e += 64) win[e] = T(0);

    // 16 rows [r0, r0+16) of the wave's columns: lane = (row li, columns q + 4i)
    auto load16 = [&](int r0, T (&v)[4]) {
        const int r = r0 + li;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int col = wcol + q + 4 * i;
            v[i] = (r < n && col < ncols) ? M[r + (i64)col * ldm] : T(0);
        }
    };
    auto put16 = [&](int r0, const T (&v)[4]) {
        const int s = (r0 + li) % WS;
#pragma unroll
        for (int i = 0; i < 4; ++i) win[s * LDW + q + 4 * i] = v[i];
    };
    auto store16 = [&](int r0) {
        const int r = r0 + li;
        if (r >= n) return;
        const int s = r % WS;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int col = wcol + q + 4 * i;
            if (col < ncols) M[r + (i64)col * ldm] = win[s * LDW + q + 4 * i];
        }
    };
    auto load_window = [&](int top) {
        for (int c = 0; c < WS; c += 16) {
            T v[4];
            load16(top + c, v);
            put16(top + c, v);
        }
    };
    auto flush_window = [&](int top) {
        for (int c = 0; c < WS; c += 16) store16(top + c);
        wave_stores_done(); // the next pass reads these rows back through other lanes
    };
    auto load_ops = [&](int g, int h, Ops &o) {
        const int s0 = g * NG;
        const T *Ug = U + (i64)h * B;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int d = 16 * (c >> 2) + MB::kidx(c & 3, q) - li, sw = s0 + li;
            o.vx[c] = (d >= 0 && d < B && sw <= last) ? Ug[d + (i64)sw * ldu] : T(0);
        }
#pragma unroll
        for (int t = 0; t < NKB; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int j = MB::kidx(r, q);
                const int d = 16 * t + li - j, sw = s0 + j;
                o.vu[t][r] = (d >= 0 && d < B && sw <= last) ? -Ug[d + (i64)sw * ldu] : T(0);
            }
        const T *Tb = Tf + (toff[g] + h) * (NG * NG);
#pragma unroll
        for (int r = 0; r < 4; ++r) o.t[r] = Tb[li + MB::kidx(r, q) * NG];
    };
    // next block in the order: hops of the group, then the group below, then the next pass
    auto next_block = [&](int &h0, int &g, int &h) -> bool {
        if (h + 1 < h0 + G && g * NG + 1 + (h + 1) * B <= n - 1) {
            ++h;
            return true;
        }
        if (g > 0) {
            --g;
            h = h0;
            return true;
        }
        h0 += G;
        const int smax = min(last, n - 2 - h0 * B);
        if (smax < 0) return false;
        g = smax / NG;
        h = h0;
        return true;
    };

    int h0 = 0, g = min(last, n - 2) / NG, h = 0;
    int top = g * NG + 1 + h0 * B; // first window row
    Ops cur, nxt;
    load_ops(g, h, cur);
    load_window(top);
    T ent[4] = {T(0), T(0), T(0), T(0)};
    for (;;) {
        int h0n = h0, gn = g, hn = h;
        const bool more = next_block(h0n, gn, hn);
        if (more) load_ops(gn, hn, nxt);
        if (h == h0 && g > 0) load16(top - 16, ent); // rows entering with the group below

        // ---- block (g, h): rows [base, base + 16 NKB) of the window
        const int base = g * NG + 1 + h * B;
        const int sb = base % WS;
        T m[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            int s = sb + 16 * (c >> 2) + MB::kidx(c & 3, q);
            if (s >= WS) s -= WS;
            m[c] = win[s * LDW + li];
        }
        V4 x = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int c = 0; c < NC; ++c) x = MB::mma(cur.vx[c], m[c], x);
        V4 y = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int r = 0; r < 4; ++r) y = MB::mma(cur.t[r], x[r], y);
#pragma unroll
        for (int t = 0; t < NKB; ++t) {
            V4 acc = {m[4 * t], m[4 * t + 1], m[4 * t + 2], m[4 * t + 3]};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc = MB::mma(cur.vu[t][r], y[r], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                int s = sb + 16 * t + MB::kidx(r, q);
                if (s >= WS) s -= WS;
                win[s * LDW + li] = acc[r];
            }
        }

        // ---- window bookkeeping for the next block
        if (!more) {
            flush_window(top);
            break;
        }
        if (h0n != h0) {
            // pass done: write the window back, load the next pass's first window
            flush_window(top);
            top = gn * NG + 1 + h0n * B;
            load_window(top);
        } else if (gn != g) {
            // group g -> g-1: 16 rows leave at the bottom, 16 enter at the top
            store16(top + WS - 16);
            put16(top - 16, ent);
            top -= 16;
        }
        h0 = h0n;
        g = gn;
        h = hn;
        cur = nxt;
    }
}

template <typename T, int B> int launch_bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm) {
    const int n = ws.n;
    if (ws.bt_blocks <= 0) return CUEV_STATUS_SUCCESS;
    {
        prof::Scope sc(ws.stream, "bt.bc_back_t", (double)ws.bt_blocks * kBtGroup * kBtGroup * B);
        hipLaunchKernelGGL((bt_tfactor_kernel<T, B>), dim3((unsigned)ceil_div(ws.bt_blocks, 4)),
                           dim3(kBTThreads), 0, ws.stream, U, ws.ldu, ws.bt_off, ws.bt_groups,
                           ws.bt_blocks, n, ws.BtT);
        CUEV_LAUNCH_CHECK();
    }
    // 4 flops per reflector entry per column: sum_s (n-1-s) entries ~ n^2/2
    prof::Scope sc(ws.stream, "bt.bc_back", 2.0 * (double)n * n * n);
    const i64 blocks = ceil_div(n, 16 * (kBBThreads / 64));
    hipLaunchKernelGGL((bc_back_kernel<T, B>), dim3((unsigned)blocks), dim3(kBBThreads), 0, ws.stream,
                       U, ws.ldu, ws.BtT, ws.bt_off, M, ldm, n, n);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

} // namespace

template <typename T> int bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm) {
    if (ldm < ws.n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back: ldm < n");
    if (ws.n < 3) return CUEV_STATUS_SUCCESS;
    switch (ws.b) {
    case 16: return launch_bc_back<T, 16>(ws, U, M, ldm);
    case 32: return launch_bc_back<T, 32>(ws, U, M, ldm);
    case 64: return launch_bc_back<T, 64>(ws, U, M, ldm);
    default: CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back: unsupported b=%d", ws.b);
    }
}

template <typename T> int sbr_back(Workspace<T> &ws, const T *Y, const T *W, T *M, i64 ldm) {
    const int n = ws.n, b = ws.b;
    const i64 ld = ws.ld;
    if (ldm < n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "sbr_back: ldm < n");
    Scratch gs{ws.gscratch, ws.gscratch_bytes};
    int jmax = -1;
    for (int j = 0; j + b < n; j += b) jmax = j;
    for (int j = jmax; j >= 0; j -= b) {
        const int rows = n - (j + b);
        const T *Yp = Y + (i64)j * ld + (j + b);
        const T *Wp = W + (i64)j * ld + (j + b);
        T *Mb = M + (j + b);
        // Zt = Y_j^T M_b (b x n) ; M_b -= W_j Zt
        prof::Scope sc(ws.stream, "bt.sbr_back", 4.0 * rows * (double)b * n);
        CUEV_TRY(gemm<T>(ws.stream, 1, 0, b, n, rows, T(1), Yp, ld, Mb, ldm, T(0), ws.Z, b, gs));
        CUEV_TRY(gemm<T>(ws.stream, 0, 0, rows, n, b, T(-1), Wp, ld, ws.Z, b, T(1), Mb, ldm, gs));
    }
    return CUEV_STATUS_SUCCESS;
}

template int bc_back<float>(Workspace<float> &, const float *, float *, i64);
template int bc_back<double>(Workspace<double> &, const double *, double *, i64);
template int sbr_back<float>(Workspace<float> &, const float *, const float *, float *, i64);
template int sbr_back<double>(Workspace<double> &, const double *, const double *, double *, i64);

} // namespace cuev
