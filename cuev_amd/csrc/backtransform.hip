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
// sbr_back (M <- Q_s M): the reference applies the DBBR panels last-first,
// M_b -= W_j (Y_j^T M_b) (backtransform.cu:210-230), one rank-b update of M per
// panel. Here the panels of one reduction block are folded into one block
// reflector first, so M is updated once per block (rank k) on the MFMA engine.
#include "backtransform.hpp"

#include <algorithm>
#include <utility>
#include <vector>

#include "gemm.hpp"
#include "profile.hpp"
#include "util.hpp"

namespace cuev {
namespace {

constexpr int kBTThreads = 256; // T-factor kernel: one wave per block reflector
constexpr int kBBThreads = 128; // apply kernel: independent waves, 2 per workgroup (LDS <= 64 KB)
// blocks per launch of the paired apply kernel (a quarter of the stream at n = 16384)
constexpr i64 kBtSegBlocks = 33000;

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

// Operands of every block (g, h), packed in the apply kernel's lane order: one
// wave per block. V's columns are the hop-h reflectors of sweeps 16g..16g+15
// (zero for sweeps past n-3 and for rows past a sweep's last hop, which U holds
// as zeros); V[i][j] = U[hb + i - j + (16g + j) ldu] for 0 <= i - j < b. The
// block is I - V T V^T with T = (I/2 + striu(V^T V))^-1 (reflectors
// H = I - 2 w w^T); TR packs the transposed block I - V T^T V^T. Per lane
// (q = lane/16, li = lane%16) the block holds 8 NKB values:
//   c <  4 NKB: V[16(c/4) + kidx(c%4, q)][li]            (A operand of X = V^T M_J)
//   c >= 4 NKB: W[16t + li][kidx(r, q)], c = 4 NKB + 4t + r (A operand of M_J += W X)
// with W = -V T (TR: -V T^T), stored as pairs: value c of lane l at
// ((c/2) 64 + l) 2 + c%2, so the apply kernel reads each pair with one 16-byte load.
template <typename T, int B, bool TR>
__global__ __launch_bounds__(kBTThreads, 2) void bt_pack_kernel(const T *__restrict__ U, i64 ldu,
                                                             const i64 *__restrict__ toff,
                                                             int ngroups, i64 nblocks, int n,
                                                             T *__restrict__ ops) {
    using MB = MfmaBt<T>;
    constexpr int NG = kBtGroup;
    constexpr int NKB = B / 16 + 1;
    constexpr int NOP = 8 * NKB;
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
    {
        // the block's NG reflectors: every load issued before its LDS store (one round trip)
        constexpr int PER = NG * B / 64;
        static_assert(NG * B % 64 == 0, "bt_pack load layout");
        T v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = lane + 64 * u, j = e / B, k = e - j * B, sw = s0 + j;
            v[u] = (act && sw <= n - 3) ? U[(i64)h * B + k + (i64)sw * ldu] : T(0);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = lane + 64 * u, j = e / B, k = e - j * B;
            sv[wave][j][k] = v[u];
        }
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
    if (!act) {
        if (id == nblocks) // an all-zero block after the last (the identity, for the paired kernel)
            for (int c = 0; c < NOP; ++c) ops[id * (i64)(NOP * 64) + c * 64 + lane] = T(0);
        return;
    }
    const int q = lane >> 4, li = lane & 15;
    T *ob = ops + id * (i64)(NOP * 64) + 2 * lane;
#pragma unroll
    for (int c = 0; c < 4 * NKB; ++c) {
        const int d = 16 * (c >> 2) + MB::kidx(c & 3, q) - li;
        ob[(c >> 1) * 128 + (c & 1)] = (d >= 0 && d < B) ? sv[wave][li][d] : T(0);
    }
#pragma unroll 1
    for (int t = 0; t < NKB; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = MB::kidx(r, q), i = 16 * t + li;
            T acc = T(0);
            for (int k = 0; k < NG; ++k) {
                const int d = i - k;
                const T v = (d >= 0 && d < B) ? sv[wave][k][d] : T(0);
                acc += v * (TR ? st[wave][j][k] : st[wave][k][j]);
            }
            const int c = 4 * NKB + 4 * t + r;
            ob[(c >> 1) * 128 + (c & 1)] = -acc;
        }
}

// Position of a block in the application order (wave-uniform), with the block's
// index in the packed operands, toff[g] + h.
struct BtPos {
    int h0, g, h; // pass start hop, sweep group, hop
    i64 id;
};

// Blocks (hops) of sweep group g: toff[g + 1] - toff[g].
template <int B> __device__ __forceinline__ int bt_hops(int g, int n) { return (n - 2 - g * kBtGroup) / B + 1; }

// Advance to the next block: hops of the group, then the group below, then the
// next pass. Returns false after the last block. The index follows from the
// previous one except at a pass start (toff read once per pass).
template <int B> __device__ __forceinline__ bool bt_next(BtPos &p, int n, const i64 *__restrict__ toff) {
    constexpr int G = 128 / B;
    if (p.h + 1 < p.h0 + G && p.g * kBtGroup + 1 + (p.h + 1) * B <= n - 1) {
        ++p.h;
        ++p.id;
        return true;
    }
    if (p.g > 0) {
        --p.g;
        p.id += p.h0 - p.h - bt_hops<B>(p.g, n); // toff[g-1] = toff[g] - hops(g-1)
        p.h = p.h0;
        return true;
    }
    p.h0 += G;
    const int smax = min(n - 3, n - 2 - p.h0 * B);
    if (smax < 0) return false;
    p.g = smax / kBtGroup;
    p.h = p.h0;
    p.id = toff[p.g] + p.h;
    return true;
}

// Step back to the previous block of that order (the order of Q_b^T M, whose blocks
// are the same reflectors in reverse, each transposed). Returns false before the first.
template <int B> __device__ __forceinline__ bool bt_prev(BtPos &p, int n) {
    constexpr int G = 128 / B;
    if (p.h > p.h0) {
        --p.h;
        --p.id;
        return true;
    }
    if (p.g < min(n - 3, n - 2 - p.h0 * B) / kBtGroup) {
        const i64 t = p.id - p.h + bt_hops<B>(p.g, n); // toff[g+1] = toff[g] + hops(g)
        ++p.g;
        p.h = min(p.h0 + G - 1, (n - 2 - p.g * kBtGroup) / B);
        p.id = t + p.h;
        return true;
    }
    if (p.h0 == 0) return false;
    p.h0 -= G;
    p.g = 0;
    p.h = min(p.h0 + G - 1, (n - 2) / B);
    p.id = p.h; // toff[0] = 0
    return true;
}

// One block's packed operands (see bt_pack_kernel), as NOP/2 pairs.
template <typename T, int NP>
__device__ __forceinline__ void bt_load_ops(T (&o)[2 * NP], const T *__restrict__ ops, i64 id, int lane) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    const T2 *p = reinterpret_cast<const T2 *>(ops + id * (i64)(2 * NP * 64)) + lane;
#pragma unroll
    for (int c = 0; c < NP; ++c) {
        const T2 v = p[c * 64];
        o[2 * c] = v.x;
        o[2 * c + 1] = v.y;
    }
}

// M <- Q_b M (TR: M <- Q_b^T M, the blocks in reverse order, each transposed) for
// columns [0, ncols): one wave per 16 columns (see the header). Each wave keeps the
// rows of the current pass's window, WS = G*b + 16 rows of its 16 columns, in a
// private LDS buffer of RS = WS + 64 rows: row r of the matrix sits in buffer row
// r - org. The window slides 16 rows per sweep group (a group costs 16 rows in and
// 16 out); when it reaches the end of the buffer, its rows move back to the other
// end in one copy (every 4 groups), so a block's 16 NKB rows are always contiguous
// and every LDS access of a block is one base address plus a constant. The next
// block's operands are loaded into the other of two register sets at the start of
// each block, and the rows entering at a group change are loaded at the previous
// change, so neither load is waited for inside the block that issues it. Rows past
// n are read from zp (zeros) and not written.
template <typename T, int B, bool TR>
__global__ __launch_bounds__(kBBThreads) void bc_back_kernel(const T *__restrict__ ops,
                                                             const i64 *__restrict__ toff,
                                                             const T *__restrict__ zp, T *M,
                                                             i64 ldm, int n, int ncols) {
    using MB = MfmaBt<T>;
    using V4 = typename MB::v4;
    constexpr int NG = kBtGroup;
    constexpr int NKB = B / 16 + 1; // 16-row tiles covering a block's b + 15 rows
    constexpr int NC = 4 * NKB;     // 4-row K chunks
    constexpr int NOP = 2 * NC;     // operand values per lane per block
    constexpr int G = 128 / B;      // hops per pass
    constexpr int WS = G * B + 16;  // window rows
    constexpr int RS = WS + 64;     // buffer rows
    constexpr int LDW = 17;         // LDS row stride: rows and 16-column reads both conflict-free
    __shared__ T win[(kBBThreads / 64) * RS * LDW];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, q = lane >> 4, li = lane & 15;
    const int wcol = (blockIdx.x * (kBBThreads / 64) + wave) * 16;
    if (wcol >= ncols) return; // whole wave
    const int wb = wave * (RS * LDW); // this wave's buffer
    const int last = n - 3;           // last sweep
    for (int e = lane; e < RS * LDW; e += 64) win[wb + e] = T(0);

// matrix rows [r0, r0+16) of the wave's 16 columns: lane = (row li, columns q + 4i)
#define BT_LOAD16(r0, v)                                                                           \
    do {                                                                                           \
        const int r_ = (r0) + li;                                                                  \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) {                                         \
            const int col_ = wcol + q + 4 * i_;                                                    \
            const bool ok_ = r_ >= 0 && r_ < n && col_ < ncols;                                    \
            (v)[i_] = *(ok_ ? M + r_ + (i64)col_ * ldm : zp);                                      \
        }                                                                                          \
    } while (0)
// buffer rows [b0, b0+16)
#define BT_PUT16(b0, v)                                                                            \
    do {                                                                                           \
        const int o_ = wb + ((b0) + li) * LDW + q;                                                 \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) win[o_ + 4 * i_] = (v)[i_];               \
    } while (0)
// matrix rows [r0, r0+16) from buffer rows [b0, b0+16)
#define BT_STORE16(r0, b0)                                                                         \
    do {                                                                                           \
        const int r_ = (r0) + li;                                                                  \
        const int o_ = wb + ((b0) + li) * LDW + q;                                                 \
        T v_[4];                                                                                   \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) v_[i_] = win[o_ + 4 * i_];                \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) {                                         \
            const int col_ = wcol + q + 4 * i_;                                                    \
            if (r_ < n && col_ < ncols) M[r_ + (i64)col_ * ldm] = v_[i_];                          \
        }                                                                                          \
    } while (0)
// buffer rows [s0, s0 + WS - 16) -> [d0, d0 + WS - 16): all read, then all written
// (so the two ranges may overlap)
#define BT_MOVE(s0, d0)                                                                            \
    do {                                                                                           \
        constexpr int NM_ = (WS - 16) / 16;                                                        \
        T v_[NM_][4];                                                                              \
        const int os_ = wb + ((s0) + li) * LDW + q, od_ = wb + ((d0) + li) * LDW + q;              \
        _Pragma("unroll") for (int c_ = 0; c_ < NM_; ++c_)                                         \
            _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) v_[c_][i_] = win[os_ + 16 * c_ * LDW + 4 * i_]; \
        _Pragma("unroll") for (int c_ = 0; c_ < NM_; ++c_)                                         \
            _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) win[od_ + 16 * c_ * LDW + 4 * i_] = v_[c_][i_]; \
    } while (0)

    BtPos p;
    if (TR) {
        p.h0 = ((n - 2) / B) / G * G; // last pass
        p.g = 0;
        p.h = min(p.h0 + G - 1, (n - 2) / B);
    } else {
        p.h0 = 0;
        p.g = min(last, n - 2) / NG;
        p.h = 0;
    }
    p.id = toff[p.g] + p.h;
    // the window moves down the buffer going up the groups (TR), up it going down
    int top = p.g * NG + 1 + p.h0 * B; // first window row
    int org = TR ? top : top - (RS - WS);
    BtPos pn = p;
    bool more = TR ? bt_prev<B>(pn, n) : bt_next<B>(pn, n, toff);
    T opa[NOP], opb[NOP];
    bt_load_ops<T, NC>(opa, ops, p.id, lane);
    for (int c = 0; c < WS; c += 16) {
        T v[4];
        BT_LOAD16(top + c, v);
        BT_PUT16(top + c - org, v);
    }
    // this lane's first element of a block tile: row kidx(0, q), column li; the
    // others are at constant offsets (kidx(r, q) = kidx(0, q) + kidx(r, 0))
    const int lofs = wb + MB::kidx(0, q) * LDW + li;
    // rows entering at the next group change (below the window going up the groups
    // (TR), above it going down), loaded as soon as the window's position is known so
    // that their latency is spread over the group's blocks
    T ent[4];
    BT_LOAD16(TR ? top + WS : top - 16, ent);

    // One block with operands cur; the next block's operands are loaded into nxt
    // first (a repeat of this block when nothing follows: no branch). Returns false
    // after the last block. Called alternately with the two register sets, so no
    // operand is copied and the loads have a whole block to arrive.
    auto step = [&](T(&cur)[NOP], T(&nxt)[NOP]) -> bool {
        bt_load_ops<T, NC>(nxt, ops, more ? pn.id : p.id, lane);
        __builtin_amdgcn_sched_barrier(0);

        // ---- block (g, h): rows [base, base + 16 NKB) of the window
        const int base = p.g * NG + 1 + p.h * B;
        const int wr = lofs + (base - org) * LDW;
        T m[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) m[c] = win[wr + (16 * (c >> 2) + MB::kidx(c & 3, 0)) * LDW];
        // X = V^T M_J in two independent chains (even and odd chunks)
        V4 x = {T(0), T(0), T(0), T(0)}, x1 = {T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int c = 0; c < NC; c += 2) {
            x = MB::mma(cur[c], m[c], x);
            x1 = MB::mma(cur[c + 1], m[c + 1], x1);
        }
        x += x1;
        // M_J += W X, W = -V T (or -V T^T): the registers of X are the B operand
#pragma unroll
        for (int t = 0; t < NKB; ++t) {
            V4 acc = {m[4 * t], m[4 * t + 1], m[4 * t + 2], m[4 * t + 3]};
#pragma unroll
            for (int r = 0; r < 4; ++r) acc = MB::mma(cur[NC + 4 * t + r], x[r], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) win[wr + (16 * t + MB::kidx(r, 0)) * LDW] = acc[r];
        }

        // ---- window bookkeeping for the next block
        if (!more || pn.h0 != p.h0) {
            // pass done: write the window back
            for (int c = 0; c < WS; c += 16) BT_STORE16(top + c, top + c - org);
            if (!more) return false;
            wave_stores_done(); // the next pass reads these rows back through other lanes
            top = pn.g * NG + 1 + pn.h0 * B;
            org = TR ? top : top - (RS - WS);
            for (int c = 0; c < WS; c += 16) {
                T v[4];
                BT_LOAD16(top + c, v);
                BT_PUT16(top + c - org, v);
            }
            BT_LOAD16(TR ? top + WS : top - 16, ent);
        } else if (pn.g != p.g) {
            if (TR) {
                // group g -> g+1: 16 rows leave at the top, 16 enter at the bottom
                BT_STORE16(top, top - org);
                top += 16;
                if (top + WS - org > RS) {
                    // rows [top, top + WS - 16) back to the start of the buffer
                    BT_MOVE(top - org, 0);
                    org = top;
                }
                BT_PUT16(top + WS - 16 - org, ent);
                BT_LOAD16(top + WS, ent);
            } else {
                // group g -> g-1: 16 rows leave at the bottom, 16 enter at the top
                BT_STORE16(top + WS - 16, top + WS - 16 - org);
                top -= 16;
                if (top < org) {
                    // rows [top + 16, top + WS) to the end of the buffer
                    const int norg = top - (RS - WS);
                    BT_MOVE(top + 16 - org, top + 16 - norg);
                    org = norg;
                }
                BT_PUT16(top - org, ent);
                BT_LOAD16(top - 16, ent);
            }
        }
        p = pn;
        more = more && (TR ? bt_prev<B>(pn, n) : bt_next<B>(pn, n, toff));
        return true;
    };
    for (;;) {
        if (!step(opa, opb)) break;
        if (!step(opb, opa)) break;
    }
#undef BT_LOAD16
#undef BT_PUT16
#undef BT_STORE16
#undef BT_MOVE
}

// Passes of two hops for the paired kernel below. The blocks of a pass, in the
// order of bc_back_kernel (groups down, hops up), taken two at a time as steps
// {(g, h0+1), (g-1, h0)}: the two blocks of a step cover disjoint rows (the first
// ends where the second starts), so they are independent and the step is the same
// sequence of blocks. In the order of Q_b^T M (TR: groups up, hops down) the steps
// are {(g, h0), (g+1, h0+1)}. Either way one step covers the rows of a window of
// 2 NKB 16-row tiles: its lower block the first NKB, its upper block the others.
struct BtStep {
    int h0, g;  // pass start hop, step group (see bt_step_blocks)
    int gA, gB; // last group with hop h0, with hop h0 + 1 (-1: none)
    i64 tg;     // forward: toff[g]; TR: toff[g + 1]
};

template <int B> __device__ __forceinline__ int bt_gmax(int n, int h) {
    const int m = min(n - 3, n - 2 - h * B);
    return m < 0 ? -1 : m / kBtGroup;
}

// the first step of the pass starting at hop h0
template <int B, bool TR>
__device__ __forceinline__ void bt_step_first(BtStep &s, int n, const i64 *__restrict__ toff, int h0) {
    s.h0 = h0;
    s.gA = bt_gmax<B>(n, s.h0);
    s.gB = bt_gmax<B>(n, s.h0 + 1);
    s.g = TR ? (s.gB >= 0 ? -1 : 0) : s.gA + 1;
    s.tg = toff[TR ? s.g + 1 : s.g];
}

// Next step; false after the last. The offsets follow from the hop counts except
// at a pass start.
template <int B, bool TR> __device__ __forceinline__ bool bt_step_next(BtStep &s, int n, const i64 *__restrict__ toff) {
    if (TR) {
        if (s.g < s.gA) {
            ++s.g;
            s.tg += bt_hops<B>(s.g, n);
            return true;
        }
        s.h0 -= 2;
        if (s.h0 < 0) return false;
        s.gA = bt_gmax<B>(n, s.h0);
        s.gB = bt_gmax<B>(n, s.h0 + 1);
        s.g = s.gB >= 0 ? -1 : 0;
        s.tg = toff[s.g + 1];
        return true;
    }
    if (s.g > (s.gB >= 0 ? 0 : 1)) {
        s.tg -= bt_hops<B>(s.g - 1, n);
        --s.g;
        return true;
    }
    s.h0 += 2;
    s.gA = bt_gmax<B>(n, s.h0);
    if (s.gA < 0) return false;
    s.gB = bt_gmax<B>(n, s.h0 + 1);
    s.g = s.gA + 1;
    s.tg = toff[s.g];
    return true;
}

// The two blocks of a step: packed index (il, iu; an absent block, at a pass end,
// gets the all-zero block nblocks, which leaves its rows unchanged) and the
// window's first row.
template <int B, bool TR>
__device__ __forceinline__ void bt_step_blocks(const BtStep &s, int n, i64 nblocks, i64 &il, i64 &iu, int &top) {
    if (TR) { // lower (g, h0), upper (g+1, h0+1)
        il = s.g >= 0 ? s.tg - bt_hops<B>(s.g, n) + s.h0 : nblocks;
        iu = s.g + 1 <= s.gB ? s.tg + s.h0 + 1 : nblocks;
        top = s.g * kBtGroup + 1 + s.h0 * B;
    } else { // lower (g-1, h0), upper (g, h0+1)
        il = s.g >= 1 ? s.tg - bt_hops<B>(s.g - 1, n) + s.h0 : nblocks;
        iu = s.g <= s.gB ? s.tg + s.h0 + 1 : nblocks;
        top = (s.g - 1) * kBtGroup + 1 + s.h0 * B;
    }
}

// One half of a block's packed operands: V (half 0) or W (half 1), NP pairs.
template <typename T, int NP>
__device__ __forceinline__ void bt_load_half(T (&o)[2 * NP], const T *__restrict__ ops, i64 id, int half, int lane) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    const T2 *p = reinterpret_cast<const T2 *>(ops + id * (i64)(4 * NP * 64)) + (half * NP) * 64 + lane;
#pragma unroll
    for (int c = 0; c < NP; ++c) {
        const T2 v = p[c * 64];
        o[2 * c] = v.x;
        o[2 * c + 1] = v.y;
    }
}

// M <- Q_b M (TR: Q_b^T M) for columns [0, ncols), one wave per 16 columns, two
// blocks per step (see BtStep), over the passes from the one starting at hop h_first
// to the one starting at hop h_last (in application order; a pass leaves all its
// rows in M, so a launch may end and the next begin at any pass): the two blocks' products are interleaved, so the
// MFMA pipe has independent work while either block's chain drains, and a step's
// fixed costs (LDS round trip, window bookkeeping) are shared by two blocks. The
// window lives in a ring of 2 NKB + 1 tiles of 16 rows per wave (24 KB at b = 64,
// so the kernel leaves LDS for the tridiagonal solver's GEMMs running beside it);
// a step slides it by one tile, storing the tile that leaves and putting the one
// that enters (loaded a step earlier) into the free slot. The next step's operands
// are loaded into the registers of the current ones as these die: V after the X
// products, W after the update.
template <typename T, int B, bool TR>
__global__ __launch_bounds__(kBBThreads) void bc_back_pair_kernel(const T *__restrict__ ops,
                                                                  const i64 *__restrict__ toff,
                                                                  i64 nblocks, const T *__restrict__ zp,
                                                                  T *M, i64 ldm, int n, int ncols, int h_first,
                                                                  int h_last) {
    using MB = MfmaBt<T>;
    using V4 = typename MB::v4;
    constexpr int NKB = B / 16 + 1; // tiles of a block
    constexpr int NC = 4 * NKB;     // 4-row K chunks of a block; values per operand half
    constexpr int NW = 2 * NKB;     // window tiles
    constexpr int NS = NW + 1;      // ring slots: the window and the entering tile
    constexpr int LDW = 17;
    constexpr int SL = 16 * LDW; // one slot
    __shared__ T win[(kBBThreads / 64) * NS * SL];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, q = lane >> 4, li = lane & 15;
    const int wcol = (blockIdx.x * (kBBThreads / 64) + wave) * 16;
    if (wcol >= ncols) return; // whole wave
    const int wb = wave * (NS * SL);
    const int lofs = wb + MB::kidx(0, q) * LDW + li; // a block tile's element of this lane
    int tb = 0;                                      // ring slot of window tile 0
    auto slot = [&](int t) { // ring slot of window tile t in [0, NS)
        const int s = tb + t;
        return s >= NS ? s - NS : s;
    };

// matrix rows [r0, r0+16) of the wave's 16 columns: lane = (row li, columns q + 4i)
#define BP_LOAD16(r0, v)                                                                           \
    do {                                                                                           \
        const int r_ = (r0) + li;                                                                  \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) {                                         \
            const int col_ = wcol + q + 4 * i_;                                                    \
            const bool ok_ = r_ >= 0 && r_ < n && col_ < ncols;                                    \
            (v)[i_] = *(ok_ ? M + r_ + (i64)col_ * ldm : zp);                                      \
        }                                                                                          \
    } while (0)
#define BP_PUT16(s, v)                                                                             \
    do {                                                                                           \
        const int o_ = wb + (s) * SL + li * LDW + q;                                               \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) win[o_ + 4 * i_] = (v)[i_];               \
    } while (0)
#define BP_STORE16(s, r0)                                                                          \
    do {                                                                                           \
        const int r_ = (r0) + li;                                                                  \
        const int o_ = wb + (s) * SL + li * LDW + q;                                               \
        T v_[4];                                                                                   \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) v_[i_] = win[o_ + 4 * i_];                \
        _Pragma("unroll") for (int i_ = 0; i_ < 4; ++i_) {                                         \
            const int col_ = wcol + q + 4 * i_;                                                    \
            if (r_ >= 0 && r_ < n && col_ < ncols) M[r_ + (i64)col_ * ldm] = v_[i_];               \
        }                                                                                          \
    } while (0)

    BtStep cur;
    bt_step_first<B, TR>(cur, n, toff, h_first);
    int top;
    i64 il, iu;
    bt_step_blocks<B, TR>(cur, n, nblocks, il, iu, top);
    T lv[NC], lw[NC], uv[NC], uw[NC];
    bt_load_half<T, NC / 2>(lv, ops, il, 0, lane);
    bt_load_half<T, NC / 2>(uv, ops, iu, 0, lane);
    bt_load_half<T, NC / 2>(lw, ops, il, 1, lane);
    bt_load_half<T, NC / 2>(uw, ops, iu, 1, lane);
    for (int t = 0; t < NW; ++t) {
        T v[4];
        BP_LOAD16(top + 16 * t, v);
        BP_PUT16(t, v);
    }
    T ent[4];
    bool putent = false;
    for (;;) {
        if (putent) BP_PUT16(slot(TR ? NW - 1 : 0), ent);
        BtStep nx = cur;
        const bool more = bt_step_next<B, TR>(nx, n, toff) && (TR ? nx.h0 >= h_last : nx.h0 <= h_last);
        int ntop = top;
        i64 nil = il, niu = iu;
        if (more) bt_step_blocks<B, TR>(nx, n, nblocks, nil, niu, ntop);
        // the tile entering at the next step, if it is in this pass
        BP_LOAD16(TR ? top + 16 * NW : top - 16, ent);

        // ---- the two blocks: lower on window tiles [0, NKB), upper on [NKB, NW)
        T ml[NC], mu[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int r = MB::kidx(c & 3, 0) * LDW;
            ml[c] = win[lofs + slot(c >> 2) * SL + r];
            mu[c] = win[lofs + slot(NKB + (c >> 2)) * SL + r];
        }
        // X = V^T M_J, two chains per block
        V4 xl = {T(0), T(0), T(0), T(0)}, xl1 = xl, xu = xl, xu1 = xl;
#pragma unroll
        for (int c = 0; c < NC; c += 2) {
            xl = MB::mma(lv[c], ml[c], xl);
            xu = MB::mma(uv[c], mu[c], xu);
            xl1 = MB::mma(lv[c + 1], ml[c + 1], xl1);
            xu1 = MB::mma(uv[c + 1], mu[c + 1], xu1);
        }
        xl += xl1;
        xu += xu1;
        // the next step's V (a repeat of this step's when nothing follows)
        bt_load_half<T, NC / 2>(lv, ops, nil, 0, lane);
        bt_load_half<T, NC / 2>(uv, ops, niu, 0, lane);
        // M_J += W X (W = -V T, or -V T^T). All products are issued before any
        // store: a store waits for its product, and in-order issue would hold the
        // remaining products behind it.
        V4 al[NKB], au[NKB];
#pragma unroll
        for (int t = 0; t < NKB; ++t) {
            al[t] = V4{ml[4 * t], ml[4 * t + 1], ml[4 * t + 2], ml[4 * t + 3]};
            au[t] = V4{mu[4 * t], mu[4 * t + 1], mu[4 * t + 2], mu[4 * t + 3]};
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int t = 0; t < NKB; ++t) {
                al[t] = MB::mma(lw[4 * t + r], xl[r], al[t]);
                au[t] = MB::mma(uw[4 * t + r], xu[r], au[t]);
            }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < NKB; ++t) {
            const int ol = lofs + slot(t) * SL, ou = lofs + slot(NKB + t) * SL;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                win[ol + MB::kidx(r, 0) * LDW] = al[t][r];
                win[ou + MB::kidx(r, 0) * LDW] = au[t][r];
            }
        }

        // ---- window
        if (!more || nx.h0 != cur.h0) {
            for (int t = 0; t < NW; ++t) BP_STORE16(slot(t), top + 16 * t);
            if (!more) break;
            wave_stores_done(); // the next pass reads these rows back through other lanes
            tb = 0;
            for (int t = 0; t < NW; ++t) {
                T v[4];
                BP_LOAD16(ntop + 16 * t, v);
                BP_PUT16(t, v);
            }
            putent = false;
        } else {
            if (TR) {
                BP_STORE16(tb, top);
                tb = tb + 1 == NS ? 0 : tb + 1;
            } else {
                BP_STORE16(slot(NW - 1), top + 16 * (NW - 1));
                tb = tb == 0 ? NS - 1 : tb - 1;
            }
            putent = true;
        }
        // the next step's W
        bt_load_half<T, NC / 2>(lw, ops, nil, 1, lane);
        bt_load_half<T, NC / 2>(uw, ops, niu, 1, lane);
        cur = nx;
        il = nil;
        iu = niu;
        top = ntop;
    }
#undef BP_LOAD16
#undef BP_PUT16
#undef BP_STORE16
}

template <typename T, int B>
int launch_bc_back(Workspace<T> &ws, hipStream_t s, const T *U, T *M, i64 ldm, bool tr) {
    const int n = ws.n;
    if (ws.bt_blocks <= 0) return CUEV_STATUS_SUCCESS;
    {
        prof::Scope sc(s, "bt.bc_pack", (double)ws.bt_blocks * kBtGroup * kBtGroup * B);
        const dim3 grid((unsigned)ceil_div(ws.bt_blocks + 1, 4)); // + the zero block
        if (tr)
            hipLaunchKernelGGL((bt_pack_kernel<T, B, true>), grid, dim3(kBTThreads), 0, s, U, ws.ldu,
                               ws.bt_off, ws.bt_groups, ws.bt_blocks, n, ws.BtT);
        else
            hipLaunchKernelGGL((bt_pack_kernel<T, B, false>), grid, dim3(kBTThreads), 0, s, U, ws.ldu,
                               ws.bt_off, ws.bt_groups, ws.bt_blocks, n, ws.BtT);
        CUEV_LAUNCH_CHECK();
    }
    // 4 flops per reflector entry per column: sum_s (n-1-s) entries ~ n^2/2
    prof::Scope sc(s, "bt.bc_back", 2.0 * (double)n * n * n);
    const T *zp = U + ws.ldu * (i64)n; // zeros past U's last column
    // Columns in chunks of one resident round each (one wave per 16 columns, waves
    // that start together): every wave streams the same block operands, and waves
    // that start together share them in L2. In one launch of more waves than fit, the
    // later rounds start as earlier waves finish, spread out along the stream, and
    // lose that sharing (n = 49152, transposed: 10.6 s in one launch, 6.9 s in three).
    int per_cu = 0;
    if constexpr (B == 64)
        CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc_back_pair_kernel<T, B, true>,
                                                              kBBThreads, 0));
    else
        CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc_back_kernel<T, B, true>,
                                                              kBBThreads, 0));
    const i64 cols_per_block = 16 * (kBBThreads / 64);
    const i64 chunk = std::max<i64>(1, (i64)per_cu * device_cu_count()) * cols_per_block;
    // The paired kernel's passes in launches of at least kBtSegBlocks blocks each (the
    // last one fewer): the waves of one launch drift apart along the block stream as
    // it goes, and a launch boundary lines them up again (measured at n = 16384,
    // 1.3e5 blocks: one launch 252-260 ms, two 245 ms, four 243.6 ms, eight 244.7 ms,
    // sixteen 245.3 ms; at n = 49152, 1.18e6 blocks per column chunk: one launch
    // 7.6-12 s, nine 6.6 s).
    std::vector<std::pair<int, int>> segs; // (first, last) pass start hops, in application order
    if constexpr (B == 64) {
        auto gmax = [&](int h) {
            const int m = std::min(n - 3, n - 2 - h * B);
            return m < 0 ? -1 : m / kBtGroup;
        };
        std::vector<int> passes;
        if (tr) {
            for (int h0 = ((n - 2) / B) / 2 * 2; h0 >= 0; h0 -= 2) passes.push_back(h0);
        } else {
            for (int h0 = 0; gmax(h0) >= 0; h0 += 2) passes.push_back(h0);
        }
        i64 acc = 0;
        for (size_t q = 0; q < passes.size(); ++q) {
            if (acc == 0) segs.emplace_back(passes[q], passes[q]);
            segs.back().second = passes[q];
            acc += (i64)(gmax(passes[q]) + 1) + (i64)(gmax(passes[q] + 1) + 1);
            if (acc >= kBtSegBlocks) acc = 0;
        }
    }
    for (i64 c0 = 0; c0 < n; c0 += chunk) {
        const int nc = (int)std::min<i64>(chunk, n - c0);
        const dim3 grid((unsigned)ceil_div(nc, cols_per_block));
        T *Mc = M + c0 * ldm;
        if constexpr (B == 64) {
            // two blocks per step (passes of two hops: the default band's window fits)
            for (const auto &sg : segs) {
                if (tr)
                    hipLaunchKernelGGL((bc_back_pair_kernel<T, B, true>), grid, dim3(kBBThreads), 0, s, ws.BtT,
                                       ws.bt_off, ws.bt_blocks, zp, Mc, ldm, n, nc, sg.first, sg.second);
                else
                    hipLaunchKernelGGL((bc_back_pair_kernel<T, B, false>), grid, dim3(kBBThreads), 0, s, ws.BtT,
                                       ws.bt_off, ws.bt_blocks, zp, Mc, ldm, n, nc, sg.first, sg.second);
                CUEV_LAUNCH_CHECK();
            }
        } else {
            if (tr)
                hipLaunchKernelGGL((bc_back_kernel<T, B, true>), grid, dim3(kBBThreads), 0, s, ws.BtT,
                                   ws.bt_off, zp, Mc, ldm, n, nc);
            else
                hipLaunchKernelGGL((bc_back_kernel<T, B, false>), grid, dim3(kBBThreads), 0, s, ws.BtT,
                                   ws.bt_off, zp, Mc, ldm, n, nc);
        }
        CUEV_LAUNCH_CHECK();
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int bc_back_any(Workspace<T> &ws, hipStream_t s, const T *U, T *M, i64 ldm, bool tr) {
    if (ldm < ws.n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back: ldm < n");
    if (ws.n < 3) return CUEV_STATUS_SUCCESS;
    switch (ws.b) {
    case 16: return launch_bc_back<T, 16>(ws, s, U, M, ldm, tr);
    case 32: return launch_bc_back<T, 32>(ws, s, U, M, ldm, tr);
    case 64: return launch_bc_back<T, 64>(ws, s, U, M, ldm, tr);
    default: CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back: unsupported b=%d", ws.b);
    }
}

// What for the block of panels i .. i+np-1 into ws.Z (rows r0 = i+b .., ld ws.ld)
template <typename T>
int build_what(Workspace<T> &ws, hipStream_t s, const T *Y, const T *W, int i, int np, Scratch gs) {
    const int n = ws.n, b = ws.b;
    const i64 ld = ws.ld;
    const int r0 = i + b;
    const i64 R0 = n - r0;
    T *Wh = ws.Z + r0;
    for (int t = 0; t < np; ++t) {
        const int j = i + t * b, rt = j + b; // panel t: columns j.., rows rt..
        T *Wt = ws.Z + (i64)t * b * ld;
        CUEV_TRY(fill2d<T>(s, Wt + r0, ld, rt - r0, b, T(0), T(0)));
        CUEV_TRY(copy2d<T>(s, W + (i64)j * ld + rt, ld, Wt + rt, ld, n - rt, b));
        if (t == 0) continue;
        // C = Yhat_{0:t}^T W_t over rows rt.. (where W_t lives), t b x b
        CUEV_TRY(gemm<T>(s, 1, 0, (i64)t * b, b, n - rt, T(1), Y + (i64)i * ld + rt, ld,
                         W + (i64)j * ld + rt, ld, T(0), ws.Dwk, (i64)t * b, gs));
        // What_t -= What_{0:t} C over rows r0..
        CUEV_TRY(gemm<T>(s, 0, 0, R0, b, (i64)t * b, T(-1), Wh, ld, ws.Dwk, (i64)t * b, T(1), Wt + r0,
                         ld, gs));
    }
    return CUEV_STATUS_SUCCESS;
}

} // namespace

template <typename T> int bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm) {
    return bc_back_any<T>(ws, ws.stream, U, M, ldm, false);
}

template <typename T> int bc_back_t(Workspace<T> &ws, hipStream_t s, const T *U, T *M, i64 ldm) {
    return bc_back_any<T>(ws, s, U, M, ldm, true);
}

template <typename T> int sbr_back(Workspace<T> &ws, const T *Y, const T *W, T *M, i64 ldm) {
    const int n = ws.n, b = ws.b, k = ws.k;
    const i64 ld = ws.ld;
    if (ldm < n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "sbr_back: ldm < n");
    Scratch gs{ws.gscratch, ws.gscratch_bytes};
    hipStream_t s = ws.stream;
    // Blocks of the reduction (k columns, np = k/b panels), last block first. Block
    // i applies H_i0 H_i1 ... = I - What Yhat^T with Yhat = [Y_i0 .. Y_inp-1] and
    // What built left to right: What_t = W_t - What_{0:t} (Yhat_{0:t}^T W_t)
    // (what panel t's W becomes once the earlier panels' reflectors are folded in).
    // Then Zt = Yhat^T M (one b-row product per panel, over that panel's rows) and
    // one rank-(np b) update M -= What Zt.
    int ilast = -1;
    for (int i = 0; i + b < n; i += k) ilast = i;
    for (int i = ilast; i >= 0; i -= k) {
        int np = 0;
        while (np * b < k && i + np * b + b < n) ++np;
        const int r0 = i + b; // first row of the block's reflectors
        const i64 R0 = n - r0;
        {
            prof::Scope sc(s, "bt.sbr_what");
            CUEV_TRY(build_what<T>(ws, s, Y, W, i, np, gs));
        }
        prof::Scope sc(s, "bt.sbr_back", 4.0 * (double)R0 * np * b * n);
        for (int t = 0; t < np; ++t) {
            const int j = i + t * b, rt = j + b;
            CUEV_TRY(gemm<T>(s, 1, 0, b, n, n - rt, T(1), Y + (i64)j * ld + rt, ld, M + rt, ldm, T(0),
                             ws.Zt + (i64)t * b, k, gs));
        }
        CUEV_TRY(gemm<T>(s, 0, 0, R0, n, (i64)np * b, T(-1), ws.Z + r0, ld, ws.Zt, k, T(1), M + r0, ldm,
                         gs));
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int sbr_back_t(Workspace<T> &ws, hipStream_t s, const T *Y, const T *W, T *M, i64 ldm) {
    const int n = ws.n, b = ws.b, k = ws.k;
    const i64 ld = ws.ld;
    if (ldm < n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "sbr_back_t: ldm < n");
    Scratch gs{ws.gscratch, ws.gscratch_bytes};
    // M = Q_s^T = Q_T^T ... Q_1^T Q_0^T with Q_t^T = I - Yhat_t What_t^T for the
    // reduction's blocks t (Yhat the block's Y columns, zero above each panel's rows,
    // copied into ws.Yh since Y there holds nothing defined). Built from the right,
    // last block first: M = I, then M <- M Q_t^T = M - (M Yhat_t) What_t^T for
    // t = T .. 0. Before block t, M differs from I only in its trailing block from
    // row and column r0_t + k on, and Yhat_t, What_t vanish above row r0_t, so the
    // update touches the trailing (n - r0_t)^2 block alone:
    //   P = M[r0:, r0:] Yhat_t[r0:],   M[r0:, r0:] -= P What_t[r0:]^T
    // (4/3 n^3 flops in all, against 2 n^3 for applying the blocks from the left to
    // every column).
    CUEV_TRY(fill2d<T>(s, M, ldm, n, n, T(0), T(1)));
    int nblk = 0;
    for (int i = 0; i + b < n; i += k) ++nblk;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int i = bi * k;
        int np = 0;
        while (np * b < k && i + np * b + b < n) ++np;
        const int r0 = i + b;
        const i64 R0 = n - r0, kb = (i64)np * b;
        {
            prof::Scope sc(s, "bt.sbr_what");
            CUEV_TRY(build_what<T>(ws, s, Y, W, i, np, gs));
            for (int t = 0; t < np; ++t) {
                const int j = i + t * b, rt = j + b;
                T *Yt = ws.Yh + (i64)t * b * ld;
                CUEV_TRY(fill2d<T>(s, Yt + r0, ld, rt - r0, b, T(0), T(0)));
                CUEV_TRY(copy2d<T>(s, Y + (i64)j * ld + rt, ld, Yt + rt, ld, n - rt, b));
            }
        }
        prof::Scope sc(s, "bt.sbr_back", 4.0 * (double)R0 * R0 * kb);
        T *Mt = M + r0 + (i64)r0 * ldm;
        // P (R0 x kb, ld n) in ws.Zt, which holds k x n elements
        CUEV_TRY(gemm<T>(s, 0, 0, R0, kb, R0, T(1), Mt, ldm, ws.Yh + r0, ld, T(0), ws.Zt, n, gs));
        CUEV_TRY(gemm<T>(s, 0, 1, R0, R0, kb, T(-1), ws.Zt, n, ws.Z + r0, ld, T(1), Mt, ldm, gs));
    }
    return CUEV_STATUS_SUCCESS;
}

template int bc_back<float>(Workspace<float> &, const float *, float *, i64);
template int bc_back<double>(Workspace<double> &, const double *, double *, i64);
template int bc_back_t<float>(Workspace<float> &, hipStream_t, const float *, float *, i64);
template int bc_back_t<double>(Workspace<double> &, hipStream_t, const double *, double *, i64);
template int sbr_back_t<float>(Workspace<float> &, hipStream_t, const float *, const float *, float *, i64);
template int sbr_back_t<double>(Workspace<double> &, hipStream_t, const double *, const double *, double *,
                                i64);
template int sbr_back<float>(Workspace<float> &, const float *, const float *, float *, i64);
template int sbr_back<double>(Workspace<double> &, const double *, const double *, double *, i64);

} // namespace cuev
