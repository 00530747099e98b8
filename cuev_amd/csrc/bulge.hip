// bulge.hip — band -> tridiagonal reduction (bulge chasing) and (d, e) extraction.
//
// Algorithm (reference: src/cuda/bulge.cu:64-224, Algorithm 2 of Wang et al.):
// sweep s = 0..n-3 eliminates column s below its subdiagonal with a Householder
// reflector on rows [s+1, s+b], then chases the bulge this creates down the band:
// hop h >= 1 eliminates the bulge column cq = s+1+(h-1)b on rows
// [cq+b, cq+2b-1]. Each hop is the similarity A <- H A H with H = I - 2 w w^T,
// ||w|| = 1, applied as A <- A - w y^T - y w^T, p = 2 A w, kappa = w^T p,
// y = p - kappa w. The reflector of hop h of sweep s occupies rows
// [s+1+hb, min(s+(h+1)b, n-1)] and is stored for the back-transform.
//
// Our design (not the reference's):
//  * Reflector convention: if the vector below the pivot is exactly zero the
//    hop is the identity (w = 0), as LAPACK's dlarfg does; the reference's
//    normalised reflector would flip the sign of the pivot instead.
//  * Each hop works on its window in LDS: the packed band columns [cq, r1]
//    (all 2b stored rows) hold every entry the hop reads or writes. The window
//    slides with the sweep: consecutive hops share b columns, so a hop loads
//    only the b new columns and writes back only the b columns the sweep has
//    finished with (a ring of 2b column slots). The Householder vector is formed
//    by wave 0 alone (L <= 64 = one wave64), with FP64 norms.
//  * Dependency: hop (s, cq, r1) touches packed columns [cq, r1] only, and the
//    previous sweep's hops touch columns [cq'', r1''] with cq'' increasing. Hop h of
//    sweep s reads columns up to r1 = r0 + b - 1, which sweep s-1 last changes in its
//    hop h+1 (the rank-2 update of that hop's R1 columns and of its column r0'' = r1,
//    and the elimination of column r1 by its hop h+2's reflector). So a hop publishes
//    progress prog[s] = r0 early: after R1 and column r0 are updated, the next hop's
//    reflector is formed from column r0, and R1's columns and column r0 are written
//    back; the rest of its rank-2 update (R2 and R3 beyond column r0) then runs beside
//    the next sweep's hop, which waits for prog[s-1] >= r1. A hop's reflector is thus
//    formed in the hop before, so a hop runs: wait for sweep s-1, load the b new
//    columns, p and kappa, R1 and column r0, the next reflector, write-back and
//    progress, the rest of R2 and R3. The next sweep's first hop waits for columns up
//    to s + 1 + b (published by hop 1), so hop 0 publishes nothing. The sweep-to-sweep
//    lag is about one hop plus the part of the next hop before its publish; the
//    reference waits for a margin of 3b columns (bulge.cu:193-199).
//  * Hand-offs use agent-scope (write-through) band accesses and flags, with no L2
//    write-back or invalidate (see ld_agent / st_agent below); every wait is bounded,
//    and the grid never exceeds the number of co-resident workgroups, so the lowest
//    unfinished sweep can always proceed.
//  * U is stored shifted: U[(r - s - 1) + s*ldu] holds row r of sweep s's
//    reflectors, so the back-transform reads aligned, contiguous segments.
#include "bulge.hpp"

#include <algorithm>

#include "profile.hpp"
#include "util.hpp"

namespace cuev {
namespace {

constexpr int kBCThreads = 512; // 8 waves: two per SIMD, so LDS latency overlaps
constexpr int kDone = 0x3fffffff; // prog value of a finished sweep
constexpr int kMaxPhases = 8;

// Sweep schedule in phases: sweeps [s0[p], s0[p+1]) go round-robin to workgroups
// 0 .. g[p]-1 (g non-increasing). A sweep of h hops is active for ~h hop times and a
// new one starts every ~2 hop times, so fewer sweeps are in flight as they shorten;
// a workgroup with no sweep left in the later phases exits and frees its CU for the
// back-transform products running beside the chase.
struct ChasePhases {
    int np;
    int s0[kMaxPhases + 1];
    int g[kMaxPhases];
};

// Band traffic between workgroups. Every global load and store of the packed band is
// an agent-scope relaxed atomic access: on gfx950 these carry the sc1 bit, so a store
// is written through to the device-coherent level (past the XCD-private L2) and a
// load reads from it. A hand-off then needs no L2 write-back and no L2 invalidate:
//   producer: band stores (sc1), s_waitcnt vmcnt(0) in every wave (the stores have
//             completed at agent scope), workgroup barrier, one flag store (sc1);
//   consumer: one thread polls the flag (sc1 loads), workgroup barrier, band loads
//             (sc1), issued only after the flag was seen.
// The reflectors (U) are read only by later kernels and use plain stores.
template <typename T> __device__ __forceinline__ T ld_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void st_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// progress flag store after wg_publish_fence (every wave's band stores complete)
__device__ __forceinline__ void flag_store_agent(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T, int B>
__global__ __launch_bounds__(kBCThreads) void bc_chase_kernel(T *Bp, T *U, i64 ldu, int n,
                                                              int *prog, int *err, const ChasePhases ph) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int b = B;
    constexpr int NEW_PER_THREAD = 2 * B * B / kBCThreads; // b new columns of 2b rows per hop
    const int ldb = 2 * b;       // packed band leading dimension
    // LDS leading dimension of a window column: odd, so that reading along a row
    // (stride WLD doubles across lanes) spreads over the banks
    const int WLD = 2 * b + 1;
    const int NSLOT = 2 * b;     // ring of window columns: matrix column j lives in slot j % NSLOT
    constexpr int NW = kBCThreads / 64;    // waves per workgroup
    // k segments of the p product: (row, segment) pairs cover the workgroup's threads
    constexpr int NSEG = (kBCThreads / (3 * B)) < 1 ? 1 : (kBCThreads / (3 * B));
    T *win = reinterpret_cast<T *>(smem_raw);          // [NSLOT][WLD]
    T *sw = win + (size_t)NSLOT * WLD;                 // [2][b] reflectors: this hop's, the next hop's
    T *sp = sw + 2 * b;                                // [3b]  p
    T *spart = sp + 3 * b;                             // [NSEG][3b] partial sums of p
    __shared__ T sred[NW];
    __shared__ double s_beta[2];
    __shared__ int s_go;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // window element A(i, j) (i >= j, both in the current window)
    auto W_ = [&](int i, int j) -> T & { return win[(j & (NSLOT - 1)) * WLD + (i - j)]; };
    // symmetric read A(i, k) = A(max, min): the address by min/max, without a branch
    auto Wsym = [&](int i, int k) -> T {
        const int hi_ = max(i, k), lo_ = min(i, k);
        return win[(lo_ & (NSLOT - 1)) * WLD + (hi_ - lo_)];
    };
    // copy packed columns [c0, c1] between global memory and their ring slots: in
    // rounds of FIRST_PER_THREAD loads per thread, all issued before any LDS store (one
    // round trip per round; a sweep's first window of b + 1 columns takes one)
    constexpr int FIRST_PER_THREAD = (2 * B * (B + 1) + kBCThreads - 1) / kBCThreads;
    auto load_cols = [&](int c0, int c1) {
        if (c1 < c0) return;
        const int cnt = (c1 - c0 + 1) * ldb;
        const T *src = Bp + (i64)c0 * ldb;
        for (int base = 0; base < cnt; base += FIRST_PER_THREAD * kBCThreads) {
            T v[FIRST_PER_THREAD];
#pragma unroll
            for (int u = 0; u < FIRST_PER_THREAD; ++u) {
                const int idx = base + tid + u * kBCThreads;
                v[u] = (idx < cnt) ? ld_agent(src + idx) : T(0);
            }
#pragma unroll
            for (int u = 0; u < FIRST_PER_THREAD; ++u) {
                const int idx = base + tid + u * kBCThreads;
                if (idx < cnt) {
                    const int jj = idx / ldb, d = idx - jj * ldb;
                    win[((c0 + jj) & (NSLOT - 1)) * WLD + d] = v[u];
                }
            }
        }
    };
    // exactly b new columns: all loads issued before any LDS store (one round trip)
    auto load_b_cols = [&](int c0) {
        const T *src = Bp + (i64)c0 * ldb;
        T v[NEW_PER_THREAD];
#pragma unroll
        for (int u = 0; u < NEW_PER_THREAD; ++u) v[u] = ld_agent(src + tid + u * kBCThreads);
#pragma unroll
        for (int u = 0; u < NEW_PER_THREAD; ++u) {
            const int idx = tid + u * kBCThreads;
            const int jj = idx / ldb, d = idx - jj * ldb;
            win[((c0 + jj) & (NSLOT - 1)) * WLD + d] = v[u];
        }
    };
    auto store_cols = [&](int c0, int c1) {
        if (c1 < c0) return;
        const int cnt = (c1 - c0 + 1) * ldb;
        T *dst = Bp + (i64)c0 * ldb;
        if (cnt <= NEW_PER_THREAD * kBCThreads) {
            // up to b columns (every hop): all LDS reads before the stores
            T v[NEW_PER_THREAD];
#pragma unroll
            for (int u = 0; u < NEW_PER_THREAD; ++u) {
                const int idx = tid + u * kBCThreads;
                const int jj = idx / ldb, d = idx - jj * ldb;
                v[u] = win[((c0 + jj) & (NSLOT - 1)) * WLD + d];
            }
#pragma unroll
            for (int u = 0; u < NEW_PER_THREAD; ++u) {
                const int idx = tid + u * kBCThreads;
                if (idx < cnt) st_agent(dst + idx, v[u]);
            }
            return;
        }
        for (int idx = tid; idx < cnt; idx += kBCThreads) {
            const int jj = idx / ldb, d = idx - jj * ldb;
            st_agent(dst + idx, win[((c0 + jj) & (NSLOT - 1)) * WLD + d]);
        }
    };

    // Householder vector of x = A[r0_ : r0_ + L_ - 1, c_] (column c_ of the window),
    // by wave 0 alone: w into wd, beta into *bd
    auto householder = [&](int r0_, int c_, int L_, T *wd, double *bd) {
        const double x = (lane < L_) ? (double)W_(r0_ + lane, c_) : 0.0;
        const double alpha = __shfl(x, 0, 64);
        const double tail = wave_sum((lane >= 1) ? x * x : 0.0);
        double wv = 0.0, beta = alpha;
        if (tail != 0.0) {
            const double nrm = sqrt(alpha * alpha + tail);
            beta = (alpha >= 0.0) ? -nrm : nrm;
            const double v0 = alpha - beta;
            const double inv = 1.0 / sqrt(v0 * v0 + tail);
            wv = ((lane == 0) ? v0 : x) * inv;
        }
        if (lane < L_) wd[lane] = (T)wv;
        if (lane == 0) *bd = beta;
    };
    // column c_ is final: write it back, the eliminated entries [o_, o_ + L_) as beta e_0
    // (skip_pivot: all but entry o_)
    auto writeback = [&](int c_, int o_, int L_, double beta_, bool skip_pivot = false) {
        T *dst = Bp + (i64)c_ * ldb;
        const T *srcw = win + (size_t)(c_ & (NSLOT - 1)) * WLD;
        for (int d = tid; d < ldb; d += kBCThreads) {
            if (skip_pivot && d == o_) continue;
            st_agent(dst + d, (d == o_) ? (T)beta_ : ((d > o_ && d < o_ + L_) ? T(0) : srcw[d]));
        }
    };

    // this workgroup's sweeps in increasing order: the next one of the current phase,
    // else its first one in a later phase that still uses it (none left: n)
    int phase = 0;
    auto settle = [&](int s_) {
        while (s_ >= ph.s0[phase + 1]) {
            if (++phase >= ph.np) return n;
            s_ = ((int)blockIdx.x < ph.g[phase]) ? ph.s0[phase] + (int)blockIdx.x : ph.s0[phase + 1];
        }
        return s_;
    };
    int s = settle(((int)blockIdx.x < ph.g[0]) ? ph.s0[0] + (int)blockIdx.x : ph.s0[1]);
    for (; s < n - 2;) {
        int held_lo = 0, held_hi = -1; // window columns currently held in LDS
        T *wc = sw, *wn = sw + b;      // this hop's reflector, the next hop's
        double *bc = &s_beta[0], *bn = &s_beta[1];
        for (int h = 0;; ++h) {
            const int cq = (h == 0) ? s : s + 1 + (h - 1) * b;
            const int r0 = (h == 0) ? s + 1 : cq + b;
            if (r0 > n - 1) break;
            const int r1 = min(r0 + b - 1, n - 1);
            const int L = r1 - r0 + 1;
            if (L < 2) {
                // nothing below the pivot: identity reflector (record it for the back-transform)
                if (tid == 0) U[(i64)(r0 - s - 1) + (i64)s * ldu] = T(0);
                break;
            }
            const int hi = min(n - 1, r1 + b);  // last row touched

            // ---- wait for sweep s-1 to be done with columns <= r1 (acquire), then bring
            // the window [cq, r1] into LDS: columns already held stay, the rest come from
            // global memory (b columns per hop after the first)
            auto wait_and_load = [&]() -> bool {
                if (s > 0) {
                    if (tid == 0) {
                        long long it = 0;
                        int go = 1;
                        while (flag_load_relaxed(prog + (s - 1)) < r1) {
                            __builtin_amdgcn_s_sleep(1);
                            if (++it > kSpinLimit) {
                                __hip_atomic_store(err, (int)DEV_WATCHDOG, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                                go = 0;
                                break;
                            }
                        }
                        s_go = go;
                    }
                    __syncthreads();
                    if (!s_go) return false;
                }
                if (held_hi < held_lo) load_cols(cq, r1);
                else if (r1 - held_hi == b) load_b_cols(held_hi + 1);
                else load_cols(max(cq, held_hi + 1), r1);
                held_lo = cq;
                held_hi = r1;
                __syncthreads();
                return true;
            };
            if (!wait_and_load()) return;
            if (h == 0) {
                // the sweep's first reflector, on column s; column s is then final (rows
                // above r0 untouched by this sweep, beta e_0 below): write it back and
                // publish progress s. Later hops find their reflector formed and their
                // column cq written back and published by the hop before.
                // No progress is published for column s (or s + 1 below): the next
                // sweep first waits for columns up to s + 1 + b, published by hop 1,
                // whose fence also covers these stores.
                if (wave == 0) householder(r0, cq, L, wc, bc);
                __syncthreads();
                writeback(cq, r0 - cq, L, *bc);
            }

            // ---- p_t = 2 sum_k A(cq+t, r0+k) w_k, t < nrow. Thread = (row t, k segment);
            // each thread walks its segment starting at a row-dependent offset, so the
            // lanes of a wave (consecutive rows) hit distinct LDS banks both for
            // A(i, k), i >= k (column k, stride 1 across lanes) and for A(k, i), i < k
            // (column i: stride WLD - 1 across lanes, which without the rotation puts
            // every lane on one bank)
            const int nrow = hi - cq + 1; // <= 3b
            const int o = r0 - cq;
            // b = 64: rows in six waves of 32 lanes, the two k segments of a row in the
            // two halves of a wave, so p_t is completed with one lane exchange and
            // kappa's partial sums come out of the same phase (one barrier)
            constexpr bool kWaveSeg = 3 * B == 192 && NSEG == 2 && kBCThreads >= 384;
            {
                const int t = kWaveSeg ? wave * 32 + (lane & 31) : tid % (3 * B);
                const int seg = kWaveSeg ? (wave < 6 ? lane >> 5 : NSEG) : tid / (3 * B);
                const int SL = (L + NSEG - 1) / NSEG;
                const int kb = seg * SL, ke = min(L, kb + SL), len = ke - kb;
                T acc = T(0);
                constexpr int FL = B / NSEG; // segment length of a full hop
                if (seg < NSEG && t < nrow && len > 0) {
                    const int i = cq + t;
                    if (B % NSEG == 0 && (FL & (FL - 1)) == 0 && len == FL) {
                        // full segment: unrolled with four partial sums, so the LDS
                        // reads of consecutive steps overlap (same rotation as below)
                        T a4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 8
                        for (int j = 0; j < FL; ++j) {
                            const int q = kb + ((t + j) & (FL - 1));
                            const int k = r0 + q;
                            a4[j & 3] += Wsym(i, k) * wc[q];
                        }
                        acc = (a4[0] + a4[1]) + (a4[2] + a4[3]);
                    } else {
                        int qq = t % len;
                        for (int j = 0; j < len; ++j) {
                            const int q = kb + qq;
                            const int k = r0 + q;
                            const T a = Wsym(i, k);
                            acc += a * wc[q];
                            qq = (qq + 1 == len) ? 0 : qq + 1;
                        }
                    }
                }
                if (kWaveSeg) {
                    // ---- p (segment 0 + segment 1, as below), and kappa = w^T p over
                    // the reflector's rows [o, o + L)
                    const T other = __shfl_xor(acc, 32, 64);
                    T kv = T(0);
                    if (seg == 0 && t < nrow) {
                        const T pv = T(2) * (acc + other);
                        sp[t] = pv;
                        if (t >= o && t < o + L) kv = pv * wc[t - o];
                    }
                    kv = wave_sum(kv);
                    if (lane == 0) sred[wave] = kv;
                } else if (seg < NSEG && t < 3 * B) {
                    spart[seg * (3 * B) + t] = acc;
                }
            }
            __syncthreads();
            if (!kWaveSeg) {
                // ---- p, and kappa = w^T p over the reflector's rows [o, o + L)
                T kv = T(0);
                if (tid < nrow) {
                    T ps = T(0);
#pragma unroll
                    for (int v = 0; v < NSEG; ++v) ps += spart[v * (3 * B) + tid];
                    const T pv = T(2) * ps;
                    sp[tid] = pv;
                    if (tid >= o && tid < o + L) kv = pv * wc[tid - o];
                }
                kv = wave_sum(kv);
                if (lane == 0) sred[wave] = kv;
                __syncthreads();
            }

            // ---- rank-2 update A -= w y^T + y w^T with y = p - kappa w, on the three
            // regions of the window where w_i or w_j is non-zero:
            //   R1: rows r0+a (a < L)      x columns cq+c, 0 <= c < o  (w_j = 0); c = 0 is the
            //       eliminated column, set to its exact value beta e_0
            //   R2: rows r0+a (a < L)      x columns r0+c, c <= a     (both non-zero)
            //   R3: rows r1+1+a (a < m3)   x columns r0+c, c < L      (w_i = 0)
            // thread = (row a, column group cg); the column value is a broadcast read.
            // part 1: R1, and column c = 0 of R2 and R3 (column r0, the next hop's
            // reflector column); part 2: the rest of R2 and R3; part 3: all of it
            auto rank2 = [&](int part) {
                const bool dR1 = part != 2, dC0 = part != 2, dRest = part != 1;
                T kap = T(0);
#pragma unroll
                for (int v = 0; v < NW; ++v) kap += sred[v];
                constexpr int NG = kBCThreads / B;
                const int a = tid % B, cg = tid / B;
                const int m3 = hi - r1;
                const T wa = (a < L) ? wc[a] : T(0);
                const T ya = (a < L) ? sp[o + a] - kap * wa : T(0);
                const T y3 = (a < m3) ? sp[o + L + a] : T(0);
                const T bet = (T)*bc;
                constexpr bool kFull = B >= NG && B % NG == 0; // whole columns per thread group
                // R1 has o < B columns in a sweep's first hop (o = 1): it then goes through
                // the general loop, and R2 and R3 through the fast path below
                const bool full23 = kFull && L == B && m3 == B;
                const bool split1 = full23 && o != B;
                if (split1 && dR1 && a < L)
                    for (int cc = cg; cc < o; cc += NG) {
                        T &e = W_(r0 + a, cq + cc);
                        if (cc == 0) e = (a == 0) ? bet : T(0);
                        else e -= wa * sp[cc];
                    }
                if (full23) {
                    const bool fR1 = dR1 && !split1;
                    // full hop (every thread in all three regions, each element read and
                    // written by its own thread only): every LDS read before any write,
                    // so the round trips overlap; same arithmetic as the general path
                    constexpr int NU = kFull ? B / NG : 1;
                    T e1[NU], e2[NU], e3[NU], p1[NU], wcv[NU], pc[NU];
#pragma unroll
                    for (int u = 0; u < NU; ++u) {
                        const int cc = cg + NG * u;
                        const bool d23 = (cc == 0) ? dC0 : dRest;
                        if (fR1) {
                            e1[u] = W_(r0 + a, cq + cc);
                            p1[u] = sp[cc];
                        }
                        if (d23) {
                            e2[u] = W_(r0 + a, r0 + min(cc, a));
                            wcv[u] = wc[cc];
                            pc[u] = sp[o + cc];
                            e3[u] = W_(r1 + 1 + a, r0 + cc);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < NU; ++u) {
                        const int cc = cg + NG * u;
                        const bool d23 = (cc == 0) ? dC0 : dRest;
                        if (fR1) W_(r0 + a, cq + cc) = (cc == 0) ? ((a == 0) ? bet : T(0)) : e1[u] - wa * p1[u];
                        if (d23) {
                            if (cc <= a) W_(r0 + a, r0 + cc) = e2[u] - (wa * (pc[u] - kap * wcv[u]) + ya * wcv[u]);
                            W_(r1 + 1 + a, r0 + cc) = e3[u] - y3 * wcv[u];
                        }
                    }
                } else {
                    if (a < L) {
                        if (dR1)
                            for (int cc = cg; cc < o; cc += NG) {
                                T &e = W_(r0 + a, cq + cc);
                                if (cc == 0) e = (a == 0) ? bet : T(0);
                                else e -= wa * sp[cc];
                            }
                        for (int cc = cg; cc <= a; cc += NG) {
                            if (!((cc == 0) ? dC0 : dRest)) continue;
                            const T wcc = wc[cc];
                            const T yc = sp[o + cc] - kap * wcc;
                            W_(r0 + a, r0 + cc) -= wa * yc + ya * wcc;
                        }
                    }
                    if (a < m3) {
                        for (int cc = cg; cc < L; cc += NG) {
                            if (!((cc == 0) ? dC0 : dRest)) continue;
                            W_(r1 + 1 + a, r0 + cc) -= y3 * wc[cc];
                        }
                    }
                }
            };

            T *ucol = U + (i64)(r0 - s - 1) + (i64)s * ldu;
            // the next hop: cq' = r0, its reflector on rows [r0 + b, ...] of column r0
            const int nr0 = r0 + b, nr1 = min(nr0 + b - 1, n - 1);
            const bool next_hh = nr0 <= n - 1 && nr1 - nr0 + 1 >= 2;
            if (next_hh) {
                // Early publish. The next sweep's hop that waits for this one reads the
                // columns up to r0 only: R1's columns [cq, r0) and column r0, which are
                // final once R1 and column r0 are updated and the next hop's reflector has
                // eliminated column r0 (this sweep later writes column r0 only with those
                // same values, and reads it only in LDS). So: R1 and column r0; the next
                // reflector (wave 0) while the other waves store R1's columns and U; column
                // r0 written back; progress r0 published; then the rest of R2 and R3,
                // beside the next sweep's hop.
                rank2(1);
                __syncthreads();
                if (wave == 0) householder(nr0, r0, nr1 - nr0 + 1, wn, bn);
                for (int t = tid; t < L; t += kBCThreads) ucol[t] = wc[t];
                store_cols(cq + 1, r0 - 1);
                // column r0 except its pivot entry (b, the only value the reflector
                // decides), then the pivot once the reflector is formed
                writeback(r0, b, nr1 - nr0 + 1, T(0), true);
                __syncthreads();
                if (tid == 0) st_agent(Bp + (i64)r0 * ldb + b, (T)*bn);
                if (h > 0) {
                    wg_publish_fence();
                    if (tid == 0) flag_store_agent(prog + s, r0);
                }
                rank2(2);
                held_lo = r0;
                // the next hop's loads overwrite the ring slots of the columns stored above,
                // and its p reads R2 and R3
                __syncthreads();
                T *tw = wc;
                wc = wn;
                wn = tw;
                double *tb = bc;
                bc = bn;
                bn = tb;
            } else {
                // the sweep's last reflector
                rank2(3);
                __syncthreads();
                for (int t = tid; t < L; t += kBCThreads) ucol[t] = wc[t];
                const int keep_lo = min(r0, r1 + 1);
                store_cols(cq + 1, keep_lo - 1);
                held_lo = keep_lo;
                __syncthreads();
            }
        }
        // the rest of the last window
        store_cols(held_lo, held_hi);
        wg_publish_fence();
        if (tid == 0) flag_store_agent(prog + s, kDone);
        s = settle(s + ph.g[phase]);
    }
}

template <typename T>
__global__ void bc_extract_kernel(const T *Bp, T *d, T *e, int n, int b) {
    const i64 ldb = 2 * b;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = Bp[(i64)i * ldb];
    if (i < n - 1) e[i] = Bp[1 + (i64)i * ldb];
    if (i == n - 1) e[i] = T(0);
}

template <typename T> size_t bc_lds_bytes(int b) {
    const size_t nseg = std::max<size_t>(1, kBCThreads / (3 * (size_t)b));
    return sizeof(T) * ((size_t)2 * b * (2 * b + 1) + 2 * b + 3 * b + nseg * 3 * b);
}

} // namespace

template <typename T, int B> int launch_chase(Workspace<T> &ws, T *Bp) {
    const int n = ws.n;
    hipStream_t s = ws.stream;
    const int nsweeps = n - 2;
    CUEV_TRY(fill_int(s, ws.prog, n, -1));
    const size_t lds = bc_lds_bytes<T>(B);
    CUEV_HIP(hipFuncSetAttribute((const void *)bc_chase_kernel<T, B>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc_chase_kernel<T, B>,
                                                          kBCThreads, lds));
    const int resident = std::max(1, per_cu * device_cu_count());
    // a sweep of H hops starts about two hops after the previous one, so up to
    // ~H/2 <= n/(2b) sweeps could be active at once; the grid stays a little below
    // that, leaving CUs to the back-transform GEMMs that run beside the chase
    // (measured at n = 16384, b = 64, with the same factor in the phases below:
    // 84 workgroups, chase 307 ms and 0.966 s per solve; 112, 276 ms and 0.931 s;
    // 136, 278 ms and 0.935 s; 168, 274 ms and 0.931 s, with Q_s^T beside the chase
    // going from 222 to 252 ms)
    const int useful = (int)ceil_div((i64)n * 13, 32 * B) + 8;
    const int grid = std::max(1, std::min({nsweeps, resident, useful}));
    // phases of equal sweep counts; phase p keeps the workgroups its sweeps' lengths
    // call for, by the same rule applied to the n - s0 columns left
    ChasePhases ph{};
    ph.np = std::min(kMaxPhases, std::max(1, nsweeps / 256));
    for (int q = 0; q <= ph.np; ++q) ph.s0[q] = (int)((i64)nsweeps * q / ph.np);
    for (int q = 0; q < ph.np; ++q) {
        const int left = n - ph.s0[q];
        const int want = (int)ceil_div((i64)left * 13, 32 * B) + 8;
        ph.g[q] = std::max(1, std::min(grid, want));
        if (q > 0) ph.g[q] = std::min(ph.g[q], ph.g[q - 1]);
    }
    // algorithmic work: ~n^2/(2b) hops, each reading and writing b columns of the
    // packed band (2 * 2b^2 S bytes) and doing ~ (2 * 3b * b + 4 * 4b^2) flops
    const double hops = (double)n * n / (2.0 * B);
    prof::Scope sc(s, "bc.chase", hops * (6.0 * B * B + 16.0 * B * B), hops * 4.0 * B * B * sizeof(T));
    hipLaunchKernelGGL((bc_chase_kernel<T, B>), dim3(grid), dim3(kBCThreads), lds, s, Bp, ws.U,
                       ws.ldu, n, ws.prog, ws.err, ph);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int bc_chase(Workspace<T> &ws, T *Bp, T *d, T *e) {
    const int n = ws.n;
    hipStream_t s = ws.stream;
    if (n - 2 > 0) {
        switch (ws.b) {
        case 16: CUEV_TRY((launch_chase<T, 16>(ws, Bp))); break;
        case 32: CUEV_TRY((launch_chase<T, 32>(ws, Bp))); break;
        case 64: CUEV_TRY((launch_chase<T, 64>(ws, Bp))); break;
        default: CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_chase: unsupported b=%d", ws.b);
        }
    }
    hipLaunchKernelGGL(bc_extract_kernel<T>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, Bp,
                       d, e, n, ws.b);
    CUEV_LAUNCH_CHECK();
    return CUEV_STATUS_SUCCESS;
}

template int bc_chase<float>(Workspace<float> &, float *, float *, float *);
template int bc_chase<double>(Workspace<double> &, double *, double *, double *);

} // namespace cuev
