// backtransform.hip — eigenvector back-transform evec <- Q_s Q_b Q_d.
//
// Reference: src/cuda/backtransform.cu:76-275 (bc_back_kernel, sbr_back,
// back_transform).
//
// bc_back (M <- Q_b M). The bulge-chasing reflectors compose to Q_b in forward
// order, so Q_b M applies them last-first. A reflector (s, h) covers rows
// [s+1+hb, s+(h+1)b]; two reflectors only need a fixed relative order when
// their row ranges overlap. Grouping hop indices into passes of G = 512/b
// (h in [pG, pG+G)) and, inside a pass, walking sweeps s from high to low while
// applying all G hops of sweep s at once is a valid reverse order (every
// overlapping pair (s,h)/(s+d,h-1), (s,h)/(s+d,h), (s,h)/(s-d,h+m) keeps its
// required order). For sweep s the G hops are rows [s+1+pGb, s+pGb+512): a
// contiguous 512-row window that moves up by exactly one row per sweep.
//
// Kernel: one wave64 per column of M; lane l keeps window rows 8l..8l+7 in
// registers. Per sweep: a reflector spans b/8 lanes (dot product = 8 FMAs plus
// a log2(b/8)-step lane reduction), then the window slides up one row (DPP
// wave shift). Rows entering and leaving the window are batched 64 at a time
// through registers so the slide never waits on memory. The reflector segment
// of a sweep is contiguous and 64-byte aligned in the shifted U layout and is
// loaded one sweep ahead. This generalises the reference's b=32, 32-lane
// sliding window (backtransform.cu:16,32-36) to wave64 and b in {16, 32, 64}.
//
// sbr_back (M <- Q_s M): DBBR panels last-first, M_b -= W_j (Y_j^T M_b), as the
// reference (backtransform.cu:210-230), on the MFMA engine.
#include "backtransform.hpp"

#include <algorithm>

#include "gemm.hpp"
#include "profile.hpp"

namespace cuev {
namespace {

constexpr int kPer = 8;               // window rows per lane
constexpr int kWin = 64 * kPer;       // window rows per wave (512)
constexpr int kBTThreads = 256;

// Value of lane l-1 (lane 0 receives 0): DPP wave_shr:1 on both 32-bit halves.
__device__ __forceinline__ double wave_shr1(double v) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffLL), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)(unsigned)lo) | ((long long)hi << 32));
}
__device__ __forceinline__ float wave_shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

template <typename T, int B>
__global__ __launch_bounds__(kBTThreads) void bc_back_kernel(const T *__restrict__ U, i64 ldu,
                                                             T *M, i64 ldm, int n, int ncols) {
    constexpr int GL = B / kPer; // lanes per reflector
    const int lane = threadIdx.x & 63;
    const int gw = (int)(((i64)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nw = (int)(((i64)gridDim.x * blockDim.x) >> 6);
    const int nsweep = n - 2; // sweeps 0..n-3

    for (int col = gw; col < ncols; col += nw) {
        T *Mc = M + (i64)col * ldm;
        for (int off = 0;; off += kWin) {
            int s = min(nsweep - 1, n - 2 - off);
            if (s < 0) break;
            int top = s + 1 + off; // first window row
            T m[kPer];
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int r = top + lane * kPer + e;
                m[e] = (r < n) ? Mc[r] : T(0);
            }
            // reflector segment of sweep s: U[off + 8 lane .. +7, s]
            const T *up = U + (i64)s * ldu + off + lane * kPer;
            T w[kPer];
#pragma unroll
            for (int e = 0; e < kPer; ++e) w[e] = up[e];

            T inc = T(0), out = T(0);
            int t = 0;           // steps taken in this pass
            int batch_top = top; // window top when the current 64-step batch started
            while (true) {
                // incoming rows for the next 64 steps: rows batch_top-1-k, k = lane
                if ((t & 63) == 0) {
                    batch_top = top;
                    const int r = top - 1 - lane;
                    inc = (r >= 1 + off && r >= 0) ? Mc[r] : T(0);
                }
                // prefetch the next sweep's reflectors
                T wn[kPer];
                if (s > 0) {
                    const T *upn = up - ldu;
#pragma unroll
                    for (int e = 0; e < kPer; ++e) wn[e] = upn[e];
                }
                // apply the G reflectors of sweep s
                T proj = T(0);
#pragma unroll
                for (int e = 0; e < kPer; ++e) proj += w[e] * m[e];
#pragma unroll
                for (int o = 1; o < GL; o <<= 1) proj += __shfl_xor(proj, o, 64);
                proj *= T(2);
#pragma unroll
                for (int e = 0; e < kPer; ++e) m[e] -= proj * w[e];

                const int k = t & 63;
                if (s == 0) break;
                // slide the window up one row: bottom row leaves, row top-1 enters
                const T bottom = m[kPer - 1];
                const T leaving = __shfl(bottom, 63, 64);
                if (lane == k) out = leaving;
                const T entering = __shfl(inc, k, 64);
                const T from_above = wave_shr1(bottom);
#pragma unroll
                for (int e = kPer - 1; e > 0; --e) m[e] = m[e - 1];
                m[0] = (lane == 0) ? entering : from_above;
                --top;
                --s;
                up -= ldu;
#pragma unroll
                for (int e = 0; e < kPer; ++e) w[e] = wn[e];
                ++t;
                if ((t & 63) == 0) {
                    // flush the 64 rows that left: step k left row batch_top + kWin - 1 - k
                    const int r = batch_top + kWin - 1 - lane;
                    if (r < n) Mc[r] = out;
                }
            }
            // flush the rows that left during the unfinished batch, then the window
            {
                const int done = t & 63;
                const int r = batch_top + kWin - 1 - lane;
                if (lane < done && r < n) Mc[r] = out;
            }
#pragma unroll
            for (int e = 0; e < kPer; ++e) {
                const int r = top + lane * kPer + e;
                if (r < n) Mc[r] = m[e];
            }
        }
    }
}

template <typename T, int B> int launch_bc_back(Workspace<T> &ws, const T *U, T *M, i64 ldm) {
    const int n = ws.n;
    int per_cu = 0;
    CUEV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bc_back_kernel<T, B>,
                                                          kBTThreads, 0));
    const int waves_per_block = kBTThreads / 64;
    const i64 max_blocks = (i64)std::max(1, per_cu) * device_cu_count();
    const i64 blocks = std::min<i64>(max_blocks, ceil_div(n, waves_per_block));
    // 4 flops per reflector entry per column: sum_s (n-1-s) entries ~ n^2/2
    prof::Scope sc(ws.stream, "bt.bc_back", 2.0 * (double)n * n * n);
    hipLaunchKernelGGL((bc_back_kernel<T, B>), dim3((unsigned)blocks), dim3(kBTThreads), 0,
                       ws.stream, U, ws.ldu, M, ldm, n, n);
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
