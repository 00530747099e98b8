// solver.cpp — orchestration of the two-stage eigensolver.
//
// Stage order follows the reference (src/cuda/solver.cu:44-57):
// DBBR -> bulge chasing -> tridiagonal D&C -> back-transform, all GPU work on
// the caller's stream and the call blocking until the results are in place.
//
// Two back-transform orders:
//  * direct (the reference's, backtransform.cu:232-275): after the host D&C,
//    evec <- Q_s (Q_b Q_d) with bc_back then sbr_back on Q_d;
//  * reordered (listed as future work in the reference README:35): while the
//    host runs dstedc, the GPU builds X = Q_s Q_b (bc_back and sbr_back applied
//    to the identity); after the D&C only one GEMM evec = X Q_d remains.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "backtransform.hpp"
#include "bulge.hpp"
#include "dbbr.hpp"
#include "dc.hpp"
#include "gemm.hpp"
#include "lapack.hpp"
#include "profile.hpp"
#include "tridi_values.hpp"
#include "util.hpp"

namespace cuev {

void options_defaults(cuev_options &o) {
    o.bandwidth = 0;
    o.block = 0;
    o.precision = CUEV_PRECISION_NATIVE;
    o.refine_sweeps = 0;
    o.bt_mode = CUEV_BT_AUTO;
    o.dc_mode = CUEV_DC_AUTO;
    o.dc_leaf = 0;
    o.refine_tol = 0.0;
}

namespace {
using clk = std::chrono::steady_clock;
double ms_since(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}
float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}
} // namespace

// Smallest merge of the device D&C that waits for the back-transform running beside
// it: the three top tree levels at n = 16384, which hold nearly all of the D&C's GEMM
// work (measured there against other levels, no gate, and no overlap at all), and
// every merge of more than 4096 rows at larger n, so the GEMM-heavy levels never share
// the GPU with bc_back (merges of up to 4096 rows leave it at its own speed at
// n = 16384). At n = 49152 one solve took 25.9 s with this gate and 27.8 s with the
// top three levels only, but bc_back alone varies there from 7.6 to 12 s between runs.
int kDcGateFrom(int n) { return std::min(n / 4, 4096); }

template <typename T>
int tridi_dc(Context<T> &ctx, T *d, T *e, T *eval, T *evec, i64 ldv, double *dc_ms) {
    Workspace<T> &ws = ctx.ws;
    const int n = ws.n;
    hipStream_t s = ws.stream;
    CUEV_HIP(hipMemcpyAsync(ctx.h_d, d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipMemcpyAsync(ctx.h_e, e, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; ++i)
        if (!std::isfinite((double)ctx.h_d[i]) || (i < n - 1 && !std::isfinite((double)ctx.h_e[i])))
            CUEV_FAIL(CUEV_STATUS_NOT_FINITE, "tridiagonal matrix has a non-finite entry at %d", i);
    auto t0 = clk::now();
    if (ctx.dc_mode == CUEV_DC_DEVICE) {
        CUEV_TRY(dc_ensure(ctx, true));
        Scratch gs{ws.gscratch, ws.gscratch_bytes};
        CUEV_TRY(dc_solve<T>(*ctx.dc, s, n, ctx.h_d, ctx.h_e, evec, ldv, eval, ctx.dcQ, ws.M, ctx.dcQn,
                             gs, ctx.dc_leaf));
        if (dc_ms) *dc_ms = ms_since(t0, clk::now());
        return CUEV_STATUS_SUCCESS;
    }
    CUEV_TRY(stedc<T>(n, ctx.h_d, ctx.h_e, ctx.h_z, n, ctx.h_work, ctx.lwork, ctx.h_iwork,
                      ctx.liwork));
    if (dc_ms) *dc_ms = ms_since(t0, clk::now());
    CUEV_HIP(hipMemcpyAsync(eval, ctx.h_d, sizeof(T) * n, hipMemcpyHostToDevice, s));
    CUEV_HIP(hipMemcpy2DAsync(evec, sizeof(T) * ldv, ctx.h_z, sizeof(T) * n, sizeof(T) * n, n,
                              hipMemcpyHostToDevice, s));
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int back_transform(Context<T> &ctx, const T *Y, const T *W, const T *U, T *evec,
                   cuev_solve_timer *timer) {
    Workspace<T> &ws = ctx.ws;
    hipStream_t s = ws.stream;
    if (timer) CUEV_HIP(hipEventRecord(ctx.ev[4], s));
    CUEV_TRY(bc_back<T>(ws, U, evec, ws.ld));
    if (timer) CUEV_HIP(hipEventRecord(ctx.ev[5], s));
    CUEV_TRY(sbr_back<T>(ws, Y, W, evec, ws.ld));
    if (timer) {
        CUEV_HIP(hipEventRecord(ctx.ev[6], s));
        CUEV_HIP(hipEventSynchronize(ctx.ev[6]));
        timer->bt_qb_ms = ev_ms(ctx.ev[4], ctx.ev[5]);
        timer->bt_qs_ms = ev_ms(ctx.ev[5], ctx.ev[6]);
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int solve_native(Context<T> &ctx, T *A, T *eval, T *evec, cuev_solve_timer *timer, int bt_mode,
                 int dc_mode, int dc_leaf) {
    Workspace<T> &ws = ctx.ws;
    const int n = ws.n;
    hipStream_t s = ws.stream;
    hipEvent_t *ev = ctx.ev;
    const auto wall0 = clk::now();
    if (bt_mode == CUEV_BT_AUTO) bt_mode = (n >= 2048) ? CUEV_BT_REORDERED : CUEV_BT_DIRECT;
    const bool reordered = bt_mode == CUEV_BT_REORDERED;
    if (dc_mode == CUEV_DC_AUTO) dc_mode = (n >= 1024) ? CUEV_DC_DEVICE : CUEV_DC_HOST;
    const bool dev_dc = dc_mode == CUEV_DC_DEVICE;

    if (reordered) CUEV_TRY(bt_ensure(ctx));
    hipStream_t bs = ctx.bt_stream;
    CUEV_HIP(hipEventRecord(ev[0], s));
    CUEV_TRY(dbbr_reduce<T>(ws, A, ws.B));
    CUEV_HIP(hipEventRecord(ev[1], s));
    CUEV_TRY(bc_chase<T>(ws, ws.B, ws.d, ws.e));
    CUEV_HIP(hipEventRecord(ev[2], s));
    // (d, e) to the host
    CUEV_HIP(hipMemcpyAsync(ctx.h_d, ws.d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipMemcpyAsync(ctx.h_e, ws.e, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipEventRecord(ev[3], s));
    if (reordered) {
        // X^T = Q_b^T Q_s^T in M on the back-transform stream: Q_s^T needs only the
        // band reduction and runs beside the bulge chasing (which occupies half the
        // CUs, LDS-bound); Q_b^T follows the chasing and runs beside the D&C.
        CUEV_HIP(hipStreamWaitEvent(bs, ev[1], 0));
        CUEV_HIP(hipEventRecord(ctx.xev[0], bs));
        CUEV_TRY(sbr_back_t<T>(ws, bs, ws.Y, ws.W, ws.M, ws.ld));
        CUEV_HIP(hipEventRecord(ctx.xev[1], bs));
        CUEV_HIP(hipStreamWaitEvent(bs, ev[2], 0));
        CUEV_HIP(hipEventRecord(ctx.xev[2], bs));
        CUEV_TRY(bc_back_t<T>(ws, bs, ws.U, ws.M, ws.ld));
        CUEV_HIP(hipEventRecord(ctx.xev[3], bs));
    }
    // wait for (d, e) only; the reordered back-transform keeps running on the GPU.
    // The device error flag is checked once the stream has drained.
    CUEV_HIP(hipEventSynchronize(ev[3]));
    for (int i = 0; i < n; ++i)
        if (!std::isfinite((double)ctx.h_d[i]) || (i < n - 1 && !std::isfinite((double)ctx.h_e[i])))
            CUEV_FAIL(CUEV_STATUS_NOT_FINITE,
                      "tridiagonal matrix has a non-finite entry at %d (input not finite?)", i);
    Scratch gs{ws.gscratch, ws.gscratch_bytes};
    if (dev_dc) CUEV_TRY(dc_ensure(ctx, reordered));
    const auto dc0 = clk::now();
    if (dev_dc) {
        // A (consumed by the reduction) and evec (written last) serve as the D&C's
        // work buffers. Reordered: the D&C runs on its own stream beside the
        // back-transform still queued on s, without split-K (its scratch is s's).
        if (reordered) {
            // small GEMM tiles: the merges then fit beside the back-transform's
            // waves; the top tree levels (nearly all of the D&C's GEMM work) wait
            // for the back-transform, so the two GPU-wide phases do not share the
            // GPU and only the small, latency-bound merges overlap it
            Scratch dcs;
            dcs.small_tiles = 1;
            CUEV_TRY(dc_solve<T>(*ctx.dc, ctx.dc_stream, n, ctx.h_d, ctx.h_e, ctx.dcQ, n, eval, A, evec,
                                 ctx.dcQn, dcs, dc_leaf, ctx.xev[3], kDcGateFrom(n)));
        }
        else
            CUEV_TRY(dc_solve<T>(*ctx.dc, s, n, ctx.h_d, ctx.h_e, evec, n, eval, A, ws.M, ctx.dcQn, gs,
                                 dc_leaf));
    } else {
        CUEV_TRY(stedc<T>(n, ctx.h_d, ctx.h_e, ctx.h_z, n, ctx.h_work, ctx.lwork, ctx.h_iwork,
                          ctx.liwork));
    }
    const auto dc1 = clk::now();
    CUEV_HIP(hipEventRecord(ev[6], s));
    if (!dev_dc) CUEV_HIP(hipMemcpyAsync(eval, ctx.h_d, sizeof(T) * n, hipMemcpyHostToDevice, s));
    if (reordered) {
        // Q_d (device D&C: already resident; host D&C: into W, free once the
        // back-transform stream is done), then evec = X Q_d = (X^T)^T Q_d
        CUEV_HIP(hipStreamWaitEvent(s, ctx.xev[3], 0));
        const T *Qd = dev_dc ? ctx.dcQ : ws.W;
        if (!dev_dc)
            CUEV_HIP(hipMemcpyAsync(ws.W, ctx.h_z, sizeof(T) * (size_t)n * n, hipMemcpyHostToDevice, s));
        CUEV_HIP(hipEventRecord(ev[7], s));
        {
            prof::Scope sc(s, "bt.gemm", 2.0 * (double)n * n * n);
            CUEV_TRY(gemm<T>(s, 1, 0, n, n, n, T(1), ws.M, ws.ld, Qd, n, T(0), evec, n, gs));
        }
        CUEV_HIP(hipEventRecord(ev[8], s));
    } else {
        if (!dev_dc)
            CUEV_HIP(hipMemcpyAsync(evec, ctx.h_z, sizeof(T) * (size_t)n * n, hipMemcpyHostToDevice, s));
        CUEV_HIP(hipEventRecord(ev[7], s));
        CUEV_TRY(bc_back<T>(ws, ws.U, evec, n));
        CUEV_HIP(hipEventRecord(ev[4], s));
        CUEV_TRY(sbr_back<T>(ws, ws.Y, ws.W, evec, n));
        CUEV_HIP(hipEventRecord(ev[5], s));
    }
    CUEV_HIP(hipStreamSynchronize(s));
    CUEV_TRY(workspace_check_error(ws));
    const auto wall1 = clk::now();
    if (timer) {
        std::memset(timer, 0, sizeof(*timer));
        timer->dbbr_ms = ev_ms(ev[0], ev[1]);
        timer->bc_ms = ev_ms(ev[1], ev[2]);
        timer->dc_ms = (float)ms_since(dc0, dc1);
        timer->total_ms = (float)ms_since(wall0, wall1);
        if (reordered) {
            timer->bt_qs_ms = ev_ms(ctx.xev[0], ctx.xev[1]);
            timer->bt_qb_ms = ev_ms(ctx.xev[2], ctx.xev[3]);
            timer->overlap_ms = timer->bt_qb_ms + timer->bt_qs_ms;
            timer->bt_copyin_ms = ev_ms(ev[6], ev[7]);
            timer->bt_copyout_ms = ev_ms(ev[7], ev[8]);
            timer->bt_ms = (float)ms_since(dc1, wall1);
        } else {
            timer->bt_copyin_ms = ev_ms(ev[6], ev[7]);
            timer->bt_qb_ms = ev_ms(ev[7], ev[4]);
            timer->bt_qs_ms = ev_ms(ev[4], ev[5]);
            timer->bt_ms = (float)ms_since(dc1, wall1);
        }
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T>
int solve_values(Context<T> &ctx, T *A, int il, int iu, T *eval, cuev_solve_timer *timer) {
    Workspace<T> &ws = ctx.ws;
    const int n = ws.n;
    hipStream_t s = ws.stream;
    hipEvent_t *ev = ctx.ev;
    const auto wall0 = clk::now();
    CUEV_HIP(hipEventRecord(ev[0], s));
    CUEV_TRY(dbbr_reduce<T>(ws, A, ws.B));
    CUEV_HIP(hipEventRecord(ev[1], s));
    CUEV_TRY(bc_chase<T>(ws, ws.B, ws.d, ws.e));
    CUEV_HIP(hipEventRecord(ev[2], s));
    {
        prof::Scope sc(s, "tv.multisect");
        CUEV_TRY(tridi_values<T>(s, ws.d, ws.e, n, il, iu, eval));
    }
    CUEV_HIP(hipEventRecord(ev[3], s));
    // (d, e) to the host for the finiteness check (the count of a NaN pivot is
    // meaningless, so a non-finite input must be reported, not returned)
    CUEV_HIP(hipMemcpyAsync(ctx.h_d, ws.d, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipMemcpyAsync(ctx.h_e, ws.e, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    CUEV_HIP(hipStreamSynchronize(s));
    CUEV_TRY(workspace_check_error(ws));
    for (int i = 0; i < n; ++i)
        if (!std::isfinite((double)ctx.h_d[i]) || (i < n - 1 && !std::isfinite((double)ctx.h_e[i])))
            CUEV_FAIL(CUEV_STATUS_NOT_FINITE,
                      "tridiagonal matrix has a non-finite entry at %d (input not finite?)", i);
    const auto wall1 = clk::now();
    if (timer) {
        std::memset(timer, 0, sizeof(*timer));
        timer->dbbr_ms = ev_ms(ev[0], ev[1]);
        timer->bc_ms = ev_ms(ev[1], ev[2]);
        timer->dc_ms = ev_ms(ev[2], ev[3]);
        timer->total_ms = (float)ms_since(wall0, wall1);
    }
    return CUEV_STATUS_SUCCESS;
}

template int solve_values<float>(Context<float> &, float *, int, int, float *, cuev_solve_timer *);
template int solve_values<double>(Context<double> &, double *, int, int, double *, cuev_solve_timer *);
template int tridi_dc<float>(Context<float> &, float *, float *, float *, float *, i64, double *);
template int tridi_dc<double>(Context<double> &, double *, double *, double *, double *, i64,
                              double *);
template int back_transform<float>(Context<float> &, const float *, const float *, const float *,
                                   float *, cuev_solve_timer *);
template int back_transform<double>(Context<double> &, const double *, const double *,
                                    const double *, double *, cuev_solve_timer *);
template int solve_native<float>(Context<float> &, float *, float *, float *, cuev_solve_timer *,
                                 int, int, int);
template int solve_native<double>(Context<double> &, double *, double *, double *,
                                  cuev_solve_timer *, int, int, int);

} // namespace cuev
