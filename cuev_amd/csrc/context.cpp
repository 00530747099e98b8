// context.cpp — context allocation and the per-device context cache.
#include "context.hpp"

#include "gemm.hpp"

#include <cstdlib>
#include <mutex>

#include "lapack.hpp"

namespace cuev {

template <typename T> int context_create(Context<T> *&out, int n, int b, int k, hipStream_t s) {
    out = nullptr;
    auto *ctx = new Context<T>();
    ctx->dtype = sizeof(T) == 8 ? CUEV_F64 : CUEV_F32;
    int st = workspace_alloc<T>(ctx->ws, n, b, k, s);
    if (st) {
        delete ctx;
        return st;
    }
    auto pin = [&](void **p, size_t bytes) -> int {
        hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            *p = nullptr;
            CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "hipHostMalloc(%zu) failed: %s", bytes,
                      hipGetErrorString(e));
        }
        return CUEV_STATUS_SUCCESS;
    };
    st = pin(reinterpret_cast<void **>(&ctx->h_d), sizeof(T) * (size_t)n);
    if (!st) st = pin(reinterpret_cast<void **>(&ctx->h_e), sizeof(T) * (size_t)n);
    if (!st) st = pin(reinterpret_cast<void **>(&ctx->h_z), sizeof(T) * (size_t)n * n);
    if (!st) {
        ctx->lwork = stedc_lwork(n);
        ctx->liwork = stedc_liwork(n);
        ctx->h_work = static_cast<T *>(std::malloc(sizeof(T) * (size_t)ctx->lwork));
        ctx->h_iwork = static_cast<int64_t *>(std::malloc(sizeof(int64_t) * (size_t)ctx->liwork));
        if (!ctx->h_work || !ctx->h_iwork) {
            detail::set_error("host malloc of the dstedc workspace failed (n=%d)", n);
            st = CUEV_STATUS_ALLOC_FAILED;
        }
    }
    for (int i = 0; i < 12 && !st; ++i) {
        if (hipEventCreate(&ctx->ev[i]) != hipSuccess) {
            detail::set_error("hipEventCreate failed");
            st = CUEV_STATUS_HIP_ERROR;
        }
    }
    if (st) {
        context_destroy(ctx);
        return st;
    }
    out = ctx;
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int dc_ensure(Context<T> &ctx, bool need_q) {
    const int n = ctx.ws.n;
    if (!ctx.dc) CUEV_TRY(dc_state_create<T>(ctx.dc, n));
    if (!ctx.dc_stream) CUEV_HIP(hipStreamCreateWithFlags(&ctx.dc_stream, hipStreamNonBlocking));
    if (!ctx.dc_pool || (need_q && !ctx.dcQ)) {
        if (ctx.dc_pool) {
            (void)hipFree(ctx.dc_pool);
            ctx.dc_pool = nullptr;
            ctx.dcQ = ctx.dcQn = nullptr;
        }
        const size_t sq = sizeof(T) * (size_t)n * n;
        const size_t bytes = (need_q ? 2 : 1) * ((sq + 255) & ~size_t(255));
        if (hipMalloc(&ctx.dc_pool, bytes) != hipSuccess) {
            ctx.dc_pool = nullptr;
            (void)hipGetLastError();
            CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "dc: hipMalloc(%zu) for the D&C buffers failed", bytes);
        }
        char *p = static_cast<char *>(ctx.dc_pool);
        ctx.dcQn = reinterpret_cast<T *>(p);
        if (need_q) ctx.dcQ = reinterpret_cast<T *>(p + ((sq + 255) & ~size_t(255)));
    }
    return CUEV_STATUS_SUCCESS;
}

template <typename T> int bt_ensure(Context<T> &ctx) {
    if (!ctx.bt_stream) CUEV_HIP(hipStreamCreateWithFlags(&ctx.bt_stream, hipStreamNonBlocking));
    for (auto &e : ctx.xev)
        if (!e) CUEV_HIP(hipEventCreate(&e));
    return CUEV_STATUS_SUCCESS;
}

template <typename T> void context_destroy(Context<T> *ctx) {
    if (!ctx) return;
    if (ctx->bt_stream) (void)hipStreamSynchronize(ctx->bt_stream);
    for (auto &e : ctx->xev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->bt_stream) (void)hipStreamDestroy(ctx->bt_stream);
    if (ctx->dc_stream) (void)hipStreamSynchronize(ctx->dc_stream);
    dc_state_destroy(ctx->dc);
    if (ctx->dc_pool) (void)hipFree(ctx->dc_pool);
    if (ctx->dc_stream) (void)hipStreamDestroy(ctx->dc_stream);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->h_d) (void)hipHostFree(ctx->h_d);
    if (ctx->h_e) (void)hipHostFree(ctx->h_e);
    if (ctx->h_z) (void)hipHostFree(ctx->h_z);
    std::free(ctx->h_work);
    std::free(ctx->h_iwork);
    workspace_free(ctx->ws);
    delete ctx;
}

namespace {
template <typename T> struct Cache {
    std::mutex mu;
    Context<T> *idle[64] = {};
};
template <typename T> Cache<T> &cache() {
    static Cache<T> c;
    return c;
}
} // namespace

template <typename T> int context_acquire(Context<T> *&ctx, int n, int b, int k, hipStream_t s) {
    CUEV_TRY(resolve_blocking(n, b, k));
    int dev = 0;
    CUEV_HIP(hipGetDevice(&dev));
    Context<T> *victim = nullptr;
    {
        auto &c = cache<T>();
        std::lock_guard<std::mutex> lk(c.mu);
        Context<T> *&slot = c.idle[dev & 63];
        if (slot && slot->ws.n == n && slot->ws.b == b && slot->ws.k == k && slot->ws.device == dev) {
            ctx = slot;
            slot = nullptr;
            ctx->ws.stream = s;
            return CUEV_STATUS_SUCCESS;
        }
        victim = slot; // a different size: free it before allocating anew
        slot = nullptr;
    }
    if (victim) {
        (void)hipDeviceSynchronize();
        context_destroy(victim);
    }
    return context_create<T>(ctx, n, b, k, s);
}

template <typename T> void context_release(Context<T> *ctx) {
    if (!ctx) return;
    auto &c = cache<T>();
    Context<T> *victim = nullptr;
    {
        std::lock_guard<std::mutex> lk(c.mu);
        Context<T> *&slot = c.idle[ctx->ws.device & 63];
        victim = slot;
        slot = ctx;
    }
    if (victim) context_destroy(victim);
}

template <typename T> void context_release_all() {
    auto &c = cache<T>();
    Context<T> *victims[64] = {};
    {
        std::lock_guard<std::mutex> lk(c.mu);
        for (int i = 0; i < 64; ++i) {
            victims[i] = c.idle[i];
            c.idle[i] = nullptr;
        }
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (Context<T> *v : victims)
        if (v) {
            (void)hipSetDevice(v->ws.device);
            (void)hipDeviceSynchronize();
            context_destroy(v);
        }
    (void)hipSetDevice(cur);
}

template void context_release_all<float>();
template void context_release_all<double>();
template int context_create<float>(Context<float> *&, int, int, int, hipStream_t);
template int context_create<double>(Context<double> *&, int, int, int, hipStream_t);
template void context_destroy<float>(Context<float> *);
template void context_destroy<double>(Context<double> *);
template int context_acquire<float>(Context<float> *&, int, int, int, hipStream_t);
template int context_acquire<double>(Context<double> *&, int, int, int, hipStream_t);
template void context_release<float>(Context<float> *);
template void context_release<double>(Context<double> *);
template int dc_ensure<float>(Context<float> &, bool);
template int bt_ensure<float>(Context<float> &);
template int bt_ensure<double>(Context<double> &);
template int dc_ensure<double>(Context<double> &, bool);

} // namespace cuev
