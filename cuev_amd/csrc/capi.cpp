// capi.cpp — the extern "C" boundary declared in include/cuev_c.h.
#include <cstring>

#include "backtransform.hpp"
#include "backward.hpp"
#include "batched.hpp"
#include "bulge.hpp"
#include "context.hpp"
#include "dbbr.hpp"
#include "gemm.hpp"
#include "lapack.hpp"
#include "refine.hpp"
#include "solver.hpp"
#include "tridi_values.hpp"
#include "util.hpp"

using namespace cuev;

struct cuev_handle_s {
    int dtype = CUEV_F64;
    Context<float> *f = nullptr;
    Context<double> *d = nullptr;
};

namespace {

template <typename T> Context<T> *ctx_of(cuev_handle_t h);
template <> Context<float> *ctx_of<float>(cuev_handle_t h) {
    return (h && h->dtype == CUEV_F32) ? h->f : nullptr;
}
template <> Context<double> *ctx_of<double>(cuev_handle_t h) {
    return (h && h->dtype == CUEV_F64) ? h->d : nullptr;
}

#define CUEV_CTX(T, h)                                                                             \
    Context<T> *ctx = ctx_of<T>(h);                                                                \
    if (!ctx)                                                                                      \
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "%s: null handle or handle of the other precision", \
                  __func__);

template <typename T>
int solve_entry(T *A, int n, T *eval, T *evec, hipStream_t s, cuev_solve_timer *timer,
                const cuev_options *opt) {
    detail::clear_error();
    cuev_options o;
    options_defaults(o);
    if (opt) o = *opt;
    if (n < 1) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: n must be >= 1 (got %d)", n);
    if (!A || !eval || !evec)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: null device pointer");
    if (o.bt_mode < CUEV_BT_AUTO || o.bt_mode > CUEV_BT_REORDERED)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: bad bt_mode %d", o.bt_mode);
    if (o.dc_mode < CUEV_DC_AUTO || o.dc_mode > CUEV_DC_DEVICE)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: bad dc_mode %d", o.dc_mode);
    if (o.dc_leaf < 0 || o.dc_leaf == 1 || o.dc_leaf > 4096)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: dc_leaf must be 0 or in [2, 4096] (got %d)",
                  o.dc_leaf);
    int b = o.bandwidth, k = o.block;
    CUEV_TRY(resolve_blocking(n, b, k));
    if (k > n) k = (int)round_up(n, b); // one block covers the matrix
    if (o.precision == CUEV_PRECISION_MIXED) {
        if constexpr (sizeof(T) == 8) {
            return solve_mixed(A, n, eval, evec, s, timer, o, b, k);
        } else {
            CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "the mixed form is defined for the double API only");
        }
    }
    if (o.precision != CUEV_PRECISION_NATIVE)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_solve: bad precision %d", o.precision);
    Context<T> *ctx = nullptr;
    CUEV_TRY(context_acquire<T>(ctx, n, b, k, s));
    const int st = solve_native<T>(*ctx, A, eval, evec, timer, o.bt_mode, o.dc_mode, o.dc_leaf);
    if (st == CUEV_STATUS_SUCCESS) {
        context_release<T>(ctx);
    } else {
        (void)hipStreamSynchronize(s);
        context_destroy<T>(ctx);
    }
    return st;
}

template <typename T>
int values_entry(T *A, int n, int il, int iu, T *eval, hipStream_t s, cuev_solve_timer *timer,
                 const cuev_options *opt) {
    detail::clear_error();
    cuev_options o;
    options_defaults(o);
    if (opt) o = *opt;
    if (n < 1) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_values: n must be >= 1 (got %d)", n);
    if (!A || !eval) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_values: null device pointer");
    if (il < 0 || iu >= n || il > iu)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "symm_eig_values: index range [%d, %d] outside [0, %d)", il, iu,
                  n);
    if (o.precision != CUEV_PRECISION_NATIVE)
        CUEV_FAIL(CUEV_STATUS_NOT_SUPPORTED, "symm_eig_values: the eigenvalue-only path runs in the API's precision");
    int b = o.bandwidth, k = o.block;
    CUEV_TRY(resolve_blocking(n, b, k));
    if (k > n) k = (int)round_up(n, b);
    Context<T> *ctx = nullptr;
    CUEV_TRY(context_acquire<T>(ctx, n, b, k, s));
    const int st = solve_values<T>(*ctx, A, il, iu, eval, timer);
    if (st == CUEV_STATUS_SUCCESS) {
        context_release<T>(ctx);
    } else {
        (void)hipStreamSynchronize(s);
        context_destroy<T>(ctx);
    }
    return st;
}

template <typename T>
int tridi_values_entry(const T *d, const T *e, int n, int il, int iu, T *w, hipStream_t s) {
    detail::clear_error();
    return tridi_values<T>(s, d, e, n, il, iu, w);
}

template <typename T>
int gemm_entry(hipStream_t s, int ta, int tb, i64 m, i64 n, i64 k, T alpha, const T *A, i64 lda,
               const T *B, i64 ldb, T beta, T *C, i64 ldc) {
    detail::clear_error();
    if (m < 0 || n < 0 || k < 0) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: negative size");
    if ((ta != 0 && ta != 1) || (tb != 0 && tb != 1))
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "gemm: trans must be 0 or 1");
    const size_t bytes = gemm_scratch_bytes(m, n, sizeof(T));
    void *scr = nullptr;
    if (hipMalloc(&scr, bytes) != hipSuccess)
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "gemm: scratch allocation failed");
    int st = gemm<T>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, Scratch{scr, bytes});
    (void)hipStreamSynchronize(s);
    (void)hipFree(scr);
    return st;
}

template <typename T>
int symm_entry(hipStream_t s, i64 m, i64 n, T alpha, const T *A, i64 lda, const T *B, i64 ldb,
               T beta, T *C, i64 ldc) {
    detail::clear_error();
    const size_t bytes = gemm_scratch_bytes(m, n, sizeof(T));
    void *scr = nullptr;
    if (hipMalloc(&scr, bytes) != hipSuccess)
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "symm: scratch allocation failed");
    int st = symm_lower<T>(s, m, n, alpha, A, lda, B, ldb, beta, C, ldc, Scratch{scr, bytes});
    (void)hipStreamSynchronize(s);
    (void)hipFree(scr);
    return st;
}

template <typename T>
int syr2k_entry(hipStream_t s, i64 n, i64 k, T alpha, const T *A, i64 lda, const T *B, i64 ldb,
                T beta, T *C, i64 ldc) {
    detail::clear_error();
    const size_t bytes = gemm_scratch_bytes(n, n, sizeof(T));
    void *scr = nullptr;
    if (hipMalloc(&scr, bytes) != hipSuccess)
        CUEV_FAIL(CUEV_STATUS_ALLOC_FAILED, "syr2k: scratch allocation failed");
    int st = syr2k_lower<T>(s, n, k, alpha, A, lda, B, ldb, beta, C, ldc, Scratch{scr, bytes});
    (void)hipStreamSynchronize(s);
    (void)hipFree(scr);
    return st;
}

// Stage entry points run asynchronously on the handle's stream (as the
// reference's kernels:: launchers do); the device error flag is checked after
// the persistent stages by synchronising.
template <typename T> int stage_pack(cuev_handle_t h, const T *A, T *B, int n, int b) {
    detail::clear_error();
    CUEV_CTX(T, h);
    return dbbr_pack<T>(ctx->ws, A, B, n, b);
}
template <typename T> int stage_panel_qr(cuev_handle_t h, T *A, T *Y, int rows, int b) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!A || rows < 1 || b < 1 || b > 64)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dbbr_panel_qr: rows=%d b=%d", rows, b);
    if (rows > ctx->ws.n) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dbbr_panel_qr: rows > handle n");
    CUEV_TRY(dbbr_panel_qr<T>(ctx->ws, A, Y, nullptr, rows, b));
    return workspace_check_error(ctx->ws);
}
template <typename T> int stage_reduce(cuev_handle_t h, T *A, T *B) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!A || !B) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "dbbr_reduce: null pointer");
    CUEV_TRY(dbbr_reduce<T>(ctx->ws, A, B));
    return workspace_check_error(ctx->ws);
}
template <typename T> int stage_chase(cuev_handle_t h, T *B, T *d, T *e) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!B || !d || !e) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_chase: null pointer");
    CUEV_TRY(bc_chase<T>(ctx->ws, B, d, e));
    return workspace_check_error(ctx->ws);
}
template <typename T> int stage_dc(cuev_handle_t h, T *d, T *e, T *eval, T *evec) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!d || !e || !eval || !evec) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "tridi_dc: null pointer");
    CUEV_TRY(tridi_dc<T>(*ctx, d, e, eval, evec, ctx->ws.n, nullptr));
    CUEV_HIP(hipStreamSynchronize(ctx->ws.stream));
    return CUEV_STATUS_SUCCESS;
}
template <typename T>
int stage_bt(cuev_handle_t h, const T *Y, const T *W, const T *U, T *evec, cuev_solve_timer *t) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!Y || !W || !U || !evec) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "back_transform: null pointer");
    return back_transform<T>(*ctx, Y, W, U, evec, t);
}
template <typename T> int stage_bc_back(cuev_handle_t h, const T *U, T *M, i64 ldm) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!U || !M) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back: null pointer");
    return bc_back<T>(ctx->ws, U, M, ldm);
}
template <typename T> int stage_bc_back_t(cuev_handle_t h, const T *U, T *M, i64 ldm) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!U || !M) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "bc_back_t: null pointer");
    CUEV_TRY(bc_back_t<T>(ctx->ws, ctx->ws.stream, U, M, ldm));
    CUEV_HIP(hipStreamSynchronize(ctx->ws.stream));
    return CUEV_STATUS_SUCCESS;
}
template <typename T> int stage_sbr_back(cuev_handle_t h, const T *Y, const T *W, T *M, i64 ldm) {
    detail::clear_error();
    CUEV_CTX(T, h);
    if (!Y || !W || !M) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "sbr_back: null pointer");
    return sbr_back<T>(ctx->ws, Y, W, M, ldm);
}

} // namespace

extern "C" {

int cuev_ssymm_eig_values(float *A, int n, float *eval, hipStream_t stream, cuev_solve_timer *timer) {
    return values_entry<float>(A, n, 0, n - 1, eval, stream, timer, nullptr);
}
int cuev_dsymm_eig_values(double *A, int n, double *eval, hipStream_t stream, cuev_solve_timer *timer) {
    return values_entry<double>(A, n, 0, n - 1, eval, stream, timer, nullptr);
}
int cuev_ssymm_eig_values_ex(float *A, int n, int il, int iu, float *eval, hipStream_t stream,
                             cuev_solve_timer *timer, const cuev_options *opt) {
    return values_entry<float>(A, n, il, iu, eval, stream, timer, opt);
}
int cuev_dsymm_eig_values_ex(double *A, int n, int il, int iu, double *eval, hipStream_t stream,
                             cuev_solve_timer *timer, const cuev_options *opt) {
    return values_entry<double>(A, n, il, iu, eval, stream, timer, opt);
}
int cuev_stridi_values(const float *d, const float *e, int n, int il, int iu, float *w, hipStream_t stream) {
    return tridi_values_entry<float>(d, e, n, il, iu, w, stream);
}
int cuev_dtridi_values(const double *d, const double *e, int n, int il, int iu, double *w,
                       hipStream_t stream) {
    return tridi_values_entry<double>(d, e, n, il, iu, w, stream);
}

int cuev_ssymm_eig_batched(float *A, int n, int batch, float *eval, float *evec, hipStream_t stream,
                           int *n_not_converged) {
    detail::clear_error();
    return eig_batched<float>(stream, A, n, batch, eval, evec, n_not_converged);
}
int cuev_dsymm_eig_batched(double *A, int n, int batch, double *eval, double *evec, hipStream_t stream,
                           int *n_not_converged) {
    detail::clear_error();
    return eig_batched<double>(stream, A, n, batch, eval, evec, n_not_converged);
}

int cuev_cache_release(void) {
    detail::clear_error();
    context_release_all<float>();
    context_release_all<double>();
    mixed_release_all();
    backward_release_all();
    return CUEV_STATUS_SUCCESS;
}

const char *cuev_status_string(int status) {
    switch (status) {
    case CUEV_STATUS_SUCCESS: return "success";
    case CUEV_STATUS_INVALID_ARGUMENT: return "invalid argument";
    case CUEV_STATUS_HIP_ERROR: return "HIP error";
    case CUEV_STATUS_ALLOC_FAILED: return "allocation failed";
    case CUEV_STATUS_LAPACK_ERROR: return "host LAPACK error";
    case CUEV_STATUS_LAPACK_UNAVAILABLE: return "host LAPACK unavailable";
    case CUEV_STATUS_WATCHDOG: return "persistent-kernel watchdog expired";
    case CUEV_STATUS_NOT_FINITE: return "non-finite values";
    case CUEV_STATUS_NOT_SUPPORTED: return "not supported";
    case CUEV_STATUS_NOT_CONVERGED: return "refinement not converged";
    default: return "unknown status";
    }
}

const char *cuev_last_error_message(void) { return detail::last_error(); }

const char *cuev_lapack_name(void) {
    const LapackFns *L = lapack();
    return L ? L->name.c_str() : "";
}

void cuev_options_init(cuev_options *opt) {
    if (opt) options_defaults(*opt);
}

int cuev_ssymm_eig_solve(float *A, int n, float *eval, float *evec, hipStream_t stream,
                         cuev_solve_timer *timer) {
    return solve_entry<float>(A, n, eval, evec, stream, timer, nullptr);
}
int cuev_dsymm_eig_solve(double *A, int n, double *eval, double *evec, hipStream_t stream,
                         cuev_solve_timer *timer) {
    return solve_entry<double>(A, n, eval, evec, stream, timer, nullptr);
}
int cuev_ssymm_eig_solve_ex(float *A, int n, float *eval, float *evec, hipStream_t stream,
                            cuev_solve_timer *timer, const cuev_options *opt) {
    return solve_entry<float>(A, n, eval, evec, stream, timer, opt);
}
int cuev_dsymm_eig_solve_ex(double *A, int n, double *eval, double *evec, hipStream_t stream,
                            cuev_solve_timer *timer, const cuev_options *opt) {
    return solve_entry<double>(A, n, eval, evec, stream, timer, opt);
}

int cuev_ssymm_eig_backward(const float *V, const float *lambda, const float *Vbar,
                            const float *lambar, int n, float *Abar, double gap_tol,
                            int *n_masked, hipStream_t stream) {
    detail::clear_error();
    return eig_backward<float>(V, lambda, Vbar, lambar, n, Abar, gap_tol, n_masked, stream);
}
int cuev_dsymm_eig_backward(const double *V, const double *lambda, const double *Vbar,
                            const double *lambar, int n, double *Abar, double gap_tol,
                            int *n_masked, hipStream_t stream) {
    detail::clear_error();
    return eig_backward<double>(V, lambda, Vbar, lambar, n, Abar, gap_tol, n_masked, stream);
}

int cuev_ssymm_eig_backward_noalloc(const float *V, const float *lambda, const float *Vbar,
                                    const float *lambar, int n, float *Abar, double gap_tol,
                                    int *n_masked, hipStream_t stream) {
    detail::clear_error();
    return eig_backward_noalloc<float>(V, lambda, Vbar, lambar, n, Abar, gap_tol, n_masked, stream);
}
int cuev_dsymm_eig_backward_noalloc(const double *V, const double *lambda, const double *Vbar,
                                    const double *lambar, int n, double *Abar, double gap_tol,
                                    int *n_masked, hipStream_t stream) {
    detail::clear_error();
    return eig_backward_noalloc<double>(V, lambda, Vbar, lambar, n, Abar, gap_tol, n_masked, stream);
}

int cuev_handle_create(cuev_handle_t *handle, int dtype, int n, int nbw, int nk, hipStream_t stream) {
    detail::clear_error();
    if (!handle) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_create: null out pointer");
    *handle = nullptr;
    if (dtype != CUEV_F32 && dtype != CUEV_F64)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_create: bad dtype %d", dtype);
    auto *h = new cuev_handle_s();
    h->dtype = dtype;
    int st = (dtype == CUEV_F32) ? context_create<float>(h->f, n, nbw, nk, stream)
                                 : context_create<double>(h->d, n, nbw, nk, stream);
    if (st) {
        delete h;
        return st;
    }
    *handle = h;
    return CUEV_STATUS_SUCCESS;
}

int cuev_handle_destroy(cuev_handle_t h) {
    if (!h) return CUEV_STATUS_SUCCESS;
    hipStream_t s = h->f ? h->f->ws.stream : (h->d ? h->d->ws.stream : nullptr);
    (void)hipStreamSynchronize(s);
    context_destroy<float>(h->f);
    context_destroy<double>(h->d);
    delete h;
    return CUEV_STATUS_SUCCESS;
}

int cuev_handle_set_dc(cuev_handle_t h, int dc_mode, int dc_leaf) {
    detail::clear_error();
    if (!h) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_set_dc: null handle");
    if (dc_mode != CUEV_DC_HOST && dc_mode != CUEV_DC_DEVICE)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_set_dc: dc_mode must be host or device (got %d)",
                  dc_mode);
    if (dc_leaf < 0 || dc_leaf == 1 || dc_leaf > 4096)
        CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_set_dc: dc_leaf must be 0 or in [2, 4096]");
    if (h->f) {
        h->f->dc_mode = dc_mode;
        h->f->dc_leaf = dc_leaf;
    }
    if (h->d) {
        h->d->dc_mode = dc_mode;
        h->d->dc_leaf = dc_leaf;
    }
    return CUEV_STATUS_SUCCESS;
}

int cuev_handle_buffer(cuev_handle_t h, int which, void **ptr, int64_t *ld) {
    detail::clear_error();
    if (!h || !ptr) CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_buffer: null argument");
    auto pick = [&](auto *ctx) -> int {
        auto &ws = ctx->ws;
        void *p = nullptr;
        int64_t l = 0;
        switch (which) {
        case CUEV_BUF_Y: p = ws.Y; l = ws.ld; break;
        case CUEV_BUF_W: p = ws.W; l = ws.ld; break;
        case CUEV_BUF_B: p = ws.B; l = 2 * ws.b; break;
        case CUEV_BUF_U: p = ws.U; l = ws.ldu; break;
        case CUEV_BUF_D: p = ws.d; l = ws.n; break;
        case CUEV_BUF_E: p = ws.e; l = ws.n; break;
        case CUEV_BUF_TRI: p = ws.Tri; l = ws.b; break;
        case CUEV_BUF_TAU: p = ws.tau; l = ws.b; break;
        case CUEV_BUF_M: p = ws.M; l = ws.ld; break;
        default: CUEV_FAIL(CUEV_STATUS_INVALID_ARGUMENT, "handle_buffer: bad buffer id %d", which);
        }
        *ptr = p;
        if (ld) *ld = l;
        return CUEV_STATUS_SUCCESS;
    };
    return h->dtype == CUEV_F32 ? pick(h->f) : pick(h->d);
}

int cuev_s_dbbr_pack(cuev_handle_t h, const float *A, float *B, int n, int b) { return stage_pack<float>(h, A, B, n, b); }
int cuev_d_dbbr_pack(cuev_handle_t h, const double *A, double *B, int n, int b) { return stage_pack<double>(h, A, B, n, b); }
int cuev_s_dbbr_panel_qr(cuev_handle_t h, float *A, float *Y, int rows, int b) { return stage_panel_qr<float>(h, A, Y, rows, b); }
int cuev_d_dbbr_panel_qr(cuev_handle_t h, double *A, double *Y, int rows, int b) { return stage_panel_qr<double>(h, A, Y, rows, b); }
int cuev_s_dbbr_reduce(cuev_handle_t h, float *A, float *B) { return stage_reduce<float>(h, A, B); }
int cuev_d_dbbr_reduce(cuev_handle_t h, double *A, double *B) { return stage_reduce<double>(h, A, B); }
int cuev_s_bc_chase(cuev_handle_t h, float *B, float *d, float *e) { return stage_chase<float>(h, B, d, e); }
int cuev_d_bc_chase(cuev_handle_t h, double *B, double *d, double *e) { return stage_chase<double>(h, B, d, e); }
int cuev_s_tridi_dc(cuev_handle_t h, float *d, float *e, float *eval, float *evec) { return stage_dc<float>(h, d, e, eval, evec); }
int cuev_d_tridi_dc(cuev_handle_t h, double *d, double *e, double *eval, double *evec) { return stage_dc<double>(h, d, e, eval, evec); }
int cuev_s_back_transform(cuev_handle_t h, const float *Y, const float *W, const float *U, float *evec, cuev_solve_timer *t) { return stage_bt<float>(h, Y, W, U, evec, t); }
int cuev_d_back_transform(cuev_handle_t h, const double *Y, const double *W, const double *U, double *evec, cuev_solve_timer *t) { return stage_bt<double>(h, Y, W, U, evec, t); }
int cuev_s_bc_back(cuev_handle_t h, const float *U, float *M, int64_t ldm) { return stage_bc_back<float>(h, U, M, ldm); }
int cuev_d_bc_back(cuev_handle_t h, const double *U, double *M, int64_t ldm) { return stage_bc_back<double>(h, U, M, ldm); }
int cuev_s_bc_back_t(cuev_handle_t h, const float *U, float *M, int64_t ldm) { return stage_bc_back_t<float>(h, U, M, ldm); }
int cuev_d_bc_back_t(cuev_handle_t h, const double *U, double *M, int64_t ldm) { return stage_bc_back_t<double>(h, U, M, ldm); }
int cuev_s_sbr_back(cuev_handle_t h, const float *Y, const float *W, float *M, int64_t ldm) { return stage_sbr_back<float>(h, Y, W, M, ldm); }
int cuev_d_sbr_back(cuev_handle_t h, const double *Y, const double *W, double *M, int64_t ldm) { return stage_sbr_back<double>(h, Y, W, M, ldm); }

int cuev_sgemm(hipStream_t s, int ta, int tb, int64_t m, int64_t n, int64_t k, float alpha,
               const float *A, int64_t lda, const float *B, int64_t ldb, float beta, float *C,
               int64_t ldc) {
    return gemm_entry<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int cuev_dgemm(hipStream_t s, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
               int64_t ldc) {
    return gemm_entry<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int cuev_dsymm_lower(hipStream_t s, int64_t m, int64_t n, double alpha, const double *A,
                     int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                     int64_t ldc) {
    return symm_entry<double>(s, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}
int cuev_ssymm_lower(hipStream_t s, int64_t m, int64_t n, float alpha, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float beta, float *C, int64_t ldc) {
    return symm_entry<float>(s, m, n, alpha, A, lda, B, ldb, beta, C, ldc);
}
int cuev_dsyr2k_lower(hipStream_t s, int64_t n, int64_t k, double alpha, const double *A,
                      int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                      int64_t ldc) {
    return syr2k_entry<double>(s, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
int cuev_ssyr2k_lower(hipStream_t s, int64_t n, int64_t k, float alpha, const float *A,
                      int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                      int64_t ldc) {
    return syr2k_entry<float>(s, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

} // extern "C"
