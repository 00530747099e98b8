/*
 * cuev_c.h — C ABI of the MI355X (gfx950) two-stage real symmetric eigensolver.
 *
 * Every entry point takes plain device pointers, sizes and a hipStream_t, and
 * returns a cuev_status_t (0 = success). No entry point calls exit(); the
 * C++ drop-in header include/cuev.h wraps these and keeps the reference's
 * void-returning, abort-on-error behaviour.
 *
 * Storage conventions (identical to the reference):
 *   - matrices are column-major with leading dimension n unless an ld is passed;
 *   - A is symmetric and only its lower triangle is read; A is overwritten;
 *   - eval receives the n eigenvalues in ascending order;
 *   - evec receives the eigenvectors as columns (column j belongs to eval[j]).
 *
 * Reference interfaces replaced (file:line under ynkrue/cuEV):
 *   cuev_{s,d}symm_eig_solve      include/cuev.h:31-33, src/cuda/solver.cu:25-79
 *   cuev_solve_timer              include/cuda/handle.h:25-37
 *   cuev_handle_create / destroy  include/cuda/handle.h:110-119, src/cuda/handle.cu:16-101
 *   cuev_{s,d}_dbbr_pack          include/cuda/kernels.cuh:46, src/cuda/dbbr.cu:116-122
 *   cuev_{s,d}_dbbr_panel_qr      include/cuda/kernels.cuh:63, src/cuda/dbbr.cu:124-143
 *   cuev_{s,d}_dbbr_reduce        include/cuda/kernels.cuh:77, src/cuda/dbbr.cu:145-229
 *   cuev_{s,d}_bc_chase           include/cuda/kernels.cuh:95, src/cuda/bulge.cu:231-247
 *   cuev_{s,d}_tridi_dc           include/cuda/kernels.cuh:114, src/cuda/tridi.cu:425-452
 *   cuev_{s,d}_back_transform     include/cuda/kernels.cuh:135-137, src/cuda/backtransform.cu:232-275
 *   cuev_{s,d}_bc_back            include/cuda/kernels.cuh:144, src/cuda/backtransform.cu:183-207
 *   cuev_{s,d}_bc_back_t          no counterpart: the transposed application of the reordered
 *                                 back-transform (listed as future work, README.md:35)
 *   cuev_{s,d}_sbr_back           include/cuda/kernels.cuh:152, src/cuda/backtransform.cu:210-230
 *   cuev_{s,d}gemm / symm / syr2k include/cuda/kernels.cuh:162-283, src/cuda/cublas.cu:15-56
 *                                 (vendor BLAS call sites, here served by our own MFMA kernels)
 * Entry points with no reference counterpart (the mixed form, the backward
 * pass, options) are marked as such below.
 */
#ifndef CUEV_C_H
#define CUEV_C_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef enum cuev_status_t {
    CUEV_STATUS_SUCCESS = 0,
    CUEV_STATUS_INVALID_ARGUMENT = 1,  /* n < 1, null pointer, bad b/k, dtype mismatch */
    CUEV_STATUS_HIP_ERROR = 2,         /* a HIP runtime call or kernel launch failed */
    CUEV_STATUS_ALLOC_FAILED = 3,      /* device or pinned host allocation failed */
    CUEV_STATUS_LAPACK_ERROR = 4,      /* host LAPACK returned info != 0 */
    CUEV_STATUS_LAPACK_UNAVAILABLE = 5,/* no host LAPACK with the needed routine found */
    CUEV_STATUS_WATCHDOG = 6,          /* a persistent kernel's bounded wait expired */
    CUEV_STATUS_NOT_FINITE = 7,        /* NaN/Inf reached the tridiagonal solver */
    CUEV_STATUS_NOT_SUPPORTED = 8,     /* option combination not implemented */
    CUEV_STATUS_NOT_CONVERGED = 9      /* mixed form: the refinement reached its sweep limit
                                          above refine_tol; eval/evec hold the last iterate */
} cuev_status_t;

/* Human-readable name of a status code. */
const char *cuev_status_string(int status);
/* Detail message of the last failure on the calling host thread ("" if none). */
const char *cuev_last_error_message(void);
/* Name of the host LAPACK library the tridiagonal stage loaded ("" if none yet). */
const char *cuev_lapack_name(void);

/* ------------------------------------------------------------------------ */
/* Per-stage timings (replaces cuev::SolveTimer, include/cuda/handle.h:25-37) */
/* The first nine fields keep the reference's order and meaning; the rest    */
/* are additions (mixed form, overlap accounting).                           */
/* ------------------------------------------------------------------------ */
typedef struct cuev_solve_timer {
    float dbbr_ms;       /* full -> band (DBBR) */
    float bc_ms;         /* band -> tridiagonal (bulge chasing, incl. bc_pack) */
    float dc_ms;         /* host tridiagonal divide and conquer (dstedc), wall time */
    float bt_ms;         /* back-transform work after the D&C returned */
    float total_ms;      /* whole-solve host wall clock */
    float bt_copyin_ms;  /* direct order: M <- Q_d upload; reordered order: Q_d upload */
    float bt_qb_ms;      /* bc_back: M <- Q_b M (reordered order: Q_b built on the identity) */
    float bt_qs_ms;      /* sbr_back: M <- Q_s M */
    float bt_copyout_ms; /* reordered order: evec <- (Q_b^T Q_s^T)^T Q_d GEMM; direct: 0 */
    /* additions */
    float refine_ms;     /* mixed form: Ogita-Aishima refinement sweeps */
    float overlap_ms;    /* GPU back-transform time hidden under the host D&C */
    int refine_sweeps;   /* mixed form: sweeps performed */
    int n_cluster_cols;  /* mixed form: eigenvector columns treated as clustered */
    float refine_resid;  /* mixed form: last measured relative residual (FP64; an estimate
                            from 16 random probes when that check ended the refinement) */
    float refine_orth;   /* mixed form: last measured orthogonality error (likewise) */
} cuev_solve_timer;

/* ------------------------------------------------------------------------ */
/* Options (no reference counterpart: the reference hard-codes b=32, k=512   */
/* at src/cuda/solver.cu:33).                                                */
/* ------------------------------------------------------------------------ */
typedef enum cuev_precision_t {
    CUEV_PRECISION_NATIVE = 0, /* every stage in the API's type */
    CUEV_PRECISION_MIXED = 1   /* double API only: FP32 stages + FP64 Ogita-Aishima refinement */
} cuev_precision_t;

typedef enum cuev_bt_mode_t {
    CUEV_BT_AUTO = 0,      /* reordered for n >= 2048, direct otherwise */
    CUEV_BT_DIRECT = 1,    /* evec <- Q_s (Q_b Q_d) after the D&C (reference order) */
    CUEV_BT_REORDERED = 2  /* build Q_b^T Q_s^T on the GPU beside the chase and the D&C, then one GEMM */
} cuev_bt_mode_t;

typedef enum cuev_dc_mode_t {
    CUEV_DC_AUTO = 0,   /* device for n >= 1024, host otherwise */
    CUEV_DC_HOST = 1,   /* host LAPACK ?stedc, Q_d uploaded (the reference's north-star form) */
    CUEV_DC_DEVICE = 2  /* divide and conquer with Q resident on the GPU: host deflation
                           bookkeeping and leaves, device secular roots, vectors and GEMMs
                           (the split of the reference's hybrid, src/cuda/tridi.cu:125-452) */
} cuev_dc_mode_t;

typedef struct cuev_options {
    int bandwidth;       /* b in {16, 32, 64}; 0 selects the default (64 for n >= 4096, else 32) */
    int block;           /* k, a multiple of b; 0 selects 512 (clamped to n) */
    int precision;       /* cuev_precision_t */
    int refine_sweeps;   /* mixed form: maximum sweeps; 0 selects 3 */
    int bt_mode;         /* cuev_bt_mode_t */
    int dc_mode;         /* cuev_dc_mode_t */
    int dc_leaf;         /* device D&C: largest leaf solved by host ?steqr; 0 selects 64 */
    double refine_tol;   /* mixed form: stop once the FP64 relative residual
                            ||AV - V diag(lambda)||_F / (n ||A||_F) is below this; 0 selects 2e-14 */
} cuev_options;

/* The drop-in entry points keep the workspace of the last problem size per
 * device and precision for the next call (the reference allocates and frees it
 * on every call, src/cuda/handle.cu:73 and solver.cu:72). This frees every idle
 * cached workspace now; it must not race with a solve in progress. No reference
 * counterpart. */
int cuev_cache_release(void);

/* Fill *opt with the defaults listed above. */
void cuev_options_init(cuev_options *opt);

/* ------------------------------------------------------------------------ */
/* Full eigensolve (drop-in for cuev::symm_eig_solve<T>, include/cuev.h:31)  */
/* Blocking: returns after all results are in eval/evec.                     */
/* ------------------------------------------------------------------------ */
int cuev_ssymm_eig_solve(float *A, int n, float *eval, float *evec, hipStream_t stream,
                         cuev_solve_timer *timer);
int cuev_dsymm_eig_solve(double *A, int n, double *eval, double *evec, hipStream_t stream,
                         cuev_solve_timer *timer);
/* Same with options (no reference counterpart). opt may be NULL. */
int cuev_ssymm_eig_solve_ex(float *A, int n, float *eval, float *evec, hipStream_t stream,
                            cuev_solve_timer *timer, const cuev_options *opt);
int cuev_dsymm_eig_solve_ex(double *A, int n, double *eval, double *evec, hipStream_t stream,
                            cuev_solve_timer *timer, const cuev_options *opt);

/* ------------------------------------------------------------------------ */
/* Eigenvalues only (no reference counterpart: the reference computes every  */
/* eigenpair, src/cuda/tridi.cu:12-13). DBBR -> bulge chasing -> Sturm-       */
/* sequence multisection on the GPU; no back-transform, no D&C. A is read     */
/* (lower triangle) and overwritten as by the full solve. Blocking.           */
/*   _values:    all n eigenvalues, ascending, into eval[0..n-1].             */
/*   _values_ex: eigenvalues il..iu (0-based, inclusive, ascending) into      */
/*               eval[0..iu-il]; opt may be NULL (bandwidth and block used;   */
/*               precision must be native).                                   */
/* ------------------------------------------------------------------------ */
int cuev_ssymm_eig_values(float *A, int n, float *eval, hipStream_t stream, cuev_solve_timer *timer);
int cuev_dsymm_eig_values(double *A, int n, double *eval, hipStream_t stream, cuev_solve_timer *timer);
int cuev_ssymm_eig_values_ex(float *A, int n, int il, int iu, float *eval, hipStream_t stream,
                             cuev_solve_timer *timer, const cuev_options *opt);
int cuev_dsymm_eig_values_ex(double *A, int n, int il, int iu, double *eval, hipStream_t stream,
                             cuev_solve_timer *timer, const cuev_options *opt);
/* Eigenvalues il..iu of the symmetric tridiagonal matrix with diagonal d (n) and
 * off-diagonal e (n-1 used), all device pointers; w[0..iu-il] ascending.
 * Asynchronous on stream; uses no scratch memory. */
int cuev_stridi_values(const float *d, const float *e, int n, int il, int iu, float *w, hipStream_t stream);
int cuev_dtridi_values(const double *d, const double *e, int n, int il, int iu, double *w,
                       hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Batched small eigenproblems (no reference counterpart: the reference      */
/* solves one matrix per call, include/cuev.h:31-33). batch matrices of order */
/* n <= 256, stored one after another (each n x n column-major, ld n; lower   */
/* triangle read; A is overwritten); eval receives batch x n eigenvalues      */
/* (ascending per matrix), evec batch eigenvector matrices (columns). One      */
/* workgroup per matrix, cyclic Jacobi. Blocking. Returns NOT_CONVERGED if a   */
/* matrix did not converge (count in *n_not_converged when not NULL).          */
/* ------------------------------------------------------------------------ */
int cuev_ssymm_eig_batched(float *A, int n, int batch, float *eval, float *evec, hipStream_t stream,
                           int *n_not_converged);
int cuev_dsymm_eig_batched(double *A, int n, int batch, double *eval, double *evec, hipStream_t stream,
                           int *n_not_converged);

/* ------------------------------------------------------------------------ */
/* Backward pass of the eigendecomposition (no reference counterpart).        */
/*   Abar = V (diag(lambar) + F o skew) V^T,  skew = (V^T Vbar - Vbar^T V)/2,  */
/*   F_ij = 1/(lambda_j - lambda_i) for i != j, 0 on the diagonal and where     */
/*   |lambda_j - lambda_i| <= gap_tol (masked pairs are counted in n_masked,   */
/*   each unordered pair once). Abar is symmetric. Vbar or lambar may be NULL  */
/*   (treated as zero). V, Vbar, Abar: n x n column-major, ld n.               */
/* ------------------------------------------------------------------------ */
int cuev_ssymm_eig_backward(const float *V, const float *lambda, const float *Vbar,
                            const float *lambar, int n, float *Abar, double gap_tol,
                            int *n_masked, hipStream_t stream);
int cuev_dsymm_eig_backward(const double *V, const double *lambda, const double *Vbar,
                            const double *lambar, int n, double *Abar, double gap_tol,
                            int *n_masked, hipStream_t stream);
/* Allocation-free variant of the two entry points above. No reference counterpart:
 * it exists for callers that cannot allocate device memory (serving and training
 * loops that forbid hipMalloc, harnesses that pre-plan every allocation). Same
 * arguments and result, limited to n <= 1024 (INVALID_ARGUMENT above that). Its
 * work buffers live in a fixed 16 MiB device array per GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU race.
 * With n_masked == NULL the call returns once its work is queued on stream
 * (asynchronous); otherwise it waits for the count, as the entry points above do. */
int cuev_ssymm_eig_backward_noalloc(const float *V, const float *lambda, const float *Vbar,
                                    const float *lambar, int n, float *Abar, double gap_tol,
                                    int *n_masked, hipStream_t stream);
int cuev_dsymm_eig_backward_noalloc(const double *V, const double *lambda, const double *Vbar,
                                    const double *lambar, int n, double *Abar, double gap_tol,
                                    int *n_masked, hipStream_t stream);

/* ------------------------------------------------------------------------ */
/* Stage-level ABI (mirrors cuev::kernels, include/cuda/kernels.cuh:28-157). */
/* A handle owns the workspace of one problem size and precision.            */
/* ------------------------------------------------------------------------ */
typedef struct cuev_handle_s *cuev_handle_t;

typedef enum cuev_dtype_t { CUEV_F32 = 0, CUEV_F64 = 1 } cuev_dtype_t;

/* Workspace buffers readable through cuev_handle_buffer(). */
typedef enum cuev_buffer_t {
    CUEV_BUF_Y = 0,   /* n x n, ld n: DBBR reflectors (explicit unit-lower panels) */
    CUEV_BUF_W = 1,   /* n x n, ld n: W = Y T per panel */
    CUEV_BUF_B = 2,   /* 2b x n, ld 2b: packed band (row i-j of column j holds A[i,j]) */
    CUEV_BUF_U = 3,   /* n x n, ld n: bulge-chasing reflectors, shifted layout:
                         U[(r - s - 1) + s*ld] = row r of the unit reflector of sweep s */
    CUEV_BUF_D = 4,   /* n: tridiagonal diagonal */
    CUEV_BUF_E = 5,   /* n: tridiagonal off-diagonal (n-1 used) */
    CUEV_BUF_TRI = 6, /* b x b, ld b: last panel's block-reflector factor T */
    CUEV_BUF_TAU = 7, /* b: last panel's Householder scalars */
    CUEV_BUF_M = 8    /* n x n, ld n: back-transform scratch (Q_b^T Q_s^T in the reordered order) */
} cuev_buffer_t;

int cuev_handle_create(cuev_handle_t *handle, int dtype, int n, int nbw, int nk,
                       hipStream_t stream);
int cuev_handle_destroy(cuev_handle_t handle);
int cuev_handle_buffer(cuev_handle_t handle, int which, void **ptr, int64_t *ld);

/* Pack the lower band of A (n x n, ld = handle n) into B (2b x n). */
int cuev_s_dbbr_pack(cuev_handle_t h, const float *A, float *B, int n, int b);
int cuev_d_dbbr_pack(cuev_handle_t h, const double *A, double *B, int n, int b);
/* Householder QR of a rows x b panel at A (ld = handle n). On return A holds R
   (upper) and the reflector tails (lower, LAPACK dgeqrf layout), Y the explicit
   unit-lower reflectors (ld = handle n), and the handle's TRI/TAU buffers hold
   T (H_0...H_{b-1} = I - Y T Y^T) and tau. */
int cuev_s_dbbr_panel_qr(cuev_handle_t h, float *A, float *Y, int rows, int b);
int cuev_d_dbbr_panel_qr(cuev_handle_t h, double *A, double *Y, int rows, int b);
/* Full -> band: A (lower) becomes Q_s^T A Q_s in its band; reflectors kept in
   the handle's Y and W; the band is packed into B. */
int cuev_s_dbbr_reduce(cuev_handle_t h, float *A, float *B);
int cuev_d_dbbr_reduce(cuev_handle_t h, double *A, double *B);
/* Band -> tridiagonal. B is overwritten; reflectors go to the handle's U. */
int cuev_s_bc_chase(cuev_handle_t h, float *B, float *d, float *e);
int cuev_d_bc_chase(cuev_handle_t h, double *B, double *d, double *e);
/* Divide and conquer on device d, e; writes eval (ascending) and the
   tridiagonal eigenvectors Q_d into evec (ld n). Host LAPACK ?stedc by default;
   cuev_handle_set_dc selects the device-resident form. */
int cuev_handle_set_dc(cuev_handle_t h, int dc_mode, int dc_leaf);
int cuev_s_tridi_dc(cuev_handle_t h, float *d, float *e, float *eval, float *evec);
int cuev_d_tridi_dc(cuev_handle_t h, double *d, double *e, double *eval, double *evec);
/* evec <- Q_s Q_b evec (evec in: Q_d). */
int cuev_s_back_transform(cuev_handle_t h, const float *Y, const float *W, const float *U,
                          float *evec, cuev_solve_timer *timer);
int cuev_d_back_transform(cuev_handle_t h, const double *Y, const double *W, const double *U,
                          double *evec, cuev_solve_timer *timer);
/* M <- Q_b M, M is n x n with leading dimension ldm >= n. */
int cuev_s_bc_back(cuev_handle_t h, const float *U, float *M, int64_t ldm);
int cuev_d_bc_back(cuev_handle_t h, const double *U, double *M, int64_t ldm);
/* M <- Q_b^T M (no reference counterpart: the reordered back-transform builds
 * Q_b^T Q_s^T with it); M is n x n with leading dimension ldm >= n. */
int cuev_s_bc_back_t(cuev_handle_t h, const float *U, float *M, int64_t ldm);
int cuev_d_bc_back_t(cuev_handle_t h, const double *U, double *M, int64_t ldm);
/* M <- Q_s M, M is n x n with leading dimension ldm >= n. */
int cuev_s_sbr_back(cuev_handle_t h, const float *Y, const float *W, float *M, int64_t ldm);
int cuev_d_sbr_back(cuev_handle_t h, const double *Y, const double *W, double *M, int64_t ldm);

/* ------------------------------------------------------------------------ */
/* MFMA matrix kernels (replace the cuBLAS call sites, src/cuda/cublas.cu).  */
/* trans: 0 = N, 1 = T. All column-major. Scratch for split-K is allocated   */
/* internally when needed.                                                   */
/* ------------------------------------------------------------------------ */
int cuev_sgemm(hipStream_t stream, int transa, int transb, int64_t m, int64_t n, int64_t k,
               float alpha, const float *A, int64_t lda, const float *B, int64_t ldb, float beta,
               float *C, int64_t ldc);
int cuev_dgemm(hipStream_t stream, int transa, int transb, int64_t m, int64_t n, int64_t k,
               double alpha, const double *A, int64_t lda, const double *B, int64_t ldb,
               double beta, double *C, int64_t ldc);
/* C <- alpha * A * B + beta * C, A m x m symmetric, lower triangle read. */
int cuev_dsymm_lower(hipStream_t stream, int64_t m, int64_t n, double alpha, const double *A,
                     int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                     int64_t ldc);
int cuev_ssymm_lower(hipStream_t stream, int64_t m, int64_t n, float alpha, const float *A,
                     int64_t lda, const float *B, int64_t ldb, float beta, float *C, int64_t ldc);
/* lower(C) <- alpha * (A B^T + B A^T) + beta * lower(C); A, B are n x k. */
int cuev_dsyr2k_lower(hipStream_t stream, int64_t n, int64_t k, double alpha, const double *A,
                      int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                      int64_t ldc);
int cuev_ssyr2k_lower(hipStream_t stream, int64_t n, int64_t k, float alpha, const float *A,
                      int64_t lda, const float *B, int64_t ldb, float beta, float *C,
                      int64_t ldc);

/* ------------------------------------------------------------------------ */
/* Opt-in per-launch timing with HIP events (no reference counterpart; the   */
/* reference times whole stages only, through SolveTimer). Also enabled by    */
/* CUEV_PROFILE=1 in the environment.                                        */
/* ------------------------------------------------------------------------ */
int cuev_profile_enable(int on);
/* Restricts recording to tags that start with one of the comma-separated
   prefixes in tags ("bc.chase,bt.gemm"); NULL or "" records every tag. Each
   recorded launch adds two event records to its stream, which separate
   back-to-back kernels by several microseconds: select the few tags a
   measurement needs when the rest of the stream is being timed. */
int cuev_profile_select(const char *tags);
/* Waits for the recorded events and writes per-tag totals as JSON,
   {"tag": {"ms", "count", "flops", "bytes", "max_ms"}, ...}, into buf (len bytes).
   Returns the length the JSON needs (including the terminating NUL).
   reset != 0 clears the records. */
int cuev_profile_read(char *buf, int len, int reset);

#ifdef __cplusplus
}
#endif
#endif /* CUEV_C_H */
