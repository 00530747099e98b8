/*
 * oracle.c — CPU restatement of ynkrue/cuEV's single-GPU two-stage eigensolver.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker the GPU library is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, never by the library (cuev_amd/) itself.
 *
 * Plain C, double precision, sequential, naive loops. Each routine states the
 * reference code it follows and how it is pinned (the checks live in
 * tests/test_oracle_*.py):
 *
 *   orc_fill_glibc      test/cuda/test.h:74-78         pinned: glibc rand() itself
 *   orc_panel_qr        src/cuda/dbbr.cu:124-143 (geqrf + extract :55-71 + syrk + larft :86-109)
 *                       pinned: LAPACK dgeqrf output (scipy) and brute-force product of H_i
 *   orc_dbbr_reduce     src/cuda/dbbr.cu:145-229        pinned: spectrum (LAPACK dsyevd),
 *                       zero below band, Q_s^T A Q_s = band with Q_s built independently
 *   orc_bc_pack         src/cuda/dbbr.cu:31-39          pinned: closed-form index map
 *   orc_bc_chase        src/cuda/bulge.cu:64-224        pinned: spectrum (dsyevd), (d,|e|) of
 *                       LAPACK dsytrd on the dense band (implicit-Q theorem), Q_b^T B Q_b = T
 *   orc_bc_back         src/cuda/backtransform.cu:76-175 (plain reverse order)
 *                       pinned: Q_b^T Band Q_b = Tri, orthogonality (test_backtransform.cu:121-125)
 *   orc_sbr_back        src/cuda/backtransform.cu:210-230
 *                       pinned: Q_s^T A Q_s = Band, orthogonality (test_backtransform.cu:115-119)
 *
 * Conventions shared with the GPU library (stated where they differ from the
 * reference): Householder vectors follow LAPACK dlarfg (tau = 0 when the tail
 * is exactly zero); the bulge-chasing reflector is the identity when the
 * vector below the pivot is zero; U is stored "shifted", U[(r-s-1) + s*ldu] =
 * row r of sweep s's reflector. dlarfg's underflow rescaling is not restated
 * (inputs of order 1 never reach it).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* test/cuda/test.h:74-78: srand(seed); x = (rand() % 2000 - 1000) / 1000 */
void orc_fill_glibc(double *out, long count, unsigned seed) {
    srand(seed);
    for (long i = 0; i < count; ++i) out[i] = (double)(rand() % 2000 - 1000) / 1000.0;
}

/* bench/bench_solver.cpp:53-58: symmetric fill (rand() % 200 - 100) / 100 over the lower triangle */
void orc_fill_glibc_symmetric(double *h, int n, unsigned seed) {
    srand(seed);
    for (int j = 0; j < n; ++j)
        for (int i = j; i < n; ++i) h[IDX(i, j, n)] = h[IDX(j, i, n)] = (double)(rand() % 200 - 100) / 100.0;
}

/* C(m x n) = alpha * op(A) op(B) + beta * C, column-major, naive */
static void gemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
                 const double *B, int ldb, double beta, double *C, int ldc) {
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) {
                const double a = ta ? A[IDX(l, i, lda)] : A[IDX(i, l, lda)];
                const double b = tb ? B[IDX(j, l, ldb)] : B[IDX(l, j, ldb)];
                s += a * b;
            }
            C[IDX(i, j, ldc)] = alpha * s + (beta == 0.0 ? 0.0 : beta * C[IDX(i, j, ldc)]);
        }
}

/*
 * Panel QR of rows x b (ld lda), LAPACK dgeqr2 + dlarft semantics.
 * On return: A holds R (upper) and the reflector tails (lower); Y (ld ldy) holds
 * the explicit unit-lower reflectors; T (b x b, ld b) with H_0..H_{b-1} = I - Y T Y^T;
 * tau (b). Follows dbbr_panel_qr (src/cuda/dbbr.cu:124-143).
 */
void orc_panel_qr(int rows, int b, double *A, int lda, double *Y, int ldy, double *T, double *tau) {
    for (int c = 0; c < b; ++c) {
        if (c >= rows) {
            tau[c] = 0.0;
            continue;
        }
        const double alpha = A[IDX(c, c, lda)];
        double xnorm2 = 0.0;
        for (int r = c + 1; r < rows; ++r) xnorm2 += A[IDX(r, c, lda)] * A[IDX(r, c, lda)];
        double beta = alpha, t = 0.0, scal = 0.0;
        if (xnorm2 != 0.0) {
            const double nrm = sqrt(alpha * alpha + xnorm2);
            beta = (alpha >= 0.0) ? -nrm : nrm;
            t = (beta - alpha) / beta;
            scal = 1.0 / (alpha - beta);
        }
        tau[c] = t;
        for (int r = c + 1; r < rows; ++r) A[IDX(r, c, lda)] *= scal;
        A[IDX(c, c, lda)] = beta;
        if (t != 0.0) {
            for (int j = c + 1; j < b; ++j) {
                double s = A[IDX(c, j, lda)];
                for (int r = c + 1; r < rows; ++r) s += A[IDX(r, c, lda)] * A[IDX(r, j, lda)];
                A[IDX(c, j, lda)] -= t * s;
                for (int r = c + 1; r < rows; ++r) A[IDX(r, j, lda)] -= t * s * A[IDX(r, c, lda)];
            }
        }
    }
    /* explicit Y (dbbr_extract_reflectors, dbbr.cu:55-71) */
    for (int c = 0; c < b; ++c)
        for (int r = 0; r < rows; ++r)
            Y[IDX(r, c, ldy)] = (r < c) ? 0.0 : (r == c ? 1.0 : A[IDX(r, c, lda)]);
    /* T by the larft recurrence on G = Y^T Y (dbbr.cu:86-109) */
    for (int i = 0; i < b * b; ++i) T[i] = 0.0;
    for (int j = 0; j < b; ++j) {
        double g[64];
        for (int q = 0; q < j; ++q) {
            double s = 0.0;
            for (int r = 0; r < rows; ++r) s += Y[IDX(r, q, ldy)] * Y[IDX(r, j, ldy)];
            g[q] = s;
        }
        for (int p = 0; p < j; ++p) {
            double s = 0.0;
            for (int q = p; q < j; ++q) s += T[IDX(p, q, b)] * g[q];
            T[IDX(p, j, b)] = -tau[j] * s;
        }
        T[IDX(j, j, b)] = tau[j];
    }
}

/* bc_pack (dbbr.cu:31-39): Bp[r + j*2b] = A[j+r, j] for r <= b, j+r < n; else 0 */
void orc_bc_pack(int n, int b, const double *A, int lda, double *Bp) {
    const int ldb = 2 * b;
    for (int j = 0; j < n; ++j)
        for (int r = 0; r < ldb; ++r)
            Bp[IDX(r, j, ldb)] = (r <= b && j + r < n) ? A[IDX(j + r, j, lda)] : 0.0;
}

/*
 * DBBR (dbbr.cu:145-229). A: n x n, lower triangle read, ld n; on return its
 * lower band is Q_s^T A Q_s. Y, W: n x n (ld n), Z: n x k (ld n), Bp: 2b x n.
 */
void orc_dbbr_reduce(int n, int b, int k, double *A, double *Y, double *W, double *Z, double *Bp) {
    const int lda = n;
    double *T = (double *)malloc(sizeof(double) * b * b);
    double *tau = (double *)malloc(sizeof(double) * b);
    double *D = (double *)malloc(sizeof(double) * (size_t)2 * k * b);
    double *C = (double *)malloc(sizeof(double) * b * b);
    for (int i = 0; i < n; i += k) {
        const int kc = (k < n - i) ? k : n - i;
        int w = 0;
        for (int j = i; j < i + kc && j + b < n; j += b) {
            const int rows = n - j - b;
            const size_t off = IDX(j + b, j, lda), tr = IDX(j + b, j + b, lda);
            double *P = Z + IDX(j + b, j - i, lda);
            double *Yb = Y + IDX(j + b, i, lda), *Zb = Z + IDX(j + b, 0, lda);
            orc_panel_qr(rows, b, A + off, lda, Y + off, lda, T, tau);
            gemm(0, 0, rows, b, b, 1.0, Y + off, lda, T, b, 0.0, W + off, lda); /* W = Y T */
            /* P = Sym(A22) W, A22 from its lower triangle */
            for (int c = 0; c < b; ++c)
                for (int r = 0; r < rows; ++r) {
                    double s = 0.0;
                    for (int q = 0; q < rows; ++q) {
                        const double a = (r >= q) ? A[tr + IDX(r, q, lda)] : A[tr + IDX(q, r, lda)];
                        s += a * W[off + IDX(q, c, lda)];
                    }
                    P[IDX(r, c, lda)] = s;
                }
            if (w > 0) {
                /* P -= Y_blk (Z_blk^T W) + Z_blk (Y_blk^T W)  (dbbr.cu:182-194) */
                gemm(1, 0, w, b, rows, 1.0, Zb, lda, W + off, lda, 0.0, D, w);
                gemm(0, 0, rows, b, w, -1.0, Yb, lda, D, w, 1.0, P, lda);
                gemm(1, 0, w, b, rows, 1.0, Yb, lda, W + off, lda, 0.0, D, w);
                gemm(0, 0, rows, b, w, -1.0, Zb, lda, D, w, 1.0, P, lda);
            }
            gemm(1, 0, b, b, rows, 1.0, W + off, lda, P, lda, 0.0, C, b);      /* C = W^T P */
            gemm(0, 0, rows, b, b, -0.5, Y + off, lda, C, b, 1.0, P, lda);      /* Z = P - Y C/2 */
            w += b;
            if (j + b < i + kc) { /* green update (dbbr.cu:208-217) */
                /* only the nc = min(b, n-j-b) columns that exist: the reference updates b
                   columns, past the end of A when the last block is ragged */
                const int nc = (b < n - j - b) ? b : n - j - b;
                gemm(0, 1, rows, nc, w, -1.0, Zb, lda, Yb, lda, 1.0, A + tr, lda);
                gemm(0, 1, rows, nc, w, -1.0, Yb, lda, Zb, lda, 1.0, A + tr, lda);
            }
        }
        const int t = i + kc;
        if (t < n && w > 0) { /* syr2k on the lower triangle (dbbr.cu:222-224) */
            const double *Yt = Y + IDX(t, i, lda), *Zt = Z + IDX(t, 0, lda);
            for (int c = 0; c < n - t; ++c)
                for (int r = c; r < n - t; ++r) {
                    double s = 0.0;
                    for (int q = 0; q < w; ++q)
                        s += Yt[IDX(r, q, lda)] * Zt[IDX(c, q, lda)] + Zt[IDX(r, q, lda)] * Yt[IDX(c, q, lda)];
                    A[IDX(t + r, t + c, lda)] -= s;
                }
        }
    }
    orc_bc_pack(n, b, A, lda, Bp);
    free(T);
    free(tau);
    free(D);
    free(C);
}

/* symmetric read of the packed band (include/common.h:57-71) */
static double *band_ref(double *Bp, int ldb, int i, int j) {
    if (i < j) {
        const int t = i;
        i = j;
        j = t;
    }
    return &Bp[IDX(i - j, j, ldb)];
}

/*
 * Bulge chasing, sequential in sweep-major order (bulge.cu:64-224 without the
 * wavefront). Bp: 2b x n packed band (overwritten). U: ldu x n, shifted layout.
 * d: n, e: n (e[n-1] = 0).
 */
void orc_bc_chase(int n, int b, double *Bp, double *U, int ldu, double *d, double *e) {
    const int ldb = 2 * b;
    double *w = (double *)malloc(sizeof(double) * b);
    double *p = (double *)malloc(sizeof(double) * 3 * b + 8);
    for (int s = 0; s < n - 2; ++s) {
        for (int h = 0;; ++h) {
            const int cq = (h == 0) ? s : s + 1 + (h - 1) * b;
            const int r0 = (h == 0) ? s + 1 : cq + b;
            if (r0 > n - 1) break;
            const int r1 = (r0 + b - 1 < n - 1) ? r0 + b - 1 : n - 1;
            const int L = r1 - r0 + 1;
            if (L < 2) {
                U[IDX(r0 - s - 1, s, ldu)] = 0.0;
                break;
            }
            const int hi = (r1 + b < n - 1) ? r1 + b : n - 1;
            /* Householder (normalised, identity for a zero tail) */
            const double alpha = *band_ref(Bp, ldb, r0, cq);
            double tail = 0.0;
            for (int t = 1; t < L; ++t) {
                const double x = *band_ref(Bp, ldb, r0 + t, cq);
                tail += x * x;
            }
            double beta = alpha;
            if (tail != 0.0) {
                const double nrm = sqrt(alpha * alpha + tail);
                beta = (alpha >= 0.0) ? -nrm : nrm;
                const double v0 = alpha - beta;
                const double inv = 1.0 / sqrt(v0 * v0 + tail);
                w[0] = v0 * inv;
                for (int t = 1; t < L; ++t) w[t] = *band_ref(Bp, ldb, r0 + t, cq) * inv;
            } else {
                for (int t = 0; t < L; ++t) w[t] = 0.0;
            }
            /* p = 2 A w over [cq, hi]; kappa = w^T p; y = p - kappa w */
            for (int i = cq; i <= hi; ++i) {
                double acc = 0.0;
                for (int t = 0; t < L; ++t) acc += *band_ref(Bp, ldb, i, r0 + t) * w[t];
                p[i - cq] = 2.0 * acc;
            }
            double kap = 0.0;
            for (int t = 0; t < L; ++t) kap += w[t] * p[r0 + t - cq];
            for (int t = 0; t < L; ++t) p[r0 + t - cq] -= kap * w[t];
            /* A <- A - w y^T - y w^T on the lower band of columns [cq, r1] */
            for (int j = cq; j <= r1; ++j)
                for (int i = j; i <= hi && i - j < ldb; ++i) {
                    const int wi = (i >= r0 && i <= r1), wj = (j >= r0);
                    if (!wi && !wj) continue;
                    double *a = &Bp[IDX(i - j, j, ldb)];
                    if (j == cq) {
                        *a = (i == r0) ? beta : 0.0;
                        continue;
                    }
                    const double wiv = wi ? w[i - r0] : 0.0, wjv = wj ? w[j - r0] : 0.0;
                    *a -= wiv * p[j - cq] + p[i - cq] * wjv;
                }
            for (int t = 0; t < L; ++t) U[IDX(r0 - s - 1 + t, s, ldu)] = w[t];
        }
    }
    for (int i = 0; i < n; ++i) d[i] = Bp[IDX(0, i, ldb)];
    for (int i = 0; i < n - 1; ++i) e[i] = Bp[IDX(1, i, ldb)];
    e[n - 1] = 0.0;
    free(w);
    free(p);
}

/*
 * M <- Q_b M (n x ncols, ld ldm). Reflectors in plain reverse order: sweeps
 * s = n-3..0, hops last..0 (the forward order is sweep-major, bulge.cu:180-182).
 * reverse_order = 0 applies them in forward order instead (Q_b^T M); the tests
 * use it to show that the invariant checks catch an order mistake.
 */
void orc_bc_back(int n, int b, const double *U, int ldu, double *M, int ldm, int ncols,
                 int reverse_order) {
    if (n < 3) return;
    const int nh = (n + b - 1) / b + 1;
    for (int si = 0; si <= n - 3; ++si) {
        const int s = reverse_order ? n - 3 - si : si;
        for (int hi_ = 0; hi_ < nh; ++hi_) {
            const int hmax = nh - 1;
            const int h = reverse_order ? hmax - hi_ : hi_;
            const int r0 = s + 1 + h * b;
            if (r0 > n - 1) continue;
            const int r1 = (s + (h + 1) * b < n - 1) ? s + (h + 1) * b : n - 1;
            for (int c = 0; c < ncols; ++c) {
                double proj = 0.0;
                for (int r = r0; r <= r1; ++r) proj += U[IDX(r - s - 1, s, ldu)] * M[IDX(r, c, ldm)];
                proj *= 2.0;
                if (proj == 0.0) continue;
                for (int r = r0; r <= r1; ++r) M[IDX(r, c, ldm)] -= proj * U[IDX(r - s - 1, s, ldu)];
            }
        }
    }
}

/* M <- Q_s M (backtransform.cu:210-230): panels last-first, M_b -= W (Y^T M_b) */
void orc_sbr_back(int n, int b, const double *Y, const double *W, int ld, double *M, int ldm,
                  int ncols) {
    int jmax = -1;
    for (int j = 0; j + b < n; j += b) jmax = j;
    double *Zt = (double *)malloc(sizeof(double) * (size_t)b * (ncols > 0 ? ncols : 1));
    for (int j = jmax; j >= 0; j -= b) {
        const int rows = n - (j + b);
        const double *Yp = Y + IDX(j + b, j, ld), *Wp = W + IDX(j + b, j, ld);
        double *Mb = M + (j + b);
        gemm(1, 0, b, ncols, rows, 1.0, Yp, ld, Mb, ldm, 0.0, Zt, b);
        gemm(0, 0, rows, ncols, b, -1.0, Wp, ld, Zt, b, 1.0, Mb, ldm);
    }
    free(Zt);
}

/* C = alpha op(A) op(B) + beta C: exported for tests that need a plain reference product */
void orc_gemm(int ta, int tb, int m, int n, int k, double alpha, const double *A, int lda,
              const double *B, int ldb, double beta, double *C, int ldc) {
    gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}
