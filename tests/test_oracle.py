"""CPU checks that pin the oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle routine is checked against something it cannot share a mistake with:
LAPACK (dgeqrf, dsyevd, dsytrd via SciPy), brute-force products, or the
reference's own invariants — and each check is shown to fail on a planted
mistake (wrong reflector order, dropped deferred term, transposed factor).
"""
import numpy as np
import pytest
import scipy.linalg as sl
from scipy.linalg import lapack

from oracle import oracle as O


def test_glibc_fill_matches_reference_formula():
    # test/cuda/test.h:74-78: srand(seed); (rand() % 2000 - 1000) / 1000
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(12345)
    expect = [(libc.rand() % 2000 - 1000) / 1000.0 for _ in range(50)]
    assert np.array_equal(O.fill_glibc(50, 12345), expect)
    A = O.fill_glibc_symmetric(6, 1234)
    assert np.array_equal(A, A.T)
    assert np.all(np.abs(A) <= 1.0)


@pytest.mark.parametrize("rows,b", [(256, 32), (512, 64), (40, 64), (64, 64), (100, 16)])
def test_panel_qr_matches_lapack_dgeqrf(rows, b):
    A0 = O.fill_glibc_matrix(rows, b, 12345)
    Af, Y, T, tau = O.panel_qr(A0, b)
    (qr, tau_l), _ = sl.qr(A0, mode="raw")
    m = min(rows, b)
    assert np.abs(np.triu(Af[:m]) - np.triu(qr[:m])).max() < 1e-13 * np.abs(A0).max() * rows
    assert np.abs(np.tril(Af, -1) - np.tril(qr, -1)).max() < 1e-12
    assert np.abs(tau[:m] - tau_l[:m]).max() < 1e-14
    # T: H_0 H_1 ... H_{b-1} = I - Y T Y^T, with the product formed explicitly
    Q = np.eye(rows)
    for c in range(m):
        v = Y[:, c]
        Q = Q @ (np.eye(rows) - tau[c] * np.outer(v, v))
    assert np.abs(Q - (np.eye(rows) - Y @ T @ Y.T)).max() < 1e-13
    # the reference's metrics (test_dbbr.cu:52-107)
    R = np.zeros_like(A0)
    R[:m] = np.triu(Af[:m])
    assert np.linalg.norm(Q @ R - A0) / np.linalg.norm(A0) < 1e-13


def test_panel_qr_zero_column_is_identity():
    A0 = O.fill_glibc_matrix(50, 8, 3)
    A0[:, 2] = 0.0
    _, _, _, tau = O.panel_qr(A0, 8)
    assert tau[2] == 0.0


def _dbbr(n, b, k, seed=7):
    A0 = O.fill_glibc_matrix(n, n, seed)
    return A0, O.dbbr_reduce(A0, b, k)


@pytest.mark.parametrize("n,b,k", [(256, 64, 256), (512, 32, 128), (384, 32, 96), (300, 16, 64), (97, 16, 32)])
def test_dbbr_oracle_invariants(n, b, k):
    A0, r = _dbbr(n, b, k)
    As = O.sym_from_lower(A0)
    Band = O.band_dense(r["A"], b)
    ev = np.linalg.eigvalsh(As)
    assert np.abs(ev - np.linalg.eigvalsh(Band)).max() / (ev[-1] - ev[0]) < 1e-13
    # Q_s formed by applying the panels' WY factors to I (independent of the band)
    Qs = O.sbr_back(r["Y"], r["W"], np.eye(n, order="F"), b)
    assert np.linalg.norm(Qs.T @ Qs - np.eye(n)) / n < 1e-14
    assert np.linalg.norm(Qs.T @ As @ Qs - Band) / np.linalg.norm(Band) < 1e-13
    # nothing but reflector storage below the band: Q_s^T A Q_s is banded
    full = Qs.T @ As @ Qs
    assert np.abs(np.tril(full, -b - 1)).max() < 1e-12 * np.abs(As).max() * n


def test_dbbr_deferred_update_is_needed():
    """Planted mistake: dropping the correction terms breaks the similarity invariant."""
    n, b, k = 192, 32, 96
    A0, r = _dbbr(n, b, k)
    As = O.sym_from_lower(A0)
    # an update with W replaced by Y (no T) is a different, wrong transform
    Qs_wrong = O.sbr_back(r["Y"], r["Y"], np.eye(n, order="F"), b)
    Band = O.band_dense(r["A"], b)
    assert np.linalg.norm(Qs_wrong.T @ As @ Qs_wrong - Band) / np.linalg.norm(Band) > 1e-2


@pytest.mark.parametrize("n,b,k", [(256, 64, 256), (512, 32, 128), (300, 16, 64), (130, 64, 128)])
def test_bc_chase_oracle(n, b, k):
    A0, r = _dbbr(n, b, k)
    As = O.sym_from_lower(A0)
    Band = O.band_from_packed(r["B"], b)
    c = O.bc_chase(r["B"], b)
    T = O.tridiag_dense(c["d"], c["e"])
    ev = np.linalg.eigvalsh(As)
    assert np.abs(ev - np.linalg.eigvalsh(T)).max() / (ev[-1] - ev[0]) < 1e-13
    # implicit-Q theorem: Q_b e_1 = e_1 for both reductions, so (d, |e|) equals
    # LAPACK dsytrd's on the dense band (lower, columns left to right)
    _, dd, ee, _, info = lapack.dsytrd(Band, lower=1)
    assert info == 0
    scale = np.abs(Band).max()
    assert np.abs(dd - c["d"]).max() / scale < 1e-9
    assert np.abs(np.abs(ee) - np.abs(c["e"][:-1])).max() / scale < 1e-9
    # Q_b^T Band Q_b = T with Q_b materialised by the back-transform oracle
    Qb = O.bc_back(c["U"], np.eye(n, order="F"), b)
    assert np.linalg.norm(Qb.T @ Band @ Qb - T) / np.linalg.norm(T) < 1e-13
    assert np.linalg.norm(Qb.T @ Qb - np.eye(n)) / n < 1e-14


def test_bc_back_order_mistake_is_caught():
    n, b, k = 200, 16, 64
    A0, r = _dbbr(n, b, k)
    Band = O.band_from_packed(r["B"], b)
    c = O.bc_chase(r["B"], b)
    T = O.tridiag_dense(c["d"], c["e"])
    Qw = O.bc_back(c["U"], np.eye(n, order="F"), b, reverse_order=False)
    assert np.linalg.norm(Qw.T @ Band @ Qw - T) / np.linalg.norm(T) > 1e-2


def test_bc_chase_tiny_brute_force():
    """n=6, b=3: the tridiagonal from the chase has the spectrum of A and Q_b e_1 = e_1."""
    rng = np.random.default_rng(3)
    n, b = 6, 3
    G = rng.standard_normal((n, n))
    A = np.tril((G + G.T) / 2)
    A[np.tril_indices(n, -b - 1)] = 0.0
    Bp = O.bc_pack(A, b)
    c = O.bc_chase(Bp, b)
    S = O.band_dense(A, b)
    assert np.allclose(np.linalg.eigvalsh(S), np.linalg.eigvalsh(O.tridiag_dense(c["d"], c["e"])), atol=1e-14)
    Qb = O.bc_back(c["U"], np.eye(n, order="F"), b)
    assert np.allclose(Qb[:, 0], np.eye(n)[:, 0], atol=1e-15)


def test_bc_chase_zero_band_is_identity():
    n, b = 40, 8
    A = np.diag(np.arange(n, dtype=float))
    c = O.bc_chase(O.bc_pack(A, b), b)
    assert np.array_equal(c["d"], np.arange(n, dtype=float))
    assert np.all(c["e"] == 0.0)
    assert np.all(c["U"] == 0.0)


def test_full_pipeline_oracle_vs_dsyevd():
    """Oracle stages + LAPACK tridiagonal solve reproduce dsyevd's eigenpairs."""
    n, b, k = 200, 16, 64
    A0, r = _dbbr(n, b, k, seed=11)
    As = O.sym_from_lower(A0)
    c = O.bc_chase(r["B"], b)
    lam, Qd = sl.eigh_tridiagonal(c["d"], c["e"][:-1])
    V = O.sbr_back(r["Y"], r["W"], O.bc_back(c["U"], np.asfortranarray(Qd), b), b)
    lam_ref = sl.eigh(As, eigvals_only=True, driver="evd")
    assert np.abs(lam - lam_ref).max() / np.abs(lam_ref).max() < 1e-13
    assert np.linalg.norm(As @ V - V * lam) / (n * np.linalg.norm(As)) < 1e-15


@pytest.mark.parametrize("n,b,k", [(250, 64, 128), (300, 16, 64), (97, 16, 32), (200, 64, 128)])
def test_dbbr_ragged_last_block_stays_in_bounds(n, b, k):
    """Regression: the next-panel update must not touch columns >= n when n - j - b < b.

    A is stored at the front of a larger buffer whose tail (where column n would
    start, ld = n) holds a sentinel; a write past column n-1 changes it. Y, W, Z
    start as garbage: the reduction may only read entries it has written, and an
    out-of-range update then writes a value that differs from the sentinel.
    """
    import ctypes
    A0 = O.fill_glibc_matrix(n, n, 7)
    flat = np.full(n * n + 2 * b * n, 12345.0)
    flat[: n * n] = A0.ravel(order="F")
    Y = np.full((n, n), 7.0, order="F")
    W = np.full((n, n), 7.0, order="F")
    Z = np.full((n, k), 7.0, order="F")
    Bp = np.zeros((2 * b, n), order="F")
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    O.lib().orc_dbbr_reduce(n, b, k, vp(flat), vp(Y), vp(W), vp(Z), vp(Bp))
    assert np.all(flat[n * n:] == 12345.0), "wrote past the end of A"
    Ar = flat[: n * n].reshape((n, n), order="F")
    ev = np.linalg.eigvalsh(O.sym_from_lower(A0))
    assert np.abs(ev - np.linalg.eigvalsh(O.band_dense(Ar, b))).max() / (ev[-1] - ev[0]) < 1e-13


def _bc_back_blocked(U, M, b, ng=16):
    """The application order of the device bc_back kernel, in numpy: passes of 128/b hop
    indices (backtransform.hip: G = 128 / B); inside a pass, groups of ng sweeps from the last down; inside a group, its
    hops upward, each as one block reflector I - V T V^T with
    T = (I/2 + striu(V^T V))^-1 (reflectors I - 2 w w^T)."""
    n = M.shape[0]
    M = M.copy()
    G, last, h0 = 128 // b, n - 3, 0
    rows_per_block = 16 * (b // 16 + 1)
    while min(last, n - 2 - h0 * b) >= 0:
        smax = min(last, n - 2 - h0 * b)
        for g in range(smax // ng, -1, -1):
            s0 = g * ng
            for h in range(h0, h0 + G):
                base = s0 + 1 + h * b
                if base > n - 1:
                    break
                V = np.zeros((rows_per_block, ng))
                for j in range(ng):
                    if s0 + j <= last:
                        V[j:j + b, j] = U[h * b:h * b + b, s0 + j]
                T = np.linalg.inv(0.5 * np.eye(ng) + np.triu(V.T @ V, 1))
                rows = np.arange(base, base + rows_per_block)
                ok = rows < n
                MJ = np.zeros((rows_per_block, M.shape[1]))
                MJ[ok] = M[rows[ok]]
                MJ -= V @ (T @ (V.T @ MJ))
                M[rows[ok]] = MJ[ok]
        h0 += G
    return M


@pytest.mark.parametrize("n,b", [(40, 16), (100, 32), (300, 64), (257, 16), (513, 32), (700, 64)])
def test_bc_back_block_schedule_equals_oracle(n, b):
    """The blocked order the device kernel uses (sweep groups, hop passes, compact-WY
    blocks) gives the same Q_b M as the oracle's reflector-by-reflector order."""
    A = O.fill_glibc_symmetric(n, seed=n)
    r = O.bc_chase(O.bc_pack(A, b), b)
    M0 = np.random.default_rng(n).standard_normal((n, 7))
    ref = O.bc_back(r["U"], M0, b)
    assert np.abs(_bc_back_blocked(r["U"], M0, b) - ref).max() < 1e-13
    # and the check has teeth: the forward (transposed) order differs
    assert np.abs(O.bc_back(r["U"], M0, b, reverse_order=False) - ref).max() > 1e-3


def _sbr_back_t_right(Y, W, b, k):
    """Q_s^T as the device's sbr_back_t builds it, in numpy: the reduction's blocks in
    compact form Q_t = I - What_t Yhat_t^T (What built left to right,
    What_t = W_t - What_{0:t} (Yhat_{0:t}^T W_t)), then M = I and, last block first,
    M[r0:, r0:] -= (M[r0:, r0:] Yhat[r0:]) What[r0:]^T (right multiplication by Q_t^T
    restricted to the trailing block)."""
    n = Y.shape[0]

    def panel(j):
        rt = j + b
        Yp = np.zeros((n, b))
        Wp = np.zeros((n, b))
        Yp[rt:] = Y[rt:, j:j + b]
        Wp[rt:] = W[rt:, j:j + b]
        return Yp, Wp

    blocks = []
    i = 0
    while i + b < n:
        Yh, Wh = [], []
        t = 0
        while t * b < k and i + t * b + b < n:
            Yp, Wp = panel(i + t * b)
            if Wh:
                Wp = Wp - np.hstack(Wh) @ (np.hstack(Yh).T @ Wp)
            Yh.append(Yp)
            Wh.append(Wp)
            t += 1
        blocks.append((i, np.hstack(Yh), np.hstack(Wh)))
        i += k
    M = np.eye(n)
    for i, Yh, Wh in reversed(blocks):
        r0 = i + b
        P = M[r0:, r0:] @ Yh[r0:]
        M[r0:, r0:] -= P @ Wh[r0:].T
    return M


@pytest.mark.parametrize("n,b,k", [(300, 16, 64), (257, 32, 96), (500, 64, 128)])
def test_sbr_back_t_right_application_equals_oracle(n, b, k):
    """The trailing-block right application the device uses for Q_s^T equals the
    transpose of the oracle's sbr_back (the reference's panel order) applied to I."""
    A0 = O.fill_glibc_matrix(n, n, n)
    r = O.dbbr_reduce(A0, b, k)
    Qs = O.sbr_back(r["Y"], r["W"], np.eye(n, order="F"), b)
    got = _sbr_back_t_right(r["Y"], r["W"], b, k)
    assert np.abs(got - Qs.T).max() < 1e-13
    # and the check has teeth: Q_s itself (no transpose) differs
    assert np.abs(got - Qs).max() > 1e-3
