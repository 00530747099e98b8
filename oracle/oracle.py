"""ctypes/numpy wrapper of the CPU oracle (oracle.c) — test infrastructure only.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the library. All matrices are numpy float64 arrays in Fortran
(column-major) order, matching the library's storage.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")


def build():
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        I, U, D, L_ = ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_long
        P = ctypes.c_void_p
        L.orc_fill_glibc.argtypes = [P, L_, U]
        L.orc_fill_glibc_symmetric.argtypes = [P, I, U]
        L.orc_panel_qr.argtypes = [I, I, P, I, P, I, P, P]
        L.orc_bc_pack.argtypes = [I, I, P, I, P]
        L.orc_dbbr_reduce.argtypes = [I, I, I, P, P, P, P, P]
        L.orc_bc_chase.argtypes = [I, I, P, P, I, P, P]
        L.orc_bc_back.argtypes = [I, I, P, I, P, I, I, I]
        L.orc_sbr_back.argtypes = [I, I, P, P, I, P, I, I]
        L.orc_gemm.argtypes = [I, I, I, I, I, D, P, I, P, I, D, P, I]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(ctypes.c_void_p)


def ldu_for(n):
    """Leading dimension of the shifted U layout used by the library (n + 512, rounded to 8)."""
    return ((n + 512 + 7) // 8) * 8


def fill_glibc(count, seed):
    """test/cuda/test.h:74-78 fill_random with glibc rand()."""
    out = np.empty(count, dtype=np.float64)
    lib().orc_fill_glibc(_p(out), count, seed)
    return out


def fill_glibc_matrix(rows, cols, seed):
    return np.asfortranarray(fill_glibc(rows * cols, seed).reshape((rows, cols), order="F"))


def fill_glibc_symmetric(n, seed=1234):
    """bench/bench_solver.cpp:53-58 fill_symmetric."""
    A = np.empty((n, n), dtype=np.float64, order="F")
    lib().orc_fill_glibc_symmetric(_p(A), n, seed)
    return A


def panel_qr(A, b):
    """Returns (A_factored, Y, T, tau) for a rows x b panel."""
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    rows = A.shape[0]
    Y = np.zeros((rows, b), order="F")
    T = np.zeros((b, b), order="F")
    tau = np.zeros(b)
    lib().orc_panel_qr(rows, b, _p(A), rows, _p(Y), rows, _p(T), _p(tau))
    return A, Y, T, tau


def bc_pack(A, b):
    n = A.shape[0]
    A = np.asfortranarray(A, dtype=np.float64)
    Bp = np.zeros((2 * b, n), order="F")
    lib().orc_bc_pack(n, b, _p(A), n, _p(Bp))
    return Bp


def dbbr_reduce(A, b, k):
    """Returns dict with A (reduced), Y, W, Z, B (packed band)."""
    n = A.shape[0]
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    Y = np.zeros((n, n), order="F")
    W = np.zeros((n, n), order="F")
    Z = np.zeros((n, k), order="F")
    Bp = np.zeros((2 * b, n), order="F")
    lib().orc_dbbr_reduce(n, b, k, _p(A), _p(Y), _p(W), _p(Z), _p(Bp))
    return dict(A=A, Y=Y, W=W, Z=Z, B=Bp)


def bc_chase(Bp, b):
    """Returns dict with B (reduced), U (shifted, ldu x n), d, e."""
    n = Bp.shape[1]
    Bp = np.asfortranarray(Bp, dtype=np.float64).copy(order="F")
    ldu = ldu_for(n)
    U = np.zeros((ldu, n), order="F")
    d = np.zeros(n)
    e = np.zeros(n)
    lib().orc_bc_chase(n, b, _p(Bp), _p(U), ldu, _p(d), _p(e))
    return dict(B=Bp, U=U, d=d, e=e)


def bc_back(U, M, b, reverse_order=True):
    """M <- Q_b M (reverse_order=False applies Q_b^T: used to show the checks catch it)."""
    n = M.shape[0]
    M = np.asfortranarray(M, dtype=np.float64).copy(order="F")
    lib().orc_bc_back(n, b, _p(U), U.shape[0], _p(M), n, M.shape[1], 1 if reverse_order else 0)
    return M


def sbr_back(Y, W, M, b):
    n = M.shape[0]
    M = np.asfortranarray(M, dtype=np.float64).copy(order="F")
    lib().orc_sbr_back(n, b, _p(Y), _p(W), n, _p(M), n, M.shape[1])
    return M


# ---- dense helpers (numpy) ------------------------------------------------------------------

def sym_from_lower(A):
    L = np.tril(A)
    return L + np.tril(A, -1).T


def band_dense(A, b):
    """Symmetric dense band from the lower band of A (test_backtransform.cu:40-46)."""
    n = A.shape[0]
    S = np.zeros((n, n))
    for j in range(n):
        hi = min(j + b, n - 1)
        S[j:hi + 1, j] = A[j:hi + 1, j]
    return S + np.tril(S, -1).T


def band_from_packed(Bp, b):
    ldb, n = Bp.shape
    S = np.zeros((n, n))
    for j in range(n):
        for r in range(ldb):
            if j + r < n:
                S[j + r, j] = Bp[r, j]
    return S + np.tril(S, -1).T


def tridiag_dense(d, e):
    n = len(d)
    T = np.diag(d)
    if n > 1:
        T += np.diag(e[:n - 1], 1) + np.diag(e[:n - 1], -1)
    return T
