"""MFMA engine (gemm / symm-lower / syr2k-lower) against float64 numpy products.

Covers both transposes, K-split shapes (few output tiles, long K), ragged sizes,
the DBBR shapes (tall-skinny, b x b with K = n) and FP32 (tolerance scaled by K).
"""
import numpy as np
import pytest

import cuev_amd as cv
from helpers import to_dev, from_dev, empty_dev

pytestmark = pytest.mark.gpu

CASES = [
    # (m, n, k, ta, tb)
    (1, 1, 1, 0, 0),
    (17, 33, 5, 0, 0),
    (128, 128, 64, 0, 0),
    (300, 200, 100, 1, 0),
    (250, 130, 70, 0, 1),
    (99, 77, 300, 1, 1),
    (64, 64, 5000, 1, 0),      # split-K: b x b = W^T P
    (64, 64, 448, 1, 0),       # split-K from K = 256 for few tiles: E = D_y^T D_z
    (130, 70, 260, 0, 0),      # the same with ragged tiles and a K tail
    (2000, 64, 2000, 0, 0),    # tall-skinny: P = A22 W
    (64, 1500, 3000, 1, 0),    # Y^T M in sbr_back
    (1024, 1024, 1024, 0, 0),
]


@pytest.mark.parametrize("m,n,k,ta,tb", CASES)
@pytest.mark.parametrize("prec", ["d", "s"])
def test_gemm(torch_dev, m, n, k, ta, tb, prec):
    torch = torch_dev
    rng = np.random.default_rng(m * 7 + n * 13 + k)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    alpha, beta = 1.25, -0.5
    dt = torch.float64 if prec == "d" else torch.float32
    dA, dB, dC = to_dev(A, dt), to_dev(B, dt), to_dev(C0, dt)
    cv.gemm(ta, tb, m, n, k, alpha, dA, A.shape[0], dB, B.shape[0], beta, dC, m)
    torch.cuda.synchronize()
    ref = alpha * ((A.T if ta else A) @ (B.T if tb else B)) + beta * C0
    got = from_dev(dC)
    scale = np.abs(A).max() * np.abs(B).max() * k + np.abs(C0).max()
    tol = (1e-14 if prec == "d" else 2e-6) * max(1.0, np.sqrt(k)) * 4
    assert np.abs(got - ref).max() / scale < tol


def test_gemm_beta_zero_ignores_nan(torch_dev):
    torch = torch_dev
    A = np.ones((40, 30))
    B = np.ones((30, 20))
    C = np.full((40, 20), np.nan)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(C)
    cv.gemm(0, 0, 40, 20, 30, 1.0, dA, 40, dB, 30, 0.0, dC, 40)
    torch.cuda.synchronize()
    assert np.allclose(from_dev(dC), 30.0)


@pytest.mark.parametrize("m,n", [(5, 3), (200, 64), (1000, 64), (3000, 64), (513, 200), (4096, 64),
                                 (8190, 64), (2050, 48)])
def test_symm_lower(torch_dev, m, n):
    torch = torch_dev
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, m))
    A_lower_only = np.tril(A) + np.triu(np.full((m, m), 1e300), 1)  # upper must not be read
    S = np.tril(A) + np.tril(A, -1).T
    B = rng.standard_normal((m, n))
    dA, dB, dC = to_dev(A_lower_only), to_dev(B), empty_dev(m, n, torch.float64)
    cv.symm_lower(m, n, 1.0, dA, m, dB, m, 0.0, dC, m)
    torch.cuda.synchronize()
    ref = S @ B
    assert np.abs(from_dev(dC) - ref).max() / (np.abs(ref).max() + 1) < 1e-13


@pytest.mark.parametrize("n,k", [(7, 16), (50, 20), (130, 32), (700, 64), (1500, 512), (2100, 96)])
def test_syr2k_lower(torch_dev, n, k):
    torch = torch_dev
    rng = np.random.default_rng(n + k)
    X = rng.standard_normal((n, k))
    Y = rng.standard_normal((n, k))
    C0 = rng.standard_normal((n, n))
    dX, dY, dC = to_dev(X), to_dev(Y), to_dev(C0)
    cv.syr2k_lower(n, k, -1.0, dX, n, dY, n, 1.0, dC, n)
    torch.cuda.synchronize()
    got = from_dev(dC)
    ref = C0 - (X @ Y.T + Y @ X.T)
    low = np.tril_indices(n)
    up = np.triu_indices(n, 1)
    assert np.abs(got[low] - ref[low]).max() / (np.abs(ref).max()) < 1e-13
    assert np.array_equal(got[up], C0[up]), "upper triangle must be untouched"


@pytest.mark.parametrize("m,n,k,ta,tb,pad", [
    (1000, 700, 999, 0, 0, 0),     # block tiles of 128 x 128: interior DMA tiles, edge tiles, K tail
    (520, 390, 700, 1, 0, 0),
    (2048, 1536, 512, 0, 1, 0),
    (700, 1100, 300, 1, 1, 0),
    (1024, 1024, 1024, 1, 0, 1),   # odd leading dimensions: the per-element path everywhere
    (777, 1024, 640, 0, 0, 2),     # even padding: aligned again
])
def test_gemm_large_tiles(torch_dev, m, n, k, ta, tb, pad):
    """FP64 products large enough for the 128 x 128 block tiles (the LDS-DMA path),
    with edges, K tails and leading dimensions larger than the matrices (odd ones
    force the per-element loads)."""
    torch = torch_dev
    rng = np.random.default_rng(m + 3 * n + 7 * k + pad)
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C0 = rng.standard_normal((m, n))
    lda, ldb, ldc = A.shape[0] + pad, B.shape[0] + pad, m + pad
    def padded(X, ld):
        Y = np.zeros((ld, X.shape[1]))
        Y[: X.shape[0]] = X
        return to_dev(Y)
    dA, dB, dC = padded(A, lda), padded(B, ldb), padded(C0, ldc)
    alpha, beta = -0.75, 1.5
    cv.gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
    torch.cuda.synchronize()
    ref = alpha * ((A.T if ta else A) @ (B.T if tb else B)) + beta * C0
    got = from_dev(dC)[:m]
    scale = np.abs(A).max() * np.abs(B).max() * k + np.abs(C0).max()
    assert np.abs(got - ref).max() / scale < 4e-14 * np.sqrt(k)
    if pad:
        assert np.all(from_dev(dC)[m:] == 0), "rows beyond m of C must be untouched"
