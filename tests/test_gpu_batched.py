"""Batched small eigenproblems: cuev_{s,d}symm_eig_batched (one workgroup per matrix,
cyclic Jacobi), against LAPACK dsyevd per matrix.

No reference counterpart (the reference solves one matrix per call,
include/cuev.h:31-33); the pins are LAPACK per matrix (eigenvalues within
1e-12 ||A||_2, BASELINE residual and orthogonality forms), closed forms (a diagonal
matrix returns itself sorted; a known Q diag(lam) Q^T returns lam), and the lower
triangle rule (garbage above the diagonal is ignored). Sizes cover both storage
variants (LDS-resident for n <= 128 in FP64, in the caller's buffer above), odd n
(the virtual padding index), n = 1 and the limit n = 256.
"""
import ctypes
import time

import numpy as np
import pytest
import scipy.linalg as sl

import cuev_amd as cv

pytestmark = pytest.mark.gpu


def _batch(rng, batch, n):
    G = rng.standard_normal((batch, n, n))
    return (G + np.transpose(G, (0, 2, 1))) / 2


def _check(A, lam, V, ev_tol=1e-12, res_tol=1e-13, orth_tol=1e-13):
    for b in range(A.shape[0]):
        ref = sl.eigh(A[b], eigvals_only=True, driver="evd")
        n = A.shape[1]
        a2 = max(np.abs(ref).max(), 1e-300)
        assert np.all(np.diff(lam[b]) >= 0), b
        assert np.abs(lam[b] - ref).max() <= ev_tol * a2, (b, np.abs(lam[b] - ref).max() / a2)
        afro = max(np.linalg.norm(A[b]), 1e-300)
        res = np.linalg.norm(A[b] @ V[b] - V[b] * lam[b]) / (n * afro)
        orth = np.linalg.norm(V[b].T @ V[b] - np.eye(n)) / n
        assert res <= res_tol and orth <= orth_tol, (b, res, orth)


@pytest.mark.parametrize("n,batch", [(1, 3), (2, 5), (7, 9), (64, 40), (127, 12), (128, 12), (129, 6),
                                     (200, 4), (256, 3)])
def test_batched_vs_lapack(torch_dev, n, batch):
    torch = torch_dev
    rng = np.random.default_rng(n * 31 + batch)
    A = _batch(rng, batch, n)
    lam, V = cv.symm_eig_batched(torch.from_numpy(A).cuda())
    torch.cuda.synchronize()
    _check(A, lam.cpu().numpy(), V.cpu().numpy())


def test_batched_float32(torch_dev):
    torch = torch_dev
    rng = np.random.default_rng(5)
    A = _batch(rng, 8, 100).astype(np.float32)
    lam, V = cv.symm_eig_batched(torch.from_numpy(A).cuda())
    torch.cuda.synchronize()
    _check(A.astype(np.float64), lam.double().cpu().numpy(), V.double().cpu().numpy(),
           ev_tol=2e-6, res_tol=1e-7, orth_tol=1e-7)


def test_batched_structured(torch_dev):
    """Diagonal (returned sorted, eigenvectors a permutation), known spectrum with a
    repeated eigenvalue, the zero matrix, and garbage above the diagonal."""
    torch = torch_dev
    rng = np.random.default_rng(9)
    n = 96
    d = rng.standard_normal(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam_k = np.sort(np.concatenate([np.full(10, 0.5), rng.uniform(-2, 2, n - 10)]))
    K = (Q * lam_k) @ Q.T
    K = (K + K.T) / 2
    G = _batch(rng, 1, n)[0]
    Ag = np.tril(G) + np.triu(np.full((n, n), 9e4), 1)
    A = np.stack([np.diag(d), K, np.zeros((n, n)), Ag])
    lam, V = cv.symm_eig_batched(torch.from_numpy(A).cuda())
    torch.cuda.synchronize()
    lam, V = lam.cpu().numpy(), V.cpu().numpy()
    assert np.array_equal(lam[0], np.sort(d))
    assert np.abs(lam[1] - lam_k).max() <= 1e-13 * 2
    assert np.all(lam[2] == 0) and np.allclose(np.abs(V[2]).sum(0), 1.0)
    A_sym = A.copy()
    A_sym[3] = G
    _check(A_sym[[0, 1, 3]], lam[[0, 1, 3]], V[[0, 1, 3]])


def test_batched_invalid_and_throughput(torch_dev):
    torch = torch_dev
    L = cv.lib()
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.cuev_dsymm_eig_batched(p, 257, 1, p, p, s, None) == 1
    assert L.cuev_dsymm_eig_batched(p, 0, 1, p, p, s, None) == 1
    assert L.cuev_dsymm_eig_batched(p, 4, 0, p, p, s, None) == 0
    # throughput at the autograd harness's scale: 1024 matrices of order 64 and 256 of 128
    for n, batch in ((64, 1024), (128, 256)):
        g = torch.Generator(device="cuda")
        g.manual_seed(n)
        G = torch.randn((batch, n, n), generator=g, dtype=torch.float64, device="cuda")
        A = (G + G.mT) * 0.5
        cv.symm_eig_batched(A[:4])
        torch.cuda.synchronize()
        t = time.perf_counter()
        lam, V = cv.symm_eig_batched(A)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        Vc = V.contiguous()
        R = A @ Vc - Vc * lam[:, None, :]
        res = (torch.linalg.matrix_norm(R) / (n * torch.linalg.matrix_norm(A))).max().item()
        print(f"\n[batched] n={n} batch={batch}: {dt * 1e3:.2f} ms, {batch / dt:.0f} matrices/s, "
              f"max residual {res:.2e}", flush=True)
        assert res <= 1e-13
