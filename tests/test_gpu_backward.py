"""Eigendecomposition backward (cuev_{s,d}symm_eig_backward) — no reference counterpart.

Pinned by FP64 central finite differences of the loss
    L(A) = sum_i w_i lambda_i + ||V_k^T B||_F^2     (k lowest eigenvectors)
along random symmetric directions (relative 1e-7), by CPU torch.linalg.eigh's
backward (symmetric parts compared), and at n = 4096 by a float64 CPU
evaluation of the same formula on the GPU's eigenvectors.
"""
import numpy as np
import pytest

import cuev_amd as cv
from cuev_amd import autograd as cag

pytestmark = pytest.mark.gpu


def make_A(n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.linspace(-1.0, 1.0, n)
    A = (Q * lam) @ Q.T
    return (A + A.T) / 2


def loss_np(A, w, B, k):
    lam, V = np.linalg.eigh(A)
    return float(w @ lam + np.sum((V[:, :k].T @ B) ** 2))


def cotangents(torch, V, w, B, k):
    """lambar = w, Vbar[:, :k] = 2 B B^T V_k, Vbar[:, k:] = 0."""
    Vb = torch.zeros_like(V)
    Bt = torch.as_tensor(B, device=V.device, dtype=V.dtype)
    Vb[:, :k] = 2.0 * Bt @ (Bt.T @ V[:, :k])
    return torch.as_tensor(w, device=V.device, dtype=V.dtype), Vb


def test_backward_finite_differences_n256(torch_dev):
    torch = torch_dev
    n, k, ndir = 256, 128, 64
    rng = np.random.default_rng(17)
    A = make_A(n, 3)
    w = rng.standard_normal(n)
    B = rng.standard_normal((n, 64))
    dA = torch.from_numpy(A).cuda()
    lam, V = cv.symm_eig_solve(dA)
    lamb, Vb = cotangents(torch, V, w, B, k)
    Abar, masked = cv.eig_backward(V, lam, Vb, lamb)
    torch.cuda.synchronize()
    G = Abar.cpu().numpy()
    assert masked == 0
    assert np.abs(G - G.T).max() < 1e-12 * np.abs(G).max()
    # Richardson-extrapolated central differences: O(h^4) truncation, so h can be
    # large enough that LAPACK's eigenvector rounding (~eps/gap) stays negligible
    h = 1e-3
    worst = 0.0
    for t in range(ndir):
        D = rng.standard_normal((n, n))
        D = (D + D.T) / 2
        D /= np.linalg.norm(D)
        c1 = (loss_np(A + h * D, w, B, k) - loss_np(A - h * D, w, B, k)) / (2 * h)
        c2 = (loss_np(A + 2 * h * D, w, B, k) - loss_np(A - 2 * h * D, w, B, k)) / (4 * h)
        fd = (4.0 * c1 - c2) / 3.0
        an = float(np.sum(G * D))
        worst = max(worst, abs(fd - an) / max(abs(fd), 1e-3))
    assert worst < 1e-7, worst


def test_backward_matches_torch_cpu_eigh(torch_dev):
    torch = torch_dev
    n, k = 192, 40
    rng = np.random.default_rng(5)
    A = make_A(n, 8)
    w = rng.standard_normal(n)
    B = rng.standard_normal((n, 64))
    At = torch.tensor(A, dtype=torch.float64, requires_grad=True)
    lam_t, V_t = torch.linalg.eigh(At)
    Bt = torch.tensor(B, dtype=torch.float64)
    L = (torch.tensor(w) * lam_t).sum() + ((V_t[:, :k].T @ Bt) ** 2).sum()
    L.backward()
    Gt = At.grad.numpy()
    Gt = (Gt + Gt.T) / 2
    dA = torch.from_numpy(A).cuda()
    lam, V, = cv.symm_eig_solve(dA)
    lamb, Vb = cotangents(torch, V, w, B, k)
    Abar, _ = cv.eig_backward(V, lam, Vb, lamb)
    G = Abar.cpu().numpy()
    assert np.abs(G - Gt).max() / np.abs(Gt).max() < 1e-10


@pytest.mark.parametrize("prec", ["d", "s"])
def test_forward_backward_n4096(torch_dev, prec):
    """n=4096 forward + backward, equispaced spectrum (min gap ~4.9e-4), k = 128."""
    torch = torch_dev
    dt = torch.float64 if prec == "d" else torch.float32
    from cuev_amd import workloads
    A, lam_exact = workloads.make("diff4k", n=4096, device="cuda")
    n, k = 4096, 128
    rng = np.random.default_rng(21)
    w = rng.standard_normal(n)
    B = rng.standard_normal((n, 64))
    Ad = A.to(dt)
    lam, V = cv.symm_eig_solve(Ad)
    lamb, Vb = cotangents(torch, V, w, B, k)
    Abar, masked = cv.eig_backward(V, lam, Vb, lamb, gap_tol=1e-6 if prec == "d" else 1e-5)
    torch.cuda.synchronize()
    assert masked == 0
    # same formula evaluated in float64 on the CPU with the GPU's (lambda, V)
    Vn = V.double().cpu().numpy()
    ln = lam.double().cpu().numpy()
    Vbn = Vb.double().cpu().numpy()
    C = Vn.T @ Vbn
    K = (C - C.T) / 2
    gap = ln[None, :] - ln[:, None]
    np.fill_diagonal(gap, 1.0)
    F = 1.0 / gap
    np.fill_diagonal(F, 0.0)
    M = np.diag(w) + F * K
    ref = Vn @ M @ Vn.T
    got = Abar.double().cpu().numpy()
    tol = 1e-10 if prec == "d" else 2e-3
    assert np.abs(got - ref).max() / np.abs(ref).max() < tol
    assert np.abs(lam.double().cpu().numpy() - lam_exact.cpu().numpy()).max() < (1e-12 if prec == "d" else 1e-5)


def test_degenerate_pairs_masked_and_counted(torch_dev):
    torch = torch_dev
    n = 120
    rng = np.random.default_rng(2)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.full(10, 0.5), np.linspace(1.0, 2.0, n - 10)])
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    lam_d, V = cv.symm_eig_solve(torch.from_numpy(A).cuda())
    Vb = torch.randn((n, n), dtype=torch.float64, device="cuda")
    Abar, masked = cv.eig_backward(V, lam_d, Vb, None, gap_tol=1e-8)
    assert masked == 10 * 9 // 2
    assert torch.isfinite(Abar).all()


def test_autograd_function(torch_dev):
    torch = torch_dev
    n, k = 96, 20
    A = torch.from_numpy(make_A(n, 4)).cuda().requires_grad_(True)
    rng = np.random.default_rng(9)
    w = torch.from_numpy(rng.standard_normal(n)).cuda()
    B = torch.from_numpy(rng.standard_normal((n, 16))).cuda()
    lam, V = cag.eigh(A)
    L = (w * lam).sum() + ((V[:, :k].T @ B) ** 2).sum()
    L.backward()
    At = A.detach().cpu().clone().requires_grad_(True)
    lt, Vt = torch.linalg.eigh(At)
    Lt = (w.cpu() * lt).sum() + ((Vt[:, :k].T @ B.cpu()) ** 2).sum()
    Lt.backward()
    Gt = (At.grad + At.grad.T) / 2
    assert (A.grad.cpu() - Gt).abs().max() / Gt.abs().max() < 1e-10
