"""Eigenvalues-only path: cuev_{s,d}symm_eig_values[_ex] (DBBR -> bulge chasing -> Sturm
multisection on the GPU) and the tridiagonal stage cuev_{s,d}tridi_values.

The reference has no such path (it always computes every eigenpair,
src/cuda/tridi.cu:12-13), so the pins are LAPACK and closed forms:
  - the tridiagonal stage against LAPACK (scipy eigvalsh_tridiagonal, dstebz/dsterf)
    on random, Wilkinson, glued-Wilkinson, clustered, graded, zero and diagonal
    matrices, and against the closed form a + 2b cos(k pi / (n + 1)) of the Toeplitz
    tridiagonal matrix;
  - the full path against dsyevd within 1e-12 ||A||_2 on the n = 1024 configuration
    and against the exact spectrum of q16k at n = 16384;
  - index ranges equal the matching slice of the full spectrum.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import cuev_amd as cv
from helpers import to_dev

pytestmark = pytest.mark.gpu


def tri_case(kind, n, rng):
    if kind == "random":
        return rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
    if kind == "wilkinson":
        m = (n - 1) / 2.0
        return np.abs(np.arange(n) - m), np.ones(max(n - 1, 0))
    if kind == "glued":
        k = 21
        d = np.tile(np.abs(np.arange(k) - 10.0), n // k + 1)[:n]
        e = np.ones(max(n - 1, 0))
        e[k - 1::k] = 1e-12
        return d, e[: max(n - 1, 0)]
    if kind == "cluster":
        d = 1.0 + 1e-10 * rng.standard_normal(n)
        return d, 1e-11 * rng.standard_normal(max(n - 1, 0))
    if kind == "graded":
        return 10.0 ** (-np.arange(n) * (12.0 / max(n, 1))), 10.0 ** (-np.arange(max(n - 1, 0)) * (12.0 / max(n, 1)) - 0.5)
    if kind == "zero":
        return np.zeros(n), np.zeros(max(n - 1, 0))
    if kind == "diagonal":
        return rng.standard_normal(n), np.zeros(max(n - 1, 0))
    raise ValueError(kind)


def tridi_gpu(torch, d, e, il=0, iu=None, dtype=None):
    dt = dtype or torch.float64
    dd = torch.tensor(d, dtype=dt, device="cuda")
    ee = torch.tensor(np.append(e, 0.0), dtype=dt, device="cuda")
    w = cv.tridi_values(dd, ee, il, iu)
    torch.cuda.synchronize()
    return w.double().cpu().numpy()


@pytest.mark.parametrize("kind", ["random", "wilkinson", "glued", "cluster", "graded", "zero", "diagonal"])
@pytest.mark.parametrize("n", [1, 2, 3, 100, 1000, 4097])
def test_tridi_values_vs_lapack(torch_dev, kind, n):
    rng = np.random.default_rng(n)
    d, e = tri_case(kind, n, rng)
    ref = sl.eigvalsh_tridiagonal(d, e) if n > 1 else d.copy()
    w = tridi_gpu(torch_dev, d, e)
    tn = max(np.abs(ref).max(), 1e-300) if n else 1.0
    assert np.all(np.diff(w) >= 0)
    # bound: a few eps ||T|| relative, or the pivot floor pivmin (~2e-308) for a zero matrix
    scale = max(tn, np.abs(d).max() + 2 * (np.abs(e).max() if n > 1 else 0))
    assert np.abs(w - ref).max() <= 1e-13 * scale + 1e-305, kind


@pytest.mark.parametrize("n,a,b", [(7, 0.0, 1.0), (500, 2.0, -0.5), (3000, -1.0, 3.0)])
def test_tridi_values_toeplitz_closed_form(torch_dev, n, a, b):
    k = np.arange(1, n + 1)
    exact = np.sort(a + 2.0 * b * np.cos(k * np.pi / (n + 1)))
    w = tridi_gpu(torch_dev, np.full(n, a), np.full(n - 1, b))
    assert np.abs(w - exact).max() <= 4e-15 * (abs(a) + 2 * abs(b)) * 8


def test_tridi_values_ranges_and_float(torch_dev):
    torch = torch_dev
    rng = np.random.default_rng(3)
    n = 777
    d, e = tri_case("random", n, rng)
    full = tridi_gpu(torch, d, e)
    for il, iu in ((0, 0), (n - 1, n - 1), (100, 400), (0, n - 1)):
        part = tridi_gpu(torch, d, e, il, iu)
        assert np.array_equal(part, full[il:iu + 1]), (il, iu)
    ref = sl.eigvalsh_tridiagonal(d, e)
    w32 = tridi_gpu(torch, d.astype(np.float32), e.astype(np.float32), dtype=torch.float32)
    ref32 = sl.eigvalsh_tridiagonal(d.astype(np.float32).astype(np.float64), e.astype(np.float32).astype(np.float64))
    assert np.abs(w32 - ref32).max() <= 1e-6 * np.abs(ref).max()


def test_tridi_values_invalid(torch_dev):
    torch = torch_dev
    L = cv.lib()
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.cuev_dtridi_values(p, p, 0, 0, 0, p, s) == 1
    assert L.cuev_dtridi_values(p, p, 4, 2, 1, p, s) == 1
    assert L.cuev_dtridi_values(p, p, 4, 0, 4, p, s) == 1


def test_values_config_n1024(torch_dev):
    """n = 1024, A = (G + G^T)/2 from seed 0, b = 32, k = 256: all eigenvalues against
    dsyevd within 1e-12 ||A||_2, and index ranges equal slices of the full set."""
    torch = torch_dev
    rng = np.random.default_rng(0)
    G = rng.standard_normal((1024, 1024))
    A = (G + G.T) / 2
    ref = sl.eigh(A, eigvals_only=True, driver="evd")
    lam, t = cv.symm_eig_values(torch.from_numpy(A).cuda(), bandwidth=32, block=256, timer=True)
    torch.cuda.synchronize()
    lam = lam.cpu().numpy()
    assert np.abs(lam - ref).max() <= 1e-12 * np.abs(ref).max()
    assert t["dbbr_ms"] > 0 and t["bc_ms"] > 0 and t["total_ms"] > 0
    part = cv.symm_eig_values(torch.from_numpy(A).cuda(), 10, 99, bandwidth=32, block=256).cpu().numpy()
    assert np.abs(part - ref[10:100]).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n,b,k", [(1, 32, 512), (2, 32, 512), (5, 16, 16), (300, 32, 96), (2500, 64, 512)])
def test_values_sizes(torch_dev, n, b, k):
    torch = torch_dev
    rng = np.random.default_rng(n + 7)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    ref = sl.eigh(A, eigvals_only=True, driver="evd")
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        lam = cv.symm_eig_values(torch.from_numpy(A).to(dt).cuda(), bandwidth=b, block=k)
        torch.cuda.synchronize()
        assert np.abs(lam.double().cpu().numpy() - ref).max() <= tol * max(np.abs(ref).max(), 1e-300)


def test_values_lower_triangle_and_errors(torch_dev):
    torch = torch_dev
    n = 200
    rng = np.random.default_rng(11)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    Ag = np.tril(A) + np.triu(np.full((n, n), 5e5), 1)
    lam = cv.symm_eig_values(torch.from_numpy(Ag).cuda()).cpu().numpy()
    ref = sl.eigh(A, eigvals_only=True)
    assert np.abs(lam - ref).max() <= 1e-12 * np.abs(ref).max()
    An = np.eye(50)
    An[10, 3] = np.nan
    with pytest.raises(cv.CuevError) as e:
        cv.symm_eig_values(torch.from_numpy(An).cuda())
    assert e.value.status == 7
    with pytest.raises(cv.CuevError) as e:
        cv.symm_eig_values(torch.from_numpy(A).cuda(), 5, 3)
    assert e.value.status == 1


def test_values_q16k_exact(torch_dev):
    """n = 16384, A = Q diag(lam) Q^T with known lam: the eigenvalue-only path within
    1e-12 ||A||_2 of the exact spectrum."""
    torch = torch_dev
    from cuev_amd import workloads
    A0, lam_exact = workloads.make("q16k", n=16384, device="cuda")
    lam, t = cv.symm_eig_values(A0, bandwidth=64, block=512, timer=True)
    torch.cuda.synchronize()
    ref = torch.sort(lam_exact).values
    err = (lam.double() - ref).abs().max().item() / ref.abs().max().item()
    print(f"\n[values q16k] {t['total_ms']:.1f} ms (dbbr {t['dbbr_ms']:.1f}, chase {t['bc_ms']:.1f}, "
          f"multisection {t['dc_ms']:.1f}) err {err:.2e}", flush=True)
    assert err <= 1e-12
