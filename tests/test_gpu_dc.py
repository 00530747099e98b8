"""Device-resident divide and conquer (tridiagonal stage) against host LAPACK.

Acceptance as the reference's D&C test (test/cuda/test_tridi.cu:22-74): residual
||T V - V diag(lam)||_F / ||T||_F and orthogonality ||V^T V - I||_F / sqrt(n), here
at the level LAPACK's own ?stedc reaches (a few units of n*eps), and the eigenvalues
against LAPACK's (dstemr via scipy, relative to ||T||_2). The matrices cover the
deflation paths: random (little deflation), Wilkinson (close pairs: Givens
deflation), glued Wilkinson (tiny couplings: z deflation), constant diagonals
(total deflation), zero off-diagonal blocks, and tiny sizes.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import cuev_amd as cv
from helpers import to_dev, from_dev, empty_dev

pytestmark = pytest.mark.gpu


def tridiag(kind, n, seed=0):
    rng = np.random.default_rng(seed)
    if kind == "random":
        return rng.standard_normal(n), rng.standard_normal(max(n - 1, 0))
    if kind == "wilkinson":
        return np.abs(np.arange(n) - (n - 1) / 2.0), np.ones(max(n - 1, 0))
    if kind == "glued":
        m = 21
        d = np.tile(np.abs(np.arange(m) - 10.0), n // m + 1)[:n]
        e = np.ones(max(n - 1, 0))
        e[m - 1::m] = 1e-9
        return d, e
    if kind == "constant":
        return np.full(n, 2.0), np.zeros(max(n - 1, 0))
    if kind == "blocks":
        e = rng.standard_normal(max(n - 1, 0))
        e[::7] = 0.0
        return rng.standard_normal(n), e
    if kind == "graded":
        return np.logspace(0, -12, n) * rng.choice([-1, 1], n), np.logspace(-1, -13, max(n - 1, 0))
    if kind == "clustered":
        return 1.0 + 1e-13 * rng.standard_normal(n), 1e-11 * rng.standard_normal(max(n - 1, 0))
    raise ValueError(kind)


def run_dc(torch, d, e, dtype, leaf=0):
    n = len(d)
    dt = torch.float64 if dtype == "d" else torch.float32
    e_full = np.concatenate([e, [0.0]])
    with cv.Handle(dt, n, 32, 512) as h:
        h.set_dc("device", leaf)
        dd, de = to_dev(d, dt), to_dev(e_full, dt)
        ev = torch.empty(n, dtype=dt, device="cuda")
        V = empty_dev(n, n, dt)
        h.tridi_dc(dd, de, ev, V)
        torch.cuda.synchronize()
        return from_dev(ev).astype(np.float64), from_dev(V).astype(np.float64)


def check(d, e, lam, V, tol_res, tol_orth, tol_ev):
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    tn = max(np.linalg.norm(T), 1e-300)
    res = np.linalg.norm(T @ V - V * lam) / tn
    orth = np.linalg.norm(V.T @ V - np.eye(n)) / np.sqrt(n)
    ref = sl.eigvalsh_tridiagonal(d, e) if n > 1 else d.copy()
    t2 = max(np.max(np.abs(ref)), 1e-300)
    ev = np.max(np.abs(lam - ref)) / t2
    assert np.all(np.diff(lam) >= 0), "eigenvalues not ascending"
    assert res < tol_res and orth < tol_orth and ev < tol_ev, (res, orth, ev)


@pytest.mark.parametrize("kind", ["random", "wilkinson", "glued", "constant", "blocks", "graded",
                                  "clustered"])
@pytest.mark.parametrize("n", [300, 2049])
def test_dc_device_kinds(torch_dev, kind, n):
    d, e = tridiag(kind, n, seed=n)
    lam, V = run_dc(torch_dev, d, e, "d")
    check(d, e, lam, V, 1e-13, 1e-13, 1e-12)


@pytest.mark.parametrize("n,leaf", [(1, 0), (2, 0), (3, 2), (17, 4), (64, 0), (65, 0), (129, 16),
                                    (1000, 256), (4097, 64)])
def test_dc_device_sizes_and_leaves(torch_dev, n, leaf):
    d, e = tridiag("random", n, seed=7 + n)
    lam, V = run_dc(torch_dev, d, e, "d", leaf)
    check(d, e, lam, V, 1e-13, 1e-13, 1e-12)


def test_dc_device_zero_matrix(torch_dev):
    n = 100
    d, e = np.zeros(n), np.zeros(n - 1)
    lam, V = run_dc(torch_dev, d, e, "d")
    assert np.all(lam == 0)
    assert np.array_equal(V, np.eye(n))


def test_dc_device_float(torch_dev):
    d, e = tridiag("random", 1500, seed=3)
    lam, V = run_dc(torch_dev, d, e, "s")
    check(d, e, lam, V, 1e-5, 1e-5, 1e-5)


def test_dc_device_matches_host_eigenvectors(torch_dev):
    """Same T through the host ?stedc stage: eigenvectors agree up to sign."""
    torch = torch_dev
    n = 1200
    d, e = tridiag("random", n, seed=11)
    lam_d, V_d = run_dc(torch, d, e, "d")
    e_full = np.concatenate([e, [0.0]])
    with cv.Handle(torch.float64, n, 32, 512) as h:
        dd, de = to_dev(d, torch.float64), to_dev(e_full, torch.float64)
        ev = torch.empty(n, dtype=torch.float64, device="cuda")
        V = empty_dev(n, n, torch.float64)
        h.tridi_dc(dd, de, ev, V)
        torch.cuda.synchronize()
        lam_h, V_h = from_dev(ev), from_dev(V)
    assert np.max(np.abs(lam_d - lam_h)) < 1e-13 * np.max(np.abs(lam_h))
    gap = np.min(np.diff(lam_h))
    if gap > 1e-8:
        dots = np.abs(np.sum(V_d * V_h, axis=0))
        assert np.max(np.abs(dots - 1)) < 1e-10


@pytest.mark.parametrize("bt", ["direct", "reordered"])
@pytest.mark.parametrize("n", [1500, 3001])
def test_solve_dc_device_vs_host(torch_dev, bt, n):
    """Full solve: device D&C and host D&C give the same spectrum and FP64 accuracy."""
    torch = torch_dev
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    dA = torch.from_numpy(A).cuda()
    out = {}
    for mode in ("device", "host"):
        lam, V, t = cv.symm_eig_solve(dA, bt_mode=bt, dc_mode=mode, timer=True)
        torch.cuda.synchronize()
        out[mode] = (lam.cpu().numpy(), V.cpu().numpy(), t)
    lam_ref = sl.eigh(A, eigvals_only=True, driver="evd")
    for mode, (lam, V, t) in out.items():
        a2 = np.max(np.abs(lam_ref))
        assert np.max(np.abs(lam - lam_ref)) / a2 < 1e-12, mode
        res = np.linalg.norm(A @ V - V * lam) / (n * np.linalg.norm(A))
        orth = np.linalg.norm(V.T @ V - np.eye(n)) / n
        assert res < 1e-14 and orth < 1e-14, (mode, res, orth, t)
