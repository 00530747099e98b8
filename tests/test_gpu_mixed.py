"""Mixed form: FP32 two-stage reduction + FP64 Ogita-Aishima refinement must reach the
FP64 tolerances (eigenvalues 1e-12 ||A||_2, residual and orthogonality 1e-13)."""
import numpy as np
import pytest
import scipy.linalg as sl

import cuev_amd as cv
from helpers import eig_metrics

pytestmark = pytest.mark.gpu


def run(torch, A, **kw):
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    lam, V, t = cv.symm_eig_solve(dA, precision="mixed", timer=True, **kw)
    torch.cuda.synchronize()
    return lam.cpu().numpy(), V.cpu().numpy(), t


@pytest.mark.parametrize("n", [120, 300, 1024, 2500])  # 120: below the 16-probe check's n >= 256
def test_mixed_random_symmetric(torch_dev, n):
    rng = np.random.default_rng(n + 1)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    lam, V, t = run(torch_dev, A)
    lam_ref = sl.eigh(A, eigvals_only=True, driver="evd")
    ev, res, orth = eig_metrics(A, lam, V, lam_ref)
    assert t["refine_sweeps"] >= 1
    assert ev < 1e-12 and res < 1e-13 and orth < 1e-13, (ev, res, orth, t)


@pytest.mark.parametrize("n", [500, 2048])
def test_mixed_geometric_with_cluster(torch_dev, n):
    """Scaled-down mixed16k: geometric spectrum 1..1e-6, random signs, 10% in a 1e-9 cluster.
    At n=2048 more than 512 columns are in clustered pairs, so the Rayleigh-Ritz step runs
    on the GPU (one FP64 solve of the projected union); at n=500 (fewer than 512 columns)
    it runs as one host dsygvd on the union."""
    torch = torch_dev
    from cuev_amd import workloads
    Q = workloads.orthogonal_q(torch, n, 1, "cuda", torch.float64)
    lam_x = workloads.lam_geometric_cluster(torch, n, 103, "cuda")
    A = workloads.sandwich(torch, Q, lam_x)
    lam, V, t = cv.symm_eig_solve(A.clone(), precision="mixed", timer=True)
    torch.cuda.synchronize()
    An = A.cpu().numpy()
    lam_n, Vn = lam.cpu().numpy(), V.cpu().numpy()
    ev, res, orth = eig_metrics(An, lam_n, Vn, lam_x.cpu().numpy())
    assert ev < 1e-12 and res < 1e-13 and orth < 1e-13, (ev, res, orth, t)
    assert (t["n_cluster_cols"] >= 512) == (n == 2048), t
    # the refinement's last check (here the 16-probe estimate, since n >= 256) agrees
    # with the residual and orthogonality measured here
    print(f"n={n}: refine_resid/res {t['refine_resid'] / res:.3f} refine_orth/orth {t['refine_orth'] / orth:.3f}")
    assert 0.1 < t["refine_resid"] / res < 10 and 0.1 < t["refine_orth"] / orth < 10, (res, orth, t)


def test_mixed_rejected_for_float(torch_dev):
    torch = torch_dev
    A = torch.eye(8, dtype=torch.float32, device="cuda")
    with pytest.raises(cv.CuevError) as e:
        cv.symm_eig_solve(A, precision="mixed")
    assert e.value.status == 8


@pytest.mark.parametrize("n", [700, 1100])
def test_mixed_reads_only_lower(torch_dev, n):
    """Only the lower triangle of A is read: with NaN above the diagonal the mixed form
    (whose refinement fills A's upper triangle from the lower one before its A X
    products) still reaches the FP64 tolerances."""
    rng = np.random.default_rng(n + 7)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    A_low = np.tril(A) + np.triu(np.full((n, n), np.nan), 1)
    # the torch API reads A[i, j] with i >= j of the tensor as given
    lam, V, t = cv.symm_eig_solve(torch_dev.from_numpy(A_low).cuda(), precision="mixed", timer=True)
    torch_dev.cuda.synchronize()
    lam, V = lam.cpu().numpy(), V.cpu().numpy()
    lam_ref = sl.eigh(A, eigvals_only=True, driver="evd")
    ev, res, orth = eig_metrics(A, lam, V, lam_ref)
    assert ev < 1e-12 and res < 1e-13 and orth < 1e-13, (ev, res, orth, t)
