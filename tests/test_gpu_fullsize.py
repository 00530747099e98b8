"""Full-size configurations, checked on the GPU in FP64 where the CPU oracle and host
LAPACK would take too long.

  g49k      n = 49152, A = (G + G^T)/2, G iid N(0, 1/n), seed 2, b = 64, k = 512.
            More rows than the reference's int32 offsets allow (src/cuda/dbbr.cu:162-164);
            bc_back runs in three column chunks, and the device D&C holds every merge
            of more than 4096 rows until the back-transform beside it is done.
  mixed16k  n = 16384, A = Q diag(lam) Q^T, lam geometric from 1 to 1e-6 with random
            signs and 10% of the values in a cluster of width 1e-9: the FP32 stages
            plus FP64 refinement, against the exact eigenvalues.

Size-independent checks: residual ||AV - V diag(lam)||_F and orthogonality
||V^T V - I||_F accumulated over column blocks of V (both normalisations: / (n ||A||_F)
and / ||A||_F), ascending order, and for g49k the trace and Frobenius-norm identities
(sum lam_i = tr A, sum lam_i^2 = ||A||_F^2) and the semicircle edge +-sqrt(2).
"""
import math

import pytest

import cuev_amd as cv
from cuev_amd import workloads

pytestmark = pytest.mark.gpu


def _free(torch):
    cv.cache_release()
    torch.cuda.empty_cache()


def test_g49k_n49152(torch_dev):
    torch = torch_dev
    _free(torch)
    n = 49152
    print(f"\n[g49k] generating n={n}", flush=True)
    A0, _ = workloads.make("g49k", n=n, device="cuda")
    tr = torch.trace(A0).item()
    fro = torch.linalg.matrix_norm(A0).item()
    work = torch.empty_like(A0)
    work.mT.copy_(A0)  # A0 is exactly symmetric: the column-major buffer holds A0
    print("[g49k] solving", flush=True)
    lam, V, t = cv.symm_eig_solve(work, overwrite=True, bandwidth=64, block=512, timer=True)
    torch.cuda.synchronize()
    del work
    print(f"[g49k] solved in {t['total_ms'] / 1e3:.1f} s; checking", flush=True)
    assert bool((lam[1:] >= lam[:-1]).all()), "eigenvalues must be ascending"
    lam64 = lam.double()
    a2 = lam64.abs().max().item()
    assert abs(lam64.sum().item() - tr) <= 1e-13 * n * a2, ("trace", lam64.sum().item(), tr)
    assert abs(math.sqrt((lam64 ** 2).sum().item()) - fro) <= 1e-12 * fro, "Frobenius norm"
    assert abs(lam64[-1].item() - math.sqrt(2.0)) < 0.02 and abs(lam64[0].item() + math.sqrt(2.0)) < 0.02
    res, orth = workloads.residual_orth(A0, V, lam64)
    print(f"[g49k] residual {res:.2e} (x n: {res * n:.2e}) orthogonality {orth:.2e}", flush=True)
    assert res <= 1e-13 and orth <= 1e-13, (res, orth)
    assert res * n <= 1e-13, ("bench_solver-form residual", res * n)
    del A0, V, lam, lam64
    _free(torch)


def test_mixed16k_full(torch_dev):
    torch = torch_dev
    _free(torch)
    n = 16384
    print(f"\n[mixed16k] generating n={n}", flush=True)
    A0, lam_exact = workloads.make("mixed16k", n=n, device="cuda")
    lam, V, t = cv.symm_eig_solve(A0, precision="mixed", timer=True)
    torch.cuda.synchronize()
    print(f"[mixed16k] solved in {t['total_ms'] / 1e3:.2f} s, {t['refine_sweeps']} sweeps, "
          f"{t['n_cluster_cols']} clustered columns", flush=True)
    ref = torch.sort(lam_exact.double()).values
    err = (lam.double() - ref).abs().max().item() / ref.abs().max().item()
    res, orth = workloads.residual_orth(A0, V, lam)
    print(f"[mixed16k] eig err {err:.2e} residual {res:.2e} (x n: {res * n:.2e}) orth {orth:.2e}", flush=True)
    assert err <= 1e-12, err
    assert res <= 1e-13 and orth <= 1e-13, (res, orth)
    assert res * n <= 1e-13, ("bench_solver-form residual", res * n)
    del A0, V
    _free(torch)


def test_mixed_not_converged_status(torch_dev):
    """The refinement stops at its sweep limit; above refine_tol it reports
    NOT_CONVERGED (and still writes its last iterate) instead of success."""
    torch = torch_dev
    n = 1024
    Q = workloads.orthogonal_q(torch, n, 1, "cuda", torch.float64)
    lam_x = workloads.lam_geometric_cluster(torch, n, 103, "cuda")
    A = workloads.sandwich(torch, Q, lam_x)
    with pytest.raises(cv.CuevError) as e:
        cv.symm_eig_solve(A, precision="mixed", refine_sweeps=1, refine_tol=1e-300)
    assert e.value.status == 9 and e.value.name == "NOT_CONVERGED"
    # a reachable tolerance with the default sweep limit converges
    lam, V = cv.symm_eig_solve(A, precision="mixed")
    torch.cuda.synchronize()
    ref = torch.sort(lam_x).values
    assert (lam.double() - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
