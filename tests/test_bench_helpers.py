"""bench.py's accuracy check: the column-blocked residual and orthogonality (used at
n = 49152, where the whole-matrix form does not fit beside the solver's workspace) must
equal the whole-matrix formulas."""
import numpy as np
import torch

import bench


def test_residual_orth_blocked_equals_whole():
    rng = np.random.default_rng(3)
    n = 300
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    lam, V = np.linalg.eigh(A)
    # perturb so both measures are far from zero and the blocks matter
    V = V + 1e-6 * rng.standard_normal((n, n))
    lam = lam + 1e-7 * rng.standard_normal(n)
    At, Vt, lt = (torch.from_numpy(x) for x in (A, V, lam))
    res, orth = bench.residual_orth(torch, At, Vt, lt, cols=64)  # 64 does not divide 300
    res_ref = np.linalg.norm(A @ V - V * lam) / (n * np.linalg.norm(A))
    orth_ref = np.linalg.norm(V.T @ V - np.eye(n)) / n
    assert abs(res - res_ref) <= 1e-12 * res_ref
    assert abs(orth - orth_ref) <= 1e-12 * orth_ref


def test_cpu_baseline_same_matrix():
    """The CPU baseline times dsyevd on the matrix it is given and reports its accuracy
    against the exact spectrum when one is known."""
    rng = np.random.default_rng(4)
    n = 120
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(-1, 1, n))
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    out = bench.cpu_baseline_same(A.copy(), lam)
    assert out["value"] > 0 and out["kind"] == "port" and out["cores"] >= 1
    assert f"n={n}" in out["sample"] and "not scaled" in out["sample"]
    assert out["eig_err_vs_exact_over_norm2"] < 1e-13
