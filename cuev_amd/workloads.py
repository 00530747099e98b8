"""Seeded synthetic inputs of the benchmark configurations (no data files, no network).

  n1024   A = (G + G^T)/2, G iid N(0,1), seed 0                         (b=32, k=256)
  q16k    A = Q diag(lam) Q^T, Q = Householder-QR Q factor of a seeded iid
          N(0,1) matrix, lam = n sorted iid U(-1,1)                      (b=64, k=512)
  g49k    A = (G + G^T)/2, G iid N(0, 1/n), seed 2                      (b=64, k=512)
  mixed16k  as q16k with lam geometric from 1 to 1e-6, random signs, 10%
          of the values replaced by a cluster of width 1e-9 around 1e-3
  diff4k  A = Q diag(lam) Q^T with lam equispaced in [-1, 1]

Large matrices are generated on the GPU (torch's seeded Philox generator and
torch.linalg.qr); the same function with device="cpu" gives the host copy used
for the LAPACK baseline at a smaller n. Seeds are fixed here so every run sees
the same matrix.
"""
import math

SEED_G = {"n1024": 0, "q16k": 1, "g49k": 2, "mixed16k": 1, "diff4k": 5}
SEED_LAM = {"q16k": 101, "mixed16k": 103, "diff4k": 105}


def _gen(torch, seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def orthogonal_q(torch, n, seed, device, dtype):
    """Q factor of the Householder QR of a seeded iid N(0,1) n x n matrix."""
    g = _gen(torch, seed, device)
    G = torch.randn((n, n), generator=g, device=device, dtype=torch.float64)
    Q, R = torch.linalg.qr(G)
    del G, R
    return Q.to(dtype)


def lam_uniform(torch, n, seed, device):
    g = _gen(torch, seed, device)
    lam = torch.rand(n, generator=g, device=device, dtype=torch.float64) * 2.0 - 1.0
    return torch.sort(lam).values


def lam_geometric_cluster(torch, n, seed, device, cluster_frac=0.10, width=1e-9, centre=1e-3):
    g = _gen(torch, seed, device)
    mag = torch.logspace(0.0, -6.0, n, device=device, dtype=torch.float64)
    sign = torch.where(torch.rand(n, generator=g, device=device, dtype=torch.float64) < 0.5, -1.0, 1.0)
    lam = mag * sign
    m = int(round(cluster_frac * n))
    idx = torch.randperm(n, generator=g, device=device)[:m]
    lam[idx] = centre + width * (torch.rand(m, generator=g, device=device, dtype=torch.float64) - 0.5)
    return torch.sort(lam).values


def sandwich(torch, Q, lam):
    A = (Q * lam) @ Q.T
    return (A + A.T) * 0.5


def residual_orth(A0, V, lam, cols=2048):
    """(||A V - V diag(lam)||_F / (n ||A||_F), ||V^T V - I||_F / n) in float64 on A0's
    device, accumulated over column blocks of V so that the check needs O(n * cols)
    extra memory (at n = 49152 the solver's cached workspace and three n x n matrices
    already fill most of HBM). Multiply the first by n for the form
    ||A V - V diag(lam)||_F / ||A||_F."""
    import torch
    n = A0.shape[0]
    A0 = A0.double()
    V = V.double()
    lam = lam.double()
    an = torch.linalg.matrix_norm(A0).item()
    an = an if an > 0 else 1.0
    r2 = o2 = 0.0
    for c0 in range(0, n, cols):
        c1 = min(n, c0 + cols)
        Vc = V[:, c0:c1]
        R = A0 @ Vc
        R.sub_(Vc * lam[c0:c1])
        r2 += torch.linalg.matrix_norm(R).item() ** 2
        G = Vc.T @ V
        G[:, c0:c1].diagonal().sub_(1.0)
        o2 += torch.linalg.matrix_norm(G).item() ** 2
        del R, G
    return r2 ** 0.5 / (n * an), o2 ** 0.5 / n


def make(name, n=None, device="cuda", dtype=None):
    """Return (A, lam_exact or None) as torch tensors on `device` (float64 unless dtype)."""
    import torch
    dt = dtype or torch.float64
    if name == "n1024":
        n = n or 1024
        g = _gen(torch, SEED_G[name], device)
        G = torch.randn((n, n), generator=g, device=device, dtype=torch.float64)
        return ((G + G.T) * 0.5).to(dt), None
    if name == "g49k":
        n = n or 49152
        g = _gen(torch, SEED_G[name], device)
        A = torch.randn((n, n), generator=g, device=device, dtype=torch.float64)
        A.mul_(1.0 / math.sqrt(n))
        A.add_(A.T.clone()).mul_(0.5)
        return A.to(dt), None
    if name in ("q16k", "mixed16k", "diff4k"):
        n = n or {"q16k": 16384, "mixed16k": 16384, "diff4k": 4096}[name]
        Q = orthogonal_q(torch, n, SEED_G[name], device, torch.float64)
        if name == "q16k":
            lam = lam_uniform(torch, n, SEED_LAM[name], device)
        elif name == "mixed16k":
            lam = lam_geometric_cluster(torch, n, SEED_LAM[name], device)
        else:
            lam = torch.linspace(-1.0, 1.0, n, device=device, dtype=torch.float64)
        A = sandwich(torch, Q, lam)
        del Q
        return A.to(dt), lam
    raise ValueError(f"unknown workload {name}")
