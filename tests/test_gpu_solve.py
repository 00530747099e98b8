"""End-to-end parity of cuev_{s,d}symm_eig_solve with host LAPACK dsyevd.

Tolerances are the north-star ones: eigenvalues within 1e-12 ||A||_2 of dsyevd,
||AV - V diag(lambda)||_F / (n ||A||_F) <= 1e-13 and ||V^T V - I||_F / n <= 1e-13.
The first form divides by n, so at large n it admits residuals far above what the
solver reaches; from n = 2048 on, fp64 solves must also meet the reference
bench_solver form ||AV - V diag(lambda)||_F / ||A||_F <= 1e-13
(bench/bench_solver.cpp:112-131).
Inputs: the bench_solver fill (bench/bench_solver.cpp:53-58), the reference's
glibc test fill, the n=1024 configuration A = (G + G^T)/2 with seed 0, edge
sizes (n <= 2, n not a multiple of b), structured matrices (zero, diagonal,
already tridiagonal, repeated eigenvalues) and every bandwidth / back-transform
order.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

import cuev_amd as cv
from helpers import to_dev, eig_metrics

pytestmark = pytest.mark.gpu


def solve_np(torch, A, prec="d", **kw):
    dt = torch.float64 if prec == "d" else torch.float32
    dA = to_dev(A, dt).mT  # the tensor A itself, stored column-major (solved in place)
    lam, V = cv.symm_eig_solve(dA, overwrite=True, **kw)
    torch.cuda.synchronize()
    return lam.double().cpu().numpy(), V.double().cpu().numpy()


def check(A, lam, V, ev_tol=1e-12, res_tol=1e-13, orth_tol=1e-13):
    lam_ref = sl.eigh(A, eigvals_only=True, driver="evd")
    ev_err, res, orth = eig_metrics(A, lam, V, lam_ref)
    n = A.shape[0]
    assert np.all(np.diff(lam) >= 0), "eigenvalues must be ascending"
    assert ev_err <= ev_tol, f"eigenvalue error {ev_err:.2e}"
    assert res <= res_tol, f"residual {res:.2e}"
    assert orth <= orth_tol, f"orthogonality {orth:.2e}"
    if n >= 2048 and res_tol <= 1e-13:
        assert res * n <= 1e-13, f"bench_solver-form residual {res * n:.2e}"
    return ev_err, res, orth


def test_config_n1024_seed0(torch_dev):
    """n=1024, A = (G + G^T)/2, G iid N(0,1) from seed 0, b=32, k=256."""
    rng = np.random.default_rng(0)
    G = rng.standard_normal((1024, 1024))
    A = (G + G.T) / 2
    lam, V = solve_np(torch_dev, A, bandwidth=32, block=256)
    check(A, lam, V)


@pytest.mark.parametrize("n", [500, 1000])
def test_bench_solver_fill(torch_dev, n):
    from oracle import oracle as O
    A = O.fill_glibc_symmetric(n, 1234)
    lam, V = solve_np(torch_dev, A)
    check(A, lam, V)


@pytest.mark.parametrize("n,b,k,bt", [(1, 32, 512, "auto"), (2, 32, 512, "auto"), (3, 16, 16, "auto"),
                                       (5, 32, 512, "direct"), (33, 32, 64, "direct"),
                                       (65, 64, 128, "reordered"), (130, 16, 32, "direct"),
                                       (257, 32, 96, "reordered"), (700, 64, 512, "direct"),
                                       (700, 64, 512, "reordered"), (2500, 64, 512, "auto"),
                                       (3001, 32, 512, "reordered")])
def test_sizes_and_orders(torch_dev, n, b, k, bt):
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    lam, V = solve_np(torch_dev, A, bandwidth=b, block=k, bt_mode=bt)
    check(A, lam, V)


def test_float32(torch_dev):
    from oracle import oracle as O
    A = O.sym_from_lower(O.fill_glibc_matrix(384, 384, 7))
    lam, V = solve_np(torch_dev, A, prec="s", bandwidth=32, block=96)
    check(A, lam, V, ev_tol=1e-5, res_tol=1e-7, orth_tol=1e-7)


@pytest.mark.parametrize("kind", ["zero", "diagonal", "tridiagonal", "repeated", "rank1"])
def test_structured(torch_dev, kind):
    n = 300
    rng = np.random.default_rng(5)
    if kind == "zero":
        A = np.zeros((n, n))
    elif kind == "diagonal":
        A = np.diag(rng.standard_normal(n))
    elif kind == "tridiagonal":
        A = np.diag(rng.standard_normal(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    elif kind == "repeated":
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = np.repeat([-1.0, 0.5, 2.0], n // 3)
        A = (Q * lam) @ Q.T
        A = (A + A.T) / 2
    else:
        v = rng.standard_normal(n)
        A = np.outer(v, v)
    lam, V = solve_np(torch_dev, A, bandwidth=32, block=128)
    check(A, lam, V)


def test_lower_triangle_only(torch_dev):
    """Only the lower triangle is read: garbage in the upper triangle must not matter."""
    n = 400
    rng = np.random.default_rng(9)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    Ag = np.tril(A) + np.triu(np.full((n, n), 7.7e5), 1)
    lam, V = solve_np(torch_dev, Ag)
    check(A, lam, V)


def test_lower_triangle_only_torch_layout(torch_dev):
    """Through the torch API with an ordinary C-contiguous tensor: like
    torch.linalg.eigh, only A[i, j] with i >= j is read, so garbage above the
    diagonal must not matter, in both the copying and the in-place call."""
    torch = torch_dev
    n = 300
    rng = np.random.default_rng(19)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    Ag = np.tril(A) + np.triu(np.full((n, n), -3.3e4), 1)
    t = torch.from_numpy(Ag).cuda()
    assert t.is_contiguous()
    lam, V = cv.symm_eig_solve(t)
    torch.cuda.synchronize()
    assert torch.equal(t.cpu(), torch.from_numpy(Ag)), "the input must not change without overwrite"
    check(A, lam.cpu().numpy(), V.cpu().numpy())
    lam2, V2 = cv.symm_eig_solve(t, overwrite=True)
    torch.cuda.synchronize()
    check(A, lam2.cpu().numpy(), V2.cpu().numpy())


@pytest.mark.parametrize("n,prec", [(700, "d"), (2100, "d"), (700, "s")])
def test_lower_triangle_only_nan_above(torch_dev, n, prec):
    """NaN above the diagonal: a read of the upper triangle that is multiplied by zero
    (which finite garbage would hide) shows as a non-finite result. n=700 takes the
    direct back-transform and the host D&C, n=2100 the reordered one and the device D&C."""
    rng = np.random.default_rng(n + 31)
    G = rng.standard_normal((n, n))
    A = (G + G.T) / 2
    An = np.tril(A) + np.triu(np.full((n, n), np.nan), 1)
    lam, V = solve_np(torch_dev, An, prec=prec)
    if prec == "d":
        check(A, lam, V)
    else:
        check(A, lam, V, ev_tol=1e-5, res_tol=1e-7, orth_tol=1e-7)


def test_invalid_arguments(torch_dev):
    torch = torch_dev
    L = cv.lib()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.cuev_dsymm_eig_solve(p, 0, p, p, s, None) == 1
    assert L.cuev_dsymm_eig_solve(None, 4, p, p, s, None) == 1
    o = cv.Options()
    L.cuev_options_init(ctypes.byref(o))
    o.bandwidth = 48
    assert L.cuev_dsymm_eig_solve_ex(p, 4, p, p, s, None, ctypes.byref(o)) == 1
    o.bandwidth, o.block = 32, 48
    assert L.cuev_dsymm_eig_solve_ex(p, 4, p, p, s, None, ctypes.byref(o)) == 1
    assert L.cuev_ssymm_eig_solve_ex(p, 4, p, p, s, None, ctypes.byref(cv.Options(precision=1))) == 8
    for field, bad in (("dc_mode", 3), ("dc_mode", -1), ("dc_leaf", 1), ("dc_leaf", 5000)):
        o = cv.Options()
        L.cuev_options_init(ctypes.byref(o))
        setattr(o, field, bad)
        assert L.cuev_dsymm_eig_solve_ex(p, 4, p, p, s, None, ctypes.byref(o)) == 1, (field, bad)


def test_non_finite_input(torch_dev):
    A = np.eye(50)
    A[10, 3] = np.nan
    with pytest.raises(cv.CuevError) as ei:
        solve_np(torch_dev, A)
    assert ei.value.status == 7


def test_timer_fields(torch_dev):
    torch = torch_dev
    rng = np.random.default_rng(1)
    G = rng.standard_normal((2048, 2048))
    A = (G + G.T) / 2
    dA = to_dev(A)
    lam, V, t = cv.symm_eig_solve(dA, timer=True, bt_mode="reordered")
    assert t["total_ms"] > 0 and t["dbbr_ms"] > 0 and t["bc_ms"] > 0 and t["dc_ms"] > 0
    assert t["total_ms"] >= t["dc_ms"]


def test_profile_select(torch_dev):
    """The event profiler records every launch site by default and only the selected
    tag prefixes after cuev_profile_select; the selected sites' totals do not depend on
    the selection, and the solve's results do not depend on the profiler."""
    torch = torch_dev
    rng = np.random.default_rng(3)
    G = rng.standard_normal((2048, 2048))
    A = (G + G.T) / 2
    try:
        lam0, V0 = solve_np(torch, A, bt_mode="reordered")
        cv.profile_enable(True)
        cv.profile_read(reset=True)
        lam1, V1 = solve_np(torch, A, bt_mode="reordered")
        every = cv.profile_read(reset=True)
        cv.profile_enable(True, only=("bc.chase", "bt.gemm"))
        lam2, V2 = solve_np(torch, A, bt_mode="reordered")
        some = cv.profile_read(reset=True)
        cv.profile_enable(True, only="dbbr.")
        solve_np(torch, A, bt_mode="reordered")
        dbbr = cv.profile_read(reset=True)
    finally:
        cv.profile_enable(False)
    assert {"bc.chase", "bt.gemm", "dbbr.symm", "dbbr.panel_qr", "dc.gemm"} <= set(every)
    assert set(some) == {"bc.chase", "bt.gemm"}
    assert set(dbbr) and all(t.startswith("dbbr.") for t in dbbr)
    for t in ("bc.chase", "bt.gemm"):
        assert some[t]["count"] == every[t]["count"] and some[t]["flops"] == every[t]["flops"]
    assert np.array_equal(lam0, lam1) and np.array_equal(lam0, lam2)
    assert np.array_equal(V0, V1) and np.array_equal(V0, V2)


def test_reordered_repeatable_n16384(torch_dev):
    """The full pipeline at n=16384 in the reordered order, where the bulge chasing
    runs beside the Q_s^T back-transform: eight solves of the same matrix must give
    bit-identical eigenvalues (every stage is deterministic, so any cross-workgroup
    ordering fault in the chase's hand-offs shows up as a difference) within 1e-12 of
    the exact spectrum."""
    torch = torch_dev
    from cuev_amd import workloads
    A0, lam_exact = workloads.make("q16k", n=16384, device="cuda")
    ref = torch.sort(lam_exact.double().cuda()).values
    work = torch.empty_like(A0)
    first = None
    for _ in range(8):
        work.copy_(A0)
        lam, _V = cv.symm_eig_solve(work, overwrite=True, bandwidth=64, block=512, bt_mode="reordered",
                                    dc_mode="device")
        torch.cuda.synchronize()
        lam = torch.sort(lam.double()).values
        assert (lam - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
        if first is None:
            first = lam.clone()
        else:
            assert torch.equal(lam, first)


@pytest.mark.parametrize("bt", ["reordered", "direct"])
def test_bc_back_in_segments_n18000(torch_dev, bt):
    """n=18000: the bc_back stream (1.6e5 block reflectors) is applied in launches of
    pass ranges (five; each launch lines its waves up again) and in two column chunks,
    in both back-transform orders (Q_b^T on M, Q_b on Q_d). Known spectrum; residual
    and orthogonality at the north-star tolerances, computed on the GPU in float64."""
    torch = torch_dev
    from cuev_amd import workloads
    n = 18000
    A0, lam_exact = workloads.make("q16k", n=n, device="cuda")
    lam, V = cv.symm_eig_solve(A0, bandwidth=64, block=512, bt_mode=bt, dc_mode="device")
    torch.cuda.synchronize()
    ref = torch.sort(lam_exact.double()).values
    assert (torch.sort(lam.double()).values - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
    res, orth = workloads.residual_orth(A0, V, lam.double())
    assert res <= 1e-13 and orth <= 1e-13, (res, orth)
    assert res * n <= 1e-13, ("bench_solver-form residual", res * n)
    del A0, V, lam


def test_reordered_above_one_round_n16500(torch_dev):
    """n=16500 is more columns than one resident round of the bc_back kernel covers
    (16384 at one wave per SIMD), so Q_b^T is applied in two column chunks. Known
    spectrum; residual and orthogonality at the north-star tolerances, computed on the
    GPU in float64."""
    torch = torch_dev
    from cuev_amd import workloads
    n = 16500
    A0, lam_exact = workloads.make("q16k", n=n, device="cuda")
    lam, V = cv.symm_eig_solve(A0, bandwidth=64, block=512, bt_mode="reordered", dc_mode="device")
    torch.cuda.synchronize()
    ref = torch.sort(lam_exact.double()).values
    assert (torch.sort(lam.double()).values - ref).abs().max().item() <= 1e-12 * ref.abs().max().item()
    V = V.double()
    R = A0 @ V - V * lam.double()
    res = torch.linalg.matrix_norm(R).item() / (n * torch.linalg.matrix_norm(A0).item())
    del R
    G = V.T @ V
    G.diagonal().sub_(1.0)
    orth = torch.linalg.matrix_norm(G).item() / n
    assert res <= 1e-13 and orth <= 1e-13, (res, orth)
    assert res * n <= 1e-13, ("bench_solver-form residual", res * n)
