"""The allocation-free backward entry point cuev_{s,d}symm_eig_backward_noalloc against
the allocating cuev_{s,d}symm_eig_backward, through the C ABI.

Two calls in a row (the second with a smaller n, so that stale contents of the shared
device array from the first call would show), each checked against the allocating
entry point on the same inputs: Abar within 1e-12 (fp64) / 1e-5 (fp32) of its largest
entry, and the same count of masked pairs (lambda has planted near-equal pairs). Also
the asynchronous form (n_masked NULL) and the size limit (n > 1024 is rejected).
"""
import ctypes

import numpy as np
import pytest

import cuev_amd as cv

pytestmark = pytest.mark.gpu


def _inputs(torch, n, seed, dtype):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.sort(rng.uniform(-1.0, 1.0, n))
    lam[5] = lam[4] + 1e-9      # near-degenerate pairs: masked at gap_tol 1e-7
    lam[n - 2] = lam[n - 3] + 5e-8
    lam = np.sort(lam)
    Vbar = rng.standard_normal((n, n))
    lambar = rng.standard_normal(n)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T if a.ndim == 2 else a)).to(dtype).cuda()
    return dev(Q), dev(lam), dev(Vbar), dev(lambar)


def _call(torch, fn, V, lam, Vbar, lambar, n, with_count=True):
    out = torch.empty((n, n), dtype=V.dtype, device="cuda")
    masked = ctypes.c_int(-1)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = fn(p(V), p(lam), p(Vbar), p(lambar), n, p(out), 1e-7, ctypes.byref(masked) if with_count else None, s)
    torch.cuda.synchronize()
    return st, out, masked.value


@pytest.mark.parametrize("pre,tol", [("d", 1e-12), ("s", 1e-5)])
def test_noalloc_matches_allocating(torch_dev, pre, tol):
    torch = torch_dev
    L = cv.lib()
    dtype = torch.float64 if pre == "d" else torch.float32
    ref_fn = getattr(L, f"cuev_{pre}symm_eig_backward")
    na_fn = getattr(L, f"cuev_{pre}symm_eig_backward_noalloc")
    for n, seed in ((700, 1), (300, 2)):
        V, lam, Vbar, lambar = _inputs(torch, n, seed, dtype)
        st_r, A_r, m_r = _call(torch, ref_fn, V, lam, Vbar, lambar, n)
        st_n, A_n, m_n = _call(torch, na_fn, V, lam, Vbar, lambar, n)
        assert st_r == 0 and st_n == 0, (st_r, st_n, L.cuev_last_error_message())
        scale = A_r.abs().max().item()
        assert (A_n - A_r).abs().max().item() <= tol * scale, n
        assert m_n == m_r and m_r >= 2, (m_n, m_r)
    # asynchronous form: no count requested
    st_a, A_a, _ = _call(torch, na_fn, V, lam, Vbar, lambar, 300, with_count=False)
    assert st_a == 0
    assert (A_a - A_r).abs().max().item() <= tol * scale


def test_noalloc_size_limit(torch_dev):
    torch = torch_dev
    L = cv.lib()
    n = 1025
    buf = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    p = ctypes.c_void_p(buf.data_ptr())
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.cuev_dsymm_eig_backward_noalloc(p, p, p, p, n, p, 0.0, None, s) == 1
    assert b"1024" in L.cuev_last_error_message()
