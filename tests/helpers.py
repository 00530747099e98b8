"""Shared helpers for the GPU tests: column-major numpy <-> device buffers and metrics."""
import numpy as np


def to_dev(A, dtype=None):
    """Device buffer holding the column-major (Fortran) contents of numpy array A.

    Returned as a C-contiguous torch tensor of A.T (so its memory is A column-major).
    """
    import torch
    A = np.asarray(A)
    dt = dtype or (torch.float64 if A.dtype == np.float64 else torch.float32)
    npdt = np.float64 if dt == torch.float64 else np.float32
    if A.ndim == 1:
        return torch.from_numpy(np.ascontiguousarray(A.astype(npdt))).cuda().to(dt)
    return torch.from_numpy(np.ascontiguousarray(A.T.astype(npdt))).cuda()


def from_dev(t, rows=None, cols=None):
    """numpy Fortran array from a device buffer written column-major by the library."""
    h = t.detach().cpu().numpy()
    if h.ndim == 1:
        return h.astype(np.float64)
    return np.asfortranarray(h.T.astype(np.float64))


def empty_dev(rows, cols, dtype):
    import torch
    return torch.zeros((cols, rows), dtype=dtype, device="cuda")


def raw_dev(ptr, count, dtype):
    """Copy `count` elements from a raw device pointer (workspace buffer) to numpy."""
    import ctypes
    import torch
    out = torch.empty(count, dtype=dtype, device="cuda")
    esize = 8 if dtype == torch.float64 else 4
    torch.cuda.synchronize()
    # a raw workspace pointer cannot be wrapped as a tensor; copy through the HIP runtime
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rc = lib.hipMemcpy(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(ptr), count * esize, 3)
    assert rc == 0, f"hipMemcpy failed {rc}"
    return out.cpu().numpy().astype(np.float64)


def raw_put(ptr, A, dtype):
    """Copy numpy array A (column-major order) into a raw device pointer (workspace buffer)."""
    import ctypes
    import torch
    npdt = np.float64 if dtype == torch.float64 else np.float32
    h = np.asfortranarray(np.asarray(A, dtype=npdt)).ravel(order="F")
    torch.cuda.synchronize()
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rc = lib.hipMemcpy(ctypes.c_void_p(ptr), h.ctypes.data_as(ctypes.c_void_p), h.nbytes, 1)
    assert rc == 0, f"hipMemcpy failed {rc}"


def eig_metrics(A, lam, V, lam_ref):
    """BASELINE metrics: eigenvalue error / ||A||_2, residual ||AV - V diag(lam)||_F/(n ||A||_F),
    orthogonality ||V^T V - I||_F / n."""
    n = A.shape[0]
    a2 = np.max(np.abs(lam_ref)) if len(lam_ref) else 0.0
    a2 = a2 if a2 > 0 else 1.0  # zero matrix: absolute errors
    afro = np.linalg.norm(A)
    afro = afro if afro > 0 else 1.0
    ev_err = np.max(np.abs(lam - lam_ref)) / a2
    res = np.linalg.norm(A @ V - V * lam) / (n * afro)
    orth = np.linalg.norm(V.T @ V - np.eye(n)) / n
    return ev_err, res, orth
