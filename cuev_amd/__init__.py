"""cuev_amd — MI355X (gfx950) two-stage symmetric eigensolver, Python side.

Thin ctypes binding of the C ABI in include/cuev_c.h (libcuev_amd.so, built by
``python -m cuev_amd.build``). PyTorch supplies device memory and streams only.

Layout: the library is column-major. A torch tensor ``t`` of shape (n, n) that
is C-contiguous holds the column-major matrix ``t.T``; one with strides (1, n)
(``t.mT`` contiguous) holds ``t`` itself. ``symm_eig_solve`` reads the lower
triangle of the tensor it is given, as ``torch.linalg.eigh`` does, and returns
eigenvectors as the columns of ``V``.

Mirrors the reference's entry point ``cuev::symm_eig_solve<T>(A, n, eval, evec,
stream, timer)`` (include/cuev.h:31-33) and the stage-level launchers of
``cuev::kernels`` (include/cuda/kernels.cuh:28-157) for tests.
"""
import ctypes
import os

__all__ = [
    "CuevError", "SolveTimer", "Options", "lib", "lib_path", "symm_eig_solve",
    "eig_backward", "Handle", "gemm", "symm_lower", "syr2k_lower", "lapack_name",
    "STATUS", "BUF", "profile_enable", "profile_read", "cache_release", "symm_eig_values",
    "tridi_values", "symm_eig_batched",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libcuev_amd.so")

STATUS = {
    0: "SUCCESS", 1: "INVALID_ARGUMENT", 2: "HIP_ERROR", 3: "ALLOC_FAILED",
    4: "LAPACK_ERROR", 5: "LAPACK_UNAVAILABLE", 6: "WATCHDOG", 7: "NOT_FINITE",
    8: "NOT_SUPPORTED", 9: "NOT_CONVERGED",
}
BUF = {"Y": 0, "W": 1, "B": 2, "U": 3, "D": 4, "E": 5, "TRI": 6, "TAU": 7, "M": 8}
PRECISION = {"native": 0, "mixed": 1}
BT_MODE = {"auto": 0, "direct": 1, "reordered": 2}
DC_MODE = {"auto": 0, "host": 1, "device": 2}


class CuevError(RuntimeError):
    def __init__(self, status, message):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"cuev error {status} ({self.name}): {message}")


class SolveTimer(ctypes.Structure):
    """cuev_solve_timer (reference SolveTimer, include/cuda/handle.h:25-37, plus additions)."""
    _fields_ = [(f, ctypes.c_float) for f in (
        "dbbr_ms", "bc_ms", "dc_ms", "bt_ms", "total_ms", "bt_copyin_ms", "bt_qb_ms",
        "bt_qs_ms", "bt_copyout_ms", "refine_ms", "overlap_ms")] + [
        ("refine_sweeps", ctypes.c_int), ("n_cluster_cols", ctypes.c_int),
        ("refine_resid", ctypes.c_float), ("refine_orth", ctypes.c_float)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Options(ctypes.Structure):
    _fields_ = [("bandwidth", ctypes.c_int), ("block", ctypes.c_int), ("precision", ctypes.c_int),
                ("refine_sweeps", ctypes.c_int), ("bt_mode", ctypes.c_int),
                ("dc_mode", ctypes.c_int), ("dc_leaf", ctypes.c_int),
                ("refine_tol", ctypes.c_double)]


_lib = None


def lib():
    """Load libcuev_amd.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise RuntimeError(
                f"{lib_path} not found: build it with `python -m cuev_amd.build` "
                "(the HIP library is the only implementation)")
        L = ctypes.CDLL(lib_path)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    P, I, I64, D, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_float
    L.cuev_status_string.restype = ctypes.c_char_p
    L.cuev_status_string.argtypes = [I]
    L.cuev_last_error_message.restype = ctypes.c_char_p
    L.cuev_lapack_name.restype = ctypes.c_char_p
    L.cuev_options_init.argtypes = [ctypes.POINTER(Options)]
    for pre in ("s", "d"):
        getattr(L, f"cuev_{pre}symm_eig_solve").argtypes = [P, I, P, P, P, P]
        getattr(L, f"cuev_{pre}symm_eig_solve_ex").argtypes = [P, I, P, P, P, P, ctypes.POINTER(Options)]
        getattr(L, f"cuev_{pre}symm_eig_values").argtypes = [P, I, P, P, P]
        getattr(L, f"cuev_{pre}symm_eig_values_ex").argtypes = [P, I, I, I, P, P, P, ctypes.POINTER(Options)]
        getattr(L, f"cuev_{pre}tridi_values").argtypes = [P, P, I, I, I, P, P]
        getattr(L, f"cuev_{pre}symm_eig_batched").argtypes = [P, I, I, P, P, P, ctypes.POINTER(I)]
        getattr(L, f"cuev_{pre}symm_eig_backward").argtypes = [P, P, P, P, I, P, D, ctypes.POINTER(I), P]
        getattr(L, f"cuev_{pre}symm_eig_backward_noalloc").argtypes = [P, P, P, P, I, P, D, ctypes.POINTER(I), P]
        getattr(L, f"cuev_{pre}_dbbr_pack").argtypes = [P, P, P, I, I]
        getattr(L, f"cuev_{pre}_dbbr_panel_qr").argtypes = [P, P, P, I, I]
        getattr(L, f"cuev_{pre}_dbbr_reduce").argtypes = [P, P, P]
        getattr(L, f"cuev_{pre}_bc_chase").argtypes = [P, P, P, P]
        getattr(L, f"cuev_{pre}_tridi_dc").argtypes = [P, P, P, P, P]
        getattr(L, f"cuev_{pre}_back_transform").argtypes = [P, P, P, P, P, P]
        getattr(L, f"cuev_{pre}_bc_back").argtypes = [P, P, P, I64]
        getattr(L, f"cuev_{pre}_bc_back_t").argtypes = [P, P, P, I64]
        getattr(L, f"cuev_{pre}_sbr_back").argtypes = [P, P, P, P, I64]
        sc = D if pre == "d" else F
        getattr(L, f"cuev_{pre}gemm").argtypes = [P, I, I, I64, I64, I64, sc, P, I64, P, I64, sc, P, I64]
        getattr(L, f"cuev_{pre}symm_lower").argtypes = [P, I64, I64, sc, P, I64, P, I64, sc, P, I64]
        getattr(L, f"cuev_{pre}syr2k_lower").argtypes = [P, I64, I64, sc, P, I64, P, I64, sc, P, I64]
    L.cuev_cache_release.argtypes = []
    L.cuev_profile_enable.argtypes = [I]
    L.cuev_profile_select.argtypes = [ctypes.c_char_p]
    L.cuev_profile_read.argtypes = [ctypes.c_char_p, I, I]
    L.cuev_handle_create.argtypes = [ctypes.POINTER(P), I, I, I, I, P]
    L.cuev_handle_destroy.argtypes = [P]
    L.cuev_handle_buffer.argtypes = [P, I, ctypes.POINTER(P), ctypes.POINTER(I64)]
    L.cuev_handle_set_dc.argtypes = [P, I, I]


def _check(st):
    if st != 0:
        msg = lib().cuev_last_error_message().decode(errors="replace")
        raise CuevError(st, msg)


def profile_enable(on=True, only=None):
    """Record per-launch HIP event timings inside the library (see profile_read).

    only: an iterable of tag prefixes (or one comma-separated string) to record;
    None records every tag. Every recorded launch puts two event records on its
    stream, a few microseconds each between back-to-back kernels, so a timed run
    should record only the tags it reports."""
    if only is not None and not isinstance(only, str):
        only = ",".join(only)
    lib().cuev_profile_select(only.encode() if only else None)
    lib().cuev_profile_enable(1 if on else 0)


def profile_read(reset=True):
    """Per-tag totals {tag: {ms, count, flops, bytes, max_ms}} of the recorded launches."""
    import json
    need = lib().cuev_profile_read(None, 0, 0)
    buf = ctypes.create_string_buffer(max(need, 2) + 64)
    lib().cuev_profile_read(buf, len(buf), 1 if reset else 0)
    return json.loads(buf.value.decode() or "{}")


def cache_release():
    """Free the workspaces the library keeps between solves (all devices)."""
    _check(lib().cuev_cache_release())


def lapack_name():
    return lib().cuev_lapack_name().decode()


def _torch():
    import torch
    return torch


def _stream_ptr(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _prefix(dtype):
    torch = _torch()
    if dtype == torch.float64:
        return "d"
    if dtype == torch.float32:
        return "s"
    raise TypeError(f"unsupported dtype {dtype} (float32/float64 only)")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def symm_eig_solve(A, *, overwrite=False, bandwidth=0, block=0, precision="native",
                   refine_sweeps=0, bt_mode="auto", dc_mode="auto", dc_leaf=0, refine_tol=0.0,
                   timer=False, stream=None):
    """Eigen-decomposition of the symmetric (n, n) CUDA tensor ``A``.

    Only the lower triangle of ``A`` (``A[i, j]`` with ``i >= j``) is read. Returns
    ``(eval, V)`` with ascending eigenvalues and eigenvectors in the columns of
    ``V``; with ``timer=True`` also a dict of per-stage times. ``A`` is consumed in
    place (as in the reference) only when ``overwrite=True`` and ``A`` is stored
    column-major (``A.mT.is_contiguous()``); otherwise a column-major copy is solved.
    A status of NOT_CONVERGED (mixed form) raises ``CuevError`` after eval and V
    were written.
    """
    torch = _torch()
    if A.dim() != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    if not A.is_cuda:
        raise ValueError("A must be a device tensor")
    n = A.shape[0]
    pre = _prefix(A.dtype)
    # the library reads the buffer column-major: storage of A.mT in row-major order
    work = A if (overwrite and A.mT.is_contiguous()) else A.mT.contiguous().mT
    ev = torch.empty(n, dtype=A.dtype, device=A.device)
    vec = torch.empty((n, n), dtype=A.dtype, device=A.device)
    t = SolveTimer()
    o = Options()
    lib().cuev_options_init(ctypes.byref(o))
    o.bandwidth, o.block = int(bandwidth), int(block)
    o.precision = PRECISION[precision]
    o.refine_sweeps = int(refine_sweeps)
    o.bt_mode = BT_MODE[bt_mode]
    o.dc_mode = DC_MODE[dc_mode]
    o.dc_leaf = int(dc_leaf)
    o.refine_tol = float(refine_tol)
    fn = getattr(lib(), f"cuev_{pre}symm_eig_solve_ex")
    _check(fn(_ptr(work), n, _ptr(ev), _ptr(vec), _stream_ptr(stream), ctypes.byref(t), ctypes.byref(o)))
    V = vec.T  # column j of the column-major result is row j of `vec`
    if timer:
        return ev, V, t.as_dict()
    return ev, V


def symm_eig_values(A, il=0, iu=None, *, overwrite=False, bandwidth=0, block=0, timer=False,
                    stream=None):
    """Eigenvalues il..iu (0-based, inclusive, ascending) of the symmetric CUDA tensor
    ``A`` (lower triangle read), without eigenvectors: band reduction, bulge chasing
    and Sturm multisection on the GPU. Returns the eigenvalues (and the stage times
    with ``timer=True``). ``A`` is consumed in place only with ``overwrite=True`` and
    column-major storage, as in ``symm_eig_solve``."""
    torch = _torch()
    if A.dim() != 2 or A.shape[0] != A.shape[1]:
        raise ValueError("A must be square")
    if not A.is_cuda:
        raise ValueError("A must be a device tensor")
    n = A.shape[0]
    iu = n - 1 if iu is None else int(iu)
    pre = _prefix(A.dtype)
    work = A if (overwrite and A.mT.is_contiguous()) else A.mT.contiguous().mT
    ev = torch.empty(max(iu - int(il) + 1, 0), dtype=A.dtype, device=A.device)
    t = SolveTimer()
    o = Options()
    lib().cuev_options_init(ctypes.byref(o))
    o.bandwidth, o.block = int(bandwidth), int(block)
    fn = getattr(lib(), f"cuev_{pre}symm_eig_values_ex")
    _check(fn(_ptr(work), n, int(il), iu, _ptr(ev), _stream_ptr(stream), ctypes.byref(t), ctypes.byref(o)))
    return (ev, t.as_dict()) if timer else ev


def symm_eig_batched(A, stream=None):
    """Eigen-decompositions of a batch of symmetric matrices, A of shape (batch, n, n)
    with n <= 256 (lower triangle of each A[b] read, as torch.linalg.eigh does).
    Returns (eval (batch, n) ascending, V (batch, n, n) with eigenvectors as columns)."""
    torch = _torch()
    if A.dim() != 3 or A.shape[1] != A.shape[2]:
        raise ValueError("A must have shape (batch, n, n)")
    batch, n = A.shape[0], A.shape[1]
    pre = _prefix(A.dtype)
    work = A.mT.contiguous()  # each matrix column-major (the library's layout)
    ev = torch.empty((batch, n), dtype=A.dtype, device=A.device)
    vec = torch.empty((batch, n, n), dtype=A.dtype, device=A.device)
    nbad = ctypes.c_int(0)
    _check(getattr(lib(), f"cuev_{pre}symm_eig_batched")(_ptr(work), n, batch, _ptr(ev), _ptr(vec),
                                                         _stream_ptr(stream), ctypes.byref(nbad)))
    return ev, vec.mT


def tridi_values(d, e, il=0, iu=None, stream=None):
    """Eigenvalues il..iu of tridiag(e, d, e) (device tensors; e has n-1 or n entries)."""
    torch = _torch()
    n = d.shape[0]
    iu = n - 1 if iu is None else int(iu)
    pre = _prefix(d.dtype)
    d = d.contiguous()
    e = e.contiguous()
    w = torch.empty(max(iu - int(il) + 1, 0), dtype=d.dtype, device=d.device)
    _check(getattr(lib(), f"cuev_{pre}tridi_values")(_ptr(d), _ptr(e), n, int(il), iu, _ptr(w),
                                                     _stream_ptr(stream)))
    return w


def eig_backward(V, lam, Vbar=None, lambar=None, gap_tol=0.0, stream=None):
    """Abar for cotangents (Vbar, lambar) of (V, lam); returns (Abar, n_masked_pairs)."""
    torch = _torch()
    n = V.shape[0]
    pre = _prefix(V.dtype)
    # column-major buffers of V and Vbar: C-contiguous tensors of their transposes
    Vc = V.T.contiguous()
    Vbc = Vbar.T.contiguous() if Vbar is not None else None
    lam = lam.contiguous()
    lbc = lambar.contiguous() if lambar is not None else None
    out = torch.empty((n, n), dtype=V.dtype, device=V.device)
    masked = ctypes.c_int(0)
    fn = getattr(lib(), f"cuev_{pre}symm_eig_backward")
    _check(fn(_ptr(Vc), _ptr(lam), _ptr(Vbc), _ptr(lbc), n, _ptr(out), float(gap_tol),
              ctypes.byref(masked), _stream_ptr(stream)))
    return out.T, masked.value  # Abar is symmetric; .T keeps the column-major meaning


class Handle:
    """Stage-level workspace (reference SolverHandle<T>, include/cuda/handle.h:61-100)."""

    def __init__(self, dtype, n, nbw, nk, stream=None):
        torch = _torch()
        self.dtype = dtype
        self.pre = _prefix(dtype)
        self.n, self.nbw, self.nk = n, nbw, nk
        self._stream = stream if stream is not None else torch.cuda.current_stream()
        h = ctypes.c_void_p(0)
        _check(lib().cuev_handle_create(ctypes.byref(h), 1 if self.pre == "d" else 0, n, nbw, nk,
                                        _stream_ptr(self._stream)))
        self.h = h

    def close(self):
        if self.h:
            lib().cuev_handle_destroy(self.h)
            self.h = ctypes.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def buffer(self, which):
        """(device pointer, leading dimension) of a workspace buffer (see BUF)."""
        p = ctypes.c_void_p(0)
        ld = ctypes.c_int64(0)
        _check(lib().cuev_handle_buffer(self.h, BUF[which], ctypes.byref(p), ctypes.byref(ld)))
        return p.value, ld.value

    def _f(self, name):
        return getattr(lib(), f"cuev_{self.pre}_{name}")

    def dbbr_pack(self, A, B, n, b):
        _check(self._f("dbbr_pack")(self.h, _ptr(A), ctypes.c_void_p(B), n, b))

    def dbbr_panel_qr(self, A, Y, rows, b):
        _check(self._f("dbbr_panel_qr")(self.h, _ptr(A), _ptr(Y), rows, b))

    def dbbr_reduce(self, A):
        _check(self._f("dbbr_reduce")(self.h, _ptr(A), ctypes.c_void_p(self.buffer("B")[0])))

    def bc_chase(self):
        B, _ = self.buffer("B")
        d, _ = self.buffer("D")
        e, _ = self.buffer("E")
        _check(self._f("bc_chase")(self.h, ctypes.c_void_p(B), ctypes.c_void_p(d), ctypes.c_void_p(e)))

    def set_dc(self, mode="host", leaf=0):
        """Tridiagonal stage: host LAPACK ?stedc ("host") or the device-resident D&C."""
        _check(lib().cuev_handle_set_dc(self.h, DC_MODE[mode], int(leaf)))

    def tridi_dc(self, d, e, ev, vec):
        _check(self._f("tridi_dc")(self.h, _ptr(d), _ptr(e), _ptr(ev), _ptr(vec)))

    def bc_back(self, M, ldm):
        U, _ = self.buffer("U")
        _check(self._f("bc_back")(self.h, ctypes.c_void_p(U), _ptr(M), ldm))

    def bc_back_t(self, M, ldm):
        """M <- Q_b^T M (the reordered back-transform's application)."""
        U, _ = self.buffer("U")
        _check(self._f("bc_back_t")(self.h, ctypes.c_void_p(U), _ptr(M), ldm))

    def sbr_back(self, M, ldm):
        Y, _ = self.buffer("Y")
        W, _ = self.buffer("W")
        _check(self._f("sbr_back")(self.h, ctypes.c_void_p(Y), ctypes.c_void_p(W), _ptr(M), ldm))

    def back_transform(self, evec):
        Y, _ = self.buffer("Y")
        W, _ = self.buffer("W")
        U, _ = self.buffer("U")
        t = SolveTimer()
        _check(self._f("back_transform")(self.h, ctypes.c_void_p(Y), ctypes.c_void_p(W),
                                         ctypes.c_void_p(U), _ptr(evec), ctypes.byref(t)))
        return t.as_dict()


def gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    """Column-major C <- alpha op(A) op(B) + beta C on the MFMA engine (raw buffers)."""
    pre = _prefix(C.dtype)
    _check(getattr(lib(), f"cuev_{pre}gemm")(_stream_ptr(stream), int(ta), int(tb), m, n, k, alpha,
                                              _ptr(A), lda, _ptr(B), ldb, beta, _ptr(C), ldc))


def symm_lower(m, n, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    pre = _prefix(C.dtype)
    _check(getattr(lib(), f"cuev_{pre}symm_lower")(_stream_ptr(stream), m, n, alpha, _ptr(A), lda,
                                                   _ptr(B), ldb, beta, _ptr(C), ldc))


def syr2k_lower(n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
    pre = _prefix(C.dtype)
    _check(getattr(lib(), f"cuev_{pre}syr2k_lower")(_stream_ptr(stream), n, k, alpha, _ptr(A), lda,
                                                    _ptr(B), ldb, beta, _ptr(C), ldc))
