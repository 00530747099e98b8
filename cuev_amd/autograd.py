"""torch.autograd wrapper: eigh forward and backward through the C ABI.

Forward: cuev_{s,d}symm_eig_solve (lower triangle of A read). Backward:
cuev_{s,d}symm_eig_backward, which needs only the saved (lambda, V):
    Abar = V (diag(lambar) + F o (V^T Vbar - Vbar^T V)/2) V^T,
    F_ij = 1/(lambda_j - lambda_i), F_ii = 0, pairs with gap <= gap_tol masked.
The returned gradient is symmetric (the derivative with respect to symmetric
perturbations of A).
"""
import torch

from . import eig_backward, symm_eig_solve


class _Eigh(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, gap_tol, precision, bandwidth):
        lam, V = symm_eig_solve(A.detach(), precision=precision, bandwidth=bandwidth)
        V = V.contiguous()
        ctx.gap_tol = gap_tol
        ctx.save_for_backward(lam, V)
        return lam, V

    @staticmethod
    def backward(ctx, glam, gV):
        lam, V = ctx.saved_tensors
        Abar, masked = eig_backward(V, lam, gV, glam, gap_tol=ctx.gap_tol)
        _Eigh.last_masked = masked
        return Abar.contiguous(), None, None, None


def eigh(A, gap_tol=0.0, precision="native", bandwidth=0):
    """(eigenvalues ascending, eigenvectors as columns) with autograd support."""
    return _Eigh.apply(A, float(gap_tol), precision, bandwidth)
