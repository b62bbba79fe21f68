"""Gram-matrix truncation on the HIP device (tnkit/linalg.py:138-204).

The Gram matrix is one libtnkit_hip K-major GEMM; the eigendecomposition of the (at most
2 chi)^2 Gram matrix is torch.linalg.eigh on the device.
"""
from __future__ import annotations

import torch

from . import kernels


def _gram_left(A: torch.Tensor, chi: int):
    """Top-chi left factorisation A ~= Q C from the smaller Gram matrix; Q an exact isometry
    (tnkit/linalg.py:138-159). Returns Q, C, eigenvalues (descending, host), k."""
    m, n = A.shape
    if m <= n:
        X = A.t().contiguous()                     # (n, m): G = X^T X = A A^T
        G = kernels.gemm_tn(X, X)
    else:
        G = kernels.gemm_tn(A.contiguous(), A.contiguous())   # A^T A
    G = 0.5 * (G + G.t())
    w, Z = torch.linalg.eigh(G)
    r = min(m, n)
    k = min(int(chi), r)
    Z = torch.flip(Z[:, r - k:], dims=[1])
    wh = torch.flip(w, dims=[0]).cpu().numpy()
    if m <= n:
        return Z, Z.t() @ A, wh, k
    Q, Rr = torch.linalg.qr(A @ Z)
    return Q, Rr @ Z.t(), wh, k


def gram_truncation(A: torch.Tensor, chi: int, iso: str = "left"):
    """(U, S, Vh, discarded weight) with the `iso` factor an exact isometry and S the norms of
    the carried factor (tnkit/linalg.py:162-204). Float64 only."""
    if A.dtype != torch.float64:
        raise ValueError(f"trunc='gram' needs double precision, got {A.dtype}")
    if iso == "right":
        Q, C, w, k = _gram_left(A.t(), chi)
        Vh, US = Q.t(), C.t()
        S = torch.linalg.norm(US, dim=0)
        Sg = torch.where(S > 0, S, torch.ones_like(S))
        return US / Sg[None, :], S, Vh.contiguous(), float(w[k:].clip(min=0.0).sum())
    if iso != "left":
        raise ValueError(f"iso must be 'left' or 'right', got {iso!r}")
    U, C, w, k = _gram_left(A, chi)
    S = torch.linalg.norm(C, dim=1)
    Sg = torch.where(S > 0, S, torch.ones_like(S))
    return U, S, C / Sg[:, None], float(w[k:].clip(min=0.0).sum())
