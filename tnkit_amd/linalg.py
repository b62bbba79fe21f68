"""Truncations on the HIP device: the Gram-matrix truncation (tnkit/linalg.py:138-204) and the
randomized range-finder SVD (tnkit/linalg.py:81-135).

The Gram matrix is one libtnkit_hip K-major GEMM; its eigendecomposition (order at most 2 chi)
is the library's own: the batched Jacobi kernel up to order 64, the divide-and-conquer solver
(`kernels.syev_top`) up to 2048, which also returns only the kept vectors. The range finder's
sketch and power products are plain dense GEMMs (torch/rocBLAS), its QRs and the small core SVD
rocSOLVER.
"""
from __future__ import annotations

import numpy as np
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
    r = min(m, n)
    k = min(int(chi), r)
    if r <= kernels.SYEV_MAX_N:
        w, Z = kernels.eigh_top(G.contiguous(), k)           # Z: the k leading, descending
    else:
        w, Z = torch.linalg.eigh(G)
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


def randomized_svd_device(A: torch.Tensor, chi: int, be=None, oversample: int = 8,
                          power_iters: int = 2, rng=None, mixed: bool = False):
    """Rank-chi (U, S, Vh, absolute discarded weight) by the Halko-Martinsson-Tropp range
    finder (tnkit/linalg.py:81-135), on the device. The Gaussian sketch (n x ell, ell =
    chi + oversample) is drawn from the host generator `rng` exactly as tnkit's backend draws it
    (real for real A, real + i real for complex), so a run with tnkit's seed uses the same
    sketch. `power_iters` rounds of subspace iteration, each product re-orthonormalised by QR,
    then the SVD of the ell x n projection Q^H A. `mixed` runs the range finder in single
    precision (the sketch and power products on the float32 matrix path) and the projection and
    core SVD in double. When min(m, n) <= ell the full SVD is used."""
    rng = np.random.default_rng() if rng is None else rng
    m, n = int(A.shape[0]), int(A.shape[1])
    if min(m, n) <= chi + oversample:
        U, S, Vh = torch.linalg.svd(A, full_matrices=False)
        keep = min(chi, int(S.shape[0]))
        disc = float(torch.sum(S[keep:] ** 2).item())
        return U[:, :keep], S[:keep], Vh[:keep], disc
    ell = int(min(chi + oversample, n))
    om = rng.standard_normal((n, ell))
    if A.is_complex():
        om = om + 1j * rng.standard_normal((n, ell))
    Om = torch.as_tensor(om).to(A.device, A.dtype)
    if mixed:
        low = {torch.float64: torch.float32, torch.complex128: torch.complex64}.get(A.dtype,
                                                                                   A.dtype)
        As = A.to(low)
        Q, _ = torch.linalg.qr(As @ Om.to(low))
        for _ in range(power_iters):
            Q, _ = torch.linalg.qr(As.mH @ Q)
            Q, _ = torch.linalg.qr(As @ Q)
        Q = Q.to(A.dtype)
    else:
        Q, _ = torch.linalg.qr(A @ Om)
        for _ in range(power_iters):
            Q, _ = torch.linalg.qr(A.mH @ Q)
            Q, _ = torch.linalg.qr(A @ Q)
    Ub, S, Vh = torch.linalg.svd(Q.mH @ A, full_matrices=False)
    U = Q @ Ub
    keep = min(chi, int(S.shape[0]))
    disc = float(torch.sum(S[keep:] ** 2).item())
    return U[:, :keep], S[:keep], Vh[:keep], disc
