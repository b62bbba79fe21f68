"""Truncations on the HIP device: the Gram-matrix truncation (tnkit/linalg.py:138-204) and the
randomized range-finder SVD (tnkit/linalg.py:81-135).

The Gram matrix is one libtnkit_hip K-major GEMM; its eigendecomposition (order at most 2 chi)
is the library's own: the batched Jacobi kernel up to order 64, the divide-and-conquer solver
(`kernels.syev_top`) up to 2048, which also returns only the kept vectors; the QR that
re-orthonormalises the projected factor on the long side is the library's (`kernels.qr`). The
range finder's
sketch and power products are the same K-major GEMM, its orthonormalisations Gram passes with
the small eigendecompositions on the same solvers, and the core SVD the one-sided Jacobi kernel
or the Gram truncation (rocSOLVER only for complex cores above order 64 and orders above 2048).
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
    AZ = A @ Z
    Q, Rr = kernels.qr(AZ) if kernels.qr_supported(AZ) else torch.linalg.qr(AZ)
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


def _kmajor(X: torch.Tensor) -> torch.Tensor:
    """X^H laid out K-major for `kernels.gemm_tn`: the conjugate of X, contiguous (X itself for
    real X). gemm_tn(_kmajor(X), Y) = X^H Y."""
    return X.conj().resolve_conj().contiguous() if X.is_complex() else X.contiguous()


def _mm(X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """X @ Y on the library's K-major GEMM (X transposed into K-major order first)."""
    return kernels.gemm_tn(X.t().contiguous(), Y.contiguous())


def _eig_desc(G: torch.Tensor):
    """All eigenpairs of a Hermitian G, eigenvalues descending: the library's solvers (the Jacobi
    kernel to order 64, real or complex; the divide-and-conquer solver for real orders to 2048);
    torch.linalg.eigh (rocSOLVER) for complex orders above 64 and real orders above 2048."""
    n = int(G.shape[0])
    if n <= kernels.EIGH_JACOBI_MAX_N and G.dtype in (torch.float64, torch.complex128):
        w, V, sw = kernels.eigh_jacobi(G.unsqueeze(0).contiguous())
        if int(sw[0].item()) >= 0:
            return torch.flip(w[0], dims=[0]), torch.flip(V[0], dims=[1])
    elif G.dtype == torch.float64 and n <= kernels.SYEV_MAX_N:
        w, Z = kernels.eigh_top(G.contiguous(), n)
        return torch.flip(w, dims=[0]), Z
    w, V = torch.linalg.eigh(G)
    return torch.flip(w, dims=[0]), torch.flip(V, dims=[1])


def _orthonormal(Y: torch.Tensor) -> torch.Tensor:
    """An orthonormal basis Q (m x l) of the span of the columns of Y (m x l, l <= m), in Y's
    dtype, by regularised Gram passes: G = Q^H Q = V diag(w) V^H on the library's GEMM and small
    eigensolvers, then Q <- Q V diag(max(w, s))^(-1/2) with s = 8 l sqrt(m) eps max(w), a bound
    on the rounding error of the computed Gram matrix (capped at max(w) / 10 in single
    precision). On the directions the Gram matrix resolves (w > s) a pass is the plain one
    (Stathopoulos and Wu's SVQB, SIAM J. Sci. Comput. 23 (2002)) and normalises them; the others
    it lifts by 1/sqrt(s) instead of dividing by a rounding error (the regularisation of shifted
    Cholesky QR, Fukaya et al., SIAM J. Sci. Comput. 42 (2020)), so each pass shrinks the
    condition number by about sqrt(s / max(w)) until the basis is orthonormal. Directions of Y
    at rounding level come out as orthonormal rounding noise, as the QR factor tnkit takes
    completes a rank-deficient sketch. Passes stop once max |Q^H Q - 1| <= 180 eps (at most
    eight). A sketch this cannot orthonormalise (exactly zero, or exactly dependent columns
    that stay dependent) gets a Householder QR (`kernels.qr`; rocSOLVER's in single precision or
    beyond its row limits)."""
    m, l = int(Y.shape[0]), int(Y.shape[1])
    eps = float(torch.finfo(Y.dtype).eps)
    wide = torch.complex128 if Y.is_complex() else torch.float64
    eye = torch.eye(l, dtype=wide, device=Y.device)
    c = 8.0 * l * np.sqrt(m) * eps
    Q = Y
    for p in range(9):
        G = kernels.gemm_tn(_kmajor(Q), Q.contiguous()).to(wide)
        G = 0.5 * (G + G.mH)
        if p >= 1 and float((G - eye).abs().max().item()) <= 180.0 * eps:
            return Q
        if p == 8:
            break
        w, V = _eig_desc(G)
        wh = w.cpu().numpy()
        if not wh[0] > 0.0:
            break
        shift = min(c * float(wh[0]), 0.1 * float(wh[0]))
        scale = torch.rsqrt(torch.clamp(w, min=shift))
        Q = _mm(Q, (V * scale[None, :].to(V.dtype)).to(Y.dtype))
    return kernels.qr(Y)[0] if kernels.qr_supported(Y) else torch.linalg.qr(Y)[0]


def _core_svd(B: torch.Tensor, keep: int):
    """(U, S, Vh, discarded weight) of the leading `keep` singular triplets of a small dense B:
    the one-sided Jacobi kernel when B fits one workgroup, the Gram truncation (GEMM + the
    library's eigensolver) for real B up to order 2048, rocSOLVER's SVD otherwise."""
    m, n = int(B.shape[0]), int(B.shape[1])
    if kernels.svd_jacobi_fits(m, n, B.dtype) or not (B.dtype == torch.float64
                                                       and min(m, n) <= kernels.SYEV_MAX_N):
        U, S, Vh = kernels.thin_svd(B)
        keep = min(keep, int(S.shape[0]))
        disc = float(torch.sum(S[keep:] ** 2).item())
        return U[:, :keep], S[:keep], Vh[:keep], disc
    return gram_truncation(B, keep, "left")


def randomized_svd_device(A: torch.Tensor, chi: int, be=None, oversample: int = 8,
                          power_iters: int = 2, rng=None, mixed: bool = False):
    """Rank-chi (U, S, Vh, absolute discarded weight) by the Halko-Martinsson-Tropp range
    finder (tnkit/linalg.py:81-135), on the library's kernels. The Gaussian sketch (n x ell,
    ell = chi + oversample) is drawn from the host generator `rng` exactly as tnkit's backend
    draws it (real for real A, real + i real for complex), so a run with tnkit's seed uses the
    same sketch. The products A Omega, A^H Q and A Q are K-major GEMMs (`kernels.gemm_tn`),
    each product orthonormalised by `_orthonormal` (same span as tnkit's QR), `power_iters`
    rounds of subspace iteration, then the SVD of the ell x n projection Q^H A by `_core_svd`.
    `mixed` runs the range finder in single precision (the float32 GEMMs and Gram passes, the
    small eigendecompositions in double) and re-orthonormalises Q once in double before the
    projection and core SVD in double. When min(m, n) <= ell the whole matrix goes to
    `_core_svd`."""
    rng = np.random.default_rng() if rng is None else rng
    m, n = int(A.shape[0]), int(A.shape[1])
    if min(m, n) <= chi + oversample:
        return _core_svd(A.contiguous(), int(chi))
    ell = int(min(chi + oversample, n))
    om = rng.standard_normal((n, ell))
    if A.is_complex():
        om = om + 1j * rng.standard_normal((n, ell))
    Om = torch.as_tensor(om).to(A.device, A.dtype)
    if mixed:
        low = {torch.float64: torch.float32, torch.complex128: torch.complex64}.get(A.dtype,
                                                                                   A.dtype)
        As, Om = A.to(low), Om.to(low)
    else:
        As = A
    At = As.t().contiguous()                  # K-major A for A X
    Ah = _kmajor(As)                          # K-major A^H for A^H X
    Q = _orthonormal(kernels.gemm_tn(At, Om))
    for _ in range(power_iters):
        Q = _orthonormal(kernels.gemm_tn(Ah, Q))
        Q = _orthonormal(kernels.gemm_tn(At, Q))
    if mixed:
        Q = _orthonormal(Q.to(A.dtype))
    Ub, S, Vh, disc = _core_svd(kernels.gemm_tn(_kmajor(Q), A.contiguous()), int(chi))
    return _mm(Q, Ub), S, Vh, disc
