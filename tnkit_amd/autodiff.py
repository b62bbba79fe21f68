"""Derivatives of the two-site path behind torch.autograd.Function, with libtnkit_hip kernels in
both directions.

tnkit has no autograd (its arrays are numpy or plain torch without gradients); these Functions
are this package's extension for the Hellmann-Feynman study of the field derivatives
dE/dg_i = <dH/dg_i> = -<Z_i> of the TFIM with per-site fields.

* `HeffMatvec`: y = H_eff theta (tnkit/_dmrg.py:183-189). Backward: theta_bar = H_eff^T y_bar, the
  same kernels on the transposed operator (bra and ket of L and R swapped, W transposed in its
  physical indices), so no intermediate is stored.
* `MPOExpectation`: E(W_0, ..., W_{N-1}) = <psi|H|psi> / <psi|psi> for a fixed MPS, forward by
  the left environment chain (tnkit/_dmrg.py:216-222) keeping only the last environment.
  Backward: the right environments are rebuilt from the right, the left ones recomputed on the
  way back to the right, and each site's MPO gradient
      dE/dW_i[w,r,s,y] = sum L_i[p,w,q] conj(A[p,s,d]) A[q,y,e] R_{i+1}[d,r,e] / <psi|psi>
  is three GEMMs on the K-major MFMA kernel. Environments are recomputed, not stored.
"""
from __future__ import annotations

from typing import List, Sequence

import torch

from . import kernels


def _transpose_operator(L, W1, W2, R):
    """The operator of H_eff^T: L and R with bra and ket swapped, W with s and s' swapped."""
    return (L.permute(2, 1, 0).contiguous(), W1.permute(0, 1, 3, 2).contiguous(),
            W2.permute(0, 1, 3, 2).contiguous(), R.permute(2, 1, 0).contiguous())


class HeffMatvec(torch.autograd.Function):
    """y = H_eff theta with theta_bar = H_eff^T y_bar in backward (environments fixed)."""

    @staticmethod
    def forward(ctx, theta, L, W1, W2, R):
        ctx.save_for_backward(L, W1, W2, R)
        op = kernels.fused_two_site(L, W1, W2, R)
        return kernels.apply_two_site_fused(theta.contiguous(), *op)

    @staticmethod
    def backward(ctx, ybar):
        L, W1, W2, R = ctx.saved_tensors
        opT = kernels.fused_two_site(*_transpose_operator(L, W1, W2, R))
        return kernels.apply_two_site_fused(ybar.contiguous(), *opT), None, None, None, None


def mpo_gradient(L: torch.Tensor, A: torch.Tensor, W: torch.Tensor, R: torch.Tensor) -> torch.Tensor:
    """dW[w,r,s,y] = sum L[p,w,q] conj(A[p,s,d]) A[q,y,e] R[d,r,e] for L (a, wl, a), A (a, d, c),
    R (c, wr, c): three K-major GEMMs (K = a, K = c, K = a c) and two layout permutations."""
    a, wl, _ = L.shape
    _, d, c = A.shape
    wr = R.shape[1]
    Ac = A.conj().resolve_conj() if A.is_complex() else A
    T1 = kernels.gemm_tn(L.reshape(a, wl * a), Ac.reshape(a, d * c).contiguous())   # [w q][s dd]
    Y = kernels.gemm_tn(A.reshape(a * d, c).t().contiguous(),
                        R.reshape(c * wr, c).t().contiguous())                       # [q y][dd r]
    X = T1.reshape(wl, a, d, c).permute(1, 3, 0, 2).reshape(a * c, wl * d).contiguous()
    Yk = Y.reshape(a, d, c, wr).permute(0, 2, 1, 3).reshape(a * c, d * wr).contiguous()
    dW = kernels.gemm_tn(X, Yk)                                                      # [w s][y r]
    return dW.reshape(wl, d, d, wr).permute(0, 3, 1, 2).contiguous()


def _identity_mpo_like(A: torch.Tensor) -> torch.Tensor:
    d = A.shape[1]
    return torch.eye(d, dtype=A.dtype, device=A.device).reshape(1, 1, d, d)


def _norm2(M: Sequence[torch.Tensor]) -> torch.Tensor:
    Ln = torch.ones((1, 1, 1), dtype=M[0].dtype, device=M[0].device)
    for A in M:
        Ln = kernels.update_left_env(Ln, A, _identity_mpo_like(A))
    return Ln.reshape(-1)[0]


class MPOExpectation(torch.autograd.Function):
    """E = <psi|H|psi> / <psi|psi> as a function of the MPO tensors, for a fixed MPS `M`."""

    @staticmethod
    def forward(ctx, M: List[torch.Tensor], *W):
        L = torch.ones((1, 1, 1), dtype=M[0].dtype, device=M[0].device)
        for A, Wi in zip(M, W):
            L = kernels.update_left_env(L, A, Wi.contiguous())
        n2 = _norm2(M)
        ctx.M = M
        ctx.save_for_backward(*W)
        return (L.reshape(-1)[0] / n2).real if L.is_complex() else L.reshape(-1)[0] / n2

    @staticmethod
    def backward(ctx, Ebar):
        M = ctx.M
        W = [w.contiguous() for w in ctx.saved_tensors]
        N = len(M)
        dev, dt = M[0].device, M[0].dtype
        R: List[torch.Tensor] = [None] * (N + 1)
        R[N] = torch.ones((1, 1, 1), dtype=dt, device=dev)
        for i in range(N - 1, 0, -1):
            R[i] = kernels.update_right_env(R[i + 1], M[i], W[i])
        scale = Ebar / _norm2(M)
        grads = []
        L = torch.ones((1, 1, 1), dtype=dt, device=dev)
        for i in range(N):
            grads.append(mpo_gradient(L, M[i], W[i], R[i + 1]) * scale)
            if i + 1 < N:
                L = kernels.update_left_env(L, M[i], W[i])
        return (None, *grads)


def field_gradient(mps: Sequence[torch.Tensor], g: torch.Tensor) -> torch.Tensor:
    """dE/dg_i of E = <psi|H(g)|psi> for the TFIM with per-site fields
    H = -sum X_i X_{i+1} - sum g_i Z_i (W_i[0, 2] = -g_i Z, tnkit/models.py:77-84), through
    `MPOExpectation`. `mps` is a device MPS (float64); returns a length-N float64 tensor."""
    dev = mps[0].device
    g = g.detach().to(dev, torch.float64).clone().requires_grad_(True)
    N = g.numel()
    I2 = torch.eye(2, dtype=torch.float64, device=dev)
    SX = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=torch.float64, device=dev)
    SZ = torch.tensor([[1.0, 0.0], [0.0, -1.0]], dtype=torch.float64, device=dev)
    Ws = []
    for i in range(N):
        base = torch.zeros((3, 3, 2, 2), dtype=torch.float64, device=dev)
        base[0, 0], base[0, 1], base[1, 2], base[2, 2] = I2, SX, -SX, I2
        field = torch.zeros((3, 3, 2, 2), dtype=torch.float64, device=dev)
        field[0, 2] = -SZ
        Wi = base + g[i] * field
        if i == 0:
            Wi = Wi[0:1]
        elif i == N - 1:
            Wi = Wi[:, 2:3]
        Ws.append(Wi)
    E = MPOExpectation.apply([t.contiguous() for t in mps], *Ws)
    (dg,) = torch.autograd.grad(E, g)
    return dg
