"""Derivatives of the two-site path behind torch.autograd.Function, with libtnkit_hip kernels in
both directions.

tnkit has no autograd (its arrays are numpy or plain torch without gradients); these Functions
are this package's extension for the Hellmann-Feynman study of the field derivatives
dE/dg_i = <dH/dg_i> = -<Z_i> of the TFIM with per-site fields.

* `HeffMatvec`: y = H_eff theta (tnkit/_dmrg.py:183-189). Backward: theta_bar = H_eff^H y_bar, the
  same kernels on the adjoint operator (bra and ket of L and R swapped, W transposed in its
  physical indices, all conjugated), so no intermediate is stored; the gradients of L, W1, W2
  and R are chains of K-major GEMMs with the MPO mix in between.
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


def _kmajor(t: torch.Tensor, k: int) -> torch.Tensor:
    """t (K, M) as given when k == 0, or t (M, K) transposed to (K, M), contiguous."""
    return (t if k == 0 else t.t()).contiguous()


def _operand_grads(z, x, L, W1, W2, R, need):
    """dF/dL, dF/dW1, dF/dW2, dF/dR of F = sum z . (H_eff(L, W1, W2, R) x), for z = conj(ybar),
    each None unless `need` asks for it. With W12[w,r,S,S'] = sum_m W1[w,m,s,s'] W2[m,r,t,t']
    (S = (s t)) the three are
        dL[p,w,q]  = sum z[p,S,b'] W12[w,r,S,S'] R[b',r,b] x[q,S',b]
        dR[b',r,b] = sum z[p,S,b'] L[p,w,q] W12[w,r,S,S'] x[q,S',b]
        dW12       = sum z[p,S,b'] L[p,w,q] R[b',r,b] x[q,S',b]
    and each is a chain of K-major GEMMs (K = a or b, then the small MPO mix, then K = d^2 a or
    d^2 b); dW1 and dW2 follow from dW12 by the other factor."""
    a, d1, d2, b = x.shape
    wl, wm, wr = W1.shape[0], W1.shape[1], W2.shape[1]
    S = d1 * d2
    W12 = torch.einsum("wmsu,mrtv->wrstuv", W1, W2).reshape(wl, wr, S, S)
    xm, zm = x.reshape(a, S, b), z.reshape(a, S, b)
    gL = gR = gW12 = None
    if need[0]:
        XR = kernels.gemm_tn(_kmajor(xm.reshape(a * S, b), 1),
                             _kmajor(R.reshape(b * wr, b), 1))           # [(q S')][(b' r)]
        XR = XR.reshape(a, S, b, wr).permute(1, 3, 0, 2).reshape(S * wr, a * b).contiguous()
        Wk = W12.permute(3, 1, 0, 2).reshape(S * wr, wl * S).contiguous()  # [(S' r)][(w S)]
        Y = kernels.gemm_tn(XR, Wk).reshape(a, b, wl, S)                   # [(q b')][(w S)]
        Yk = Y.permute(3, 1, 0, 2).reshape(S * b, a * wl).contiguous()     # [(S b')][(q w)]
        gL = kernels.gemm_tn(_kmajor(zm.reshape(a, S * b), 1), Yk)         # [p][(q w)]
        gL = gL.reshape(a, a, wl).permute(0, 2, 1).contiguous()
    if need[1] or need[2] or need[3]:
        LX = kernels.gemm_tn(_kmajor(L.reshape(a * wl, a), 1),
                             xm.reshape(a, S * b).contiguous())          # [(p w)][(S' b)]
        LX = LX.reshape(a, wl, S, b)
        if need[3]:
            LXk = LX.permute(1, 2, 0, 3).reshape(wl * S, a * b).contiguous()  # [(w S')][(p b)]
            Wk = W12.permute(0, 3, 2, 1).reshape(wl * S, S * wr).contiguous()  # [(w S')][(S r)]
            Y2 = kernels.gemm_tn(LXk, Wk).reshape(a, b, S, wr)             # [(p b)][(S r)]
            Y2k = Y2.permute(0, 2, 3, 1).reshape(a * S, wr * b).contiguous()  # [(p S)][(r b)]
            gR = kernels.gemm_tn(zm.reshape(a * S, b).contiguous(), Y2k).reshape(b, wr, b)
        if need[1] or need[2]:
            ZR = kernels.gemm_tn(_kmajor(zm.reshape(a * S, b), 1),
                                 R.reshape(b, wr * b).contiguous())        # [(p S)][(r b)]
            ZRk = ZR.reshape(a, S, wr, b).permute(0, 3, 1, 2).reshape(a * b, S * wr).contiguous()
            LXk = LX.permute(0, 3, 1, 2).reshape(a * b, wl * S).contiguous()  # [(p b)][(w S')]
            G = kernels.gemm_tn(LXk, ZRk).reshape(wl, S, S, wr)            # [(w S')][(S r)]
            gW12 = G.permute(0, 3, 2, 1).reshape(wl, wr, d1, d2, d1, d2)
    gW1 = gW2 = None
    if need[1]:
        gW1 = torch.einsum("wrstuv,mrtv->wmsu", gW12, W2)
    if need[2]:
        gW2 = torch.einsum("wrstuv,wmsu->mrtv", gW12, W1)
    return gL, gW1, gW2, gR


class HeffMatvec(torch.autograd.Function):
    """y = H_eff theta (environments and MPO as inputs). Backward: theta_bar = H_eff^H y_bar by
    the forward kernels on the adjoint operator, and, for the inputs that require them, the
    gradients of L, W1, W2 and R (torch's convention: the conjugate of dF/dT for
    F = sum conj(y_bar) y, which is linear in each)."""

    @staticmethod
    def forward(ctx, theta, L, W1, W2, R):
        theta = theta.contiguous()
        ctx.save_for_backward(theta, L, W1, W2, R)
        op = kernels.fused_two_site(L, W1, W2, R)
        return kernels.apply_two_site_fused(theta, *op)

    @staticmethod
    def backward(ctx, ybar):
        theta, L, W1, W2, R = ctx.saved_tensors
        ybar = ybar.contiguous()
        gth = None
        if ctx.needs_input_grad[0]:
            gth = kernels.matvec2_bwd(L.contiguous(), W1.contiguous(), W2.contiguous(),
                                      R.contiguous(), ybar)
        need = ctx.needs_input_grad[1:]
        if not any(need):
            return gth, None, None, None, None
        z = ybar.conj().resolve_conj() if ybar.is_complex() else ybar
        gL, gW1, gW2, gR = _operand_grads(z, theta, L.contiguous(), W1.contiguous(),
                                          W2.contiguous(), R.contiguous(),
                                          (need[0], need[1], need[2], need[3]))
        if ybar.is_complex():
            gL, gW1, gW2, gR = (g.conj().resolve_conj() if g is not None else None
                                for g in (gL, gW1, gW2, gR))
        return gth, gL, gW1, gW2, gR


def _cj(t: torch.Tensor) -> torch.Tensor:
    return t.conj().resolve_conj() if t.is_complex() else t


class LeftEnvUpdate(torch.autograd.Function):
    """L' = update_left_env(L, A, W) (tnkit/_dmrg.py:216-222). Backward: the gradient of L by
    `tnh_env_left_bwd` (a right-environment update of the cotangent) and of W by
    `mpo_gradient`; A is treated as fixed (its gradient is not provided and raises)."""

    @staticmethod
    def forward(ctx, L, A, W):
        ctx.save_for_backward(L, A, W)
        return kernels.update_left_env(L.contiguous(), A.contiguous(), W.contiguous())

    @staticmethod
    def backward(ctx, gout):
        L, A, W = ctx.saved_tensors
        if ctx.needs_input_grad[1]:
            raise NotImplementedError("LeftEnvUpdate: the gradient of the MPS tensor is not "
                                      "provided")
        gout = gout.contiguous()
        gL = kernels.env_left_bwd(gout, A.contiguous(), W.contiguous()) \
            if ctx.needs_input_grad[0] else None
        gW = None
        if ctx.needs_input_grad[2]:
            # dF/dW = sum L conj(A) A conj(gout); torch's gradient is its conjugate
            gW = _cj(mpo_gradient(L.contiguous(), A.contiguous(), W, _cj(gout)))
        return gL, None, gW


class RightEnvUpdate(torch.autograd.Function):
    """R' = update_right_env(R, B, W) (tnkit/_dmrg.py:225-231), the mirror of LeftEnvUpdate:
    the gradient of R by `tnh_env_right_bwd`, of W by `mpo_gradient` on the mirrored network."""

    @staticmethod
    def forward(ctx, R, B, W):
        ctx.save_for_backward(R, B, W)
        return kernels.update_right_env(R.contiguous(), B.contiguous(), W.contiguous())

    @staticmethod
    def backward(ctx, gout):
        R, B, W = ctx.saved_tensors
        if ctx.needs_input_grad[1]:
            raise NotImplementedError("RightEnvUpdate: the gradient of the MPS tensor is not "
                                      "provided")
        gout = gout.contiguous()
        gR = kernels.env_right_bwd(gout, B.contiguous(), W.contiguous()) \
            if ctx.needs_input_grad[0] else None
        gW = None
        if ctx.needs_input_grad[2]:
            # R'[u,w,v] = sum R[x,r,y] conj(B[u,s,x]) W[w,r,s,p] B[v,p,y]: dF/dW is the MPO
            # gradient with the cotangent as the left environment and R as the right one
            gW = _cj(mpo_gradient(_cj(gout), B.contiguous(), W, R.contiguous()))
        return gR, None, gW


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
            gi = mpo_gradient(L, M[i], W[i], R[i + 1])
            # torch's gradient of the real E is the conjugate of dE/dW for complex tensors
            grads.append((gi.conj().resolve_conj() if gi.is_complex() else gi) * scale)
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
