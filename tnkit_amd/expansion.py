"""Local subspace expansion of two-site DMRG on the HIP device (tnkit/expansion.py:173-213,
after Hubig, McCulloch, Schollwoeck and Wolf, Phys. Rev. B 91, 155115 (2015)).

The optimised centre x is split at the bond the sweep crosses, with the new bond basis the `chi`
leading eigenvectors of x x^H + alpha^2 P P^H on the kept legs, where P is H_eff applied to x on
the kept side only (tnkit/_dmrg.py:700-705) less its part in the span of x. The states x holds
keep their weights; P competes only for the rest, so the basis gains the directions the
Hamiltonian couples to, and on block tensors the charge sectors that x lacks.

Dense tensors are torch tensors (the factorisations are rocSOLVER through torch); block tensors
are `block_ops.BlockTensor`, whose contractions run as grouped GEMM launches of libtnkit_hip and
whose factorisations are `block_linalg.truncated_svd` with tnkit's zero_tol rule.
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import block_linalg as BL
from . import kernels
from .block_ops import BlockTensor, einsum as beinsum


def _top(mat: torch.Tensor, k: int, rel_tol: float):
    """U, S, Vh of the k largest singular values above rel_tol times the largest, at least one
    (tnkit/expansion.py:35-50, dense branch)."""
    U, S, Vh = kernels.thin_svd(mat)
    s = S.cpu()
    n = min(int(k), max(1, int((s > rel_tol * s[0]).sum())))
    return U[:, :n], S[:n], Vh[:n]


def _axpy(x: BlockTensor, y: BlockTensor, a: float) -> BlockTensor:
    """x + a y for block tensors on the same legs and charge, in the canonical layout."""
    x, y = x.canonical(), y.canonical()
    if x.legs != y.legs or x.charge != y.charge:
        raise ValueError("block tensors on different legs or charges")
    return x.with_data(x.data + a * y.data)


def _split_dense(x, extra, alpha, chi, right, rel_tol):
    a, s1, s2, b = (int(v) for v in x.shape)
    X = x.reshape(a * s1, s2 * b)
    if right:
        U, _, _ = _top(X, min(X.shape), rel_tol)
        E = extra.reshape(a * s1, -1)
        E = E - U @ (U.t() @ E)
        rho = X @ X.t() + alpha ** 2 * (E @ E.t())
        Ur, _, _ = _top(rho, chi, rel_tol ** 2)
        A = Ur.reshape(a, s1, -1)
        C = (Ur.t() @ X).reshape(-1, s2, b)
    else:
        _, _, Vh = _top(X, min(X.shape), rel_tol)
        E = extra.reshape(-1, s2 * b)
        E = E - (E @ Vh.t()) @ Vh
        rho = X.t() @ X + alpha ** 2 * (E.t() @ E)
        _, _, Vr = _top(rho, chi, rel_tol ** 2)
        A = Vr.reshape(-1, s2, b)
        C = (X @ Vr.t()).reshape(a, s1, -1)
    nx, nc = float(torch.linalg.norm(x)), float(torch.linalg.norm(C))
    return A.contiguous(), C.contiguous(), max(nx * nx - nc * nc, 0.0)


def _split_block(x: BlockTensor, extra: BlockTensor, alpha, chi, right, rel_tol):
    groups = ((0, 1), (2, 3))
    mat, rec = BL.fuse(x, groups)
    U, _, Vh, _ = BL.truncated_svd(mat, min(mat.shape), zero_tol=rel_tol)
    if right:
        Q = BL.split(U, rec, legs=(0,))
        c = beinsum("abK,abcde->Kcde", Q.conj(), extra)
        extra = _axpy(extra, beinsum("abK,Kcde->abcde", Q, c), -1.0)
        rho = _axpy(beinsum("abcd,ABcd->abAB", x, x.conj()),
                    beinsum("abcde,ABcde->abAB", extra, extra.conj()), alpha ** 2)
    else:
        Q = BL.split(Vh, rec, legs=(1,))
        c = beinsum("abcde,Kde->abcK", extra, Q.conj())
        extra = _axpy(extra, beinsum("abcK,Kde->abcde", c, Q), -1.0)
        rho = _axpy(beinsum("abCD,abcd->CDcd", x.conj(), x),
                    beinsum("abcDE,abcde->DEde", extra.conj(), extra), alpha ** 2)
    m2, rec2 = BL.fuse(rho, groups)
    U, _, Vh, _ = BL.truncated_svd(m2, chi, zero_tol=rel_tol ** 2)
    if right:
        A = BL.split(U, rec2, legs=(0,))
        C = beinsum("abK,abcd->Kcd", A.conj(), x)
    else:
        A = BL.split(Vh, rec2, legs=(1,))
        C = beinsum("abcd,Kcd->abK", x, A.conj())
    nx, nc = BL.norm(x), BL.norm(C)
    return A, C, max(nx * nx - nc * nc, 0.0)


def enriched_split(x, extra, alpha: float, chi: int, be=None, right: bool = True,
                   rel_tol: float = 1e-14) -> Tuple[object, object, float]:
    """Split the centre x (a, s1, s2, b) into an isometry and the part that moves on
    (tnkit/expansion.py:173-213). right=True: A (a, s1, k) spans the first two legs and
    C = A^H x (k, s2, b); right=False: A (k, s2, b) spans the last two and C = x A^H
    (a, s1, k). `extra` has the same kept legs as x. Returns (A, C, |x|^2 - |C|^2 clipped at 0).
    Singular values of x below rel_tol times the largest, and eigenvalues of the density below
    rel_tol^2 times the largest, are dropped. `be` is accepted for tnkit's signature."""
    if isinstance(x, BlockTensor):
        return _split_block(x, extra, alpha, chi, right, rel_tol)
    return _split_dense(x, extra, alpha, chi, right, rel_tol)


def expansion_term(x, L, W1, W2, R, right: bool):
    """P of tnkit/_dmrg.py:700-705: H_eff applied to x on the kept side only,
    (a, s1, w, s2, b) with w the MPO channel between the two sites."""
    if isinstance(x, BlockTensor):
        if right:
            return beinsum("wmsi,pwijb->psmjb", W1, beinsum("pwq,qijb->pwijb", L, x))
        return beinsum("wrjt,astdr->aswjd", W2, beinsum("drb,astb->astdr", R, x))
    if right:
        t = torch.einsum("pwq,qijb->pwijb", L, x)
        return torch.einsum("wmsi,pwijb->psmjb", W1, t)
    t = torch.einsum("drb,astb->astdr", R, x)
    return torch.einsum("wrjt,astdr->aswjd", W2, t)
