"""Local subspace expansion of two-site DMRG on the HIP device (tnkit/expansion.py:173-213,
after Hubig, McCulloch, Schollwoeck and Wolf, Phys. Rev. B 91, 155115 (2015)).

The optimised centre x is split at the bond the sweep crosses, with the new bond basis the `chi`
leading eigenvectors of x x^H + alpha^2 P P^H on the kept legs, where P is H_eff applied to x on
the kept side only (tnkit/_dmrg.py:700-705) less its part in the span of x. The states x holds
keep their weights; P competes only for the rest, so the basis gains the directions the
Hamiltonian couples to, and on block tensors the charge sectors that x lacks.

Dense tensors are torch tensors: their contractions and products run on the library's K-major
GEMM (`contract.einsum`, `linalg._mm`), the thin SVD of the centre on the Jacobi kernel where it
fits one workgroup (rocSOLVER above), and the density matrix's leading eigenvectors on the
library's eigensolvers (`linalg._eig_desc`: Jacobi to order 64, divide and conquer for real
orders to 2048). Block tensors are `block_ops.BlockTensor`, whose contractions run as grouped
GEMM launches of libtnkit_hip and whose factorisations are `block_linalg.truncated_svd` with
tnkit's zero_tol rule.
"""
from __future__ import annotations

from typing import Tuple

import torch

from . import block_linalg as BL
from . import contract, kernels
from .block_ops import BlockTensor, einsum as beinsum
from .linalg import _eig_desc, _kmajor, _mm


def _top(mat: torch.Tensor, k: int, rel_tol: float):
    """U, S, Vh of the k largest singular values above rel_tol times the largest, at least one
    (tnkit/expansion.py:35-50, dense branch)."""
    U, S, Vh = kernels.thin_svd(mat)
    s = S.cpu()
    n = min(int(k), max(1, int((s > rel_tol * s[0]).sum())))
    return U[:, :n], S[:n], Vh[:n]


def _top_psd(rho: torch.Tensor, k: int, rel_tol: float):
    """The k leading eigenvectors of a Hermitian positive semidefinite rho whose eigenvalues
    exceed rel_tol times the largest, at least one: the leading left singular vectors of rho,
    as `_top` takes them, from its eigendecomposition (eigenvalues by magnitude, as the
    singular values of a Hermitian matrix are)."""
    w, V = _eig_desc(0.5 * (rho + rho.mH))
    a = w.abs()
    order = torch.argsort(a, descending=True)
    s = a[order].cpu()
    n = min(int(k), max(1, int((s > rel_tol * s[0]).sum())))
    return V[:, order[:n]]


def _hmm(X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """X^H @ Y on the library's K-major GEMM."""
    return kernels.gemm_tn(_kmajor(X), Y.contiguous())


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
        E = E - _mm(U, _hmm(U, E))
        Xt, Et = X.t().contiguous(), E.t().contiguous()
        rho = _hmm(Xt, Xt) + alpha ** 2 * _hmm(Et, Et)             # X X^T + a^2 E E^T
        Ur = _top_psd(rho, chi, rel_tol ** 2)
        A = Ur.reshape(a, s1, -1)
        C = _hmm(Ur, X).reshape(-1, s2, b)
    else:
        _, _, Vh = _top(X, min(X.shape), rel_tol)
        E = extra.reshape(-1, s2 * b)
        E = E - _mm(_mm(E, Vh.t()), Vh)
        rho = _hmm(X, X) + alpha ** 2 * _hmm(E, E)                 # X^T X + a^2 E^T E
        Vr = _top_psd(rho, chi, rel_tol ** 2).t()
        A = Vr.reshape(-1, s2, b)
        C = _mm(X, Vr.t()).reshape(a, s1, -1)
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
        t = contract.einsum("pwq,qijb->pwijb", L, x)
        return contract.einsum("wmsi,pwijb->psmjb", W1, t)
    t = contract.einsum("drb,astb->astdr", R, x)
    return contract.einsum("wrjt,astdr->aswjd", W2, t)
