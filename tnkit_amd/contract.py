"""tnkit's einsum contractions on libtnkit_hip kernels: the `einsum` of `HipBackend`.

tnkit's drivers reach every tensor contraction through `be.einsum(sub, *ops)` with a fixed set of
two-operand equations (tnkit/_dmrg.py:165-231, :677; tnkit/batch_tdvp.py:211-243, :535-623;
tnkit/tdvp.py:590-910) and reach the Krylov projections through matrix-vector equations
(tnkit/eigensolver.py:160-161, :567, :588-595, :634, :672). Here an equation is classified once
(cached) and run as:

* two operands, each label in at most one axis of an operand and every label either in both
  operands or in the output: one K-major GEMM (`tnh_gemm_tn`), batched over the labels common
  to both operands and the output. Each operand is laid out as (batch, contracted, free) with
  one permuting copy when it is not already in that order; the (batch, free_a, free_b) result is
  returned as a permuted view in the requested output order.
* a matrix and a vector sharing one label: `tnh_rows_dot` (the vector contracts the matrix's
  last axis) or `tnh_rows_combine` (its first axis); both stream the matrix once.
* one operand whose output is a permutation of its labels: a permuted view.

Anything else (three or more operands, traces, diagonals, sums over a label of one operand) goes
to torch.einsum, as tnkit's torch backend does for every equation (tnkit/backend.py:588-590).
"""
from __future__ import annotations

from functools import lru_cache
from typing import Optional, Tuple

import torch

from . import kernels

_GEMM_DTYPES = (torch.float64, torch.float32, torch.complex128, torch.complex64)


@lru_cache(maxsize=512)
def _plan(sub: str, ndims: Tuple[int, ...]):
    """('perm', perm) | ('rows_dot', mat_first) | ('rows_combine', vec_first) |
    ('gemm', (perm_a, perm_b, nbatch, nk, n_fa, n_fb, out_perm)) | ('torch', None)."""
    s = sub.replace(" ", "")
    if "->" not in s or "." in s:
        return "torch", None
    lhs, out = s.split("->")
    terms = lhs.split(",")
    if len(terms) != len(ndims) or any(len(t) != n for t, n in zip(terms, ndims)):
        return "torch", None
    if any(len(set(t)) != len(t) for t in terms) or len(set(out)) != len(out):
        return "torch", None
    if len(terms) == 1:
        (t,) = terms
        if sorted(t) != sorted(out):
            return "torch", None
        return "perm", tuple(t.index(c) for c in out)
    if len(terms) != 2:
        return "torch", None
    ta, tb = terms
    if set(out) - (set(ta) | set(tb)):
        return "torch", None
    for c in set(ta) ^ set(tb):                 # a label of one operand must survive
        if c not in out:
            return "torch", None
    # matrix . vector on one shared label, the vector's label summed
    if {len(ta), len(tb)} == {1, 2}:
        mat, vec, mat_first = (ta, tb, True) if len(ta) == 2 else (tb, ta, False)
        v = vec[0]
        if v not in out and v in mat:
            keep = mat.replace(v, "")
            if out == keep:
                return ("rows_dot", mat_first) if mat[1] == v else ("rows_combine", mat_first)
    batch = [c for c in ta if c in tb and c in out]
    contr = [c for c in ta if c in tb and c not in out]
    fa = [c for c in ta if c not in tb]
    fb = [c for c in tb if c not in ta]
    perm_a = tuple(ta.index(c) for c in batch + contr + fa)
    perm_b = tuple(tb.index(c) for c in batch + contr + fb)
    natural = batch + fa + fb
    out_perm = tuple(natural.index(c) for c in out)
    return "gemm", (perm_a, perm_b, len(batch), len(contr), len(fa), len(fb), out_perm)


def _on_device(ops) -> bool:
    return all(isinstance(o, torch.Tensor) and o.is_cuda for o in ops)


def _layout(x: torch.Tensor, perm) -> torch.Tensor:
    if tuple(perm) != tuple(range(x.dim())):
        x = x.permute(*perm)
    x = x.contiguous()
    return x.resolve_conj() if x.is_conj() else x


def einsum(sub: str, *ops: torch.Tensor) -> torch.Tensor:
    """`torch.einsum(sub, *ops)` for HIP tensors, on libtnkit_hip kernels where the equation is
    one of the forms above."""
    if not _on_device(ops):
        return torch.einsum(sub, *ops)
    kind, arg = _plan(sub, tuple(o.dim() for o in ops))
    if kind == "perm":
        return ops[0].permute(*arg)
    if kind == "torch" or any(o.numel() == 0 for o in ops):
        return torch.einsum(sub, *ops)
    dt = torch.result_type(ops[0], ops[1])
    if dt not in _GEMM_DTYPES:
        return torch.einsum(sub, *ops)
    a, b = (o if o.dtype == dt else o.to(dt) for o in ops)
    if kind == "rows_dot":
        mat, vec = (a, b) if arg else (b, a)
        return kernels.rows_dot(_layout(mat, (0, 1)), _layout(vec, (0,)))
    if kind == "rows_combine":
        mat, vec = (a, b) if arg else (b, a)
        return kernels.rows_combine(_layout(vec, (0,)), _layout(mat, (0, 1)))
    perm_a, perm_b, nb, nk, nfa, nfb, out_perm = arg
    A = _layout(a, perm_a)
    B = _layout(b, perm_b)
    bshape = tuple(A.shape[:nb])
    kshape = tuple(A.shape[nb:nb + nk])
    if tuple(B.shape[:nb]) != bshape or tuple(B.shape[nb:nb + nk]) != kshape:
        raise ValueError(f"einsum {sub}: operand shapes {tuple(a.shape)} and {tuple(b.shape)} "
                         "do not agree")
    fa_shape = tuple(A.shape[nb + nk:])
    fb_shape = tuple(B.shape[nb + nk:])
    Z, K = _prod(bshape), _prod(kshape)
    M, N = _prod(fa_shape), _prod(fb_shape)
    C = kernels.gemm_tn_batched(A.reshape(Z, K, M), B.reshape(Z, K, N))
    C = C.reshape(bshape + fa_shape + fb_shape)
    return C.permute(*out_perm) if out_perm != tuple(range(C.dim())) else C


def _prod(shape) -> int:
    n = 1
    for s in shape:
        n *= int(s)
    return n


def plan_of(sub: str, ndims: Tuple[int, ...]) -> Optional[str]:
    """Which route an equation takes ('gemm', 'rows_dot', 'rows_combine', 'perm', 'torch')."""
    return _plan(sub, tuple(ndims))[0]
