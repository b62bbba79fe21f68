"""The Krylov exponential on U(1) block tensors (tnkit/block_lanczos.py:76-121).

tnkit runs the exponential's vector algebra on the flat buffers of BlockTensors sharing one
layout, with alpha and beta read once per iteration. Here the Krylov basis is the flat
canonical buffer on the device (`eigensolver.lanczos_expm_flat`, one device step and one host
read per iteration), and the block operator, whose grouped kernels are real, is applied to the
real and imaginary parts of each complex basis vector. That is exact for the real symmetric
block Hamiltonians the block path builds (U(1) XXZ), and it is the same number of real
multiply-adds as a complex product with a real operator.
"""
from __future__ import annotations

import torch

from . import eigensolver
from .block_ops import BlockTensor

__all__ = ["lanczos_expm"]


def lanczos_expm(matvec, v: BlockTensor, tau, ops=None, m=30, tol=1e-12, adaptive=None,
                 info=None, est_ref_j=2) -> BlockTensor:
    """exp(tau H) v for a BlockTensor v: `matvec` maps a real BlockTensor in v's canonical
    layout to a BlockTensor of the same charge sector (returned in any layout; it is brought to
    canonical order). Returns a complex128 BlockTensor in the canonical layout; `info` as in
    `eigensolver.lanczos_expm`."""
    v = v.canonical()
    s = v.struct

    def real_mv(part: torch.Tensor) -> torch.Tensor:
        y = matvec(BlockTensor(s, part.contiguous()))
        if y.struct is not s:
            y = y.canonical()
        if y.struct is not s:
            raise ValueError("block lanczos_expm: the operator changes the block layout")
        return y.data.to(torch.float64)

    def cmv(x: torch.Tensor) -> torch.Tensor:
        return torch.complex(real_mv(x.real), real_mv(x.imag))

    if v.data.is_complex():
        x = v.data.to(torch.complex128)
    else:
        x = v.data.to(torch.float64).to(torch.complex128)
    out = eigensolver.lanczos_expm_flat(cmv, x, tau, m=m, tol=tol, adaptive=adaptive, info=info,
                                        est_ref_j=est_ref_j)
    return BlockTensor(s, out)
