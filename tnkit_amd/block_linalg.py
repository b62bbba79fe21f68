"""Fusing, splitting, truncated SVD and LQ of U(1) block tensors on the HIP device.

The layouts and leg orders are tnkit's (`blocks.fuse_map`, `blocks.split_map`, and the new-leg
rule of tnkit/blocks.py:1380-1462), so the factors land in the same buffers tnkit's would. Data
moves through the library's gather kernels. The per-block factorisations are the dense ones
of this package: an SVD (rocSOLVER through torch) or the Gram-matrix route (`linalg`), one block
of one charge sector at a time; the cut across sectors is one host read of all singular values.
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple

import numpy as np
import torch

from . import blocks as B
from . import kernels
from .block_ops import BlockTensor, gather
from .tdvp import _eigh_robust


def fuse(x: BlockTensor, groups) -> Tuple[BlockTensor, B.FuseRecord]:
    """Fuse each group of legs into one leg of sign +1 (tnkit/blocks.py:822-825)."""
    groups = tuple(tuple(int(a) for a in g) for g in groups)
    rec = B.fuse_record(x.legs, x.charge, groups)
    out, g = B.fuse_map(x.struct, rec)
    return BlockTensor(out, gather(g, x.data)), rec


def split(y: BlockTensor, rec: B.FuseRecord, legs: Optional[Sequence[int]] = None
          ) -> BlockTensor:
    """Undo `fuse` for the fused legs `legs` (all by default) (tnkit/blocks.py:827-838)."""
    todo = tuple(range(len(rec.groups))) if legs is None else tuple(sorted(int(g) for g in legs))
    if legs is not None and rec.perm != tuple(range(len(rec.legs))):
        raise ValueError(f"a partial split needs an unpermuted fuse; this one used {rec.perm}")
    if y.ndim != len(rec.groups) or any(y.legs[g] != rec.fused[g] for g in todo):
        raise ValueError("the record does not describe this tensor's fused legs")
    out, g = B.split_map(y.struct, rec, todo)
    return BlockTensor(out, gather(g, y.data))


def scale(x: BlockTensor, s: torch.Tensor, axis: int) -> BlockTensor:
    """x with leg `axis` multiplied by the vector s (sector order) (tnkit/blocks.py:816-820)."""
    sv = gather(B.leg_positions(x.struct, axis), s.to(x.dtype).contiguous())
    return BlockTensor(x.struct, x.data * sv)


def norm(x) -> float:
    data = x.data if isinstance(x, BlockTensor) else x
    if data.dtype == torch.float64:
        return float(kernels.norm(data.reshape(-1).contiguous()).item())
    return float(torch.linalg.norm(data))


def _matrix_blocks(m: BlockTensor):
    """(row sector, column sector, R x C view) of a canonical two-leg block tensor."""
    s = m.struct
    for k, (r, c) in enumerate(s.blocks):
        R, C = (int(v) for v in s.shp[k])
        o = int(s.off[k])
        yield r, c, m.data[o:o + R * C].reshape(R, C)


def _eigh_desc(G: torch.Tensor):
    """(w, V) of one symmetric float64 Gram matrix, both descending: the library's Jacobi
    kernel (order <= 64) or divide-and-conquer solver (<= 2048), the vendor's above."""
    G = 0.5 * (G + G.t())
    n = int(G.shape[0])
    if n <= kernels.SYEV_MAX_N:
        w, V = kernels.eigh_top(G.contiguous(), n)
        return torch.flip(w, (0,)), V
    w, V = torch.linalg.eigh(G)
    return torch.flip(w, (0,)), torch.flip(V, (-1,))


def _svd(M: torch.Tensor, method: str):
    if method == "svd":
        return kernels.thin_svd(M)
    if method != "gram":
        raise ValueError(f"unknown truncation {method!r}; expected 'svd' or 'gram'")
    # Gram route (tnkit/blocks.py:1345-1377): eigh of the smaller Gram matrix, then a QR of the
    # projected factor re-orthonormalises the singular vectors
    R, C = M.shape
    if R >= C:
        _, V = _eigh_desc(kernels.gemm_tn(M.contiguous(), M.contiguous()))
        Q, Rr = torch.linalg.qr(M @ V)
    else:
        Mt = M.t().contiguous()
        _, U = _eigh_desc(kernels.gemm_tn(Mt, Mt))
        Q, Rr = torch.linalg.qr(M.t() @ U)
    d = Rr.diagonal()
    S = d.abs()
    Q = Q * (d / (S + (S == 0)))
    o = torch.argsort(S, descending=True)
    S, Q = S[o], Q[:, o]
    if R >= C:
        return Q, S, V[:, o].t()
    return U[:, o], S, Q.t()


def truncated_svd(m: BlockTensor, chi: int, method: str = "svd", zero_tol: float = 0.0):
    """m ~ U diag(S) Vh keeping the chi largest singular values across all sectors
    (tnkit/blocks.py:1380-1462 with svd_tol = degenerate_tol = 0, bucket = 1). zero_tol > 0
    drops singular values at or below zero_tol times the largest before the cap (at least one
    is kept).
    Returns (U, S, Vh, disc): U with legs (m.legs[0], k-), Vh with (k+, m.legs[1]), S a device
    vector over the new leg in sector order, disc the discarded weight sum(sigma_dropped^2).
    The new leg has one sector per block with kept values, labelled by the row sector's charge,
    in charge order."""
    if m.ndim != 2:
        raise ValueError(f"truncated_svd needs a two-leg tensor, got {m.ndim} legs")
    if chi < 1:
        raise ValueError(f"chi must be at least 1, got {chi}")
    m = m.canonical()
    lr, lc = m.legs
    parts = [(r, c) + tuple(_svd(M, method)) for r, c, M in _matrix_blocks(m)]
    if not parts:
        raise ValueError("the matrix has no allowed blocks")
    s_host = torch.cat([p[3] for p in parts]).cpu().numpy().astype(np.float64)
    owner = np.concatenate([np.full(len(p[3]), j) for j, p in enumerate(parts)])
    order = np.lexsort((np.arange(s_host.size), -s_host))
    live = s_host.size
    if zero_tol > 0:
        live = max(1, int((s_host > zero_tol * s_host.max()).sum()))
    k = min(int(chi), live)
    kept = np.bincount(owner[order[:k]], minlength=len(parts))
    disc = float((s_host[order[k:]] ** 2).sum())
    label = (lambda r: lr.charges[r]) if lr.sign > 0 else (lambda r: -lr.charges[r])
    new = sorted(((label(r), r, U[:, :n], S[:n], Vh[:n])
                  for (r, c, U, S, Vh), n in zip(parts, map(int, kept)) if n),
                 key=lambda x: (x[0], x[1]))
    bond = B.Leg(tuple(q for q, *_ in new), tuple(int(S.numel()) for _, _, _, S, _ in new), 1)
    U = torch.cat([u.reshape(-1) for _, _, u, _, _ in sorted(new, key=lambda x: x[1])])
    S = torch.cat([s for _, _, _, s, _ in new])
    Vh = torch.cat([v.reshape(-1) for _, _, _, _, v in new])
    return (BlockTensor(B.canonical((lr, bond.flip()), 0), U.contiguous()), S,
            BlockTensor(B.canonical((bond, lc), m.charge), Vh.contiguous()), disc)


def qr(m: BlockTensor, right: bool = False) -> Tuple[BlockTensor, BlockTensor]:
    """m = Q R per block (right=False) or m = L Q with orthonormal rows of Q (right=True); the new
    leg has one sector per block of size min(rows, cols), labelled as in `truncated_svd`
    (tnkit/blocks.py:1463-1493)."""
    if m.ndim != 2:
        raise ValueError(f"qr needs a two-leg tensor, got {m.ndim} legs")
    m = m.canonical()
    lr, lc = m.legs
    new = []
    for r, c, M in _matrix_blocks(m):
        if right:
            q, t = torch.linalg.qr(M.t())
            a, b = t.t(), q.t()
        else:
            a, b = torch.linalg.qr(M)
        new.append((lr.charges[r] if lr.sign > 0 else -lr.charges[r], r, a, b))
    if not new:
        raise ValueError("the matrix has no allowed blocks")
    new.sort(key=lambda x: x[0])
    bond = B.Leg(tuple(q for q, *_ in new), tuple(int(a.shape[1]) for _, _, a, _ in new), 1)
    A = torch.cat([a.reshape(-1) for _, _, a, _ in sorted(new, key=lambda x: x[1])])
    Bq = torch.cat([b.reshape(-1) for _, _, _, b in new])
    return (BlockTensor(B.canonical((lr, bond.flip()), 0), A.contiguous()),
            BlockTensor(B.canonical((bond, lc), m.charge), Bq.contiguous()))


def truncated_gram(m: BlockTensor, chi: int, iso: str = "left"):
    """`truncated_svd` through the Gram matrix of each block on the `iso` side: 'left' gives
    U = the eigenvectors of M M^T (an exact left isometry) and S Vh = U^T M; 'right' gives
    Vh = the eigenvectors of M^T M (an exact right isometry) and U S = M Vh^T. The per-block
    Gram matrices (K-major MFMA GEMMs) of order up to 64 are padded to a power-of-two size (at
    least 16) with distinct values in [-2, -1) on the padding diagonal, which keeps the padding
    apart from the nonnegative Gram spectrum and free of degeneracy, and diagonalised in batched
    Jacobi calls; larger ones go one by one to the library's divide-and-conquer solver. One host
    read of all eigenvalues makes the global cut (chi largest, as `truncated_svd`). S is taken as the norms of the carried
    factor, as tnkit's dense Gram truncation does (tnkit/linalg.py:138-204). Same return
    values and new-leg rule as `truncated_svd`."""
    if iso not in ("left", "right"):
        raise ValueError(f"iso must be 'left' or 'right', got {iso!r}")
    if m.ndim != 2 or m.dtype != torch.float64:
        raise ValueError("truncated_gram needs a two-leg float64 block tensor")
    m = m.canonical()
    lr, lc = m.legs
    blocks = list(_matrix_blocks(m))
    if not blocks:
        raise ValueError("the matrix has no allowed blocks")
    dev = m.data.device
    mats, grams = [], []
    for r, c, M in blocks:
        Mt = M.t().contiguous()
        mats.append((M, Mt))
        grams.append(kernels.gemm_tn(Mt, Mt) if iso == "left" else kernels.gemm_tn(M, M))
    sizes = [g.shape[0] for g in grams]
    # batched eigh per power-of-two size class (classes up to 64 run on the Jacobi kernel)
    cls = {}
    for j, n in enumerate(sizes):
        cls.setdefault(max(16, 1 << (int(n) - 1).bit_length()), []).append(j)
    w_all = [None] * len(grams)
    v_all = [None] * len(grams)
    for nmax, idx in cls.items():
        if nmax > kernels.EIGH_JACOBI_MAX_N:
            # one divide-and-conquer solve per block (the library's solver, no padding)
            for j in idx:
                w_all[j], v_all[j] = _eigh_desc(grams[j])
            continue
        Gp = torch.zeros((len(idx), nmax, nmax), dtype=torch.float64, device=dev)
        diag = (-(1.0 + torch.arange(nmax, dtype=torch.float64, device=dev) / nmax))[None, :] \
            .repeat(len(idx), 1)
        for t, j in enumerate(idx):
            n = sizes[j]
            Gp[t, :n, :n] = grams[j]
            diag[t, :n] = 0.0
        w, V = _eigh_robust(Gp, diag)
        for t, j in enumerate(idx):
            n = sizes[j]
            # the n real eigenpairs are the largest n (the padding sits below -1), descending
            w_all[j] = torch.flip(w[t, nmax - n:], (0,))
            v_all[j] = torch.flip(V[t, :n, nmax - n:], (1,))
    w_host = torch.cat(w_all).cpu().numpy()
    s_host = np.sqrt(np.clip(w_host, 0.0, None))
    owner = np.concatenate([np.full(n, j) for j, n in enumerate(sizes)])
    order = np.lexsort((np.arange(s_host.size), -s_host))
    k = min(int(chi), s_host.size)
    kept = np.bincount(owner[order[:k]], minlength=len(blocks))
    parts = []
    for j, ((r, c, _), (M, Mt), n) in enumerate(zip(blocks, mats, kept)):
        n = int(n)
        if n == 0:
            continue
        Q = v_all[j][:, :n].contiguous()                   # exact isometry, (rows or cols) x n
        if iso == "left":
            carried = kernels.gemm_tn(Q, M)                # S Vh = U^T M   (n x C)
            S = torch.linalg.norm(carried, dim=1)
            Vh = carried / torch.where(S > 0, S, torch.ones_like(S))[:, None]
            parts.append((r, Q, S, Vh))
        else:
            carried = kernels.gemm_tn(Mt, Q)               # U S = M Vh^T   (R x n)
            S = torch.linalg.norm(carried, dim=0)
            U = carried / torch.where(S > 0, S, torch.ones_like(S))[None, :]
            parts.append((r, U, S, Q.t()))
    disc = float((s_host[order[k:]] ** 2).sum())
    label = (lambda r: lr.charges[r]) if lr.sign > 0 else (lambda r: -lr.charges[r])
    new = sorted(((label(r), r, U, S, Vh) for r, U, S, Vh in parts), key=lambda x: (x[0], x[1]))
    bond = B.Leg(tuple(q for q, *_ in new), tuple(int(S.numel()) for _, _, _, S, _ in new), 1)
    U = torch.cat([u.reshape(-1) for _, _, u, _, _ in sorted(new, key=lambda x: x[1])])
    S = torch.cat([s for _, _, _, s, _ in new])
    Vh = torch.cat([v.reshape(-1) for _, _, _, _, v in new])
    return (BlockTensor(B.canonical((lr, bond.flip()), 0), U.contiguous()), S,
            BlockTensor(B.canonical((bond, lc), m.charge), Vh.contiguous()), disc)
