"""Two-site DMRG on U(1) block MPOs and MPSs, on the HIP device.

The sweep of tnkit/_dmrg.py:560-799 taking its `block` branches: a random block MPS (`block_mps.random_block_mps`) made right-canonical by an LQ per
block (tnkit/_dmrg.py:295-306), then per bond theta = M[i] M[i+1], the lowest eigenpair of the
block H_eff (`block_ops.BlockTwoSite`, all sector GEMMs of each contraction in one grouped
launch) by thick-restart or mixed-precision Lanczos on the flat canonical buffer
(tnkit/_dmrg.py:638-644), the truncated SVD across sectors, and the block environment update.
Subspace expansion runs in the opening sweeps as in tnkit (by default min(4, n_sweeps - 2) of
them on block MPOs): the centre is split by `expansion.enriched_split`, which fills charge
sectors the bond lacks.
"""
from __future__ import annotations

import time
from typing import Any, Dict, List, Optional, Sequence

import numpy as np
import torch

from . import block_linalg as BL
from . import block_mps as BM
from .block_ops import BlockTensor, BlockTwoSite, einsum
from .expansion import enriched_split, expansion_term
from ._dmrg import expand_sweeps
from .eigensolver import (check_eig_options, lanczos_ground, lanczos_ground_mixed,
                          lanczos_ground_restart)
from .mps import MPS


def right_canonicalize(M: List[BlockTensor]) -> List[BlockTensor]:
    """Every tensor right-canonical by an LQ per block, centre at site 0, the running centre
    renormalised at each step (tnkit/_dmrg.py:295-306)."""
    for i in range(len(M) - 1, 0, -1):
        f, rec = BL.fuse(M[i], ((0,), (1, 2)))
        Lm, Q = BL.qr(f, right=True)
        M[i], Lm = BL.split(Q, rec, legs=(1,)), BL.split(Lm, rec, legs=(0,))
        nxt = einsum("asb,bc->asc", M[i - 1], Lm)
        M[i - 1] = nxt.with_data(nxt.data / BL.norm(nxt))
    M[0] = M[0].with_data(M[0].data / BL.norm(M[0]))
    return M


def _check(o, mpo) -> None:
    if o.sites != 2:
        raise ValueError(f"sites must be 2 on the HIP backend, got {o.sites!r}")
    if o.eig not in ("lanczos", "lanczos-restart", "lanczos-mixed"):
        raise ValueError(f"unknown eig solver {o.eig!r}; expected one of "
                         "['lanczos', 'lanczos-mixed', 'lanczos-restart']")
    check_eig_options(o)
    if o.trunc not in ("svd", "gram"):
        raise ValueError(f"unknown trunc {o.trunc!r}; expected 'svd' or 'gram'")
    if (o.charge is None) == (o.init_mps is None):
        raise ValueError("a block MPO needs exactly one of charge= (the total charge of the state "
                         "sought) and a block init_mps (which fixes the charge)")
    if o.init_mps is not None and not getattr(o.init_mps[0], "is_block", False):
        raise ValueError("init_mps for a block MPO must hold block tensors")
    if any(W.dtype != torch.float64 for W in mpo):
        raise TypeError("the HIP block sweep takes float64 block MPOs")


class BlockDMRGEngine:
    """State of one block two-site DMRG run: MPS, environments, MPO, counters."""

    def __init__(self, mpo: Sequence[BlockTensor], chi: int, o):
        _check(o, mpo)
        self.o, self.chi, self.N = o, int(chi), len(mpo)
        self.mpo = list(mpo)
        if o.init_mps is not None:
            self.M = right_canonicalize([t for t in o.init_mps])
        else:
            self.M = right_canonicalize(BM.random_block_mps(self.mpo, chi, o.charge,
                                                            np.random.default_rng(o.seed)))
        self.rng = np.random.default_rng(o.seed + 1)
        self.mixed_counts = {"lo": 0, "hi": 0}
        N = self.N
        self.L: List[Optional[BlockTensor]] = [None] * (N + 1)
        self.R: List[Optional[BlockTensor]] = [None] * (N + 1)
        self.L[0], self.R[N] = BM.edge_envs(self.M, self.mpo)
        for i in range(N - 1, 0, -1):
            self.R[i] = BM.update_right_env(self.R[i + 1], self.M[i], self.mpo[i])
        self.total_disc = 0.0
        self.matvec_calls = 0
        self.phase = {"eig": 0.0, "trunc": 0.0, "env": 0.0}

    def eigsolve(self, theta0: BlockTensor, op: BlockTwoSite, op_lo, cur_tol: float):
        o = self.o
        x0 = theta0.canonical()
        count = {"n": 0}

        def matvec(x):
            count["n"] += 1
            return op.apply(x)

        def matvec_lo(x):
            count["n"] += 1
            return op_lo.apply(x)

        n = x0.struct.size
        v0 = x0.data if o.warm_start else torch.from_numpy(
            self.rng.standard_normal(n)).to(x0.data.device)
        mi = o.max_iter if o.max_iter is not None else min(n, max(200, 4 * self.chi))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if o.eig == "lanczos-mixed":
            e, x, _ = lanczos_ground_mixed(matvec_lo, matvec, v0.clone(), tol=cur_tol,
                                           max_iter=mi, m=o.restart_m, k=o.restart_k,
                                           tol_lo=o.mixed_tol_lo, counts=self.mixed_counts)
        elif o.eig == "lanczos":
            e, x, _ = lanczos_ground(matvec, v0.clone(), tol=cur_tol, max_iter=mi,
                                     reorth=o.reorth, stop=o.eig_stop)
        else:
            e, x, _ = lanczos_ground_restart(matvec, v0.clone(), tol=cur_tol, max_iter=mi,
                                             m=o.restart_m, k=o.restart_k)
        torch.cuda.synchronize()
        self.phase["eig"] += time.perf_counter() - t0
        self.matvec_calls += count["n"]
        return e, BlockTensor(x0.struct, x.reshape(-1).contiguous())

    def solve_block(self, i: int, direction: str, cur_tol: float, alpha: float = 0.0) -> float:
        o, M, L, R = self.o, self.M, self.L, self.R
        W1, W2 = self.mpo[i], self.mpo[i + 1]
        theta0 = einsum("asc,ctb->astb", M[i], M[i + 1])
        op = BlockTwoSite(L[i], W1, W2, R[i + 2], theta_legs=theta0.legs,
                          charge=theta0.charge)
        op_lo = (BlockTwoSite(L[i], W1, W2, R[i + 2], theta_legs=theta0.legs,
                              charge=theta0.charge, dtype=torch.float32)
                 if o.eig == "lanczos-mixed" else None)
        e, x = self.eigsolve(theta0, op, op_lo, cur_tol)
        del op, op_lo
        t0 = time.perf_counter()
        if alpha:
            right = direction == "right"
            P = expansion_term(x, L[i], W1, W2, R[i + 2], right)
            A, C, disc = enriched_split(x, P, alpha, self.chi, None, right)
            C = C.with_data(C.data / BL.norm(C))
            M[i], M[i + 1] = (A, C) if right else (C, A)
            torch.cuda.synchronize()
            self.phase["trunc"] += time.perf_counter() - t0
            self.total_disc += disc
            self._env(i, direction)
            return e
        mat, rec = BL.fuse(x, ((0, 1), (2, 3)))
        if o.trunc == "gram":
            U, S, Vh, disc = BL.truncated_gram(mat, self.chi,
                                               "left" if direction == "right" else "right")
        else:
            U, S, Vh, disc = BL.truncated_svd(mat, self.chi)
        S = S / torch.linalg.norm(S)
        if direction == "right":
            M[i] = BL.split(U, rec, legs=(0,))
            M[i + 1] = BL.split(BL.scale(Vh, S, 0), rec, legs=(1,))
        else:
            M[i] = BL.split(BL.scale(U, S, 1), rec, legs=(0,))
            M[i + 1] = BL.split(Vh, rec, legs=(1,))
        torch.cuda.synchronize()
        self.phase["trunc"] += time.perf_counter() - t0
        self.total_disc += disc
        self._env(i, direction)
        return e

    def _env(self, i: int, direction: str) -> None:
        t0 = time.perf_counter()
        if direction == "right":
            self.L[i + 1] = BM.update_left_env(self.L[i], self.M[i], self.mpo[i])
        else:
            self.R[i + 1] = BM.update_right_env(self.R[i + 2], self.M[i + 1], self.mpo[i + 1])
        torch.cuda.synchronize()
        self.phase["env"] += time.perf_counter() - t0

    def sweep(self, cur_tol: float, alpha: float = 0.0) -> float:
        E = None
        for i in range(0, self.N - 1):
            E = self.solve_block(i, "right", cur_tol, alpha)
        for i in range(self.N - 2, -1, -1):
            E = self.solve_block(i, "left", cur_tol, alpha)
        return E


def block_dmrg(mpo: Sequence[BlockTensor], chi: int, o) -> tuple:
    """The sweep loop of tnkit/_dmrg.py:767-799 on block tensors; returns (E, MPS, info)."""
    expand = expand_sweeps(o, True)
    eng = BlockDMRGEngine(mpo, chi, o)
    energies: List[float] = []
    sweep_s: List[float] = []
    alphas: List[float] = []
    prev_E, E, sweep = None, None, -1
    for sweep in range(o.n_sweeps):
        t0 = time.perf_counter()
        cur_tol = (o.eig_tol_warmup if (o.eig_tol_warmup is not None and sweep < o.warmup_sweeps)
                   else o.eig_tol)
        alpha = o.expand_alpha * o.expand_decay ** sweep if sweep < expand else 0.0
        alphas.append(alpha)
        E = eng.sweep(cur_tol, alpha)
        energies.append(E)
        torch.cuda.synchronize()
        sweep_s.append(time.perf_counter() - t0)
        if o.verbose:
            print(f"  sweep {sweep + 1:2d}: E = {E:.12f}  (chi={chi}, blocks, hip)")
        if prev_E is not None and abs(E - prev_E) < max(o.tol, o.rtol * abs(E)):
            break
        prev_E = None if alpha else E
    info: Dict[str, Any] = {
        "sweeps": sweep + 1, "energies": energies, "backend": "hip-real", "blocks": True,
        "eig": o.eig, "discarded_weight": eng.total_disc, "matvec_calls": eng.matvec_calls,
        "phase_time": eng.phase, "mixed_counts": eng.mixed_counts, "sites": 2,
        "expand": expand, "expand_alphas": alphas, "sweep_s": sweep_s}
    return E, MPS(eng.M, center=0), info
