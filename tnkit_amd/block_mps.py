"""Block MPOs, block MPSs and block environments on the HIP device (U(1)).

* `mpo_charges`, `to_block_mpo`: the channel charges q(wr) = q(wl) + q(s) - q(s') of a dense MPO
  and its block form with legs (wl+, wr-, s+, s'-) (tnkit/block_backend.py:147-256).
* `random_block_mps`: tnkit's random start state (tnkit/block_backend.py:259-311), same bond
  sectors, same dimensions, same random numbers for the same generator.
* `update_left_env`, `update_right_env`, `edge_envs`: the environment updates of
  tnkit/_dmrg.py:216-247 as three block contractions each (grouped GEMMs).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import blocks as B
from .block_ops import BlockTensor, einsum


def _live_channels(mpo: Sequence[np.ndarray], tol: float):
    """Per MPO bond (len(mpo) + 1 of them), a boolean mask of the channels that carry part of
    the operator: those joined by entries above tol both to some left channel of the first
    tensor and to some right channel of the last (tnkit/block_backend.py:162-173). Reachability
    is propagated with 0/1 matrix products of the per-site channel adjacency."""
    adj = [(np.abs(np.asarray(W)) > tol).any(axis=(2, 3)).astype(np.int64) for W in mpo]
    ahead = [np.ones(adj[0].shape[0], np.int64)]
    for A in adj:
        ahead.append((ahead[-1] @ A > 0).astype(np.int64))
    behind = [np.ones(adj[-1].shape[1], np.int64)]
    for A in adj[::-1]:
        behind.insert(0, (A @ behind[0] > 0).astype(np.int64))
    return [(f > 0) & (b > 0) for f, b in zip(ahead, behind)]


def mpo_charges(mpo: Sequence[np.ndarray], site_charges, tol: float = 0.0,
                left: Optional[Sequence[int]] = None) -> List[np.ndarray]:
    """The U(1) charge of every channel of every MPO bond (len(mpo) + 1 int64 arrays): an entry
    W[i, j, s, s'] above tol between live channels sends charge q(i) + q(s) - q(s') to right
    channel j, and every live channel must receive exactly one value; channels that are not
    live get 0. `left` gives the first bond's charges (default 0). Raises ValueError when a
    channel would carry two charges (tnkit/block_backend.py:176-208)."""
    live = _live_channels(mpo, tol)
    q = (np.zeros(np.shape(mpo[0])[0], np.int64) if left is None
         else np.asarray(left, np.int64))
    charges = [q]
    for site, (W, qs) in enumerate(zip(mpo, site_charges)):
        W = np.asarray(W)
        qs = np.asarray([int(x) for x in qs], np.int64)
        i, j, s, t = np.nonzero(np.abs(W) > tol)
        on = live[site][i] & live[site + 1][j]
        sent = q[i[on]] + qs[s[on]] - qs[t[on]]
        dest = j[on]
        nxt = np.zeros(W.shape[1], np.int64)
        for ch in np.flatnonzero(live[site + 1]):
            vals = np.unique(sent[dest == ch])
            if vals.size != 1:
                raise ValueError(f"MPO site {site}, right channel {ch}: charges "
                                 f"{vals.tolist()} (an operator that mixes charges?)")
            nxt[ch] = vals[0]
        charges.append(nxt)
        q = nxt
    return charges


def from_dense(x: np.ndarray, legs: Sequence[B.Leg], charge: int = 0,
               orders: Optional[Sequence[Optional[np.ndarray]]] = None):
    """(canonical buffer, squared norm of the dropped entries) of a dense host array whose leg i
    is in dense order orders[i] (tnkit/blocks.py:841-861)."""
    x = np.asarray(x)
    for i, order in enumerate(orders or ()):
        if order is not None:
            x = np.take(x, np.asarray(order), axis=i)
    s = B.canonical(tuple(legs), int(charge))
    starts = [np.concatenate([[0], np.cumsum(leg.dims)[:-1]]).astype(np.int64) for leg in legs]
    data = np.empty(s.size, dtype=x.dtype)
    kept = np.zeros(x.shape, bool)
    for k, b in enumerate(s.blocks):
        sl = tuple(slice(starts[i][sec], starts[i][sec] + legs[i].dims[sec])
                   for i, sec in enumerate(b))
        n = int(np.prod(s.shp[k]))
        data[s.off[k]:s.off[k] + n] = x[sl].reshape(-1)
        kept[sl] = True
    rest = x[~kept]
    return data, float(np.vdot(rest, rest).real)


def to_block_mpo(mpo: Sequence[np.ndarray], site_charges, device, dtype=torch.float64,
                 tol: float = 0.0):
    """A dense MPO (wl, wr, s, s') as block tensors with legs (wl+, wr-, s+, s'-), and for each
    site the dense-index orders of those legs (tnkit/block_backend.py:211-256)."""
    chans = mpo_charges(mpo, site_charges, tol)
    out, orders = [], []
    for n, W in enumerate(mpo):
        cl, ol = B.Leg.from_charges(chans[n])
        cr, orr = B.Leg.from_charges(chans[n + 1])
        p, op = B.Leg.from_charges(site_charges[n])
        legs = (cl, cr.flip(), p, p.flip())
        data, _ = from_dense(np.asarray(W), legs, 0, (ol, orr, op, op))
        out.append(BlockTensor(B.canonical(legs, 0), torch.from_numpy(data).to(device, dtype)))
        orders.append((ol, orr, op, op))
    return out, orders


def _state_counts(phys: Sequence[B.Leg], first: int, sign: int):
    """Per prefix of the sites (0..len(phys) of them): (lowest charge, counts) where counts[k]
    is the number of basis states of those sites whose charges, each times `sign`, add to
    first + k. Python-int counts (object arrays): they reach 2^N."""
    lo, cnt = first, np.array([1], dtype=object)
    out = [(lo, cnt)]
    for leg in phys:
        qs = [sign * c for c in leg.charges]
        hist = np.zeros(max(qs) - min(qs) + 1, dtype=object)
        hist[:] = 0
        for c, d in zip(qs, leg.dims):
            hist[c - min(qs)] += d
        lo, cnt = lo + min(qs), np.convolve(cnt, hist)
        out.append((lo, cnt))
    return out


def _count(table, c: int) -> int:
    lo, cnt = table
    k = c - lo
    return int(cnt[k]) if 0 <= k < len(cnt) else 0


def _trim(dims: List[int], chi: int) -> List[int]:
    """Bring sum(dims) down to chi by taking one state at a time from the largest entry (the
    first one on ties). Closed form: cap every entry at the lowest level T whose capping
    removes no more than the excess, then take one more state from each of the first r entries
    left at T, r being what the capping did not remove."""
    excess = sum(dims) - chi
    if excess <= 0:
        return list(dims)
    removed = lambda level: sum(max(x - level, 0) for x in dims)  # noqa: E731
    level = max(dims)
    while removed(level - 1) <= excess:
        level -= 1
    r = excess - removed(level)
    out = []
    for x in dims:
        y = min(x, level)
        if y == level and r > 0:
            y, r = y - 1, r - 1
        out.append(y)
    return out


def random_block_mps(mpo: Sequence[BlockTensor], chi: int, charge: int,
                     rng: Optional[np.random.Generator] = None, device=None,
                     dtype=torch.float64) -> List[BlockTensor]:
    """tnkit's random block MPS (tnkit/block_backend.py:259-311), same bond sectors, dimensions
    and random numbers for the same generator. Bond n offers the sectors that site n - 1's
    charges carry the kept sectors of bond n - 1 into (bond 0: charge 0), among those from which
    the right end can still reach `charge`. A sector's weight w is the smaller of its basis-state
    counts to the left and to the right; the chi heaviest are kept (smaller charge first on
    ties) and share chi states in proportion to w, at least 1 and at most w each, then trimmed
    to chi (`_trim`). Sectors that no charge of the next site carries into a kept sector of the
    next bond are dropped, right to left. Entries are standard normal, one draw of the whole
    canonical buffer per site tensor, site by site."""
    rng = np.random.default_rng(0) if rng is None else rng
    device = device if device is not None else mpo[0].data.device
    phys = [W.legs[2] for W in mpo]
    N, target = len(phys), int(charge)
    from_left = _state_counts(phys, 0, 1)
    from_right = _state_counts(phys[::-1], target, -1)[::-1]
    if _count(from_left[N], target) == 0:
        raise ValueError(f"no state of these sites has total charge {target}")
    bonds: List[Dict[int, int]] = []
    offered = [0]
    for n in range(N + 1):
        weight = {}
        for c in offered:
            r = _count(from_right[n], c)
            if r:
                weight[c] = min(_count(from_left[n], c), r)
        kept = sorted(weight, key=lambda c: (-weight[c], c))[:chi]
        total = sum(weight[c] for c in kept)
        share = [min(weight[c], max(1, int(chi * weight[c] / total))) for c in kept]
        bonds.append(dict(zip(kept, _trim(share, chi))))
        if n < N:
            offered = sorted({c + a for c in kept for a in phys[n].charges})
    for n in range(N - 1, -1, -1):
        nxt = bonds[n + 1]
        bonds[n] = {c: d for c, d in bonds[n].items()
                    if any(c + a in nxt for a in phys[n].charges)}
    legs = [B.Leg(tuple(sorted(b)), tuple(b[c] for c in sorted(b)), 1) for b in bonds]
    out = []
    for n, p in enumerate(phys):
        s = B.canonical((legs[n], p, legs[n + 1].flip()), 0)
        data = torch.from_numpy(rng.standard_normal(s.size)).to(device, dtype)
        out.append(BlockTensor(s, data))
    return out


def update_left_env(L: BlockTensor, A: BlockTensor, W: BlockTensor) -> BlockTensor:
    """L' = sum L A* W A past A (a, s, c) (tnkit/_dmrg.py:216-222), real data."""
    t1 = einsum("pwq,psd->wqsd", L, A.conj())
    t2 = einsum("wqsd,wrsy->qdry", t1, W)
    return einsum("qdry,qye->dre", t2, A)


def update_right_env(R: BlockTensor, Bt: BlockTensor, W: BlockTensor) -> BlockTensor:
    """R' = sum R B* W B past B (u, s, v) (tnkit/_dmrg.py:225-231), real data."""
    t1 = einsum("xry,usx->ryus", R, Bt.conj())
    t2 = einsum("ryus,wrsp->yuwp", t1, W)
    return einsum("yuwp,vpy->uwv", t2, Bt)


def edge_envs(M: Sequence[BlockTensor], mpo: Sequence[BlockTensor]) -> Tuple[BlockTensor,
                                                                             BlockTensor]:
    """The one-element boundary environments with the legs of the outer bonds and MPO channels
    (tnkit/_dmrg.py:234-247)."""
    dev, dt = M[0].data.device, M[0].dtype
    a, b = M[0].legs[0], M[-1].legs[2]
    La = (a, mpo[0].legs[0].flip(), a.flip())
    Rb = (b, mpo[-1].legs[1].flip(), b.flip())
    one = lambda legs: BlockTensor(B.canonical(legs, 0),
                                   torch.ones(B.canonical(legs, 0).size, dtype=dt, device=dev))
    for legs in (La, Rb):
        if B.canonical(legs, 0).size != 1:
            raise ValueError("the outer bonds and MPO channels must be one-dimensional")
    return one(La), one(Rb)
