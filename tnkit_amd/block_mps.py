"""Block MPOs, block MPSs and block environments on the HIP device (U(1)).

* `mpo_charges`, `to_block_mpo`: the channel charges q(wr) = q(wl) + q(s) - q(s') of a dense MPO
  and its block form with legs (wl+, wr-, s+, s'-) (tnkit/block_backend.py:147-256).
* `random_block_mps`: tnkit's random start state (tnkit/block_backend.py:259-311), same bond
  sectors, same dimensions, same random numbers for the same generator.
* `update_left_env`, `update_right_env`, `edge_envs`: the environment updates of
  tnkit/_dmrg.py:216-247 as three block contractions each (grouped GEMMs).
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import blocks as B
from .block_ops import BlockTensor, einsum


def _live_channels(mpo: Sequence[np.ndarray], tol: float):
    """Per MPO bond, the channels on a path of entries above tol from the first tensor's left
    channels to the last tensor's right channels (tnkit/block_backend.py:147-159)."""
    nz = [(np.abs(np.asarray(W)) > tol).any(axis=(2, 3)) for W in mpo]
    fwd = [np.ones(nz[0].shape[0], bool)]
    for m in nz:
        fwd.append(m[fwd[-1]].any(0))
    bwd = [np.ones(nz[-1].shape[1], bool)]
    for m in reversed(nz):
        bwd.append(m[:, bwd[-1]].any(1))
    return [f & b for f, b in zip(fwd, reversed(bwd))]


def mpo_charges(mpo: Sequence[np.ndarray], site_charges, tol: float = 0.0,
                left: Optional[Sequence[int]] = None) -> List[np.ndarray]:
    """The U(1) charge of every channel of every MPO bond (len(mpo) + 1 arrays): q(wr) = q(wl) +
    q(s) - q(s') over the nonzero live entries; dead channels get charge 0. Raises when a channel
    would carry two charges (tnkit/block_backend.py:161-208)."""
    live = _live_channels(mpo, tol)
    out = [np.zeros(np.shape(mpo[0])[0], np.int64) if left is None
           else np.asarray(left, np.int64)]
    for n, (W, q) in enumerate(zip(mpo, site_charges)):
        W, q = np.asarray(W), [int(x) for x in q]
        a, b, s, t = np.nonzero(np.abs(W) > tol)
        on = live[n][a] & live[n + 1][b]
        found = {int(j): set() for j in np.flatnonzero(live[n + 1])}
        for i, j, x, y in zip(a[on], b[on], s[on], t[on]):
            found[int(j)].add(int(out[-1][i]) + q[x] - q[y])
        qr = np.zeros(W.shape[1], np.int64)
        for j, vals in found.items():
            if len(vals) != 1:
                raise ValueError(f"MPO site {n}, right channel {j}: charges {sorted(vals)} (an "
                                 "operator that mixes charges?)")
            qr[j] = vals.pop()
        out.append(qr)
    return out


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


def random_block_mps(mpo: Sequence[BlockTensor], chi: int, charge: int,
                     rng: Optional[np.random.Generator] = None, device=None,
                     dtype=torch.float64) -> List[BlockTensor]:
    """tnkit's random block MPS (tnkit/block_backend.py:259-311): bond n keeps the sectors
    reachable from the left and, backwards, from `charge` at the right end, each with states in
    proportion to w = min(basis states left, basis states right) (at least 1, at most w); with
    more than chi candidates the largest w are kept; sectors without a kept successor are
    dropped. Entries are standard normal, drawn from `rng` block tensor by block tensor in
    canonical order."""
    rng = np.random.default_rng(0) if rng is None else rng
    device = device if device is not None else mpo[0].data.device
    phys = [W.legs[2] for W in mpo]
    target = int(charge)

    def grow(counts, leg, sign):
        out = {}
        for c, n in counts.items():
            for a, d in zip(leg.charges, leg.dims):
                c2 = c + (a if sign > 0 else -a)
                out[c2] = out.get(c2, 0) + n * d
        return out

    left, right = [{0: 1}], [{target: 1}]
    for p in phys:
        left.append(grow(left[-1], p, 1))
    for p in reversed(phys):
        right.append(grow(right[-1], p, -1))
    right.reverse()
    if target not in left[-1]:
        raise ValueError(f"no state of these sites has total charge {target}")
    bonds, cand = [], {0}
    for n, (lc, rc) in enumerate(zip(left, right)):
        w = {c: min(lc[c], rc[c]) for c in cand if c in rc}
        keep = sorted(w, key=lambda c: (-w[c], c))[:chi]
        tot = sum(w[c] for c in keep)
        dims = {c: min(w[c], max(1, int(chi * w[c] / tot))) for c in keep}
        while sum(dims.values()) > chi:
            dims[max(dims, key=dims.get)] -= 1
        bonds.append(dims)
        if n < len(phys):
            cand = {c + a for c in keep for a in phys[n].charges}
    for n in range(len(phys) - 1, -1, -1):
        prv = {c - a for c in bonds[n + 1] for a in phys[n].charges}
        bonds[n] = {c: d for c, d in bonds[n].items() if c in prv}
    legs = []
    for b in bonds:
        cs = sorted(b)
        legs.append(B.Leg(tuple(cs), tuple(b[c] for c in cs), 1))
    out = []
    for n, p in enumerate(phys):
        s = B.canonical((legs[n], p, legs[n + 1].flip()), 0)
        out.append(BlockTensor(s, torch.from_numpy(rng.standard_normal(s.size)).to(device,
                                                                                  dtype)))
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
