"""MPS and MPO files: npz archives that load without executing anything (allow_pickle=False).

tnkit moves states as pickles of `MPS` (tnkit/mps.py:95-96, `__reduce__` returns the tensor list
and the centre) and converts block states through dense arrays
(tnkit/block_backend.py:462-466). Pickles run code when loaded, so this package writes a plain
numeric archive instead, and `MPS.__reduce__` stays tnkit's for in-process copies.

Layout of an archive (every entry a numeric array):
  kind          0 = dense MPS, 1 = block MPS, 2 = dense MPO, 3 = block MPO
  n             number of sites;  center  the orthogonality centre, -1 if unknown
  t{i}          dense site tensor i, in its own dtype and shape
  For block tensors, per site i: d{i} the flat canonical buffer, q{i} the total charge and, per
  leg k, l{i}_{k}_c / l{i}_{k}_d / l{i}_{k}_s the leg's charges, sector sizes and sign. The
  buffer is tnkit's canonical layout (lexicographic sector order, row-major blocks), so it is
  interchangeable with tnkit's `BlockTensor.data` for the same legs.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from .mps import MPS

_KINDS = {"dense-mps": 0, "block-mps": 1, "dense-mpo": 2, "block-mpo": 3}


def _is_block(t) -> bool:
    return getattr(t, "is_block", False)


def _host(t) -> np.ndarray:
    if hasattr(t, "detach"):
        return np.ascontiguousarray(t.detach().cpu().numpy())
    return np.ascontiguousarray(np.asarray(t))


def _save(path, tensors: Sequence, kind: str, center: Optional[int]) -> None:
    block = _is_block(tensors[0])
    if any(_is_block(t) != block for t in tensors):
        raise TypeError("a file holds either dense or block tensors, not both")
    out = {"kind": np.array(_KINDS[("block-" if block else "dense-") + kind]),
           "n": np.array(len(tensors)), "center": np.array(-1 if center is None else int(center))}
    for i, t in enumerate(tensors):
        if block:
            out[f"d{i}"] = _host(t.canonical().data)
            out[f"q{i}"] = np.array(int(t.charge))
            for k, leg in enumerate(t.legs):
                out[f"l{i}_{k}_c"] = np.asarray(leg.charges, dtype=np.int64)
                out[f"l{i}_{k}_d"] = np.asarray(leg.dims, dtype=np.int64)
                out[f"l{i}_{k}_s"] = np.array(int(leg.sign))
        else:
            out[f"t{i}"] = _host(t)
    np.savez(path, **out)


def _load(path, kind: str, device):
    with np.load(path, allow_pickle=False) as z:
        code = int(z["kind"])
        block = code in (1, 3)
        if code != _KINDS[("block-" if block else "dense-") + kind]:
            raise ValueError(f"{path}: not an {kind.upper()} archive (kind code {code})")
        n, center = int(z["n"]), int(z["center"])
        tensors = []
        for i in range(n):
            if block:
                import torch
                from . import blocks as B
                from .block_ops import BlockTensor
                legs = []
                k = 0
                while f"l{i}_{k}_c" in z:
                    legs.append(B.Leg(tuple(int(v) for v in z[f"l{i}_{k}_c"]),
                                      tuple(int(v) for v in z[f"l{i}_{k}_d"]),
                                      int(z[f"l{i}_{k}_s"])))
                    k += 1
                s = B.canonical(tuple(legs), int(z[f"q{i}"]))
                dev = device if device is not None else torch.device("cuda")
                tensors.append(BlockTensor(s, torch.from_numpy(np.array(z[f"d{i}"])).to(dev)))
            else:
                a = np.array(z[f"t{i}"])
                if device is not None:
                    import torch
                    a = torch.from_numpy(a).to(device)
                tensors.append(a)
    return tensors, (None if center < 0 else center)


def save_mps(path, mps) -> None:
    """Write a dense or U(1) block MPS (device or host tensors) and its centre."""
    _save(path, list(mps), "mps", getattr(mps, "center", None))


def load_mps(path, device=None) -> MPS:
    """Read an MPS archive. Dense tensors come back as numpy arrays, or torch tensors on
    `device`; block tensors as `BlockTensor`s on `device` (default the current HIP device)."""
    tensors, center = _load(path, "mps", device)
    return MPS(tensors, center=center)


def save_mpo(path, mpo) -> None:
    """Write a dense or U(1) block MPO, axes (wl, wr, s_out, s_in)."""
    _save(path, list(mpo), "mpo", None)


def load_mpo(path, device=None) -> list:
    """Read an MPO archive (as `load_mps`)."""
    return _load(path, "mpo", device)[0]
