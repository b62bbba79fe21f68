"""U(1) block-sparse layouts by brute force (TEST INFRASTRUCTURE ONLY).

Only `tests/`, `__graft_entry__.smoke()` and `bench.py`'s `cpu_baseline` leg use this module.

A leg is (charges, dims, sign); a block is one sector index per leg, allowed when the signed
charges sum to the total charge. tnkit stores every allowed block in lexicographic sector order,
row-major, at the exclusive prefix sum of block sizes (tnkit/blocks.py:220-253); the dense array
has every leg in sector order (tnkit/blocks.py:762-773). Here both are enumerated element by
element, without pruning or vectorisation, so the result does not share code with the product's
maps; they are pinned against tnkit's recorded layouts in tests/test_block_maps.py.
"""
from __future__ import annotations

import itertools

import numpy as np


def allowed_blocks(legs, charge=0):
    """Every allowed block in lexicographic sector order (full product, filtered)."""
    ranges = [range(len(cs)) for cs, _, _ in legs]
    return [t for t in itertools.product(*ranges)
            if sum(sg * cs[s] for (cs, _, sg), s in zip(legs, t)) == charge]


def canonical_layout(legs, charge=0):
    """(blocks, offsets, shapes, size) of the canonical layout."""
    blocks = allowed_blocks(legs, charge)
    shapes = [tuple(dims[s] for (_, dims, _), s in zip(legs, b)) for b in blocks]
    off, pos = [], 0
    for shp in shapes:
        off.append(pos)
        pos += int(np.prod(shp, dtype=np.int64))
    return blocks, off, shapes, pos


def to_dense(legs, data, blocks, off, strides):
    """The dense array (legs in sector order) of a block buffer with the given layout."""
    starts = [np.concatenate([[0], np.cumsum(dims)[:-1]]).astype(int) for _, dims, _ in legs]
    x = np.zeros(tuple(sum(dims) for _, dims, _ in legs), dtype=np.asarray(data).dtype)
    for b, o, st in zip(blocks, off, strides):
        shp = tuple(legs[i][1][s] for i, s in enumerate(b))
        for idx in np.ndindex(*shp):
            pos = int(o) + sum(int(i) * int(t) for i, t in zip(idx, st))
            x[tuple(starts[k][s] + i for k, (s, i) in enumerate(zip(b, idx)))] = data[pos]
    return x


def canonical_to_dense(legs, data, charge=0):
    blocks, off, shapes, _ = canonical_layout(legs, charge)
    strides = []
    for shp in shapes:
        st = [1] * len(shp)
        for j in range(len(shp) - 2, -1, -1):
            st[j] = st[j + 1] * shp[j + 1]
        strides.append(st)
    return to_dense(legs, data, blocks, off, strides)


def dense_to_canonical(x, legs, charge=0):
    """The canonical buffer of a dense array (entries outside allowed blocks dropped)."""
    blocks, off, shapes, size = canonical_layout(legs, charge)
    starts = [np.concatenate([[0], np.cumsum(dims)[:-1]]).astype(int) for _, dims, _ in legs]
    out = np.zeros(size, dtype=x.dtype)
    for b, o, shp in zip(blocks, off, shapes):
        sl = tuple(slice(starts[k][s], starts[k][s] + n) for k, (s, n) in enumerate(zip(b, shp)))
        out[o:o + int(np.prod(shp))] = x[sl].reshape(-1)
    return out
