"""U(1) block-sparse layouts and contraction plans: the host-side integer maps of the block path.

A block tensor stores every allowed block (one sector per leg, signed charges summing to the
tensor's total charge) in one flat device buffer; where each block sits is host metadata in a
`Struct`. The maps here are those of tnkit's block layer, so a buffer written by this package and
one written by tnkit hold the same numbers at the same positions:

* `canonical(legs, charge)`: allowed blocks in lexicographic sector order, each packed row-major
  at the exclusive prefix sum of the block sizes (tnkit/blocks.py:220-253).
* `plan(sa, sb, ax_a, ax_b)`: per contracted charge (taken in ascending order) one GEMM
  (pa, pb, po, R, K, C) between the stacked matrices of both operands, the output layout made of
  strided views into the GEMM results, and, per operand, the gather that stacks it
  (tnkit/blocks.py:952-1100).

What differs is how the device consumes the maps. tnkit gathers both operands into (R, K) and
(K, C) matrices with an expanded position index and then issues one matmul per charge. Here a
gather is a list of runs (arithmetic progressions, destination-contiguous) that one kernel walks
without an index array, operand a is stacked K-major with padded leading dimensions, and all
charge groups of a contraction run in one persistent grouped-GEMM launch (`tnkit_amd.block_ops`).
The reference layout of every gather (`Gather.expand`) is kept for parity checks.

Only U(1) is supported: labels are Python ints, fused by addition, sorted ascending.
"""
from __future__ import annotations

import itertools
import weakref
from dataclasses import dataclass
from functools import lru_cache
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

__all__ = ["Leg", "Struct", "canonical", "permuted", "dual", "plan", "Plan", "Gather",
           "box_runs", "expand_runs"]


@dataclass(frozen=True)
class Leg:
    """The sectors of one leg: distinct U(1) charges in ascending order, their dimensions, and the
    sign (+1 or -1) with which the leg enters the conservation rule (tnkit/blocks.py:79-118)."""
    charges: Tuple[int, ...]
    dims: Tuple[int, ...]
    sign: int = 1

    def __post_init__(self):
        cs = [int(c) for c in self.charges]
        if (cs != sorted(set(cs)) or len(self.dims) != len(cs) or any(d < 1 for d in self.dims)
                or self.sign not in (1, -1)):
            raise ValueError(f"invalid leg {self}")
        object.__setattr__(self, "charges", tuple(cs))
        object.__setattr__(self, "dims", tuple(int(d) for d in self.dims))

    @property
    def size(self) -> int:
        return sum(self.dims)

    def flip(self) -> "Leg":
        return Leg(self.charges, self.dims, -self.sign)

    @classmethod
    def from_charges(cls, q, sign: int = 1) -> Tuple["Leg", np.ndarray]:
        """The leg whose dense index i has charge q[i], and the dense indices in sector order
        (stable within a sector)."""
        q = np.asarray(q, dtype=np.int64)
        c, n = np.unique(q, return_counts=True)
        return cls(tuple(int(x) for x in c), tuple(int(x) for x in n), sign), \
            np.argsort(q, kind="stable")


def _row_major_rows(shp: np.ndarray) -> np.ndarray:
    """Row-major element strides of every row of a (boxes, axes) array of shapes."""
    st = np.ones_like(shp)
    for j in range(shp.shape[1] - 2, -1, -1):
        st[:, j] = st[:, j + 1] * shp[:, j + 1]
    return st


class Struct:
    """Where each block sits: block k (one sector index per leg, `blocks[k]`) starts at `off[k]`
    with element strides `st[k]` and shape `shp[k]`. One object per layout content while it
    lives (`_intern`), so plans can be cached by identity."""
    __slots__ = ("legs", "charge", "blocks", "off", "st", "shp", "size", "_index", "__weakref__")

    def __init__(self, legs, charge, blocks, off, st, shp, size):
        self.legs, self.charge, self.blocks = tuple(legs), int(charge), tuple(blocks)
        m = len(self.legs)
        self.off = np.asarray(off, np.int64).reshape(len(self.blocks))
        self.st = np.asarray(st, np.int64).reshape(len(self.blocks), m)
        self.shp = np.asarray(shp, np.int64).reshape(len(self.blocks), m)
        self.size = int(size)
        self._index = None

    @property
    def index(self) -> Dict[Tuple[int, ...], int]:
        if self._index is None:
            self._index = {b: k for k, b in enumerate(self.blocks)}
        return self._index

    @property
    def labels(self):
        """Everything but the sector sizes: the key of symbolic plans."""
        return (tuple((leg.charges, leg.sign) for leg in self.legs), self.charge, self.blocks)

    @property
    def shape(self) -> Tuple[int, ...]:
        return tuple(leg.size for leg in self.legs)

    def __repr__(self):
        return f"Struct(legs={len(self.legs)}, blocks={len(self.blocks)}, size={self.size})"


_LAYOUTS: "weakref.WeakValueDictionary" = weakref.WeakValueDictionary()


def _intern(legs, charge, blocks, off, st, shp, size) -> Struct:
    s = Struct(legs, charge, blocks, off, st, shp, size)
    key = (s.labels, tuple(leg.dims for leg in s.legs), s.size, s.off.tobytes(), s.st.tobytes())
    return _LAYOUTS.setdefault(key, s)


@lru_cache(maxsize=4096)
def _allowed(labels, charge: int):
    """Every allowed block in lexicographic sector order (tnkit/blocks.py:220-241): prefixes are
    extended one leg at a time, each keeping only those whose missing charge the remaining legs
    can still supply; the last leg's sector is then determined."""
    signed = [np.array([c * sg for c in cs], np.int64) for cs, sg in labels]
    n = len(signed)
    if n == 0:
        return ((),) if charge == 0 else ()
    reach = [None] * n               # reach[j]: totals the legs after j can supply
    tot = set(signed[-1].tolist())
    for j in range(n - 2, -1, -1):
        reach[j] = np.array(sorted(tot), np.int64)
        tot = {a + b for a in signed[j].tolist() for b in tot}
    pre = np.zeros((1, 0), np.int64)
    acc = np.zeros(1, np.int64)
    for j in range(n - 1):
        k = len(signed[j])
        pre = np.concatenate([np.repeat(pre, k, 0), np.tile(np.arange(k), len(acc))[:, None]], 1)
        acc = np.repeat(acc, k) + np.tile(signed[j], len(acc))
        keep = np.isin(charge - acc, reach[j])
        pre, acc = pre[keep], acc[keep]
    last = {int(c): i for i, c in enumerate(signed[-1])}
    need = charge - acc
    blocks = tuple(tuple(int(x) for x in p) + (last[int(q)],)
                   for p, q in zip(pre, need) if int(q) in last)
    return blocks


def _block_shapes(legs: Sequence[Leg], blocks) -> np.ndarray:
    sec = np.array(blocks, np.int64).reshape(len(blocks), len(legs))
    shp = np.empty(sec.shape, np.int64)
    for j, leg in enumerate(legs):
        shp[:, j] = np.asarray(leg.dims, np.int64)[sec[:, j]]
    return shp


@lru_cache(maxsize=4096)
def canonical(legs: Tuple[Leg, ...], charge: int) -> Struct:
    """Every allowed block in lexicographic sector order, each packed row-major
    (tnkit/blocks.py:243-253)."""
    legs = tuple(legs)
    blocks = _allowed(tuple((leg.charges, leg.sign) for leg in legs), int(charge))
    shp = _block_shapes(legs, blocks)
    n = shp.prod(1)
    return _intern(legs, charge, blocks, np.cumsum(n) - n, _row_major_rows(shp), shp,
                   int(n.sum()))


@lru_cache(maxsize=4096)
def permuted(s: Struct, perm: Tuple[int, ...]) -> Struct:
    """The legs permuted; no element moves (tnkit/blocks.py:256-262)."""
    p = list(perm)
    return _intern(tuple(s.legs[i] for i in perm), s.charge,
                   tuple(tuple(b[i] for i in perm) for b in s.blocks), s.off, s.st[:, p],
                   s.shp[:, p], s.size)


@lru_cache(maxsize=4096)
def dual(s: Struct) -> Struct:
    """Every leg's sign and the total charge flipped (tnkit/blocks.py:265-270)."""
    return _intern(tuple(leg.flip() for leg in s.legs), -s.charge, s.blocks, s.off, s.st, s.shp,
                   s.size)


# ---------------------------------------------------------------------------------------------
# runs: gathers as arithmetic progressions
# ---------------------------------------------------------------------------------------------


def _run_shape(shp, ss, ds):
    """Per box: which axes form its runs, the run length and the run's source step. A run starts
    at the innermost axis of extent > 1 and merges outwards (axes of extent 1 are transparent)
    while the destination stays contiguous and the source an arithmetic progression; a box whose
    innermost axis is not destination-contiguous has runs of one element."""
    B, m = shp.shape
    rows = np.arange(B)
    big = shp > 1
    inner = np.where(big.any(1), m - 1 - np.argmax(big[:, ::-1], 1), m - 1)
    ok = ds[rows, inner] == 1
    tail = np.where(ok, shp[rows, inner], 1)
    step = np.where(ok, ss[rows, inner], 1)
    in_run = np.zeros((B, m), bool)
    in_run[rows, inner] = ok
    chain = ok.copy()
    for j in range(m - 1, -1, -1):
        act = chain & (j < inner)
        unit = shp[:, j] == 1
        grow = act & ~unit & (ds[:, j] == tail) & (ss[:, j] == tail * step)
        in_run[:, j] |= act & (unit | grow)
        chain &= ~act | unit | grow
        tail = np.where(grow, tail * shp[:, j], tail)
    return in_run, tail, step, inner


def _enumerate(src0, ss, dst0, ds, outer):
    """(box, src start, dst start) of every multi-index over the `outer` extents of each box."""
    B, m = outer.shape
    nrun = outer.prod(1)
    box = np.repeat(np.arange(B), nrun)
    local = np.arange(box.size) - np.repeat(np.cumsum(nrun) - nrun, nrun)
    src, dst = src0[box].copy(), dst0[box].copy()
    for k in range(m - 1, -1, -1):
        d = outer[box, k]
        if not (d > 1).any():
            continue
        i = local % d
        local //= d
        src += i * ss[box, k]
        dst += i * ds[box, k]
    return box, src, dst


def box_runs(src0, shp, ss, dst0, ds):
    """Runs of a set of boxes. Box b covers multi-indices i < shp[b], read at src0[b] + i.ss[b]
    and written at dst0[b] + i.ds[b]; each box is cut into runs as `_run_shape` says. Returns
    (src start, src step, dst start, length), sorted by destination."""
    src0, dst0 = np.asarray(src0, np.int64), np.asarray(dst0, np.int64)
    shp, ss, ds = (np.asarray(x, np.int64) for x in (shp, ss, ds))
    B, m = shp.shape
    if B == 0:
        z = np.zeros(0, np.int64)
        return z, z, z, z
    if m == 0:
        one = np.ones(B, np.int64)
        return src0, one, dst0, one
    in_run, tail, step, _ = _run_shape(shp, ss, ds)
    box, src, dst = _enumerate(src0, ss, dst0, ds, np.where(in_run, 1, shp))
    o = np.argsort(dst, kind="stable")
    return src[o], step[box][o], dst[o], tail[box][o]


def expand_runs(runs, n: int, fill: int = -1) -> np.ndarray:
    """The position index idx[dst] = src of a set of runs over a destination of n elements;
    positions no run covers hold `fill`."""
    src, sstep, dst, length = runs
    out = np.full(n, fill, np.int64)
    tot = int(length.sum())
    if tot:
        rid = np.repeat(np.arange(len(length)), length)
        local = np.arange(tot) - np.repeat(np.cumsum(length) - length, length)
        out[dst[rid] + local] = src[rid] + local * sstep[rid]
    return out


TILE = 64          # edge of the 2-D tiles
TILE_MIN = 8       # smallest extent worth a tile
SHORT_RUN = 32     # runs shorter than this are grouped into row tiles


def _tiles(sb, db, srow, drow, scol, nrows, ncols):
    """Cut 2-D slabs (i < nrows, j < ncols; src = sb + i srow + j scol, dst = db + i drow + j)
    into TILE x TILE tiles (src0, srow, scol, dst0, drow, nrows, ncols)."""
    ti, tj = -(-nrows // TILE), -(-ncols // TILE)
    nt = ti * tj
    slab = np.repeat(np.arange(sb.size), nt)
    loc = np.arange(slab.size) - np.repeat(np.cumsum(nt) - nt, nt)
    i0 = (loc // tj[slab]) * TILE
    j0 = (loc % tj[slab]) * TILE
    return np.stack([sb[slab] + i0 * srow[slab] + j0 * scol[slab], srow[slab], scol[slab],
                     db[slab] + i0 * drow[slab] + j0, drow[slab],
                     np.minimum(TILE, nrows[slab] - i0), np.minimum(TILE, ncols[slab] - j0)],
                    1).astype(np.int64)


def box_moves(src0, shp, ss, dst0, ds):
    """Split boxes between the gather kernels. Tiles (src0, srow, scol, dst0, drow, nrows, ncols)
    move dst[dst0 + i drow + j] = src[src0 + i srow + j scol] for i < nrows, j < ncols (at most
    TILE each); everything else becomes runs (`box_runs`). Two kinds of box are tiled:
      * transposes: the destination-contiguous axis a_d is strided in the source and another
        axis a_s is source-contiguous, both of extent >= TILE_MIN (i over a_s, j over a_d);
      * short runs: the run is shorter than SHORT_RUN and an axis outside it has extent
        >= TILE_MIN (i over that axis, j along the run), so many short runs move as one tile.
    Returns (runs, tiles) with tiles a (T, 7) int64 array."""
    src0, dst0 = np.asarray(src0, np.int64), np.asarray(dst0, np.int64)
    shp, ss, ds = (np.asarray(x, np.int64) for x in (shp, ss, ds))
    nb, m = shp.shape
    empty = np.zeros((0, 7), np.int64)
    if nb == 0 or m < 2:
        return box_runs(src0, shp, ss, dst0, ds), empty
    rows = np.arange(nb)
    in_run, tail, step, a_d = _run_shape(shp, ss, ds)
    ok_d = ds[rows, a_d] == 1
    cand = (ss == 1) & (shp >= TILE_MIN)
    cand[rows, a_d] = False
    a_s = np.argmax(cand, 1)
    trans = (cand.any(1) & ok_d & (ss[rows, a_d] != 1) & (shp[rows, a_d] >= TILE_MIN))
    out_ext = np.where(in_run, 0, shp)
    a_o = np.argmax(out_ext, 1)
    short = (~trans & ok_d & (tail < SHORT_RUN) & (out_ext[rows, a_o] >= TILE_MIN))
    rest = ~(trans | short)
    runs = box_runs(src0[rest], shp[rest], ss[rest], dst0[rest], ds[rest])
    parts = []
    if trans.any():
        b = np.flatnonzero(trans)
        k = np.arange(len(b))
        outer = shp[b].copy()
        outer[k, a_d[b]] = 1
        outer[k, a_s[b]] = 1
        box, sb, db = _enumerate(src0[b], ss[b], dst0[b], ds[b], outer)
        bb = b[box]
        parts.append(_tiles(sb, db, np.ones_like(sb), ds[bb, a_s[bb]], ss[bb, a_d[bb]],
                            shp[bb, a_s[bb]], shp[bb, a_d[bb]]))
    if short.any():
        b = np.flatnonzero(short)
        k = np.arange(len(b))
        outer = np.where(in_run[b], 1, shp[b])
        outer[k, a_o[b]] = 1
        box, sb, db = _enumerate(src0[b], ss[b], dst0[b], ds[b], outer)
        bb = b[box]
        parts.append(_tiles(sb, db, ss[bb, a_o[bb]], ds[bb, a_o[bb]], step[bb],
                            shp[bb, a_o[bb]], tail[bb]))
    tiles = np.concatenate(parts) if parts else empty
    return runs, tiles


class Gather:
    """One operand's stacking into GEMM matrices over a destination of `n` elements: runs
    (destination-sorted, see `box_runs`) plus transposing tiles (see `box_moves`); positions
    covered by neither are padding. `identity` says the stacking is the buffer itself."""
    __slots__ = ("runs", "tiles", "n", "identity", "_dev")

    def __init__(self, runs, n: int, src_size: int, tiles=None):
        self.runs, self.n = runs, int(n)
        self.tiles = np.zeros((0, 7), np.int64) if tiles is None else tiles
        src, sstep, dst, length = runs
        self.identity = bool(len(self.tiles) == 0 and self.n == src_size
                             and int(length.sum()) == self.n and (src == dst).all()
                             and ((sstep == 1) | (length == 1)).all())
        self._dev = {}

    def expand(self) -> np.ndarray:
        """idx[dst] = src over the destination, -1 at padding."""
        out = expand_runs(self.runs, self.n)
        for s0, sr, sc, d0, dr, nr, nc in self.tiles.tolist():
            i = np.arange(nr)[:, None]
            j = np.arange(nc)[None, :]
            out[(d0 + i * dr + j).ravel()] = (s0 + i * sr + j * sc).ravel()
        return out


# ---------------------------------------------------------------------------------------------
# contraction plans
# ---------------------------------------------------------------------------------------------


def _label_groups(labels, axes) -> Dict[int, List[Tuple[int, ...]]]:
    """Sector tuples over `axes`, in lexicographic order, grouped by their signed charge total
    (tnkit/blocks.py:908-915); groups appear in order of first occurrence."""
    groups: Dict[int, list] = {}
    for t in itertools.product(*(range(len(labels[a][0])) for a in axes)):
        q = sum(labels[a][1] * labels[a][0][s] for a, s in zip(axes, t))
        groups.setdefault(q, []).append(t)
    return groups


def _compose(n: int, parts) -> Tuple[int, ...]:
    b = [0] * n
    for axes, t in parts:
        for a, s in zip(axes, t):
            b[a] = s
    return tuple(b)


class SymPlan:
    """The size-free part of a plan (tnkit/blocks.py:941-997): per contracted charge c (a group,
    ascending), its row items (sector tuples over a's free legs), k items (over the contracted
    legs) and column items (over b's free legs), numbered globally; `a_box`/`b_box` name which
    block of each operand fills which (row, k) / (k, column) item pair, `out_rc` which (row,
    column) pair is which output block."""
    __slots__ = ("r_sec", "k_sec", "c_sec", "r_grp", "k_grp", "c_grp", "r_beg", "k_beg",
                 "c_beg", "r_end", "k_end", "c_end", "a_box", "b_box", "out_rc", "out_blocks",
                 "charge", "ngroups")


@lru_cache(maxsize=4096)
def _sym_plan(la, lb, ax_a: Tuple[int, ...], ax_b: Tuple[int, ...]) -> SymPlan:
    (legs_a, qa, blocks_a), (legs_b, qb, blocks_b) = la, lb
    fa = tuple(i for i in range(len(legs_a)) if i not in ax_a)
    fb = tuple(j for j in range(len(legs_b)) if j not in ax_b)
    rows, ks, cols = (_label_groups(legs_a, fa), _label_groups(legs_a, ax_a),
                      _label_groups(legs_b, fb))
    ia = {b: k for k, b in enumerate(blocks_a)}
    ib = {b: k for k, b in enumerate(blocks_b)}
    charges = sorted({qa - f for f in rows} & {g - qb for g in cols})
    items = {"r": [], "k": [], "c": []}
    grp = {"r": [], "k": [], "c": []}
    a_box, b_box, out_rc, out_blocks = [], [], [], []
    na, nb = len(legs_a), len(legs_b)
    for gi, c in enumerate(charges):
        r_it, k_it, c_it = rows[qa - c], ks.get(c, []), cols[qb + c]
        r0, k0, c0 = len(items["r"]), len(items["k"]), len(items["c"])
        for key, its in (("r", r_it), ("k", k_it), ("c", c_it)):
            items[key] += its
            grp[key] += [gi] * len(its)
        for kj, kt in enumerate(k_it):
            for ri, rt in enumerate(r_it):
                a_box.append((ia[_compose(na, ((fa, rt), (ax_a, kt)))], r0 + ri, k0 + kj))
            for ci, ct in enumerate(c_it):
                b_box.append((ib[_compose(nb, ((ax_b, kt), (fb, ct)))], k0 + kj, c0 + ci))
        for ri, rt in enumerate(r_it):
            for ci, ct in enumerate(c_it):
                out_rc.append((r0 + ri, c0 + ci))
                out_blocks.append(rt + ct)
    p = SymPlan()
    G = len(charges)
    for key, axes in (("r", fa), ("k", ax_a), ("c", fb)):
        sec = np.array(items[key], np.int64).reshape(len(items[key]), len(axes))
        g = np.array(grp[key], np.int64)
        setattr(p, key + "_sec", sec)
        setattr(p, key + "_grp", g)
        setattr(p, key + "_beg", np.searchsorted(g, np.arange(G), side="left"))
        setattr(p, key + "_end", np.searchsorted(g, np.arange(G), side="right"))
    p.a_box = np.array(a_box, np.int64).reshape(len(a_box), 3)
    p.b_box = np.array(b_box, np.int64).reshape(len(b_box), 3)
    p.out_rc = np.array(out_rc, np.int64).reshape(len(out_rc), 2)
    p.out_blocks = tuple(out_blocks)
    p.charge = qa + qb
    p.ngroups = G
    return p


def _group_offsets(n, grp, beg, end):
    """For items of sizes n listed group by group: each item's offset inside its group and each
    group's total (tnkit/blocks.py:1022-1028)."""
    cum = np.zeros(len(n) + 1, np.int64)
    np.cumsum(n, out=cum[1:])
    first = cum[beg]
    return cum[:-1] - first[grp], cum[end] - first


def _item_shapes(sec, axes, legs):
    shp = np.empty(sec.shape, np.int64)
    for j, a in enumerate(axes):
        shp[:, j] = np.asarray(legs[a].dims, np.int64)[sec[:, j]]
    return shp


def _excl(n):
    return np.cumsum(n) - n


def _pad(n, q: int = 4):
    return (np.asarray(n, np.int64) + q - 1) // q * q


# K rows of the device stacks of long-K plans are padded with zero rows to a multiple of this
# (the grouped GEMM's direct-to-LDS form reads whole slabs of up to 32 rows)
K_PAD = 32


class Plan:
    """A contraction of two layouts (tnkit/blocks.py:1073-1100).

    gemms  : (G, 6) int64, per contracted charge (pa, pb, po, R, K, C) in tnkit's stacking;
             K = 0 marks an output region with no contribution (zero-filled).
    macs   : sum of R K C; nmm the number of groups with K > 0.
    out    : the output Struct (free legs of a, then of b) whose blocks are strided views of the
             GEMM results, stacked at po with row length C.
    ref_a, ref_b : tnkit's gathers into (R, K) and (K, C) matrices (parity checks).
    dev_a  : operand a stacked K-major, (K, R) per group with leading dimension lda (a multiple
             of 4) at offset qa, followed by zero rows up to a multiple of kpad rows.
    dev_b  : operand b stacked (K, C) per group with leading dimension ldb (a multiple of 4) at
             offset qb, followed by zero rows up to a multiple of kpad rows.
    kpad   : K_PAD when some group has more than K_PAD rows of K (the plans the tiled kernels
             serve), else 1 (short-K plans, served by the memory-bound kernels).
    desc   : (G, 8) int64 rows (qa, lda, qb, ldb, po, R, K, C) for the grouped GEMM.
    """
    __slots__ = ("gemms", "macs", "nmm", "out", "dev_a", "dev_b", "desc", "size_a", "size_b",
                 "kpad", "_cache", "_ref")

    @property
    def ref_a(self) -> "Gather":
        return self._reference(0)

    @property
    def ref_b(self) -> "Gather":
        return self._reference(1)

    def _reference(self, side: int) -> "Gather":
        """tnkit's own stacking of operand `side` (built on demand: the device path does not
        use it)."""
        key = ("ref", side)
        g = self._cache.get(key)
        if g is None:
            sp, s, axes = self._ref[side]
            moves, n, *_ = _stack(sp, s, axes, side, False)
            g = self._cache[key] = Gather(moves[0], n, s.size)
        return g


def _k_totals(sp: SymPlan, s: Struct, axes) -> np.ndarray:
    """Per group, the number of contracted (K) rows."""
    return _group_offsets(_item_shapes(sp.k_sec, axes, s.legs).prod(1), sp.k_grp, sp.k_beg,
                          sp.k_end)[1]


def _stack(sp: SymPlan, s: Struct, axes, side: int, ld_pad: bool, kpad: int = 1):
    """Runs stacking operand `side` (0: a, 1: b) of `sp`. For side 0 with ld_pad=False: tnkit's
    (R, K) matrices; for side 0 with ld_pad: K-major (K, R) matrices with padded leading
    dimension; for side 1: (K, C) matrices, padded when ld_pad. Returns (runs, total size,
    per-group offsets, per-group leading dimension, outer shapes/offsets/totals, K totals)."""
    sec, grp, beg, end = ((sp.r_sec, sp.r_grp, sp.r_beg, sp.r_end) if side == 0 else
                          (sp.c_sec, sp.c_grp, sp.c_beg, sp.c_end))
    free = tuple(i for i in range(len(s.legs)) if i not in axes)
    oshp = _item_shapes(sec, free, s.legs)
    o, n = _group_offsets(oshp.prod(1), grp, beg, end)
    kshp = _item_shapes(sp.k_sec, axes, s.legs)
    ko, K = _group_offsets(kshp.prod(1), sp.k_grp, sp.k_beg, sp.k_end)
    box = sp.a_box if side == 0 else sp.b_box
    blk = box[:, 0]
    if side == 0:
        oi, ki = box[:, 1], box[:, 2]
    else:
        ki, oi = box[:, 1], box[:, 2]
    g = sp.k_grp[ki]
    if side == 0 and not ld_pad:              # tnkit: (R, K) row-major, free legs first
        ld = K
        gsz = n * K
        base = _excl(gsz)
        perm = list(free + tuple(axes))
        nrow = len(free)
        dst0 = base[g] + o[oi] * ld[g] + ko[ki]
    else:                                     # (K, outer): contracted legs first
        ld = _pad(n) if ld_pad else n
        # device stacks of plans with long K carry zero rows up to a multiple of kpad, so a tile
        # loop reads whole K slabs
        gsz = (_pad(K, kpad) if ld_pad else K) * ld
        base = _excl(_pad(gsz)) if ld_pad else _excl(gsz)
        perm = list(tuple(axes) + free)
        nrow = len(axes)
        dst0 = base[g] + ko[ki] * ld[g] + o[oi]
    bs, bst = s.shp[blk][:, perm], s.st[blk][:, perm]
    ds = np.concatenate([_row_major_rows(bs[:, :nrow]) * ld[g][:, None],
                         _row_major_rows(bs[:, nrow:])], 1)
    total = int(_pad(gsz).sum()) if ld_pad else int(gsz.sum())
    if ld_pad:
        moves = box_moves(s.off[blk], bs, bst, dst0, ds)
    else:
        moves = (box_runs(s.off[blk], bs, bst, dst0, ds), None)
    return moves, total, base, ld, (oshp, o, n), K


_PLANS: Dict = {}


def plan(sa: Struct, sb: Struct, ax_a: Tuple[int, ...], ax_b: Tuple[int, ...]) -> Plan:
    """The plan contracting legs ax_a of layout sa with legs ax_b of layout sb (cached)."""
    key = (sa, sb, tuple(ax_a), tuple(ax_b))
    p = _PLANS.get(key)
    if p is not None:
        return p
    for i, j in zip(ax_a, ax_b):
        la, lb = sa.legs[i], sb.legs[j]
        if la.charges != lb.charges or la.dims != lb.dims or la.sign != -lb.sign:
            raise ValueError(f"cannot contract leg {i} {la} with leg {j} {lb}: the sectors must "
                             "match and the signs be opposite")
    sp = _sym_plan(sa.labels, sb.labels, tuple(ax_a), tuple(ax_b))
    p = Plan()
    Kg = _k_totals(sp, sa, ax_a)
    p.kpad = K_PAD if len(Kg) and int(Kg.max()) > K_PAD else 1
    da, p.size_a, qa, lda, (r_shp, ro, R), K = _stack(sp, sa, ax_a, 0, True, p.kpad)
    db, p.size_b, qb, ldb, (c_shp, co, C), _ = _stack(sp, sb, ax_b, 1, True, p.kpad)
    p.dev_a = Gather(da[0], p.size_a, sa.size, da[1])
    p.dev_b = Gather(db[0], p.size_b, sb.size, db[1])
    pa, pb, po = _excl(R * K), _excl(K * C), _excl(R * C)
    p.gemms = np.stack([pa, pb, po, R, K, C], 1).astype(np.int64).reshape(len(R), 6)
    p.desc = np.stack([qa, lda, qb, ldb, po, R, K, C], 1).astype(np.int64).reshape(len(R), 8)
    p.macs = int((R * K * C).sum())
    p.nmm = int(np.count_nonzero(K))
    orr, oc = sp.out_rc[:, 0], sp.out_rc[:, 1]
    go = sp.r_grp[orr]
    st = np.concatenate([_row_major_rows(r_shp[orr]) * C[go][:, None],
                         _row_major_rows(c_shp[oc])], 1)
    p.out = _intern(tuple(sa.legs[i] for i in range(len(sa.legs)) if i not in ax_a)
                    + tuple(sb.legs[j] for j in range(len(sb.legs)) if j not in ax_b),
                    sp.charge, sp.out_blocks, po[go] + ro[orr] * C[go] + co[oc], st,
                    np.concatenate([r_shp[orr], c_shp[oc]], 1), int((R * C).sum()))
    p._cache = {}
    p._ref = ((sp, sa, tuple(ax_a)), (sp, sb, tuple(ax_b)))
    _PLANS[key] = p
    return p


@lru_cache(maxsize=4096)
def to_canonical(s: Struct) -> Gather:
    """The gather of layout s into canonical(s.legs, s.charge) (tnkit/blocks.py:648-655)."""
    c = canonical(s.legs, s.charge)
    k = np.array([s.index[b] for b in c.blocks], np.int64)
    if len(k) == 0:
        z = np.zeros(0, np.int64)
        return Gather((z, z, z, z), c.size, s.size)
    runs, tiles = box_moves(s.off[k], c.shp, s.st[k], c.off, c.st)
    return Gather(runs, c.size, s.size, tiles)


def clear_cache() -> None:
    _PLANS.clear()
    for f in (canonical, permuted, dual, to_canonical, _sym_plan, _allowed):
        f.cache_clear()


def parse_einsum(sub: str):
    """(ax_a, ax_b, perm) of a two-operand contraction without batch indices
    (tnkit/block_backend.py:47-68): contracted labels in the order they appear in the first
    operand; the result holds a's free legs then b's, permuted by `perm` into the output order."""
    lhs, out = sub.replace(" ", "").split("->")
    ta, tb = lhs.split(",")
    shared = [c for c in ta if c in tb]
    if set(shared) & set(out) or set(out) != set(ta) ^ set(tb):
        raise ValueError(f"not a plain two-operand contraction: {sub}")
    free = [c for c in ta if c not in shared] + [c for c in tb if c not in shared]
    return (tuple(ta.index(c) for c in shared), tuple(tb.index(c) for c in shared),
            tuple(free.index(c) for c in out))


# ---------------------------------------------------------------------------------------------
# fusing and splitting legs (the matrix views of truncation and canonicalisation)
# ---------------------------------------------------------------------------------------------


@dataclass(frozen=True)
class FuseRecord:
    """How groups of legs were fused (tnkit/blocks.py:1197-1224). Fused leg g is fused[g]
    (sign +1); its sector k stacks the sub-blocks sectors[g][k], each (sector tuple over
    groups[g], row offset, size), in lexicographic order."""
    legs: Tuple[Leg, ...]
    charge: int
    groups: Tuple[Tuple[int, ...], ...]
    fused: Tuple[Leg, ...]
    sectors: Tuple

    @property
    def perm(self) -> Tuple[int, ...]:
        return tuple(a for g in self.groups for a in g)


@lru_cache(maxsize=4096)
def fuse_record(legs: Tuple[Leg, ...], charge: int, groups) -> FuseRecord:
    labels = tuple((leg.charges, leg.sign) for leg in legs)
    fused, sectors = [], []
    for g in groups:
        by_q = _label_groups(labels, g)
        qs = sorted(by_q)
        secs, dims = [], []
        for q in qs:
            items, off = [], 0
            for t in by_q[q]:
                n = 1
                for a, s in zip(g, t):
                    n *= legs[a].dims[s]
                items.append((t, off, n))
                off += n
            secs.append(tuple(items))
            dims.append(off)
        sectors.append(tuple(secs))
        fused.append(Leg(tuple(qs), tuple(dims), 1))
    return FuseRecord(tuple(legs), int(charge), tuple(tuple(g) for g in groups), tuple(fused),
                      tuple(sectors))


def _row_major(shape) -> List[int]:
    st, n = [], 1
    for d in reversed(shape):
        st.append(n)
        n *= d
    return list(reversed(st))


def _gather_of(out: Struct, src_size: int, boxes) -> Gather:
    """The gather of a list of boxes (src start, shape, src strides, dst start, dst strides),
    each given in the destination's axis order."""
    if not boxes:
        z = np.zeros(0, np.int64)
        return Gather((z, z, z, z), out.size, src_size)
    src0, shp, ss, d0, ds = (np.array(x, np.int64) for x in zip(*boxes))
    runs, tiles = box_moves(src0, shp.reshape(len(boxes), -1), ss.reshape(len(boxes), -1), d0,
                            ds.reshape(len(boxes), -1))
    return Gather(runs, out.size, src_size, tiles)


@lru_cache(maxsize=4096)
def fuse_map(s: Struct, rec: FuseRecord) -> Tuple[Struct, Gather]:
    """The fused layout (canonical over rec.fused) and the gather into it from layout s
    (tnkit/blocks.py:1227-1241). Every block of s lands whole in one fused block: group g's
    sub-sector tuple names the fused sector and the row offset inside it (rec.sectors), and the
    block's axes of group g are laid out row-major inside the fused row. One box per source
    block."""
    out = canonical(rec.fused, rec.charge)
    where = [{t: (k, off) for k, items in enumerate(rec.sectors[g]) for t, off, _ in items}
             for g in range(len(rec.groups))]
    perm = list(rec.perm)
    boxes = []
    for kb, blk in enumerate(s.blocks):
        hit = [where[g][tuple(blk[a] for a in grp)] for g, grp in enumerate(rec.groups)]
        kd = out.index[tuple(k for k, _ in hit)]
        ost = out.st[kd].tolist()
        dst = [r * ost[g] for g, grp in enumerate(rec.groups)
               for r in _row_major([int(s.shp[kb][a]) for a in grp])]
        d0 = int(out.off[kd]) + sum(off * ost[g] for g, (_, off) in enumerate(hit))
        boxes.append((int(s.off[kb]), s.shp[kb][perm], s.st[kb][perm], d0, dst))
    return out, _gather_of(out, s.size, boxes)


@lru_cache(maxsize=4096)
def split_map(s: Struct, rec: FuseRecord, todo: Tuple[int, ...]) -> Tuple[Struct, Gather]:
    """Split the fused legs `todo` of a tensor laid out as s (its leg g is rec.fused[g] for g in
    todo), keeping its other legs; a full split also undoes the fuse's permutation
    (tnkit/blocks.py:1244-1286). Returns the canonical output layout and the gather into it. A
    source block's fused sector on leg g stacks the sub-blocks of rec.sectors[g]; every choice
    of one sub-block per split leg is one output block, read at its row offsets with the split
    axes row-major inside the fused row."""
    legs = []
    for g in range(len(s.legs)):
        legs += [rec.legs[a] for a in rec.groups[g]] if g in todo else [s.legs[g]]
    full = len(todo) == len(rec.groups)
    order = [int(i) for i in np.argsort(rec.perm)] if full else list(range(len(legs)))
    out = canonical(tuple(legs[i] for i in order), s.charge)
    boxes = []
    for kb, blk in enumerate(s.blocks):
        sst = s.st[kb].tolist()
        for pick in itertools.product(*(rec.sectors[g][blk[g]] for g in todo)):
            sub = dict(zip(todo, pick))
            sec, st = [], []
            src = int(s.off[kb])
            for g in range(len(s.legs)):
                if g not in sub:
                    sec.append(blk[g])
                    st.append(sst[g])
                    continue
                t, off, _ = sub[g]
                src += off * sst[g]
                sec += list(t)
                st += [r * sst[g] for r in _row_major([rec.legs[a].dims[t[i]]
                                                        for i, a in enumerate(rec.groups[g])])]
            kd = out.index[tuple(sec[i] for i in order)]
            boxes.append((src, out.shp[kd], [st[i] for i in order], int(out.off[kd]),
                          out.st[kd]))
    return out, _gather_of(out, s.size, boxes)


@lru_cache(maxsize=4096)
def leg_positions(s: Struct, axis: int) -> Gather:
    """For every element of layout s, its sector-ordered index along leg `axis`, as a gather
    from a vector over that leg (tnkit/blocks.py:670-679); used to scale a leg by a vector."""
    if not len(s.off):
        z = np.zeros(0, np.int64)
        return Gather((z, z, z, z), s.size, s.legs[axis].size)
    start = np.concatenate([[0], np.cumsum(s.legs[axis].dims)[:-1]]).astype(np.int64)
    sec = np.array(s.blocks, np.int64).reshape(len(s.off), -1)[:, axis]
    step = np.zeros_like(s.st)
    step[:, axis] = 1
    # destination positions: the buffer itself; its row-major order for canonical layouts
    return Gather(box_runs(start[sec], s.shp, step, s.off, s.st), s.size, s.legs[axis].size)
