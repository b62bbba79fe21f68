"""U(1) block-sparse tensors on the HIP device: run gathers, the persistent grouped GEMM, its tile
autotuner, `tensordot`, and the block two-site operator with hipGraph replay.

A `BlockTensor` is a host `blocks.Struct` plus a flat, contiguous device buffer (float64 or
float32), the same buffer layout tnkit uses (tnkit/blocks.py:697-760). A contraction
(tnkit/blocks.py:1156-1192) is here: one run gather per operand into K-major, row-padded
per-charge matrices, then all charge groups in one grouped-GEMM launch whose output is the
result buffer in tnkit's layout.

`BlockTwoSite` is the effective Hamiltonian of tnkit's block DMRG (`apply_two_site` through the
BlockBackend einsum, tnkit/_dmrg.py:162-169, tnkit/block_backend.py:121-131, and the flattening
of tnkit/_dmrg.py:638-644): the operator tensors L, W1, W2, R are stacked once per local
problem (W1 and W2 contracted into one operator), every buffer is allocated once, and each apply
is four gathers and three grouped products on the current stream, so it can be captured in a
hipGraph (`graph()`), which replaces tnkit's block_graph.MatvecGraph
(tnkit/block_graph.py:89-180).
"""
from __future__ import annotations

import math
import os
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from . import blocks as B
from ._lib import check, dcode, ptr

_REAL = (torch.float64, torch.float32)
TWO_SITE_SUBS = ("pwq,qijb->pwijb", "wmsi,pwijb->pmsjb", "mrtj,pmsjb->pstrb", "drb,pstrb->pstd")
# the applied form: the two MPO tensors contracted into one (w, s1, s2) -> (s1', s2', w') operator
MPO_PAIR_SUB = "wmsi,mrtj->wijstr"
APPLY_SUBS = ("pwq,qijb->pwijb", "wijstr,pwijb->pstrb", "drb,pstrb->pstd")
TILE_CONFIGS = (1, 2, 3, 4, 5, 11, 12, 13, 15)   # 11..15: direct-to-LDS forms of 1, 2, 3, 5
DMA_CONFIGS = (11, 12, 13, 15)
SMALL = 9              # the memory-bound kernel for groups with R, K <= 8 (not a tile shape)
MID = 10               # the same with R, K <= 32 (the operand staged in LDS)
_SMALL_COLS = 1024


def _stream(device) -> int:
    return _lib.stream_of(device)


# ---------------------------------------------------------------------------------------------
# gathers
# ---------------------------------------------------------------------------------------------


class _DevRuns:
    __slots__ = ("runs", "chunks", "nruns", "n", "tiles", "ntiles")


def _device_runs(g: B.Gather, device, src_size: int) -> _DevRuns:
    """The runs of `g` on `device` (uploaded once), after checking on the host that every read
    stays inside a source of `src_size` elements and every write inside the destination."""
    key = (device, src_size)
    d = g._dev.get(key)
    if d is not None:
        return d
    src, sstep, dst, length = g.runs
    if len(length):
        last = src + (length - 1) * sstep
        if (src.min() < 0 or last.max() >= src_size or (length < 1).any() or dst.min() < 0
                or (dst + length).max() > g.n or (dst[1:] < dst[:-1] + length[:-1]).any()):
            raise ValueError("gather runs outside their buffers or overlapping")
    chunk = _lib.load().tnh_gather_chunk()
    nch = max(1, -(-g.n // chunk))
    starts = np.arange(nch, dtype=np.int64) * chunk
    ends = np.minimum(starts + chunk, g.n) - 1
    lo = np.maximum(np.searchsorted(dst, starts, side="right") - 1, 0)
    hi = np.searchsorted(dst, ends, side="right")
    lo = np.minimum(lo, hi)
    d = _DevRuns()
    d.nruns = len(length)
    d.n = g.n
    host = np.concatenate([src, sstep, dst, length]).astype(np.int64)
    d.runs = torch.from_numpy(host if host.size else np.zeros(4, np.int64)).to(device)
    d.chunks = torch.from_numpy(np.concatenate([lo, hi]).astype(np.int32)).to(device)
    t = g.tiles
    d.ntiles = len(t)
    if d.ntiles:
        s0, sr, sc, d0, dr, nr, nc = t.T
        if ((nr < 1) | (nc < 1) | (nr > B.TILE) | (nc > B.TILE)).any() or s0.min() < 0 \
                or (sr < 0).any() or (sc < 0).any() or (dr < 0).any() \
                or (s0 + (nr - 1) * sr + (nc - 1) * sc).max() >= src_size or d0.min() < 0 \
                or (d0 + (nr - 1) * dr + nc - 1).max() >= g.n:
            raise ValueError("gather tiles outside their buffers")
        d.tiles = torch.from_numpy(np.ascontiguousarray(t, np.int64).reshape(-1)).to(device)
    else:
        d.tiles = None
    g._dev[key] = d
    return d


def gather(g: B.Gather, src: torch.Tensor, dtype: Optional[torch.dtype] = None,
           out: Optional[torch.Tensor] = None, prezeroed: bool = False) -> torch.Tensor:
    """out[j] = src[position j of g] (zero where g covers nothing), converted to `dtype`.
    `prezeroed`: out's uncovered positions already hold zeros (a buffer allocated with zeros
    and only ever written by this gather), so they are not written again."""
    if src.device.type != "cuda" or src.dtype not in _REAL or src.dim() != 1 \
            or not src.is_contiguous():
        raise ValueError(f"gather: need a contiguous 1-D float64/float32 device buffer, got "
                         f"{src.dtype} {tuple(src.shape)} on {src.device}")
    dtype = dtype or src.dtype
    d = _device_runs(g, src.device, src.numel())
    if out is None:   # with tiles the run pass leaves padding alone: start from zeros
        out = (torch.zeros if d.ntiles else torch.empty)(g.n, dtype=dtype, device=src.device)
    if out.numel() != g.n or out.dtype != dtype or not out.is_contiguous():
        raise ValueError("gather: output buffer does not match")
    lib, st = _lib.load(), _stream(src.device)
    # with tiles the run pass leaves uncovered positions alone (tiles and padding): the padding
    # of a reused buffer keeps what it was allocated with (zeros, see BlockTwoSite), and the
    # zero K-padding rows are the only padding the grouped GEMM sums into a stored result
    check(lib.tnh_gather_runs(dcode(src.dtype), dcode(dtype), g.n, ptr(src), ptr(out), d.nruns,
                              ptr(d.runs), ptr(d.chunks), int(d.ntiles == 0 and not prezeroed),
                              st), "gather_runs")
    if d.ntiles:
        check(lib.tnh_gather_tiles(dcode(src.dtype), dcode(dtype), d.ntiles, ptr(src), ptr(out),
                                   ptr(d.tiles), st), "gather_tiles")
    return out


# ---------------------------------------------------------------------------------------------
# grouped GEMM: tile tables and the autotuner
# ---------------------------------------------------------------------------------------------

_TILE_INFO: Dict = {}


def tile_info(dtype: torch.dtype, cfg: int, device) -> Tuple[int, int, int]:
    """(bm, bn, persistent grid) of grouped tile config `cfg`."""
    key = (dtype, cfg, device)
    if key not in _TILE_INFO:
        import ctypes
        bm, bn, grid = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        with torch.cuda.device(device):
            check(_lib.load().tnh_grouped_tile(dcode(dtype), cfg, ctypes.byref(bm),
                                               ctypes.byref(bn), ctypes.byref(grid)),
                  "grouped_tile")
        _TILE_INFO[key] = (bm.value, bn.value, grid.value)
    return _TILE_INFO[key]


def tile_table(desc: np.ndarray, bm: int, bn: int, panel: int = 8) -> np.ndarray:
    """Positions of the grouped launch: (group, tile row, tile column) per position, -1 rows as
    padding, length a multiple of 8. Tiles are cut into panels of up to `panel` tiles along one
    tile row (they share that row's A block), panels sorted by descending K and dealt round-robin
    to the 8 XCD queues; the j-th tile of queue x sits at position 8 j + x, which a grid that is a
    multiple of 8 runs on XCD x."""
    R, K, C = desc[:, 5], desc[:, 6], desc[:, 7]
    live = (R > 0) & (C > 0)
    g_ids = np.flatnonzero(live)
    nbm = -(-R[g_ids] // bm)
    nbn = -(-C[g_ids] // bn)
    npan = -(-nbn // panel)                       # panels per tile row
    # one entry per panel: group, tile row, first column, number of tiles
    rows_per_g = nbm
    pg = np.repeat(g_ids, rows_per_g * npan)
    if pg.size == 0:
        return np.full((8, 3), -1, np.int32)
    idx_in_g = np.arange(pg.size) - np.repeat(np.cumsum(rows_per_g * npan) - rows_per_g * npan,
                                              rows_per_g * npan)
    npan_g = np.repeat(npan, rows_per_g * npan)
    nbn_g = np.repeat(nbn, rows_per_g * npan)
    ptm = idx_in_g // npan_g
    pc0 = (idx_in_g % npan_g) * panel
    plen = np.minimum(panel, nbn_g - pc0)
    order = np.argsort(-K[pg], kind="stable")
    pg, ptm, pc0, plen = pg[order], ptm[order], pc0[order], plen[order]
    xcd = np.arange(pg.size) % 8
    tg = np.repeat(pg, plen)
    ttm = np.repeat(ptm, plen)
    tx = np.repeat(xcd, plen)
    ttn = np.repeat(pc0, plen) + (np.arange(tg.size) - np.repeat(np.cumsum(plen) - plen, plen))
    rank = np.zeros(tg.size, np.int64)
    for x in range(8):                            # rank of each tile inside its XCD queue
        sel = np.flatnonzero(tx == x)
        rank[sel] = np.arange(sel.size)
    depth = int(rank.max()) + 1
    table = np.full((8 * depth, 3), -1, np.int32)
    pos = 8 * rank + tx
    table[pos, 0], table[pos, 1], table[pos, 2] = tg, ttm, ttn
    return table


class _Grouped:
    """A plan's grouped-GEMM state on one device: descriptors, and per tile config its tile table
    and grid."""
    __slots__ = ("desc", "desc_np", "tables", "cfg", "size_a", "size_b", "size_out", "macs",
                 "dma")

    def __init__(self, p: B.Plan, device):
        d = p.desc
        for qa, lda, qb, ldb, po, R, K, C in d.tolist():
            kp = -(-K // p.kpad) * p.kpad
            if not (qa % 4 == 0 and qb % 4 == 0 and lda % 4 == 0 and ldb % 4 == 0
                    and lda >= R and ldb >= C and qa + kp * lda <= p.size_a
                    and qb + kp * ldb <= p.size_b and po + R * C <= p.out.size):
                raise ValueError("grouped GEMM descriptor outside its buffers")
        self.desc_np = d
        self.desc = torch.from_numpy(np.ascontiguousarray(d, np.int64)).to(device)
        self.tables = {}
        self.cfg = {}
        self.size_a, self.size_b, self.size_out = p.size_a, p.size_b, p.out.size
        self.macs = p.macs
        self.dma = p.kpad == B.K_PAD

    def small_ok(self, limit: int = 8) -> bool:
        d = self.desc_np
        return len(d) > 0 and int(d[:, 5].max()) <= limit and int(d[:, 6].max()) <= limit

    def mid_ok(self) -> bool:
        return self.small_ok(32)

    def table(self, dtype, cfg, device):
        key = (dtype, cfg)
        t = self.tables.get(key)
        if t is None and cfg in (SMALL, MID):
            C = self.desc_np[:, 7]
            nch = -(-C // _SMALL_COLS)
            g = np.repeat(np.arange(len(C)), nch)
            c0 = (np.arange(g.size) - np.repeat(np.cumsum(nch) - nch, nch)) * _SMALL_COLS
            tab = np.stack([g, c0], 1).astype(np.int32)
            t = (torch.from_numpy(tab.reshape(-1)).to(device), len(tab), 0)
            self.tables[key] = t
        elif t is None:
            bm, bn, gmax = tile_info(dtype, cfg, device)
            tab = tile_table(self.desc_np, bm, bn)
            grid = int(min(gmax, len(tab)))
            t = (torch.from_numpy(tab.reshape(-1)).to(device), len(tab), grid)
            self.tables[key] = t
        return t


def _grouped(p: B.Plan, device) -> _Grouped:
    key = ("grouped", device)
    g = p._cache.get(key)
    if g is None:
        g = p._cache[key] = _Grouped(p, device)
    return g


def grouped_gemm(p: B.Plan, A: torch.Tensor, Bm: torch.Tensor, out: torch.Tensor,
                 cfg: Optional[int] = None) -> torch.Tensor:
    """out (tnkit's stacked result of plan p) from the stacked operands A, Bm, all charge groups
    in one launch with tile config `cfg` (default: the autotuned one)."""
    dev = A.device
    for t, n, nm in ((A, p.size_a, "A"), (Bm, p.size_b, "B"), (out, p.out.size, "out")):
        if t.device != dev or t.dtype != A.dtype or t.numel() != n or not t.is_contiguous() \
                or t.dtype not in _REAL:
            raise ValueError(f"grouped_gemm: {nm} buffer {t.dtype} {tuple(t.shape)} on "
                             f"{t.device} does not match the plan ({n} elements)")
    gs = _grouped(p, dev)
    if cfg is None:
        cfg = autotune(p, A, Bm, out)
    if cfg == SMALL and not gs.small_ok():
        raise ValueError("grouped_gemm: the small kernel needs R, K <= 8 in every group")
    if cfg == MID and not gs.mid_ok():
        raise ValueError("grouped_gemm: the mid kernel needs R, K <= 32 in every group")
    if cfg in DMA_CONFIGS and not gs.dma:
        raise ValueError("grouped_gemm: the direct-to-LDS configs need a plan with K padding "
                         "(some group with K > 32)")
    tiles, ntiles, grid = gs.table(A.dtype, cfg, dev)
    if cfg == SMALL:
        check(_lib.load().tnh_grouped_small(dcode(A.dtype), ptr(gs.desc), ntiles, ptr(tiles),
                                            ptr(A), ptr(Bm), ptr(out), _stream(dev)),
              "grouped_small")
        return out
    if cfg == MID:
        check(_lib.load().tnh_grouped_mid(dcode(A.dtype), ptr(gs.desc), ntiles, ptr(tiles),
                                          ptr(A), ptr(Bm), ptr(out), _stream(dev)),
              "grouped_mid")
        return out
    check(_lib.load().tnh_grouped_gemm(dcode(A.dtype), ptr(gs.desc), ntiles, ptr(tiles), ptr(A),
                                       ptr(Bm), ptr(out), cfg, grid, _stream(dev)),
          "grouped_gemm")
    return out


_TUNED: Dict = {}
TUNE_LOG = []


def _signature(p: B.Plan, dtype) -> tuple:
    d = p.desc
    lg = lambda x: int(math.ceil(math.log2(max(int(x), 1))))
    return (dtype, lg(p.macs) if p.macs else -1, lg(d[:, 5].max(initial=1)),
            lg(d[:, 6].max(initial=1)), lg(d[:, 7].max(initial=1)))


def heuristic_cfg(p: B.Plan, dtype) -> int:
    """The tile config without timing: the memory-bound kernels when every group has R, K <= 8
    (SMALL) or <= 32 (MID), the skinny 16-row tile when every group has at most 16 rows, else
    64x64 for float64 and 128x128 for float32 when there are enough tiles to fill the chip, else
    32x32."""
    d = p.desc
    R, C = d[:, 5], d[:, 7]
    if R.max(initial=0) <= 8 and d[:, 6].max(initial=0) <= 8:
        return SMALL
    if R.max(initial=0) <= 32 and d[:, 6].max(initial=0) <= 32:
        return MID
    if R.max(initial=0) <= 16:
        return 4
    t64 = int(((-(-R // 64)) * (-(-C // 64))).sum())
    t128 = int(((-(-R // 128)) * (-(-C // 128))).sum())
    if dtype == torch.float32 and t128 >= 512:
        return 1
    return 2 if t64 >= 256 else 3


def autotune(p: B.Plan, A, Bm, out, reps: int = 3) -> int:
    """The fastest tile config for this plan's shape class (cached by `_signature`): each
    candidate is launched once to warm up and then timed over `reps` launches with HIP events on
    the current stream, in two interleaved rounds (the better one counts).
    TNKIT_AMD_AUTOTUNE=0 uses `heuristic_cfg` instead."""
    sig = _signature(p, A.dtype)
    cfg = _TUNED.get(sig)
    if cfg is not None:
        return cfg
    if os.environ.get("TNKIT_AMD_AUTOTUNE", "1") == "0" or p.macs == 0:
        cfg = heuristic_cfg(p, A.dtype)
        _TUNED[sig] = cfg
        return cfg
    R = p.desc[:, 5]
    cands = [c for c in TILE_CONFIGS if not (c == 4 and R.max(initial=0) > 64)
             and (c not in DMA_CONFIGS or _grouped(p, A.device).dma)]
    if _grouped(p, A.device).small_ok():
        cands.append(SMALL)
    if _grouped(p, A.device).mid_ok():
        cands.append(MID)
    if torch.cuda.is_current_stream_capturing():
        cfg = heuristic_cfg(p, A.dtype)
        return cfg
    times = {}
    st = torch.cuda.current_stream(A.device)
    for c in cands:
        grouped_gemm(p, A, Bm, out, cfg=c)
    # two interleaved rounds, best of each candidate: one slow round (host launch gaps, a
    # profiler) does not decide the choice
    for _ in range(2):
        for c in cands:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(st)
            for _ in range(reps):
                grouped_gemm(p, A, Bm, out, cfg=c)
            e1.record(st)
            e1.synchronize()
            t = e0.elapsed_time(e1) / reps
            times[c] = min(times.get(c, t), t)
    cfg = min(times, key=times.get)
    _TUNED[sig] = cfg
    TUNE_LOG.append((sig, times, cfg))
    return cfg


# ---------------------------------------------------------------------------------------------
# block tensors
# ---------------------------------------------------------------------------------------------


class BlockTensor:
    """A U(1) block tensor on the device: host layout `struct`, flat buffer `data`."""
    __slots__ = ("struct", "data")
    is_block = True

    def __init__(self, struct: B.Struct, data: torch.Tensor):
        if data.dim() != 1 or data.numel() != struct.size or not data.is_contiguous():
            raise ValueError(f"buffer of shape {tuple(data.shape)} for a layout of size "
                             f"{struct.size}")
        self.struct, self.data = struct, data

    @property
    def legs(self):
        return self.struct.legs

    @property
    def charge(self) -> int:
        return self.struct.charge

    @property
    def ndim(self) -> int:
        return len(self.struct.legs)

    @property
    def shape(self):
        return self.struct.shape

    @property
    def dtype(self):
        return self.data.dtype

    def transpose(self, perm: Sequence[int]) -> "BlockTensor":
        perm = tuple(int(i) for i in perm)
        if perm == tuple(range(self.ndim)):
            return self
        return BlockTensor(B.permuted(self.struct, perm), self.data)

    def canonical(self, dtype: Optional[torch.dtype] = None) -> "BlockTensor":
        c = B.canonical(self.legs, self.charge)
        if self.struct is c and (dtype is None or dtype == self.dtype):
            return self
        return BlockTensor(c, gather(B.to_canonical(self.struct), self.data, dtype))

    def conj(self) -> "BlockTensor":
        """Complex conjugate of real data: the dual layout (every sign and the charge flipped)."""
        return BlockTensor(B.dual(self.struct), self.data)

    def with_data(self, data: torch.Tensor) -> "BlockTensor":
        return BlockTensor(self.struct, data)

    def __repr__(self):
        return (f"BlockTensor(shape={self.shape}, charge={self.charge}, "
                f"blocks={len(self.struct.blocks)}, stored={self.struct.size}, {self.dtype})")


def random(legs, charge: int, rng: np.random.Generator, device, dtype=torch.float64):
    """Standard-normal entries in every allowed block (drawn on the host from `rng`)."""
    s = B.canonical(tuple(legs), int(charge))
    return BlockTensor(s, torch.from_numpy(rng.standard_normal(s.size)).to(device, dtype))


def tensordot(a: BlockTensor, b: BlockTensor, axes, cfg: Optional[int] = None) -> BlockTensor:
    """Contract legs axes[0] of a with legs axes[1] of b (tnkit/blocks.py:1156-1192): the result
    has a's free legs then b's, total charge a.charge + b.charge, in tnkit's output layout."""
    if a.dtype != b.dtype:
        raise TypeError(f"tensordot: dtypes {a.dtype} and {b.dtype} differ")
    ax_a, ax_b = tuple(int(i) for i in axes[0]), tuple(int(j) for j in axes[1])
    p = B.plan(a.struct, b.struct, ax_a, ax_b)
    A = gather(p.dev_a, a.data)
    Bm = gather(p.dev_b, b.data)
    out = torch.empty(p.out.size, dtype=a.dtype, device=a.data.device)
    grouped_gemm(p, A, Bm, out, cfg)
    return BlockTensor(p.out, out)


def einsum(sub: str, a: BlockTensor, b: BlockTensor) -> BlockTensor:
    """A two-operand einsum without batch indices on block tensors
    (tnkit/block_backend.py:121-131)."""
    ax_a, ax_b, perm = B.parse_einsum(sub)
    return tensordot(a, b, (ax_a, ax_b)).transpose(perm)


# ---------------------------------------------------------------------------------------------
# the two-site operator
# ---------------------------------------------------------------------------------------------


class BlockTwoSite:
    """H_eff of one two-site block problem on canonical theta buffers.

    `dtype` is the arithmetic type (float64, or float32 for the FP32 matrix-core operator); the
    input and output buffers keep their own dtype (float64 Krylov vectors), converted inside the
    first and last gathers. The operator is tnkit's four contractions (`TWO_SITE_SUBS`); it is
    applied as three (`APPLY_SUBS`): W1 and W2 are contracted once, at construction, into one
    six-leg operator (w, s1, s2) -> (s1', s2', w'), so the memory-bound middle of the matvec is
    one pass over the w-wide intermediate instead of two passes and a gather between them.
    `flops` = 2 * sum of the macs of tnkit's four contractions (the count the TF/s figures use);
    `stages` are the three applied ones."""

    def __init__(self, L: BlockTensor, W1: BlockTensor, W2: BlockTensor, R: BlockTensor,
                 theta_legs=None, charge: int = 0, dtype: Optional[torch.dtype] = None,
                 io_dtype: torch.dtype = torch.float64):
        dev = L.data.device
        self.device = dev
        self.dtype = dtype or L.dtype
        self.io_dtype = io_dtype
        if self.dtype not in _REAL or io_dtype not in _REAL:
            raise TypeError("BlockTwoSite: float64 or float32 only")
        if theta_legs is None:
            theta_legs = (L.legs[2].flip(), W1.legs[3].flip(), W2.legs[3].flip(),
                          R.legs[2].flip())
        self.in_struct = B.canonical(tuple(theta_legs), int(charge))
        x = self.in_struct
        self.macs = []
        for sub, T in zip(TWO_SITE_SUBS, (L, W1, W2, R)):
            ax_a, ax_b, perm = B.parse_einsum(sub)
            p = B.plan(T.struct, x, ax_a, ax_b)
            self.macs.append(p.macs)
            x = B.permuted(p.out, perm)
        W12 = einsum(MPO_PAIR_SUB, W1, W2)
        x = self.in_struct
        self.stages = []
        for sub, T in zip(APPLY_SUBS, (L, W12, R)):
            ax_a, ax_b, perm = B.parse_einsum(sub)
            p = B.plan(T.struct, x, ax_a, ax_b)
            # zeros: the K-padding rows of the stacks are never written by a gather
            A = gather(p.dev_a, T.data, self.dtype,
                       torch.zeros(p.size_a, dtype=self.dtype, device=dev))
            Bbuf = torch.zeros(p.size_b, dtype=self.dtype, device=dev)
            out = torch.empty(p.out.size, dtype=self.dtype, device=dev)
            self.stages.append([p, A, Bbuf, out, None])
            x = B.permuted(p.out, perm)
        self.out_struct = B.canonical(x.legs, x.charge)
        if self.out_struct is not self.in_struct:
            raise ValueError("BlockTwoSite: the output legs differ from theta's")
        self.final = B.to_canonical(x)
        self.flops = 2 * sum(self.macs)
        self.size = self.in_struct.size
        self._graph = None

    def tune(self, reps: int = 3) -> None:
        """Autotune every stage's tile config on the operator's own buffers."""
        for st in self.stages:
            p, A, Bbuf, out, _ = st
            st[4] = autotune(p, A, Bbuf, out, reps)

    def apply(self, theta: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """out = H_eff theta for a flat canonical theta (io dtype)."""
        if theta.numel() != self.size or theta.dtype != self.io_dtype:
            raise ValueError(f"BlockTwoSite.apply: theta of {theta.numel()} {theta.dtype}, "
                             f"expected {self.size} {self.io_dtype}")
        src = theta.reshape(-1)
        for st in self.stages:
            p, A, Bbuf, o, cfg = st
            gather(p.dev_b, src, self.dtype, Bbuf, prezeroed=True)
            if cfg is None:
                cfg = st[4] = autotune(p, A, Bbuf, o)
            grouped_gemm(p, A, Bbuf, o, cfg)
            src = o
        if out is None:
            out = torch.empty(self.size, dtype=self.io_dtype, device=self.device)
        return gather(self.final, src, self.io_dtype, out)

    __call__ = apply

    def graph(self):
        """A replayable hipGraph of `apply` with static input and output buffers (run `apply`
        once eagerly first; this tunes every stage). Returns f(theta) -> new output tensor: copy
        in, replay, clone out, as tnkit's block_graph.solve does (tnkit/block_graph.py:137-151)."""
        if self._graph is None:
            x = torch.zeros(self.size, dtype=self.io_dtype, device=self.device)
            self.apply(x)                      # tune and warm outside the capture
            static_in = torch.zeros_like(x)
            static_out = torch.empty_like(x)
            torch.cuda.synchronize(self.device)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.apply(static_in, static_out)
            self._graph = (g, static_in, static_out)
        g, si, so = self._graph

        def run(theta: torch.Tensor) -> torch.Tensor:
            si.copy_(theta.reshape(-1))
            g.replay()
            return so.clone()
        return run
