"""Two-site TDVP on the HIP device, single (`tdvp2_evolve`) and batched over k trajectories
(`evolve_batch`), after tnkit/tdvp.py:555-800 and tnkit/batch_tdvp.py:572-1002.

One implementation serves both: a trajectory is a batch of one. The k legs are stacked on a
leading axis and zero-padded at each bond position (tnkit's three ragged modes: 'pad' to the
widest bond the cap allows, 'tight' to the widest rank kept on that step, 'group' per entry bond
profile); every
contraction is one batched libtnkit_hip launch for all legs (`kernels.*_b`), the Krylov
exponential keeps its scalars on the device and reads them back once per solve, and the
truncation is one batched Gram eigendecomposition with per-leg ranks and masks, so padding
stays exact zeros. Per-leg bond caps (`chi_max` as a list) extend tnkit's single cap.

The local step (tnkit/tdvp.py:555-650): theta = M[i] M[i+1]; theta <- exp(-i H2 dt/2) theta;
cut to the leg's rank; renormalise S; update the environment; back-evolve the one-site centre
with exp(+i H1 dt/2) unless it sits at the chain end. Sweep right then left (2nd-order Strang).
"""
from __future__ import annotations

import dataclasses
import os
from collections import OrderedDict
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import kernels
from .mps import MPS

KRYLOV_TOL_DEFAULT = 1e-10


class BatchGeometryError(ValueError):
    """The legs are not one program on different data (tnkit/batch_tdvp.py:93-101)."""


class BatchKrylovError(RuntimeError):
    """A leg's Krylov residual exceeds the tolerance at the largest allowed m
    (tnkit/batch_tdvp.py:104-116)."""


@dataclasses.dataclass
class BatchTDVPOptions:
    """Settings of `evolve_batch` (tnkit/batch_tdvp.py:134-201). Served: evolver '2tdvp',
    ragged 'pad' / 'tight' / 'group', expm_where 'eigh'; `chi_max` may be one cap or one per
    leg."""

    backend: Any = "hip"
    evolver: str = "2tdvp"
    krylov_m: Union[int, str] = 20
    krylov_tol: float = KRYLOV_TOL_DEFAULT
    krylov_m_max: Optional[int] = None
    expm_where: str = "eigh"
    ragged: str = "pad"
    chi_max: Optional[Union[int, Sequence[int]]] = None
    svd_tol: float = 0.0
    renorm: bool = True
    canonicalize: bool = True
    keep_device: bool = False
    measure: Optional[Callable[[Any], Any]] = None
    meas_every: int = 1
    stats: Optional[Dict[str, Any]] = None
    low_precision: bool = False        # complex64 matvec with complex128 Krylov vectors
    progress: Optional[Callable[[int, float], None]] = None   # called after each time step


# --------------------------------------------------------------------------------------------
# the batched Krylov exponential
# --------------------------------------------------------------------------------------------


_GRAPHS: "OrderedDict" = OrderedDict()       # key -> (graph, static op, V, ab)
_GRAPH_SEEN: set = set()
_GRAPH_MAX_BYTES = 256 << 20                 # largest Krylov basis captured (bytes)
_GRAPH_CACHE_BYTES = 2048 << 20              # all captured bases together
_graph_bytes = [0]


def graphs_enabled() -> bool:
    return os.environ.get("TNKIT_AMD_GRAPHS", "1") != "0"


def clear_graphs() -> None:
    _GRAPHS.clear()
    _GRAPH_SEEN.clear()
    _graph_bytes[0] = 0


def _krylov_graph(key, op, make_mv, k: int, m: int, n: int, dev, rag=None):
    """The captured Krylov loop for `key`, or None to run eagerly. As tnkit's site-step graphs
    (tnkit/cudagraph.py:87-342, block_graph.py:137-151) a shape signature runs eagerly the first
    time and is captured the second: the m matvecs and orthogonalisation steps become one
    hipGraph over static copies of the operator, the basis V and the coefficients ab. Bases
    above _GRAPH_MAX_BYTES are not captured (their kernels are long enough on their own); the
    cache keeps the most recent signatures within _GRAPH_CACHE_BYTES."""
    g = _GRAPHS.get(key)
    if g is not None:
        _GRAPHS.move_to_end(key)
        return g
    nbytes = k * m * n * 16
    if nbytes > _GRAPH_MAX_BYTES or torch.cuda.is_current_stream_capturing():
        return None
    if key not in _GRAPH_SEEN:
        _GRAPH_SEEN.add(key)
        return None
    while _GRAPHS and _graph_bytes[0] + nbytes > _GRAPH_CACHE_BYTES:
        _, old = _GRAPHS.popitem(last=False)
        _graph_bytes[0] -= old[4]
    sop = tuple(t.clone() for t in op)
    V = torch.zeros((k, m, n), dtype=torch.complex128, device=dev)
    ab = torch.zeros((k, 2 * m), dtype=torch.float64, device=dev)
    ws = kernels.expm_ws(k, dev)
    mv = make_mv(sop)
    torch.cuda.synchronize(dev)
    graph = torch.cuda.CUDAGraph()
    srag = None if rag is None else (sop[rag[0]], rag[1], rag[2])
    with torch.cuda.graph(graph):
        for j in range(m):
            w = mv(V[:, j]).reshape(k, n)
            kernels.expm_step_c(V, w, ab, j, ws, ragged=srag)
    entry = (graph, sop, V, ab, nbytes, ws)
    _GRAPHS[key] = entry
    _graph_bytes[0] += nbytes
    return entry


def _expm_tridiag(ab: torch.Tensor, beta0: torch.Tensor, tau: complex, m: int):
    """Coefficients c (k, m) of exp(tau T_leg) beta0 e_0 in the Krylov basis and each leg's Saad
    estimate |beta_m tau [e^{tau T}]_{m-1,0}| beta0, from the k tridiagonals T (alphas then betas
    in `ab`), all on the device: one batched eigendecomposition of the k tridiagonals
    (tnkit/batch_tdvp.py:164-168, 'eigh' on the device; the Jacobi kernel up to m = 64). A
    tridiagonal the kernel leaves unconverged gets an infinite estimate, which the per-step
    residual check turns into a larger m or an error."""
    k = ab.shape[0]
    al, be = ab[:, :m], ab[:, m:2 * m]
    T = torch.diag_embed(al)
    if m > 1:
        off = be[:, :m - 1]
        T = T + torch.diag_embed(off, 1) + torch.diag_embed(off, -1)
    T = T.contiguous()
    if m <= kernels.EIGH_JACOBI_MAX_N and _JACOBI:
        lam, Z, sw = kernels.eigh_jacobi(T)
        ok = sw > 0
    else:
        lam, Z = torch.linalg.eigh(T)
        ok = torch.ones(k, dtype=torch.bool, device=T.device)
    Zc = Z.to(torch.complex128)
    ph = torch.exp(tau * lam.to(torch.complex128))
    e0 = torch.matmul(Zc, (ph * Zc[:, 0, :]).unsqueeze(-1))[..., 0]      # [e^{tau T}]_{:, 0}
    c = beta0.to(torch.complex128)[:, None] * e0
    est = be[:, m - 1] * abs(tau) * e0[:, m - 1].abs() * beta0
    est = torch.where(ok, est, torch.full_like(est, float("inf")))
    return c, est


def lanczos_expm_fixed_batched(matvec_b, v: torch.Tensor, tau: complex, m: int,
                               resid=None, graph=None) -> torch.Tensor:
    """exp(tau H_leg) v_leg for k legs at once, from a fixed m Krylov steps
    (tnkit/batch_tdvp.py:247-324, tnkit/eigensolver.py:95-217): CGS2 with alpha from the first
    projection, clamped division instead of a breakdown branch, exp of the k tridiagonals by
    their eigendecompositions on the device, so a solve reads nothing back. `v` is (k, n)
    complex128. `resid` (length k: a device float64 tensor, or a numpy array, which costs a host
    read) is raised to each leg's Saad a-posteriori error estimate
    |beta_m tau [e^{tau T}]_{m-1,0}| beta0. `graph` = (key, operator tensors, make_mv[, rag])
    lets the Krylov loop run as a captured hipGraph; rag = (index of the legs' live extents among
    the operator tensors, rowlen, bpad) lets its orthogonalisation skip each leg's padding."""
    k, n = v.shape
    m = max(1, min(int(m), n))
    dev = v.device
    beta0 = torch.linalg.vector_norm(v, dim=1)
    entry = None
    rag = None
    if graph is not None and len(graph) > 3 and graph[3] is not None:
        rag = (graph[1][graph[3][0]], graph[3][1], graph[3][2])
    if graph is not None and graphs_enabled():
        key, op, make_mv = graph[:3]
        entry = _krylov_graph((key, m), op, make_mv, k, m, n, dev,
                              None if len(graph) < 4 else graph[3])
    if entry is not None:
        g, sop, V, ab, _, _ = entry
        for st, t in zip(sop, op):
            st.copy_(t)
        V[:, 0].copy_(v / torch.clamp(beta0, min=1e-200)[:, None])
        g.replay()
    else:
        V = torch.zeros((k, m, n), dtype=torch.complex128, device=dev)
        ab = torch.zeros((k, 2 * m), dtype=torch.float64, device=dev)
        V[:, 0] = v / torch.clamp(beta0, min=1e-200)[:, None]
        ws = kernels.expm_ws(k, dev)
        for j in range(m):
            w = matvec_b(V[:, j]).reshape(k, n).to(torch.complex128).contiguous()
            kernels.expm_step_c(V, w, ab, j, ws, ragged=rag)
    c, est = _expm_tridiag(ab, beta0, tau, m)
    if isinstance(resid, torch.Tensor):
        torch.maximum(resid, est, out=resid)
    elif resid is not None:
        resid[:] = np.maximum(resid, est.cpu().numpy())
    out = torch.empty((k, n), dtype=torch.complex128, device=dev)
    kernels.lincomb_c(V, c, out, m, ragged=rag)
    return out


# --------------------------------------------------------------------------------------------
# batch state
# --------------------------------------------------------------------------------------------


def _pad3(t: torch.Tensor, a: int, b: int) -> torch.Tensor:
    out = torch.zeros((a, t.shape[1], b), dtype=t.dtype, device=t.device)
    out[:t.shape[0], :, :t.shape[2]] = t
    return out


def widest_profile(N: int, d: int, cap: int) -> List[int]:
    """The widest each of the N + 1 bonds of an open chain (outer bonds 1) can be under a
    two-site sweep capped at `cap`: min(cap, d^i, d^(N-i)) (tnkit/batch_tdvp.py:897-911)."""
    return [int(min(cap, d ** min(i, 40), d ** min(N - i, 40))) for i in range(N + 1)]


class _Batch:
    """k legs on a leading axis, padded to the widest bond at each position, with the true
    per-leg ranks (tnkit/batch_tdvp.py:403-496)."""

    def __init__(self, states, mpos, dev, caps, ragged: str = "tight"):
        self.k = len(states)
        self.N = len(states[0])
        self.d = int(states[0][0].shape[1])
        self.caps = np.asarray(caps, dtype=np.int64)
        self.ragged = ragged
        ranks = np.array([[1] + [int(t.shape[2]) for t in s[:-1]] + [1] for s in states])
        self.rank = [ranks[:, i] for i in range(self.N + 1)]
        self.dims = [int(ranks[:, i].max()) for i in range(self.N + 1)]
        if ragged == "pad":                 # fixed shapes: the widest each bond can ever be
            self.dims = [max(w, x) for w, x in zip(self.dims, widest_profile(
                self.N, self.d, int(self.caps.max())))]
        self.M = []
        for i in range(self.N):
            self.M.append(torch.stack([_pad3(s[i], self.dims[i], self.dims[i + 1])
                                       for s in states]).contiguous())
        self.W = [torch.stack([w[i] for w in mpos]).contiguous() for i in range(self.N)]
        self.L: List[Optional[torch.Tensor]] = [None] * (self.N + 1)
        self.R: List[Optional[torch.Tensor]] = [None] * (self.N + 1)
        one = torch.ones((self.k, 1, 1, 1), dtype=torch.complex128, device=dev)
        self.L[0], self.R[self.N] = one, one.clone()

    def mask(self, bond: int, nk: int) -> torch.Tensor:
        r = torch.as_tensor(self.rank[bond], device=self.M[0].device)
        return (torch.arange(nk, device=r.device)[None, :] < r[:, None])


def _right_canonical(M: List[torch.Tensor]) -> List[torch.Tensor]:
    """Right-canonical form of one complex MPS by LQ per site (the library's QR), centre at 0,
    normalised."""
    return _right_canonical_all([M])[0]


def _right_canonical_all(states: Sequence[Sequence[torch.Tensor]]) -> List[List[torch.Tensor]]:
    """`_right_canonical` of several MPS: states whose site shapes all agree go through each
    site's LQ together, one batched QR per site (a batched factorisation equals the single ones
    bit for bit)."""
    groups = {}
    for q, M in enumerate(states):
        groups.setdefault(tuple(tuple(t.shape) for t in M), []).append(q)
    out: List[Optional[List[torch.Tensor]]] = [None] * len(states)
    for shapes, idx in groups.items():
        Ms = [torch.stack([states[q][i] for q in idx]) for i in range(len(shapes))]
        for i in range(len(Ms) - 1, 0, -1):
            k, Dl, d, Dr = Ms[i].shape
            Q, Rr = kernels.qr_any(Ms[i].reshape(k, Dl, d * Dr).mT)
            Ms[i] = Q.mT.reshape(k, -1, d, Dr).contiguous()
            Ms[i - 1] = torch.matmul(Ms[i - 1], Rr.mT.unsqueeze(1)).contiguous()
            Ms[i - 1] = Ms[i - 1] / torch.linalg.vector_norm(
                Ms[i - 1].reshape(k, -1), dim=1)[:, None, None, None]
        Ms[0] = Ms[0] / torch.linalg.vector_norm(Ms[0].reshape(len(idx), -1), dim=1)[:, None, None,
                                                                                      None]
        for j, q in enumerate(idx):
            out[q] = [t[j].contiguous() for t in Ms]
    return out


def _eigh_robust(G: torch.Tensor, add: Optional[torch.Tensor] = None, deferred: bool = False,
                 k: Optional[int] = None):
    """Batched eigh (ascending) of the Hermitian positive semidefinite G + diag(add) on the
    device. Orders up to 64 go to the one-workgroup-per-matrix Jacobi kernel
    (`kernels.eigh_jacobi`); complex orders 65..1024 to the library's batched divide-and-conquer
    solver (`kernels.heev_top_async`), which returns only the k leading eigenvectors when k is
    given (Z then has k columns, those of the k largest eigenvalues, ascending like the rest);
    other orders, and any matrix a library solver leaves unconverged or non-finite, go to
    rocSOLVER's batched eigh. That one returns non-finite eigenvectors for complex matrices with
    exactly decoupled rows (a leg's padded rows), so the matrices it gets carry a 1e-30 coupling
    on the first off-diagonals (`_couple`); a leg that still comes back non-finite is decomposed
    by an SVD of the unshifted G (G = U diag(S) U^H for a PSD matrix), whose eigenvalues it
    returns ascending like eigh. Raises if a leg stays non-finite. With `deferred`, a library
    result is returned with the device mask of legs to redo, (w, Z, redo), and the caller, which
    reads w back anyway, finishes with `_eigh_redo`; otherwise (w, Z) after the check here."""
    r = int(G.shape[-1])
    if (r <= kernels.EIGH_JACOBI_MAX_N and G.dtype in (torch.complex128, torch.float64)
            and _JACOBI):
        w, Z, sw = kernels.eigh_jacobi(G.contiguous(), None if add is None
                                       else add.to(torch.float64).contiguous())
        redo = (sw < 0) | (~torch.isfinite(Z).flatten(1).all(1)) | (~torch.isfinite(w).all(1))
    elif G.dtype == torch.complex128 and r <= kernels.HEEV_MAX_N and _HEEV:
        kk = r if k is None else max(1, min(int(k), r))
        Gs = G if add is None else G + torch.diag_embed(add.to(G.dtype))
        w, Zt, info = kernels.heev_top_async(Gs.contiguous(), kk)
        Z = torch.flip(Zt, dims=[2])
        redo = ((info != 0).expand(G.shape[0]) | (~torch.isfinite(Z).flatten(1).all(1))
                | (~torch.isfinite(w).all(1)))
    else:
        w, Z = _eigh_robust_vendor(G, add)
        return (w, Z, None) if deferred else (w, Z)
    if deferred:
        return w, Z, redo
    if not bool(redo.any()):
        return w, Z
    idx = torch.nonzero(redo).flatten()
    wv, Zv = _eigh_robust_vendor(G[idx], None if add is None else add[idx])
    w[idx], Z[idx] = wv, Zv[:, :, r - Z.shape[2]:]
    return w, Z


def _eigh_redo(G: torch.Tensor, add: Optional[torch.Tensor], w: torch.Tensor, Z: torch.Tensor,
               redo: np.ndarray) -> None:
    """Replace in place the legs of a deferred `_eigh_robust` that a library solver left
    unconverged or non-finite (host mask) by the vendor eigensolver's."""
    idx = torch.as_tensor(np.nonzero(redo)[0], device=G.device)
    wv, Zv = _eigh_robust_vendor(G[idx], None if add is None else add[idx])
    w[idx], Z[idx] = wv, Zv[:, :, Zv.shape[2] - Z.shape[2]:]


_JACOBI = os.environ.get("TNKIT_AMD_JACOBI", "1") != "0"
_HEEV = os.environ.get("TNKIT_AMD_HEEV", "1") != "0"


def _couple(Gs: torch.Tensor) -> torch.Tensor:
    """Gs plus 1e-30 of its largest entry on the first off-diagonals. The batched complex eigh
    (rocSOLVER) returns non-finite vectors for a matrix with exactly decoupled rows (the padded
    rows of a leg narrower than the batch); this coupling is below the rounding of every nonzero
    entry, so it only replaces exact zeros and moves no eigenvalue by more than 1e-30 of the
    spectrum's scale."""
    r = Gs.shape[-1]
    if r < 2:
        return Gs
    eps = 1e-30 * Gs.abs().amax(dim=(-2, -1))
    band = torch.diag_embed(eps[:, None].expand(-1, r - 1).to(Gs.dtype), offset=1)
    return Gs + band + band.mH


def _eigh_robust_vendor(G: torch.Tensor, add: Optional[torch.Tensor] = None):
    Gs = G if add is None else G + torch.diag_embed(add.to(G.dtype))
    w, Z = torch.linalg.eigh(_couple(Gs))
    bad = (~torch.isfinite(Z).flatten(1).all(1)) | (~torch.isfinite(w).all(1))
    if not bool(bad.any()):
        return w, Z
    idx = torch.nonzero(bad).flatten()
    w[idx], Z[idx] = _eigh_by_svd(G[idx], None if add is None else add[idx])
    if not (bool(torch.isfinite(Z).all()) and bool(torch.isfinite(w).all())):
        raise FloatingPointError("Hermitian eigendecomposition returned non-finite values")
    return w, Z


def _eigh_by_svd(G: torch.Tensor, add: Optional[torch.Tensor]):
    """Ascending eigenpairs of a batch (k, r, r) of PSD Gram matrices plus their padding shifts
    `add` (negative on exactly zero rows, else 0), by one batched SVD: the padded diagonal is
    lifted to distinct values above the whole spectrum (2 (tr G + 1) + j / r), so the padding's
    unit vectors come out as the leading singular vectors, apart from the live null directions
    (a plain SVD of G would mix the two inside its zero singular subspace). Those leading
    triplets are then given back their negative shift values and everything is sorted
    ascending, as eigh would return it; for a PSD matrix G = U diag(S) U^H."""
    k, r, _ = G.shape
    dev = G.device
    if add is None:
        add = torch.zeros((k, r), dtype=torch.float64, device=dev)
    dead = add != 0
    tr = torch.diagonal(G, dim1=1, dim2=2).real.sum(1)
    lift = 2.0 * (tr[:, None] + 1.0) + torch.arange(r, device=dev, dtype=torch.float64)[None, :] / r
    Gp = G + torch.diag_embed(torch.where(dead, lift, torch.zeros_like(lift)).to(G.dtype))
    U, S, _ = torch.linalg.svd(Gp)
    # a leading triplet belongs to the padding: its vector is the unit vector of a dead index
    col = torch.argmax(U.abs(), dim=1)                                  # (k, r) index per vector
    is_dead = torch.gather(dead, 1, col)
    vals = torch.where(is_dead, torch.gather(add.to(torch.float64), 1, col), S.to(torch.float64))
    order = torch.argsort(vals, dim=1)
    w = torch.gather(vals, 1, order)
    Z = torch.gather(U, 2, order[:, None, :].expand(k, r, r))
    return w, Z


def _gram_cut(mat: torch.Tensor, ranks_cap: np.ndarray, ceiling: np.ndarray, svd_tol: float,
              iso: str, width: int = 0, live=None):
    """Batched truncation of mat (k, m, n) by the Gram matrix of its smaller side: per-leg rank =
    the fewest values whose relative dropped weight stays under svd_tol, capped by the leg's
    cap and structural ceiling (tnkit/batch_tdvp.py:572-599). Returns U (k, m, nk), S (k, nk),
    Vh (k, nk, n), ranks, relative discarded weight per leg. The `iso` factor is an exact
    isometry on each leg's kept columns (tnkit/linalg.py:162-204). `live` = (row indices, column
    indices) of mat that any leg can fill (the union of the legs' unpadded extents), or None:
    the Gram matrix is formed over those only, so a bond held wider than every leg's rank (the
    fixed shapes of 'pad') costs the rank, not the width; the other indices are exact zeros of
    every leg and get zero rows in the factors."""
    k, m, n = mat.shape
    if iso == "right":
        lv = None if live is None else (live[1], live[0])
        U, S, Vh, ranks, disc = _gram_cut(mat.mH.resolve_conj().contiguous(), ranks_cap, ceiling,
                                          svd_tol, "left", width, lv)
        return (Vh.mH.resolve_conj().contiguous(), S, U.mH.resolve_conj().contiguous(), ranks,
                disc)
    tall = m > n                                        # the smaller Gram side, as tnkit does
    side = None if live is None else live[1 if tall else 0]
    if side is not None and int(side.numel()) == (n if tall else m):
        side = None
    X = mat if side is None else mat.index_select(2 if tall else 1, side)
    G = (torch.matmul(X.mH.resolve_conj().contiguous(), X) if tall
         else torch.matmul(X, X.mH.resolve_conj().contiguous()))       # (k, r, r)
    G = 0.5 * (G + G.mH)
    # padded legs leave exact zero rows; a large exactly degenerate zero block makes the batched
    # eigensolver return NaN, so those rows get distinct negative diagonal entries, which keeps
    # them below the nonnegative Gram spectrum and out of every kept set
    r_ = G.shape[1]
    dg = torch.diagonal(G, dim1=1, dim2=2).real
    shift = -(1.0 + torch.arange(r_, device=G.device, dtype=torch.float64) / r_)
    add = torch.where(dg == 0, shift[None, :], torch.zeros_like(dg))
    # only the leading vectors any leg can keep are formed (the stacked width below is at most
    # this many)
    kz = min(r_, max(int(np.minimum(ranks_cap, ceiling).max()), int(width), 1))
    w, Z, redo = _eigh_robust(G, add, deferred=True, k=kz)
    # one host read: the spectra, and which legs the Jacobi kernel asks to redo
    both = torch.cat([w, redo.to(w.dtype)[:, None]], 1) if redo is not None else w
    both = both.cpu().numpy()
    if redo is not None and both[:, -1].any():
        _eigh_redo(G, add, w, Z, both[:, -1] > 0)
        both = w.cpu().numpy()
    elif redo is not None:
        both = both[:, :-1]
    w = torch.flip(w, dims=[1]).clamp(min=0.0)
    Z = torch.flip(Z, dims=[2])
    r = int(Z.shape[1])
    wh = np.clip(both[:, ::-1], 0.0, None)
    tot = wh.sum(axis=1)
    tot = np.where(tot > 0, tot, 1.0)
    tail = np.cumsum(wh[:, ::-1], axis=1)[:, ::-1]
    ranks = np.empty(k, dtype=np.int64)
    disc = np.empty(k)
    for leg in range(k):
        lim = max(1, min(int(ranks_cap[leg]), int(ceiling[leg]), r))
        drop = np.append(tail[leg] / tot[leg], 0.0)
        if svd_tol > 0.0:
            ok = np.nonzero(drop[1:lim + 1] <= svd_tol)[0]
            ranks[leg] = int(ok[0]) + 1 if len(ok) else lim
        else:
            # no budget: keep the cap. Gram eigenvalues of Schmidt values below ~1e-8 round to
            # zero, and a zero-weight test on them would drop directions an SVD keeps.
            ranks[leg] = lim
        disc[leg] = float(drop[ranks[leg]])
    nk = min(max(int(ranks.max()), int(width)), min(m, n))   # stacked width: the ragged mode's
    mk = torch.arange(nk, device=mat.device)[None, :] < torch.as_tensor(ranks, device=mat.device)[:, None]
    nz = min(nk, r)
    if side is None and nz == nk:
        Zk = Z[:, :, :nk] * mk[:, None, :]
    else:                       # back to the full side, zero rows (and masked columns) elsewhere
        Zk = torch.zeros((k, n if tall else m, nk), dtype=Z.dtype, device=Z.device)
        rows = slice(None) if side is None else side
        Zk[:, rows, :nz] = Z[:, :, :nz]
        Zk = Zk * mk[:, None, :]
    if tall:                                            # A Z = Q R, C = R Z^H (tnkit/linalg.py:155-159)
        # a leg keeping fewer than nk columns has exact zero columns at the end of A Z, which
        # rocSOLVER's batched QR (the fallback beyond the library QR's limits) turns into NaN;
        # they are filled with fixed Gaussian columns, on either QR, so both give the same
        # factors. The kept columns come first, so their Q and R do not see the fill, and the
        # fill's rows of R meet the zero rows of Z^H below.
        AZ = torch.matmul(mat, Zk)
        if bool((ranks < nk).any()):                   # host ranks: no device read
            gen = torch.Generator(device=mat.device)
            gen.manual_seed(12345)
            fill = torch.randn(AZ.shape, generator=gen, device=mat.device, dtype=torch.float64)
            AZ = torch.where(mk[:, None, :], AZ, fill.to(AZ.dtype))
        Q, Rr = kernels.qr(AZ) if kernels.qr_supported(AZ) else torch.linalg.qr(AZ)
        U = Q * mk[:, None, :]
        C = torch.matmul(Rr, Zk.mH.resolve_conj().contiguous())        # (k, nk, n) carried factor
    else:
        U = Zk
        C = torch.matmul(U.mH.resolve_conj().contiguous(), mat)
    S = torch.linalg.vector_norm(C, dim=2)
    Vh = C / torch.where(S > 0, S, torch.ones_like(S))[:, :, None]
    return U, S.to(torch.float64), Vh, ranks, disc


# --------------------------------------------------------------------------------------------
# the sweep
# --------------------------------------------------------------------------------------------


def _live(Bt: _Batch, left: int, right: int) -> torch.Tensor:
    """(k, 2) int32 device tensor of each leg's true bonds at positions left and right: the live
    block of the padded batch, outside which the legs' tensors are zero."""
    lv = np.stack([Bt.rank[left], Bt.rank[right]], axis=1).astype(np.int32)
    return torch.as_tensor(lv, device=Bt.M[0].device)


def _matvec_two(Bt: _Batch, i: int, shape, low: bool):
    """(matvec, graph context) of the batched two-site H_eff at bond i. The legs' true bonds ride
    along (`_live`), so the matvec's products skip what lies outside each leg's own block."""
    op = kernels.fused_two_site_b(Bt.L[i], Bt.W[i], Bt.W[i + 1], Bt.R[i + 2],
                                  dtype=torch.complex64 if low else None)
    op = (*op, _live(Bt, i, i + 2))

    def make(o):
        def mv(x):
            return kernels.apply_two_site_b(x.reshape(shape), *o[:3], live=o[3]).reshape(
                shape[0], -1)
        return mv
    key = ("two", tuple(shape), tuple(tuple(t.shape) for t in op), low)
    k, a, d1, d2, b = shape
    return make(op), (key, op, make, (3, d1 * d2 * b, b))


def _matvec_one(Bt: _Batch, L, W, R, shape, low: bool, live: torch.Tensor):
    """One-site H_eff (tnkit/_dmrg.py:192-199) as the two-site kernel with d2 = 1 and the
    identity of the right MPO bond as the second MPO tensor: S[(s r)][(w s')] = W[w][r][s][s'].
    `live` = the legs' true bonds either side (see `_live`). Returns (matvec, graph context)."""
    k, a, d, b = shape
    wr = int(W.shape[2])
    Wid = torch.eye(wr, dtype=W.dtype, device=W.device).reshape(1, wr, wr, 1, 1)
    Wid = Wid.expand(k, -1, -1, -1, -1).contiguous()
    op = kernels.fused_two_site_b(L, W, Wid, R, dtype=torch.complex64 if low else None)
    op = (*op, live)

    def make(o):
        def mv(x):
            return kernels.apply_two_site_b(x.reshape(k, a, d, 1, b), *o[:3],
                                            live=o[3]).reshape(k, -1)
        return mv
    key = ("one", tuple(shape), tuple(tuple(t.shape) for t in op), low)
    return make(op), (key, op, make, (3, d * b, b))


def _step(Bt: _Batch, i: int, dt: float, m: int, svd_tol: float, renorm: bool, moving: str,
          resid: torch.Tensor, low: bool) -> np.ndarray:
    k, d = Bt.k, Bt.d
    A, Bm = Bt.M[i], Bt.M[i + 1]
    a, b = A.shape[1], Bm.shape[3]
    theta = torch.matmul(A.reshape(k, a * d, -1), Bm.reshape(k, Bm.shape[1], -1))
    shape = (k, a, d, d, b)
    mv, gctx = _matvec_two(Bt, i, shape, low)
    theta = lanczos_expm_fixed_batched(mv, theta.reshape(k, -1), -0.5j * dt, m, resid, gctx)
    mat = theta.reshape(k, a * d, d * b)
    ceiling = np.minimum(Bt.rank[i] * d, Bt.rank[i + 2] * d)
    # how wide the stacked bond is held (tnkit/batch_tdvp.py:602-628): 'pad' at its fixed
    # widest, 'tight' at the widest rank any leg kept, 'group' at what the spectrum and the cap
    # allow; the rank masks carry each leg's own rank in all three
    if Bt.ragged == "pad":
        width = Bt.dims[i + 1]
    elif Bt.ragged == "group":
        width = int(min(Bt.caps.max(), a * d, d * b))
    else:
        width = 0
    # the rows (a, s1) and columns (s2, b) that some leg fills: a below the widest left rank,
    # b below the widest right rank
    ra, rb = int(Bt.rank[i].max()), int(Bt.rank[i + 2].max())
    live = None
    if ra < a or rb < b:
        dev = mat.device
        rows = torch.arange(ra * d, device=dev)
        cols = (torch.arange(d, device=dev)[:, None] * b + torch.arange(rb, device=dev)[None, :])
        live = (rows, cols.reshape(-1))
    U, S, Vh, ranks, disc = _gram_cut(mat, Bt.caps, ceiling, svd_tol,
                                      "left" if moving == "right" else "right", width, live)
    nk = U.shape[2]
    Bt.rank[i + 1] = ranks
    Bt.dims[i + 1] = nk
    if renorm:
        S = S / torch.linalg.vector_norm(S, dim=1, keepdim=True)
    Sc = S.to(torch.complex128)
    if moving == "right":
        Bt.M[i] = U.reshape(k, a, d, nk).contiguous()
        Bt.L[i + 1] = kernels.update_left_env_b(Bt.L[i], Bt.M[i], Bt.W[i])
        C = (Sc[:, :, None] * Vh).reshape(k, nk, d, b).contiguous()
        if i < Bt.N - 2:
            mv, gctx = _matvec_one(Bt, Bt.L[i + 1], Bt.W[i + 1], Bt.R[i + 2], (k, nk, d, b), low,
                                   _live(Bt, i + 1, i + 2))
            C = lanczos_expm_fixed_batched(mv, C.reshape(k, -1), 0.5j * dt, m, resid,
                                           gctx).reshape(k, nk, d, b)
        Bt.M[i + 1] = C.contiguous()
    else:
        Bt.M[i + 1] = Vh.reshape(k, nk, d, b).contiguous()
        Bt.R[i + 1] = kernels.update_right_env_b(Bt.R[i + 2], Bt.M[i + 1], Bt.W[i + 1])
        C = (U * Sc[:, None, :]).reshape(k, a, d, nk).contiguous()
        if i > 0:
            mv, gctx = _matvec_one(Bt, Bt.L[i], Bt.W[i], Bt.R[i + 1], (k, a, d, nk), low,
                                   _live(Bt, i, i + 1))
            C = lanczos_expm_fixed_batched(mv, C.reshape(k, -1), 0.5j * dt, m, resid,
                                           gctx).reshape(k, a, d, nk)
        Bt.M[i] = C.contiguous()
    return disc


def _trim(Bt: _Batch, leg: int) -> MPS:
    out = []
    for i in range(Bt.N):
        a, b = int(Bt.rank[i][leg]), int(Bt.rank[i + 1][leg])
        out.append(Bt.M[i][leg, :a, :, :b].contiguous())
    return MPS(out, center=0)


def evolve_batch(states: Sequence[Sequence[Any]], hams: Any, dt: float, n_steps: int,
                 options: Optional[Union[BatchTDVPOptions, Dict[str, Any]]] = None,
                 **kwargs: Any) -> Tuple[List[MPS], List[Any]]:
    """Evolve k MPS of one geometry as one batched program (tnkit/batch_tdvp.py:721-857).

    `hams` is one MPO for every leg or a list of k MPOs with the same bond structure. `ragged`
    says how legs of different bonds share the stacked shapes: 'pad' holds every bond at the
    widest the cap allows (fixed shapes), 'tight' at the widest rank any leg kept on that step,
    'group' steps the legs of each entry bond profile as their own batch. Returns (states,
    records): each leg trimmed to its own bonds, and [(t, measure(states_t)), ...], one such
    list per group when 'group' made several."""
    if options is None:
        o = BatchTDVPOptions(**kwargs)
    elif isinstance(options, dict):
        o = BatchTDVPOptions(**{**options, **kwargs})
    else:
        o = dataclasses.replace(options, **kwargs) if kwargs else options
    if o.evolver != "2tdvp":
        raise ValueError(f"evolver {o.evolver!r}: the HIP batch serves '2tdvp'")
    if o.ragged not in RAGGED_MODES:
        raise ValueError(f"ragged must be one of {RAGGED_MODES}, got {o.ragged!r}")
    if o.chi_max is None:
        raise ValueError("'2tdvp' needs chi_max")
    if not isinstance(o.krylov_m, (int, np.integer)):
        raise ValueError("krylov_m must be an int on the HIP backend")
    k = len(states)
    if k == 0:
        return [], []
    N = len(states[0])
    if any(len(s) != N for s in states):
        raise BatchGeometryError("legs differ in their number of sites")
    mpos = hams if (isinstance(hams, (list, tuple)) and len(hams) == k
                    and isinstance(hams[0], (list, tuple))) else [hams] * k
    if len(mpos) != k or any(len(h) != N for h in mpos):
        raise BatchGeometryError("one MPO per leg, each with one tensor per site")
    for i in range(N):
        if len({tuple(h[i].shape) for h in mpos}) != 1:
            raise BatchGeometryError(f"MPO bond structure differs between legs at site {i}")
    caps = (np.full(k, int(o.chi_max)) if np.isscalar(o.chi_max)
            else np.asarray(list(o.chi_max), dtype=np.int64))
    if len(caps) != k:
        raise ValueError("chi_max: one cap, or one per leg")
    dev = torch.device("cuda", torch.cuda.current_device())
    c = lambda x: torch.as_tensor(np.asarray(x) if not isinstance(x, torch.Tensor) else x) \
        .to(dev, torch.complex128)  # noqa: E731
    st = [[c(t) for t in s] for s in states]
    if o.canonicalize:
        st = _right_canonical_all(st)
    wd = {id(h): [c(w) for w in h] for h in mpos}     # one device copy per distinct MPO
    mp = [wd[id(h)] for h in mpos]
    groups = _partition(st, o.ragged)
    out_legs: List[Any] = [None] * k
    all_records = []
    merged: Dict[str, Any] = {"krylov_m": int(o.krylov_m), "resid_max": [0.0] * k,
                              "disc_per_leg": [0.0] * k, "bonds": None,
                              "groups": [len(g) for g in groups], "n_groups": len(groups),
                              "ragged": o.ragged, "fill": []}
    bonds = [[0] * k for _ in range(N + 1)]
    for g in groups:
        legs, recs, gst = _run_group([st[n] for n in g], [mp[n] for n in g], caps[g], dt,
                                     n_steps, o, dev)
        all_records.append(recs)
        merged["krylov_m"] = max(merged["krylov_m"], gst["krylov_m"])
        merged["fill"].append(gst["fill"])
        for slot, n in enumerate(g):
            out_legs[n] = legs[slot]
            merged["resid_max"][n] = gst["resid_max"][slot]
            merged["disc_per_leg"][n] = gst["disc_per_leg"][slot]
            for i in range(N + 1):
                bonds[i][n] = gst["bonds"][i][slot]
    merged["bonds"] = bonds
    merged["fill"] = float(np.mean(merged["fill"]))
    if o.stats is not None:
        o.stats.update(merged)
    return out_legs, (all_records[0] if len(groups) == 1 else all_records)


RAGGED_MODES = ("pad", "tight", "group")


def _bond_profile(s) -> Tuple[int, ...]:
    return tuple([1] + [int(t.shape[2]) for t in s[:-1]] + [1])


def _partition(states, ragged: str) -> List[List[int]]:
    """Leg indices that step as one batch: every leg for 'pad' and 'tight', the legs of equal
    entry bond profile for 'group' (in order of first appearance; tnkit/batch_tdvp.py:887-895)."""
    if ragged != "group":
        return [list(range(len(states)))]
    seen: Dict[Tuple[int, ...], List[int]] = {}
    for n, s in enumerate(states):
        seen.setdefault(_bond_profile(s), []).append(n)
    return list(seen.values())


def _run_group(st, mpos, caps, dt: float, n_steps: int, o: "BatchTDVPOptions", dev):
    """Step one batch of legs for n_steps (tnkit/batch_tdvp.py:937-1002): records, the legs
    trimmed to their own bonds, and this group's statistics."""
    k, N = len(st), len(st[0])
    Bt = _Batch(st, mpos, dev, caps, o.ragged)
    for i in range(N - 1, 0, -1):
        Bt.R[i] = kernels.update_right_env_b(Bt.R[i + 1], Bt.M[i], Bt.W[i])
    m = int(o.krylov_m)
    m_max = int(o.krylov_m_max) if o.krylov_m_max is not None else 2 * m
    records = []

    def snapshot():
        legs = [_trim(Bt, j) for j in range(k)]
        return legs if o.keep_device else [MPS([t.cpu().numpy() for t in s], 0) for s in legs]

    if o.measure is not None:
        records.append((0.0, o.measure(snapshot())))
    disc_tot = np.zeros(k)
    resid_max = np.zeros(k)
    fills = []
    for step in range(n_steps):
        resid_d = torch.zeros(k, dtype=torch.float64, device=dev)
        for i in range(0, N - 1):
            disc_tot += _step(Bt, i, dt, m, o.svd_tol, o.renorm, "right", resid_d, o.low_precision)
        for i in range(N - 2, -1, -1):
            disc_tot += _step(Bt, i, dt, m, o.svd_tol, o.renorm, "left", resid_d, o.low_precision)
        resid = resid_d.cpu().numpy()                 # the step's one read of the residuals
        resid_max = np.maximum(resid_max, resid)
        fills.append(float(np.mean([np.mean(Bt.rank[i] / float(Bt.dims[i]))
                                    for i in range(N + 1)])))
        bad = np.nonzero(resid > o.krylov_tol)[0]
        if len(bad):
            if m >= m_max:
                raise BatchKrylovError(f"leg {int(bad[0])}: Krylov residual {resid[bad[0]]:.2e} "
                                       f"above krylov_tol={o.krylov_tol:g} at krylov_m={m} "
                                       f"(krylov_m_max={m_max})")
            m = min(m_max, m + 2)
        if o.measure is not None and (step + 1) % o.meas_every == 0:
            records.append(((step + 1) * dt, o.measure(snapshot())))
        if o.progress is not None:
            o.progress(step + 1, (step + 1) * dt)
    gst = {"krylov_m": m, "resid_max": resid_max.tolist(), "disc_per_leg": disc_tot.tolist(),
           "bonds": [list(map(int, r)) for r in Bt.rank],
           "fill": float(np.mean(fills)) if fills else 1.0}
    return [_trim(Bt, j) for j in range(k)], records, gst


def tdvp2_evolve(mps: Sequence[Any], mpo: Sequence[Any], chi: int, dt: float, n_steps: int,
                 krylov_m: int = 20, measure: Optional[Callable] = None, meas_every: int = 1,
                 svd_tol: float = 0.0, renorm: bool = True, canonicalize: bool = True,
                 low_precision: bool = False) -> Tuple[MPS, List[Tuple[float, Any]]]:
    """Two-site TDVP of one state (tnkit/tdvp.py:653-800): a batch of one."""
    meas = None if measure is None else (lambda legs: measure(legs[0]))
    out, rec = evolve_batch([list(mps)], [list(mpo)], dt, n_steps, chi_max=int(chi),
                            krylov_m=krylov_m, measure=meas, meas_every=meas_every,
                            svd_tol=svd_tol, renorm=renorm, canonicalize=canonicalize,
                            low_precision=low_precision, keep_device=True)
    return out[0], rec
