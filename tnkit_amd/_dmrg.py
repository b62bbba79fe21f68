"""Two-site DMRG on the HIP backend: tnkit's dense sweep (tnkit/_dmrg.py:439-799) with its
contractions, environment updates and Krylov-vector work in libtnkit_hip.

Tensor axes follow tnkit (tnkit/_dmrg.py:7-11):
  MPS  M[i]: (a, s, b);   MPO W[i]: (wl, wr, s_out, s_in);   L[i], R[i]: (bra, mpo, ket)

The front door `dmrg(mpo, chi, options=None, **kwargs)` and `DMRGOptions` keep tnkit's names,
defaults and return values (E, MPS, info). What is served: dense real MPOs, two-site steps,
eig in {'lanczos', 'lanczos-restart', 'lanczos-mixed'}, trunc in {'svd', 'gram', 'randomized',
'randomized-mixed'}, and subspace expansion in the opening sweeps (expand, expand_alpha,
expand_decay). Other options of tnkit raise ValueError naming what is not provided.

'lanczos-mixed' here runs stage 1 on a float32 operator on the float32 matrix cores (real, not
tnkit's complex64 cast; tnkit/backend.py:554-563) with float64 Krylov vectors, and stage 2 in
float64 (tnkit/eigensolver.py:677-725).
"""
from __future__ import annotations

import dataclasses
import time
from typing import Any, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import kernels
from .backend import HipBackend, get_backend
from .eigensolver import (check_eig_options, lanczos_ground, lanczos_ground_mixed,
                          lanczos_ground_restart)
from .expansion import enriched_split, expansion_term
from .linalg import gram_truncation, randomized_svd_device
from .mps import MPS


# --------------------------------------------------------------------------------------------
# contractions (tnkit/_dmrg.py:162-231), HIP only
# --------------------------------------------------------------------------------------------


def fused_two_site(L, W1, W2, R, be=None, dtype=None):
    """The two-site operator (LT, S, RT) built once per local problem (tnkit/_dmrg.py:172-180).
    The matrices are in the device layouts of libtnkit_hip (see `kernels.fused_two_site`)."""
    return kernels.fused_two_site(L, W1, W2, R, dtype=dtype)


def apply_two_site_fused(theta, LT, S, RT):
    """H_eff theta from the operator of `fused_two_site` (tnkit/_dmrg.py:183-189)."""
    return kernels.apply_two_site_fused(theta, LT, S, RT)


def apply_two_site(theta, L, W1, W2, R, be=None):
    """H_eff theta (tnkit/_dmrg.py:162-169); on this backend the fused kernels compute it."""
    return kernels.apply_two_site_fused(theta, *kernels.fused_two_site(L, W1, W2, R))


def update_left_env(L, A, W, be=None):
    """Grow the left environment past a left-canonical A (tnkit/_dmrg.py:216-222)."""
    return kernels.update_left_env(L, A, W)


def update_right_env(R, B, W, be=None):
    """Grow the right environment past a right-canonical B (tnkit/_dmrg.py:225-231)."""
    return kernels.update_right_env(R, B, W)


# --------------------------------------------------------------------------------------------
# states
# --------------------------------------------------------------------------------------------


def random_mps(N: int, chi: int, d: int = 2, rng=None, be=None) -> List[torch.Tensor]:
    """Gaussian MPS drawn on the host from `rng` in site order, then moved to the device, so
    the state equals tnkit's for the same generator (tnkit/_dmrg.py:253-261)."""
    be = be or get_backend("hip-real")
    rng = np.random.default_rng(0) if rng is None else rng
    dims = [1] + [min(chi, d ** min(i, N - i)) for i in range(1, N)] + [1]
    return [be.randn(rng, (dims[i], d, dims[i + 1])) for i in range(N)]


def right_canonicalize(M, be=None):
    """Right-canonical form with the running centre renormalised at every site; centre at
    site 0 (tnkit/_dmrg.py:264-292). Mutates and returns M."""
    be = be or get_backend("hip-real")
    N = len(M)
    for i in range(N - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        U, S, Vh = be.svd(M[i].reshape(Dl, d * Dr))
        M[i] = Vh.reshape(-1, d, Dr).contiguous()
        nxt = torch.matmul(M[i - 1].reshape(-1, Dl), U * S).reshape(M[i - 1].shape[0], d, -1)
        M[i - 1] = (nxt / torch.linalg.norm(nxt)).contiguous()
    M[0] = (M[0] / torch.linalg.norm(M[0])).contiguous()
    if isinstance(M, MPS):
        M.center = 0
    return M


def right_canonicalize_lq(M, be=None):
    """Right-canonical form by an LQ factorisation per site (QR of the transpose) instead of an
    SVD; the same state up to a gauge on each bond, as tnkit's block path does
    (tnkit/_dmrg.py:295-306). Much cheaper than a device SVD at large bond dimension."""
    be = be or get_backend("hip-real")
    N = len(M)
    for i in range(N - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        Q, Rr = torch.linalg.qr(M[i].reshape(Dl, d * Dr).t())      # M^T = Q R  ->  M = R^T Q^T
        M[i] = Q.t().reshape(-1, d, Dr).contiguous()
        nxt = torch.matmul(M[i - 1].reshape(-1, Dl), Rr.t()).reshape(M[i - 1].shape[0], d, -1)
        M[i - 1] = (nxt / torch.linalg.norm(nxt)).contiguous()
    M[0] = (M[0] / torch.linalg.norm(M[0])).contiguous()
    if isinstance(M, MPS):
        M.center = 0
    return M


# --------------------------------------------------------------------------------------------
# options and the front door
# --------------------------------------------------------------------------------------------


@dataclasses.dataclass
class DMRGOptions:
    """Settings of `dmrg` apart from `mpo` and `chi`; tnkit's fields and defaults
    (tnkit/_dmrg.py:312-436). Fields of tnkit that this backend does not serve keep their
    defaults and are refused when changed."""

    n_sweeps: int = 8
    tol: float = 1e-10
    trunc: str = "svd"
    eig_tol: float = 1e-9
    eig: str = "lanczos-restart"
    warm_start: bool = True
    reorth: Union[str, bool] = "full"
    eig_tol_warmup: Optional[float] = None
    warmup_sweeps: int = 2
    seed: int = 0
    init_mps: Optional[Any] = None
    backend: Any = "auto"
    max_iter: Optional[int] = None
    verbose: bool = False
    restart_m: int = 40
    restart_k: int = 10
    mixed_tol_lo: float = 1e-4
    eig_stop: str = "residual"
    reorth_alg: str = "cgs2"
    charge: Any = None
    sites: int = 2
    expand: Optional[int] = None
    expand_alpha: float = 3e-3
    expand_decay: float = 0.3
    rtol: float = 1e-10


def _resolve(options, kwargs) -> DMRGOptions:
    """Flat keywords override the fields of `options` (tnkit/_options.py:16-50)."""
    if options is None:
        o = DMRGOptions()
    elif isinstance(options, DMRGOptions):
        o = options
    elif isinstance(options, dict):
        o = DMRGOptions(**options)
    else:
        raise TypeError(f"options must be DMRGOptions or dict, got {type(options).__name__}")
    names = {f.name for f in dataclasses.fields(DMRGOptions)}
    bad = set(kwargs) - names
    if bad:
        raise TypeError(f"dmrg() got unexpected keyword argument(s) {sorted(bad)}")
    return dataclasses.replace(o, **kwargs) if kwargs else o


def dmrg(mpo: Sequence[Any], chi: Optional[int] = None,
         options: Optional[Union[DMRGOptions, Dict[str, Any]]] = None,
         **legacy_kwargs: Any) -> Tuple[float, MPS, Dict[str, Any]]:
    """Ground state of a dense MPO by two-site DMRG on the HIP device (tnkit/_dmrg.py:439-522).

    Returns (energy, mps, info) with tnkit's info keys: sweeps, energies, sweep_s, backend,
    blocks, eig, discarded_weight, matvec_calls, phase_time, mixed_counts, sites, expand,
    expand_alphas."""
    if chi is not None and "chi" in legacy_kwargs:
        raise TypeError("dmrg() got multiple values for argument 'chi'")
    if chi is None and "chi" not in legacy_kwargs:
        raise TypeError("dmrg() missing required argument 'chi' (bond dimension)")
    chi = legacy_kwargs.pop("chi", chi)
    o = _resolve(options, legacy_kwargs)
    return _dmrg(mpo, int(chi), o)


def _check(o: DMRGOptions, mpo) -> None:
    if getattr(mpo[0], "is_block", False):
        raise ValueError("block MPOs are not served by the dense HIP sweep")
    if o.sites != 2:
        raise ValueError(f"sites must be 2 on the HIP backend, got {o.sites!r}")
    if o.eig not in ("lanczos", "lanczos-restart", "lanczos-mixed"):
        raise ValueError(f"unknown eig solver {o.eig!r}; expected one of "
                         "['lanczos', 'lanczos-mixed', 'lanczos-restart']")
    if o.trunc not in ("svd", "gram", "randomized", "randomized-mixed"):
        raise ValueError(f"unknown trunc {o.trunc!r} for the HIP backend; expected 'svd', "
                         "'gram', 'randomized' or 'randomized-mixed'")
    if o.charge is not None:
        raise ValueError("charge= needs a block MPO")
    check_eig_options(o)


BLOCK_EXPAND = 4    # opening expansion sweeps on block MPOs when expand=None (tnkit/_dmrg.py:45)


def expand_sweeps(o, block: bool) -> int:
    """The number of opening sweeps with subspace expansion, with tnkit's rules and errors
    (tnkit/_dmrg.py:579-590): None means min(4, n_sweeps - 2) on block MPOs and 0 on dense ones;
    the last sweep is always without expansion."""
    expand = o.expand
    if expand is not None and (isinstance(expand, bool)
                               or not isinstance(expand, (int, np.integer))):
        raise TypeError(f"expand must be an int or None, got {expand!r}")
    if expand is None:
        expand = min(BLOCK_EXPAND, max(0, o.n_sweeps - 2)) if block else 0
    if not 0 <= expand < o.n_sweeps:
        raise ValueError(f"expand must be in [0, n_sweeps), so that the last sweep is unbiased; "
                         f"got expand={expand!r}, n_sweeps={o.n_sweeps!r}")
    if expand and not (o.expand_alpha > 0 and 0 < o.expand_decay <= 1):
        raise ValueError("expand needs expand_alpha > 0 and 0 < expand_decay <= 1 "
                         "(pass expand=0 to turn expansion off)")
    return int(expand)



class DMRGEngine:
    """The state of one dense two-site DMRG run on the HIP device: MPS, environments, MPO and
    counters, with `solve_block` the local step of tnkit/_dmrg.py:677-727 and `sweep` the
    right-then-left pass of tnkit/_dmrg.py:775-793.

    The start is made right-canonical by LQ factorisations (`canonical='lq'`, default) rather
    than tnkit's SVDs: the same state in another bond gauge, which two-site DMRG does not see
    (its local problems are covariant under the gauge), at a fraction of the device cost of an
    SVD at large bond dimension. `canonical='svd'` follows tnkit/_dmrg.py:264-292 exactly."""

    def __init__(self, mpo, chi: int, o: DMRGOptions, canonical: str = "lq"):
        _check(o, mpo)
        be = get_backend(o.backend, mpo=mpo, init_mps=o.init_mps)
        if not isinstance(be, HipBackend):
            raise ValueError(f"backend {o.backend!r} is not a HIP backend")
        if be.is_complex:
            raise ValueError("the HIP DMRG sweep serves real Hamiltonians (backend 'hip-real'); "
                             "this MPO has a significant imaginary part")
        self.be, self.o, self.chi = be, o, int(chi)
        N = len(mpo)
        self.N = N
        d = int(mpo[0].shape[2])
        self.mpo = [be.array(W) for W in mpo]
        canon = right_canonicalize if canonical == "svd" else right_canonicalize_lq
        if o.init_mps is not None:
            self.M = canon([be.array(t) for t in o.init_mps], be)
        else:
            self.M = canon(random_mps(N, chi, d, np.random.default_rng(o.seed), be), be)
        self.mixed_counts = {"lo": 0, "hi": 0}
        self.rng = np.random.default_rng(o.seed + 1)
        self.L: List[Optional[torch.Tensor]] = [None] * (N + 1)
        self.R: List[Optional[torch.Tensor]] = [None] * (N + 1)
        self.L[0] = be.ones((1, 1, 1))
        self.R[N] = be.ones((1, 1, 1))
        for i in range(N - 1, 0, -1):
            self.R[i] = update_right_env(self.R[i + 1], self.M[i], self.mpo[i])
        self.total_disc = 0.0
        self.matvec_calls = 0
        self.phase = {"eig": 0.0, "trunc": 0.0, "env": 0.0}
        self.last_eig_info: Dict[str, Any] = {}

    @classmethod
    def from_state(cls, mpo, chi: int, o: DMRGOptions, M, Ls: Dict[int, torch.Tensor],
                   Rs: Sequence[Optional[torch.Tensor]]) -> "DMRGEngine":
        """An engine over a given mixed-canonical state and the environments it already has
        (device tensors), for resuming or timing a part of a sweep."""
        eng = cls.__new__(cls)
        _check(o, mpo)
        eng.be = get_backend(o.backend, mpo=mpo)
        eng.o, eng.chi, eng.N = o, int(chi), len(mpo)
        eng.mpo = [eng.be.array(W) for W in mpo]
        eng.M = list(M)
        eng.L = [None] * (eng.N + 1)
        for k, v in Ls.items():
            eng.L[k] = v
        eng.R = list(Rs)
        eng.mixed_counts = {"lo": 0, "hi": 0}
        eng.rng = np.random.default_rng(o.seed + 1)
        eng.total_disc, eng.matvec_calls = 0.0, 0
        eng.phase = {"eig": 0.0, "trunc": 0.0, "env": 0.0}
        eng.last_eig_info = {}
        return eng

    def eigsolve(self, theta0, op, op_lo, cur_tol):
        o, be = self.o, self.be
        shape = tuple(int(s) for s in theta0.shape)
        n = int(np.prod(shape))
        count = {"n": 0}
        ws = kernels.two_site_ws(theta0, *op)

        def matvec(x):
            count["n"] += 1
            return kernels.apply_two_site_fused(x.reshape(shape), *op, ws=ws).reshape(-1)

        def matvec_lo(x):
            count["n"] += 1
            return kernels.apply_two_site_fused(x.reshape(shape), *op_lo, ws=ws).reshape(-1)

        v0 = theta0.reshape(-1) if o.warm_start else be.randn(self.rng, n)
        mi = o.max_iter if o.max_iter is not None else min(n, max(200, 4 * self.chi))
        be.sync()
        t0 = time.perf_counter()
        if o.eig == "lanczos-mixed":
            e, x, inf = lanczos_ground_mixed(matvec_lo, matvec, v0, be, tol=cur_tol, max_iter=mi,
                                             m=o.restart_m, k=o.restart_k, tol_lo=o.mixed_tol_lo,
                                             counts=self.mixed_counts)
        elif o.eig == "lanczos":
            e, x, inf = lanczos_ground(matvec, v0, be, tol=cur_tol, max_iter=mi, reorth=o.reorth,
                                       stop=o.eig_stop)
        else:
            e, x, inf = lanczos_ground_restart(matvec, v0, be, tol=cur_tol, max_iter=mi,
                                               m=o.restart_m, k=o.restart_k,
                                               reorth_alg=o.reorth_alg)
        be.sync()
        self.phase["eig"] += time.perf_counter() - t0
        self.matvec_calls += count["n"]
        self.last_eig_info = dict(inf, matvecs=count["n"])
        return e, x.reshape(shape)

    def solve_block(self, i: int, direction: str, cur_tol: float, alpha: float = 0.0) -> float:
        """One two-site step; alpha > 0 splits the centre with subspace expansion
        (tnkit/_dmrg.py:700-705) instead of the truncation."""
        o, be, M, L, R, chi = self.o, self.be, self.M, self.L, self.R, self.chi
        W1, W2 = self.mpo[i], self.mpo[i + 1]
        a, s1 = M[i].shape[0], M[i].shape[1]
        s2, b = M[i + 1].shape[1], M[i + 1].shape[2]
        theta0 = torch.matmul(M[i].reshape(a * s1, -1), M[i + 1].reshape(M[i + 1].shape[0], -1))
        theta0 = theta0.reshape(a, s1, s2, b)
        op = fused_two_site(L[i], W1, W2, R[i + 2])
        op_lo = fused_two_site(L[i], W1, W2, R[i + 2], dtype=torch.float32) \
            if o.eig == "lanczos-mixed" else None
        e, x = self.eigsolve(theta0, op, op_lo, cur_tol)
        del op, op_lo
        t0 = time.perf_counter()
        if alpha:
            right = direction == "right"
            P = expansion_term(x, L[i], W1, W2, R[i + 2], right)
            A, C, disc = enriched_split(x, P, alpha, chi, be, right)
            C = (C / torch.linalg.norm(C)).contiguous()
            M[i], M[i + 1] = (A, C) if right else (C, A)
            be.sync()
            self.phase["trunc"] += time.perf_counter() - t0
            self.total_disc += disc
            self._env(i, direction)
            return e
        mat = x.reshape(a * s1, s2 * b)
        if o.trunc == "svd":
            U, S, Vh = be.svd(mat)
            keep = min(chi, int(S.shape[0]))
            disc = float(torch.sum(S[keep:] ** 2).item())
            U, S, Vh = U[:, :keep], S[:keep], Vh[:keep]
        elif o.trunc == "gram":
            U, S, Vh, disc = gram_truncation(mat, chi, "left" if direction == "right" else "right")
        else:
            U, S, Vh, disc = randomized_svd_device(mat, chi, be, rng=self.rng,
                                                   mixed=o.trunc == "randomized-mixed")
        S = S / torch.linalg.norm(S)
        if direction == "right":
            M[i] = U.reshape(a, s1, -1).contiguous()
            M[i + 1] = (S[:, None] * Vh).reshape(-1, s2, b).contiguous()
        else:
            M[i] = (U * S).reshape(a, s1, -1).contiguous()
            M[i + 1] = Vh.reshape(-1, s2, b).contiguous()
        be.sync()
        self.phase["trunc"] += time.perf_counter() - t0
        self.total_disc += disc
        self._env(i, direction)
        return e

    def _env(self, i: int, direction: str) -> None:
        t0 = time.perf_counter()
        if direction == "right":
            self.L[i + 1] = update_left_env(self.L[i], self.M[i], self.mpo[i])
        else:
            self.R[i + 1] = update_right_env(self.R[i + 2], self.M[i + 1], self.mpo[i + 1])
        self.be.sync()
        self.phase["env"] += time.perf_counter() - t0

    def sweep(self, cur_tol: float, alpha: float = 0.0) -> float:
        E = None
        for i in range(0, self.N - 1):
            E = self.solve_block(i, "right", cur_tol, alpha)
        for i in range(self.N - 2, -1, -1):
            E = self.solve_block(i, "left", cur_tol, alpha)
        return E


def _dmrg(mpo, chi: int, o: DMRGOptions):
    if getattr(mpo[0], "is_block", False):          # U(1) block tensors: the block sweep
        from .block_dmrg import block_dmrg
        return block_dmrg(mpo, chi, o)
    expand = expand_sweeps(o, False)
    eng = DMRGEngine(mpo, chi, o)
    be = eng.be
    energies: List[float] = []
    sweep_s: List[float] = []
    alphas: List[float] = []
    prev_E = None
    E = None
    sweep = -1
    for sweep in range(o.n_sweeps):
        t_sweep = time.perf_counter()
        cur_tol = (o.eig_tol_warmup if (o.eig_tol_warmup is not None and sweep < o.warmup_sweeps)
                   else o.eig_tol)
        alpha = o.expand_alpha * o.expand_decay ** sweep if sweep < expand else 0.0
        alphas.append(alpha)
        E = eng.sweep(cur_tol, alpha)
        energies.append(E)
        be.sync()
        sweep_s.append(time.perf_counter() - t_sweep)
        if o.verbose:
            print(f"  sweep {sweep + 1:2d}: E = {E:.12f}  (chi={chi}, trunc={o.trunc}, {be.name})")
        if prev_E is not None and abs(E - prev_E) < max(o.tol, o.rtol * abs(E)):
            break
        prev_E = None if alpha else E
    info = {"sweeps": sweep + 1, "energies": energies, "backend": be.name, "blocks": False,
            "eig": o.eig, "discarded_weight": eng.total_disc, "matvec_calls": eng.matvec_calls,
            "phase_time": eng.phase, "mixed_counts": eng.mixed_counts, "sites": 2,
            "expand": expand, "expand_alphas": alphas, "sweep_s": sweep_s}
    return E, MPS(eng.M, center=0), info
