"""Excited states by penalised two-site DMRG on the HIP device (tnkit/dmrg_excited.py).

`dmrg_excited` finds the lowest eigenstate of H + w sum_i |psi_i><psi_i| for previously found
states psi_i (tnkit/dmrg_excited.py:545-836, dense branch). The penalty never becomes an MPO:
each previous state carries its overlap environments EL[i] (sites < i) and ER[i] (sites >= i),
and in the two-site block at (i, i+1) it is the block vector

    V[a, s1, s2, b] = EL[i][a, x] P[i][x, s1, z] P[i+1][z, s2, y] ER[i+2][b, y],

so the local operator is H_eff + w sum_j V_j V_j^T. Its matvec is the fused two-site kernel
followed by one `tnh_rank_update` launch pair (dot products against every V_j reduced on the
device, then the rank-k update), with no host read inside the Krylov loop. deflate='project'
solves Q H_eff Q + sigma (1 - Q) instead, Q the projector off the block vectors
(tnkit/dmrg_excited.py:384-418), applied by the same kernel with weight -1 on an orthonormal
basis of their span.

The overlap environments and block vectors are small (chi^2 d per state) and use the vendor
GEMM (torch.matmul); the H_eff matvec, the environment updates and the Krylov-vector work are
the libtnkit_hip kernels of the ground-state sweep.

`dmrg_tower` climbs the spectrum by repeated calls with tnkit's adaptive penalty and overlap
retry (tnkit/dmrg_excited.py:1062-1128); `mps_inner` and `mpo_expectation` are tnkit's
observables of the same module (:140-211) on device tensors.
"""
from __future__ import annotations

import dataclasses
import math
import time
import warnings
from typing import Any, List, Optional, Sequence

import numpy as np
import torch

from . import kernels
from ._dmrg import DMRGEngine, DMRGOptions
from .backend import HipBackend, get_backend
from .eigensolver import check_reorth_periodic, lanczos_ground, lanczos_ground_restart
from .mps import MPS

__all__ = ["mps_inner", "mpo_expectation", "dmrg_excited", "DMRGExcitedOptions", "dmrg_tower",
           "DeflationWarning", "OVERLAP_WARN"]

PROJ_RTOL = 1e-5     # singular values of the stacked block vectors kept in the projector basis
PROJ_KEEP = 0.5      # least norm a projected solution must keep in the range of Q
OVERLAP_WARN = 0.5   # |<psi_i|M>|^2 above which the state found is most likely psi_i again


class DeflationWarning(RuntimeWarning):
    """An excited-state solve returned a state that overlaps a state it was deflated against."""


def _dev(be, x):
    return x if isinstance(x, torch.Tensor) and x.device == be.device and x.dtype == be.dtype \
        else be.array(x)


# --------------------------------------------------------------------------------------------
# overlaps and expectation values (tnkit/dmrg_excited.py:140-211)
# --------------------------------------------------------------------------------------------


def _grow_EL(E, A, P):
    """E'[b, y] = sum E[a, x] conj(A)[a, s, b] P[x, s, y] (tnkit/dmrg_excited.py:359-361)."""
    a, s, b = A.shape
    x, _, y = P.shape
    t = torch.matmul(E.t(), A.conj().reshape(a, s * b))                   # (x, s b)
    return torch.matmul(t.reshape(x * s, b).t(), P.reshape(x * s, y)).contiguous()


def _grow_ER(E, A, P):
    """E'[a, x] = sum conj(A)[a, s, b] E[b, y] P[x, s, y] (tnkit/dmrg_excited.py:364-366)."""
    a, s, b = A.shape
    x, _, y = P.shape
    t = torch.matmul(A.conj().reshape(a * s, b), E)                       # (a s, y)
    return torch.matmul(t.reshape(a, s * y), P.reshape(x, s * y).t()).contiguous()


def _block_vector(EL, P1, P2, ER):
    """V[a, s1, s2, b] of one previous state in the current block (tnkit/dmrg_excited.py:369-374)."""
    x, s1, z = P1.shape
    _, s2, y = P2.shape
    t = torch.matmul(EL, P1.reshape(x, s1 * z)).reshape(-1, z)            # (a s1, z)
    t = torch.matmul(t, P2.reshape(z, s2 * y)).reshape(-1, y)             # (a s1 s2, y)
    return torch.matmul(t, ER.t()).reshape(-1)                            # (a s1 s2 b)


def mps_inner(A, B, be=None) -> complex:
    """<A|B> of two dense MPSs in any gauge, A conjugated (tnkit/dmrg_excited.py:140-175)."""
    be = be or get_backend("hip-real", mpo=None, init_mps=list(A) + list(B))
    A = [_dev(be, t) for t in A]
    B = [_dev(be, t) for t in B]
    E = be.ones((1, 1))
    for a, b in zip(A, B):
        E = _grow_EL(E, a, b)
    return complex(E.reshape(-1)[0].item())


def mpo_expectation(mps, mpo, be=None) -> complex:
    """<psi|W|psi> / <psi|psi> of an MPO by the left-environment kernels
    (tnkit/dmrg_excited.py:185-211); the state need not be normalised."""
    be = be or get_backend("hip-real", mpo=mpo, init_mps=mps)
    M = [_dev(be, t) for t in mps]
    F = be.ones((1, 1, 1))
    for A, W in zip(M, mpo):
        F = kernels.update_left_env(F, A, _dev(be, W))
    return complex(F.reshape(-1)[0].item()) / mps_inner(M, M, be)


# --------------------------------------------------------------------------------------------
# options and the front door
# --------------------------------------------------------------------------------------------


@dataclasses.dataclass
class DMRGExcitedOptions:
    """Settings of `dmrg_excited` apart from `mpo` and `chi`: tnkit's fields and defaults
    (tnkit/dmrg_excited.py:421-501). `charge` (block and symmetric MPOs) is refused."""

    prev_states: Sequence[Any] = ()
    penalty: float = 1.0
    n_sweeps: int = 12
    tol: float = 1e-11
    eig_tol: float = 1e-9
    trunc: str = "svd"
    seed: int = 0
    init_mps: Optional[Any] = None
    backend: Any = "auto"
    verbose: bool = False
    eig_max_iter: Optional[int] = None
    reorth: str = "full"
    eig: str = "lanczos-restart"
    restart_m: int = 40
    restart_k: int = 10
    reorth_alg: str = "cgs2"
    deflate: str = "penalty"
    charge: Any = None
    time_budget_s: Optional[float] = None
    start_noise: float = 0.0


def _resolve(options, kwargs) -> DMRGExcitedOptions:
    if options is None:
        o = DMRGExcitedOptions()
    elif isinstance(options, DMRGExcitedOptions):
        o = options
    elif isinstance(options, dict):
        o = DMRGExcitedOptions(**options)
    else:
        raise TypeError("options must be DMRGExcitedOptions or dict, got "
                        f"{type(options).__name__}")
    names = {f.name for f in dataclasses.fields(DMRGExcitedOptions)}
    bad = set(kwargs) - names
    if bad:
        raise TypeError(f"dmrg_excited() got unexpected keyword argument(s) {sorted(bad)}")
    return dataclasses.replace(o, **kwargs) if kwargs else o


def _warn_overlaps(overlaps, penalty, deflate, where, retry=""):
    worst = max(overlaps, default=0.0) ** 2
    if worst > OVERLAP_WARN:
        fix = ("raise `penalty` above the spread of the levels" + retry
               + (", or use deflate='project'" if deflate == "penalty" else ""))
        warnings.warn(
            f"{where}: the state found has |<psi_i|psi>|^2 = {worst:.3g} > {OVERLAP_WARN} with a "
            f"previous state, so it is most likely that state again. The penalty ({penalty:g}) "
            f"is probably below the gap to the level sought: {fix}.",
            DeflationWarning, stacklevel=3)


def dmrg_excited(mpo, chi, options=None, **legacy_kwargs):
    """The lowest eigenstate of a dense MPO orthogonal to `prev_states`, by penalised two-site
    DMRG on the HIP device (tnkit/dmrg_excited.py:545-593).

    Returns (E, M, info): E = <M|H|M> without the penalty; M an MPS with its centre at site 0;
    info with tnkit's keys (sweeps, energies, E_penalised, penalty_overlaps, discarded_weight,
    eig_time, converged, cut, sweep_s, wall_s, expand, max_bond, project_fallbacks, ...) and
    'matvec_calls' (H_eff applies per sweep in 'matvecs_per_sweep')."""
    o = _resolve(options, legacy_kwargs)
    out = _dmrg_excited(mpo, int(chi), o)
    _warn_overlaps(out[2]["penalty_overlaps"], o.penalty, o.deflate, "dmrg_excited")
    return out


class _ExcitedEngine(DMRGEngine):
    """The ground-state engine's sweep with the local operator replaced by the penalised (or
    projected) one, and the overlap environments of every previous state carried along."""

    def __init__(self, mpo, chi, o: DMRGExcitedOptions, be: HipBackend):
        base = DMRGOptions(n_sweeps=o.n_sweeps, tol=o.tol, trunc=o.trunc, eig_tol=o.eig_tol,
                           eig=o.eig, reorth=o.reorth, seed=o.seed, init_mps=o.init_mps,
                           backend=be, max_iter=o.eig_max_iter, restart_m=o.restart_m,
                           restart_k=o.restart_k, reorth_alg=o.reorth_alg)
        super().__init__(mpo, chi, base)
        self.xo = o
        self.prev = [[_dev(be, x) for x in P] for P in o.prev_states]
        N = self.N
        for k, P in enumerate(self.prev):
            if len(P) != N:
                raise ValueError(f"prev_states[{k}] has {len(P)} sites, the MPO {N}")
        self.ELs, self.ERs = [], []
        for P in self.prev:
            el: List[Optional[torch.Tensor]] = [None] * (N + 1)
            er: List[Optional[torch.Tensor]] = [None] * (N + 1)
            el[0] = be.ones((1, 1))
            er[N] = be.ones((1, 1))
            for i in range(N - 1, -1, -1):
                er[i] = _grow_ER(er[i + 1], self.M[i], P[i])
            self.ELs.append(el)
            self.ERs.append(er)
        self.noise_rng = np.random.default_rng(o.seed + 2) if o.start_noise else None
        self.noise = False
        self.fallbacks = 0
        self.t_eig = 0.0
        self._V: Optional[torch.Tensor] = None

    def solve_block(self, i: int, direction: str, cur_tol: float, alpha: float = 0.0) -> float:
        if self.prev:
            self._V = torch.stack([_block_vector(self.ELs[j][i], P[i], P[i + 1],
                                                 self.ERs[j][i + 2])
                                   for j, P in enumerate(self.prev)]).contiguous()
        else:
            self._V = None
        e = super().solve_block(i, direction, cur_tol, alpha)
        for j, P in enumerate(self.prev):
            if direction == "right":
                self.ELs[j][i + 1] = _grow_EL(self.ELs[j][i], self.M[i], P[i])
            else:
                self.ERs[j][i + 1] = _grow_ER(self.ERs[j][i + 2], self.M[i + 1], P[i + 1])
        return e

    def eigsolve(self, theta0, op, op_lo, cur_tol):
        o, xo, be = self.o, self.xo, self.be
        shape = tuple(int(s) for s in theta0.shape)
        dim = int(np.prod(shape))
        count = {"n": 0}
        ws = kernels.two_site_ws(theta0, *op)
        kw = kernels.KrylovWorkspace(dim, theta0.device)
        V = self._V

        def h(x):
            count["n"] += 1
            return kernels.apply_two_site_fused(x.reshape(shape), *op, ws=ws).reshape(-1)

        def penalised(x):
            y = h(x)
            if V is not None:
                kernels.rank_update(V, x, xo.penalty, y, kw)
            return y

        mi = min(dim, o.max_iter if o.max_iter is not None else 100)

        def eigsolve(matvec, v):
            if o.eig == "lanczos-restart":
                e, x, _ = lanczos_ground_restart(matvec, v, be, tol=cur_tol, max_iter=mi,
                                                 m=o.restart_m, k=o.restart_k,
                                                 reorth_alg=o.reorth_alg)
            else:
                e, x, _ = lanczos_ground(matvec, v, be, tol=cur_tol, max_iter=mi,
                                         reorth=o.reorth)
            return float(e), x

        def nrm(x):
            return float(kernels.norm(x, kw).item())

        v0 = theta0.reshape(-1).contiguous()
        if self.noise:
            r = be.randn(self.noise_rng, dim)
            v0 = v0 + (xo.start_noise * nrm(v0) / nrm(r)) * r
        be.sync()
        t0 = time.perf_counter()
        x = None
        if xo.deflate == "project":
            r, B = 0, None
            if V is not None:
                _, s, Vh = torch.linalg.svd(V, full_matrices=False)
                sv = s.cpu().numpy()
                r = int(np.sum(sv > PROJ_RTOL * sv[0])) if sv.size and sv[0] > 0 else 0
                B = Vh[:r].contiguous()

            def proj(z):
                pz = z.clone()
                return kernels.rank_update(B, pz, -1.0, pz, kw) if r else pz

            if r < dim:
                v = proj(v0)
                if nrm(v) <= 1e-6 * max(nrm(v0), 1e-300):
                    v = proj(be.array(self.rng.standard_normal(dim)))
                hv = proj(h(v))
                sigma = xo.penalty + float(kernels.dot(v, hv, kw).item()) / nrm(v) ** 2

                def lifted(z):
                    pz = proj(z)
                    y = proj(h(pz))
                    return y.add_(z - pz, alpha=sigma)

                e, x = eigsolve(lifted, v)
                x = proj(x)
                nx = nrm(x)
                x = x / nx if nx > PROJ_KEEP else None
            self.fallbacks += x is None
        if x is None:
            e, x = eigsolve(penalised, v0)
        be.sync()
        dt = time.perf_counter() - t0
        self.phase["eig"] += dt
        self.t_eig += dt
        self.matvec_calls += count["n"]
        self.last_eig_info = {"matvecs": count["n"]}
        return e, x.reshape(shape)


def _check(mpo, o: DMRGExcitedOptions) -> None:
    if o.eig not in ("lanczos", "lanczos-restart"):
        raise ValueError(f"eig must be 'lanczos' or 'lanczos-restart', got {o.eig!r}")
    if o.deflate not in ("penalty", "project"):
        raise ValueError(f"deflate must be 'penalty' or 'project', got {o.deflate!r}")
    if o.eig == "lanczos-restart":
        check_reorth_periodic(o.reorth, "residual")
    if not o.start_noise >= 0:
        raise ValueError(f"start_noise must be >= 0, got {o.start_noise!r}")
    if o.time_budget_s is not None and math.isnan(o.time_budget_s):
        raise ValueError("time_budget_s must be a number of seconds or None, got nan")
    if o.charge is not None or getattr(mpo[0], "is_block", False):
        raise ValueError("dmrg_excited on the HIP backend takes dense MPOs (no charge=)")
    if len(o.prev_states) > 64:
        raise ValueError("at most 64 previous states")


def _dmrg_excited(mpo, chi: int, o: DMRGExcitedOptions):
    t_call = time.perf_counter()
    _check(mpo, o)
    be = get_backend(o.backend, mpo=mpo, init_mps=o.init_mps)
    if not isinstance(be, HipBackend):
        raise ValueError(f"backend {o.backend!r} is not a HIP backend")
    eng = _ExcitedEngine(mpo, chi, o, be)
    deadline = None if o.time_budget_s is None else t_call + o.time_budget_s
    energies: List[float] = []
    sweep_s: List[float] = []
    per_sweep: List[int] = []
    prev_E = None
    E = None
    converged = cut = False
    for sweep in range(o.n_sweeps):
        if deadline is not None and sweep_s and time.perf_counter() + max(sweep_s) > deadline:
            cut = True
            break
        t_sweep = time.perf_counter()
        eng.noise = bool(o.start_noise) and sweep == 0
        c0 = eng.matvec_calls
        E = eng.sweep(o.eig_tol)
        energies.append(E)
        per_sweep.append(eng.matvec_calls - c0)
        be.sync()
        sweep_s.append(time.perf_counter() - t_sweep)
        if o.verbose:
            print(f"    sweep {sweep + 1:2d}: E_pen = {E:.12f}", flush=True)
        if prev_E is not None and abs(E - prev_E) < o.tol:
            converged = True
            break
        prev_E = E
    M = eng.M
    ov = [abs(mps_inner(M, P, be)) for P in eng.prev]
    E_phys = float(np.real(mpo_expectation(M, eng.mpo, be)))
    bonds = [int(m.shape[2]) for m in M[:-1]] or [1]
    info = {"sweeps": len(energies), "energies": energies, "E_penalised": E,
            "penalty_overlaps": ov, "discarded_weight": eng.total_disc, "eig_time": eng.t_eig,
            "chi": chi, "eig": o.eig, "backend": be.name, "deflate": o.deflate,
            "project_fallbacks": eng.fallbacks,
            "restart_m": o.restart_m if o.eig == "lanczos-restart" else None,
            "restart_k": o.restart_k if o.eig == "lanczos-restart" else None,
            "converged": converged, "cut": cut, "sweep_s": sweep_s, "expand": 0,
            "max_bond": max(bonds), "max_bond_dense": max(bonds),
            "matvec_calls": eng.matvec_calls, "matvecs_per_sweep": per_sweep,
            "phase_time": eng.phase}
    info["wall_s"] = time.perf_counter() - t_call
    return E_phys, MPS(M, center=0), info


# --------------------------------------------------------------------------------------------
# the tower (tnkit/dmrg_excited.py:1055-1128, without pair_solve)
# --------------------------------------------------------------------------------------------


def _tower_penalty(energies, penalty):
    """3x the spread of the levels found so far plus `penalty` (tnkit/dmrg_excited.py:1055-1059)."""
    if not energies:
        return penalty
    return max(penalty, 3.0 * (max(energies) - min(energies)) + penalty)


def _tower_level(mpo, chi, n, states, w0, seed, ov_tol, max_retries, warm, deflate, verbose,
                 kw):
    """Level n of the tower (tnkit/dmrg_excited.py:1097-1128). Try k = 0, 1, ... runs
    dmrg_excited with penalty w0 4^k and seed seed + 17 n + 101 k; only try 0 may start from the
    level below (`warm`). A try whose largest overlap with the levels below is under ov_tol is
    kept; otherwise the next try follows, and the last try is kept whatever its overlap. Its
    info gains `penalty`, `attempt` and the matvec counts of every try."""
    calls: List[int] = []
    for attempt in range(max_retries):
        w = w0 * 4.0 ** attempt
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", DeflationWarning)
            E, M, info = dmrg_excited(mpo, chi, prev_states=states, penalty=w,
                                      seed=seed + 17 * n + 101 * attempt,
                                      init_mps=states[-1] if warm and attempt == 0 else None,
                                      verbose=False, deflate=deflate, **kw)
        calls += info["matvecs_per_sweep"]
        info.update(penalty=w, attempt=attempt)
        ov = max(info["penalty_overlaps"], default=0.0)
        if ov < ov_tol:
            break
        if verbose and attempt + 1 < max_retries:
            print(f"  level {n}: overlap {ov:.2e} > {ov_tol:g}; retry with penalty {4.0 * w:g}",
                  flush=True)
    info["matvecs_all_attempts"] = calls
    if verbose:
        print(f"  level {n}: E = {E:.12f}   sweeps={info['sweeps']:2d} "
              f"max|<prev|psi>|={ov:.2e}  w={info['penalty']:g}", flush=True)
    _warn_overlaps(info["penalty_overlaps"], info["penalty"], deflate,
                   f"dmrg_tower level {n}", " or `max_retries`")
    return E, M, info


def dmrg_tower(mpo, chi, nlevels, penalty=1.0, seed=0, verbose=False, ov_tol=1e-5,
               max_retries=2, warm_chain=True, deflate="penalty", pair_solve=False, **kw):
    """The lowest `nlevels` eigenstates by repeated penalised DMRG with tnkit's adaptive
    penalty (3x the spread found so far) and its overlap retry (penalty x4, another seed).
    Returns (energies, states, infos); states are host numpy arrays. `pair_solve` (the
    near-degenerate block repair of tnkit) is not provided here and raises."""
    if pair_solve:
        raise ValueError("dmrg_tower(pair_solve=True) is not served by the HIP backend")
    if "options" in kw:
        raise TypeError("dmrg_tower takes dmrg_excited settings as flat keywords, not options=")
    if kw.get("time_budget_s") is not None:
        raise ValueError("dmrg_tower takes no time_budget_s: it would apply to each solve, not "
                         "to the tower")
    states: List[List[np.ndarray]] = []
    energies: List[float] = []
    infos = []
    for n in range(nlevels):
        E, M, info = _tower_level(mpo, chi, n, states, _tower_penalty(energies, penalty), seed,
                                  ov_tol, max_retries,
                                  warm_chain and deflate == "penalty" and bool(states),
                                  deflate, verbose, kw)
        states.append([t.detach().cpu().numpy() for t in M])
        energies.append(E)
        infos.append(info)
    return np.array(energies), states, infos
