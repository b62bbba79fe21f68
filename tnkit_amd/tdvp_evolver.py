"""Stateful TDVP on the HIP device: tnkit's `TDVPEvolver` (tnkit/tdvp_evolver.py:65-364).

The evolver keeps the device state, the environments and the schedule between calls, so
`run(n)` continues where the last call stopped (the state ends every step right-canonical with
its centre at site 0 and the right environments current). Three schedules:

* '1site': tnkit's one-site step on every site (bond dimensions fixed; `tdvp1.step_right` /
  `step_left`, tnkit/tdvp.py:804-850);
* '2site': the two-site step on every bond (tnkit/tdvp.py:555-650): theta evolved by
  exp(-i H2 dt/2), cut by an SVD at the cap with tnkit's `svd_tol` / `degenerate_tol` rule
  (tnkit/_dmrg.py:77-113), the centre back-evolved by exp(+i H1 dt/2);
* '12site': per bond, decided at the start of every step, two-site while the bond is below the
  cap (and was not pinned below it by `degenerate_tol`), one-site once it is full; every
  `two_site_every`-th step is two-site everywhere (tnkit/tdvp_evolver.py:240-276, :313-325).

H2, H1 and H0 are the complex128 fused two-site kernels of libtnkit_hip (H1 and H0 with identity
MPO legs, as in `tdvp1`). The Krylov exponential is the adaptive `lanczos_expm_flat`
(fast_expm=None) or the fixed schedule `tdvp1.expm_fixed` (fast_expm=m, no host read).
Subspace expansion, graphs and boundary environments are not served here and raise.
"""
from __future__ import annotations

import dataclasses
from typing import Any, Callable, Dict, List, Optional, Tuple, Union

import numpy as np
import torch

from . import kernels
from .mps import MPS
from .tdvp import _right_canonical
from .tdvp1 import C128, _matvec, _solver, one_site_op, step_left, step_right

METHODS = ("1site", "2site", "12site")


def _group_start(s: np.ndarray, k: int, dtol: float) -> int:
    """Lowest index of the run of values just above position k that lie within dtol (relative
    to themselves) of s[k]: s[i] - s[k] <= dtol s[i] for every i in it."""
    ok = (s[:k] - s[k]) <= dtol * s[:k]
    bad = np.flatnonzero(~ok)
    return int(bad[-1]) + 1 if bad.size else 0


def _group_end(s: np.ndarray, c: int, k: int, dtol: float) -> int:
    """One past the run of values s[c:k] that lie within dtol (relative to s[c-1]) of s[c-1]."""
    ok = (s[c - 1] - s[c:k]) <= dtol * s[c - 1]
    bad = np.flatnonzero(~ok)
    return c + (int(bad[0]) if bad.size else k - c)


def cut(s: np.ndarray, chi: int, svd_tol: float = 0.0, degenerate_tol: float = 0.0,
        disc0: float = 0.0, tot: Optional[float] = None) -> Tuple[int, bool, float]:
    """(k, pinned, relative dropped weight) for a descending spectrum s (tnkit/_dmrg.py:77-113).

    * Cap: keep min(chi, len(s)). With degenerate_tol > 0 a cap cut that would split a group of
      near-equal values (within degenerate_tol, relative to the larger) moves down to the
      group's start, unless the group starts at s[0] or s[k] is zero; `pinned` says the cap cut
      moved.
    * svd_tol > 0: the fewest kept values, at least one, whose relative dropped weight
      disc0 + sum(s[j:]^2) / tot stays within svd_tol (tail sums accumulated from the small
      end); that cut moves up to the end of the group of the last kept value, and is taken
      only when it lies below the cap cut (then nothing is pinned).
    tot defaults to sum(s^2) (1 for an empty or zero spectrum)."""
    n = len(s)
    tot = tot or float(np.sum(s ** 2)) or 1.0
    cap = min(chi, n)
    k = cap
    if degenerate_tol > 0 and k < n and s[k] > 0:
        k = _group_start(s, k, degenerate_tol) or k
    pinned = k < cap
    disc = disc0 + float(np.sum(s[k:] ** 2)) / tot
    if svd_tol <= 0.0 or k <= 1:
        return k, pinned, disc
    loss = np.cumsum((s[:k] ** 2)[::-1])[::-1] / tot + disc   # loss[j]: keep j values
    within = np.flatnonzero(loss[1:] <= svd_tol)
    if not within.size:
        return k, pinned, disc
    j = int(within[0]) + 1
    if degenerate_tol > 0:
        j = _group_end(s, j, k, degenerate_tol)
    if j < k:
        return j, False, float(loss[j])
    return k, pinned, disc


def two_site_step(M, L, R, W, i: int, cap: int, dt: float, direction: str, solve,
                  svd_tol: float = 0.0, degenerate_tol: float = 0.0,
                  renorm: bool = True) -> Tuple[bool, float]:
    """In place on M, L / R: the two-site step at bond (i, i+1) (tnkit/tdvp.py:555-650).
    Returns (pinned, relative discarded weight)."""
    N = len(M)
    a, s1, c = (int(v) for v in M[i].shape)
    s2, b = int(M[i + 1].shape[1]), int(M[i + 1].shape[2])
    th = (M[i].reshape(a * s1, c) @ M[i + 1].reshape(c, s2 * b)).reshape(a, s1, s2, b)
    mv = _matvec(kernels.fused_two_site(L[i], W[i], W[i + 1], R[i + 2]), (a, s1, s2, b))
    th = solve(mv, th.contiguous(), -0.5j * dt).reshape(a * s1, s2 * b)
    U, S, Vh = kernels.thin_svd(th)
    s_host = S.cpu().numpy().astype(np.float64)
    tot = float(np.sum(s_host ** 2)) or 1.0
    k, pinned, disc = cut(s_host, cap, svd_tol, degenerate_tol, 0.0, tot)
    # torch may hand back Vh as a lazily conjugated view; the kernels read memory as stored
    U, S, Vh = U[:, :k].resolve_conj(), S[:k], Vh[:k].resolve_conj()
    if renorm:
        S = S / torch.linalg.norm(S)
    Sc = S.to(C128)
    if direction == "right":
        M[i] = U.reshape(a, s1, k).contiguous()
        L[i + 1] = kernels.update_left_env(L[i], M[i], W[i])
        C = (Sc[:, None] * Vh).reshape(k, s2, b).contiguous()
        if i < N - 2:
            mv1 = _matvec(one_site_op(L[i + 1], W[i + 1], R[i + 2]), (k, s2, 1, b))
            C = solve(mv1, C, 0.5j * dt).reshape(k, s2, b)
        M[i + 1] = C.contiguous()
    else:
        M[i + 1] = Vh.reshape(k, s2, b).contiguous()
        R[i + 1] = kernels.update_right_env(R[i + 2], M[i + 1], W[i + 1])
        C = (U * Sc[None, :]).reshape(a, s1, k).contiguous()
        if i > 0:
            mv1 = _matvec(one_site_op(L[i], W[i], R[i + 1]), (a, s1, 1, k))
            C = solve(mv1, C, 0.5j * dt).reshape(a, s1, k)
        M[i] = C.contiguous()
    return pinned, disc


@dataclasses.dataclass
class EvolverOptions:
    """The fields of tnkit's TDVPOptions (tnkit/tdvp.py:77-222) the evolver takes here. Fields
    of tnkit that are not served keep their defaults and raise when set."""

    measure: Optional[Callable[[Any], Any]] = None
    meas_every: int = 1
    canonicalize: bool = True
    keep_device: bool = False
    fast_expm: Optional[int] = None
    krylov_m: int = 30
    step_callback: Optional[Callable[..., Any]] = None
    trunc: str = "svd"
    chi_max: Optional[int] = None
    svd_tol: float = 0.0
    renorm: bool = True
    seed: int = 0
    profile: Optional[Dict[str, Any]] = None
    degenerate_tol: float = 0.0
    expand: int = 0
    backend: Any = "hip"
    graph: str = "auto"
    cuda_graph: Any = None
    prefix_graph: Any = None
    L0: Any = None
    RN: Any = None
    adaptive_dt: Any = None


def _resolve(options, kwargs) -> EvolverOptions:
    if options is None:
        o = EvolverOptions()
    elif isinstance(options, EvolverOptions):
        o = options
    elif isinstance(options, dict):
        o = EvolverOptions(**options)
    else:
        raise TypeError(f"options must be EvolverOptions or dict, got {type(options).__name__}")
    names = {f.name for f in dataclasses.fields(EvolverOptions)}
    bad = set(kwargs) - names
    if bad:
        raise TypeError(f"TDVPEvolver() got unexpected keyword argument(s) {sorted(bad)}")
    return dataclasses.replace(o, **kwargs) if kwargs else o


class TDVPEvolver:
    """Stateful TDVP on a fixed MPO and step dt (tnkit/tdvp_evolver.py:65-364). `chi` is the
    two-site bond cap (`options.chi_max` overrides it); None is allowed for '1site' runs only."""

    def __init__(self, mps, mpo, dt: float, chi: Optional[int] = None,
                 options: Optional[Union[EvolverOptions, Dict[str, Any]]] = None,
                 step0: int = 0, **kwargs: Any) -> None:
        o = _resolve(options, kwargs)
        if o.adaptive_dt is not None:
            raise ValueError("TDVPEvolver does not support adaptive_dt; use tdvp*_evolve")
        if callable(mpo):
            raise ValueError("TDVPEvolver needs a fixed MPO; use tdvp*_evolve for a callable")
        if o.krylov_m < 1:
            raise ValueError(f"krylov_m must be a positive Krylov dimension, got {o.krylov_m!r}")
        if o.meas_every < 1:
            raise ValueError("meas_every must be >= 1")
        if o.degenerate_tol < 0:
            raise ValueError("degenerate_tol must be >= 0")
        if o.fast_expm is not None and (isinstance(o.fast_expm, bool)
                                        or not isinstance(o.fast_expm, (int, np.integer))
                                        or o.fast_expm < 1):
            raise ValueError("fast_expm must be None or a Krylov dimension >= 1 here")
        if o.trunc != "svd":
            raise ValueError(f"the HIP evolver truncates with trunc='svd', got {o.trunc!r}")
        if not isinstance(o.expand, int) or isinstance(o.expand, bool) or o.expand < 0:
            raise ValueError(f"expand must be an int >= 0, got {o.expand!r}")
        if o.expand:
            raise ValueError("subspace expansion of the evolver is not served by the HIP backend")
        if o.cuda_graph is not None or o.prefix_graph is not None or o.graph == "on":
            raise ValueError("the HIP evolver runs its site steps eagerly (use tdvp1_evolve "
                             "with cuda_graph for captured one-site steps)")
        if o.L0 is not None or o.RN is not None:
            raise ValueError("the HIP evolver takes open boundaries (L0 = RN = None)")
        if not torch.cuda.is_available():
            raise RuntimeError("TDVPEvolver: the HIP backend needs the GPU")
        self.o, self.dt = o, float(dt)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self.mpo = [torch.as_tensor(np.asarray(w) if not isinstance(w, torch.Tensor) else w)
                    .to(self.dev, C128).contiguous() for w in mpo]
        cap = chi if o.chi_max is None else o.chi_max
        self.cap = None if cap is None else int(cap)
        self.solve = _solver(o.fast_expm, o.krylov_m)
        self.step, self.t = int(step0), int(step0) * self.dt
        self.reset(mps)

    # ---- state ----------------------------------------------------------------------------
    @property
    def mps(self) -> MPS:
        """The live state on the device; right-canonical with centre 0 after `run`."""
        return self._mps

    @property
    def center(self) -> Optional[int]:
        return self._mps.center

    def reset(self, mps) -> None:
        """Take a new state; it is rebuilt at the next `run`. Step count and time carry on."""
        self._mps = MPS([torch.as_tensor(np.asarray(t) if not isinstance(t, torch.Tensor)
                                         else t).to(self.dev, C128).contiguous() for t in mps])
        self._pinned = [False] * max(len(self._mps) - 1, 0)

    def prepare(self) -> None:
        """Canonicalise (if `canonicalize`) and build the environments now."""
        if self._mps.center != 0:
            self._rebuild()
            self._mps.center = 0

    def invalidate(self) -> None:
        """Force a rebuild at the next `run` (after editing a tensor's data in place)."""
        self._mps.center = None

    def _rebuild(self) -> None:
        M, W = self._mps, self.mpo
        self._pinned = [False] * max(len(M) - 1, 0)
        T = [t.contiguous() for t in M]
        if self.o.canonicalize:
            T = _right_canonical(T)
        M[:] = T
        N = len(M)
        self.L: List[Optional[torch.Tensor]] = [None] * (N + 1)
        self.R: List[Optional[torch.Tensor]] = [None] * (N + 1)
        self.L[0] = torch.ones((1, 1, 1), dtype=C128, device=self.dev)
        self.R[N] = torch.ones((1, 1, 1), dtype=C128, device=self.dev)
        for i in range(N - 1, 0, -1):
            self.R[i] = kernels.update_right_env(self.R[i + 1], M[i], W[i])

    # ---- steps ----------------------------------------------------------------------------
    def _one(self, i: int, direction: str) -> None:
        M, L, R, W = self._mps, self.L, self.R, self.mpo
        N = len(M)
        if direction == "right":
            last = i == N - 1
            Mi, Lo, Mn = step_right(M[i], None if last else M[i + 1], L[i], R[i + 1], W[i],
                                    self.dt, self.solve, last)
            M[i] = Mi
            if not last:
                L[i + 1], M[i + 1] = Lo, Mn
        else:
            first = i == 0
            Mi, Ro, Mp = step_left(M[i], None if first else M[i - 1], L[i], R[i + 1], W[i],
                                   self.dt, self.solve, first)
            M[i] = Mi
            if not first:
                R[i], M[i - 1] = Ro, Mp

    def _two(self, i: int, direction: str) -> float:
        o = self.o
        pinned, disc = two_site_step(self._mps, self.L, self.R, self.mpo, i, self.cap, self.dt,
                                     direction, self.solve, o.svd_tol, o.degenerate_tol,
                                     o.renorm)
        self._pinned[i] = pinned
        return disc

    def _sweep(self, two: List[bool]) -> float:
        """One step on a per-bond schedule; two[i] says whether bond (i, i+1) is two-site
        (tnkit/tdvp_evolver.py:240-276)."""
        N, disc = len(self._mps), 0.0
        for i in range(N):
            if i < N - 1 and two[i]:
                disc += self._two(i, "right")
            elif not (i == N - 1 and N > 1 and two[N - 2]):
                self._one(i, "right")
        for i in range(N - 1, -1, -1):
            if i > 0 and two[i - 1]:
                disc += self._two(i - 1, "left")
            elif not (i == 0 and N > 1 and two[0]):
                self._one(i, "left")
        return disc

    def run(self, n_steps: int, method: str = "2site",
            step_callback: Optional[Callable[..., Any]] = None,
            two_site_every: int = 0) -> List[Tuple[float, Any]]:
        """Take n_steps steps of `method`; returns the records of `options.measure` taken every
        `meas_every` steps of this call (tnkit/tdvp_evolver.py:278-364)."""
        if method not in METHODS:
            raise ValueError(f"method must be one of {METHODS}, got {method!r}")
        if two_site_every and method != "12site":
            raise ValueError("two_site_every applies to method='12site' only")
        if two_site_every < 0:
            raise ValueError(f"two_site_every must be >= 0, got {two_site_every!r}")
        if method != "1site" and self.cap is None:
            raise ValueError("two-site steps need a bond cap: pass chi or options.chi_max")
        o = self.o
        if self._mps.center != 0:
            self._rebuild()
        cb = step_callback if step_callback is not None else o.step_callback
        prof = o.profile
        if prof is not None:
            prof.setdefault("steps", [])
        M = self._mps
        records = []
        for k in range(n_steps):
            N = len(M)
            n_two = None
            if method == "12site":
                every = two_site_every and (self.step + 1) % two_site_every == 0
                two = [bool(every) or (int(M[i].shape[2]) != self.cap and not self._pinned[i])
                       for i in range(N - 1)]
                n_two = sum(two)
            else:
                two = [method == "2site"] * (N - 1)
            disc = self._sweep(two)
            self.step += 1
            self.t += self.dt
            d = {"step": self.step, "t": self.t, "dt": self.dt, "disc": disc,
                 "max_chi": max((int(m.shape[2]) for m in M[:-1]), default=1)}
            if n_two is not None:
                d["n_two"] = n_two
            if prof is not None:
                prof["steps"].append(d)
            rec = None
            if o.measure is not None and (k + 1) % o.meas_every == 0:
                snap = list(M) if o.keep_device else [t.cpu().numpy() for t in M]
                rec = (self.t, o.measure(snap))
                records.append(rec)
            if cb is not None:
                cb(d, rec)
        if n_steps > 0:
            M.center = 0
        return records
