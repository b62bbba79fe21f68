"""Thick-restart Lanczos for the lowest eigenpair, with the Krylov basis on the HIP device.

The algorithm is tnkit's `lanczos_ground_restart` (tnkit/eigensolver.py:507-674) and
`lanczos_ground_mixed` (tnkit/eigensolver.py:677-725); the same arguments, return values and
stopping rule. What changes is where the vector work runs: the basis is one preallocated
(m+1, n) float64 device buffer, and each iteration's subtraction, alpha, two Gram-Schmidt passes
and beta are one libtnkit_hip call (`kernels.krylov_orth`), after which the host reads the two
scalars (alpha, beta) in a single transfer. The projected (m+1)^2 matrix stays on the host, as
in tnkit.
"""
from __future__ import annotations

import os
import time
import warnings

import numpy as np
import torch

from . import kernels

try:
    from scipy.linalg import eigh_tridiagonal as _eigh_tridiagonal
except ImportError:  # pragma: no cover
    _eigh_tridiagonal = None


def check_reorth_periodic(reorth, stop) -> None:
    """tnkit's guard on reorth='periodic' (tnkit/eigensolver.py:27-58): it raises with
    stop='residual' and warns with stop='energy'; other reorth values pass."""
    if reorth != "periodic":
        return
    if stop == "residual":
        raise ValueError("reorth='periodic' with stop='residual' is not supported: periodic "
                         "reorthogonalisation loses basis orthogonality on long local solves, "
                         "and the residual rule then iterates against a tridiagonal that no "
                         "longer describes the basis. Use reorth='full'.")
    warnings.warn("reorth='periodic' loses orthogonality on long local solves (large bond "
                  "dimension near a critical point) and can stall DMRG above the ground state. "
                  "Use reorth='full'.", RuntimeWarning, stacklevel=3)


def check_eig_options(o) -> None:
    """tnkit's rules for eig_stop and reorth (tnkit/_dmrg.py:603-611): 'lanczos' takes both;
    the other solvers ignore them, but reorth='periodic' is still refused with
    eig_stop='residual' (and warned about with 'energy')."""
    if o.eig_stop not in ("residual", "energy"):
        raise ValueError(f"eig_stop must be 'residual' or 'energy', got {o.eig_stop!r}")
    if o.eig != "lanczos":
        check_reorth_periodic(o.reorth, o.eig_stop)


def _tridiag_ground(alphas, betas, k, vector=True):
    """(lowest eigenvalue, its eigenvector or None) of the k x k symmetric tridiagonal with
    diagonal alphas[:k] and off-diagonal betas[:k-1] (tnkit/eigensolver.py:61-92, 352-371)."""
    d = np.asarray(alphas[:k], dtype=np.float64)
    if k == 1:
        return float(d[0]), (np.array([1.0]) if vector else None)
    e = np.asarray(betas[:k - 1], dtype=np.float64)
    if _eigh_tridiagonal is not None:
        try:
            if vector:
                vals, vecs = _eigh_tridiagonal(d, e, select="i", select_range=(0, 0))
                return float(vals[0]), vecs[:, 0]
            vals = _eigh_tridiagonal(d, e, eigvals_only=True, select="i", select_range=(0, 0))
            return float(vals[0]), None
        except Exception:  # pragma: no cover - LAPACK stebz/stein failure: dense fallback
            pass
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    vals, vecs = np.linalg.eigh(T)
    return float(vals[0]), (vecs[:, 0] if vector else None)


_CHUNK = 64  # basis rows per device buffer of lanczos_ground (the orthogonalisation window)


def lanczos_ground(matvec, v0, ops=None, tol=1e-8, max_iter=200, reorth="full",
                   reorth_period=5, timers=None, stop="residual"):
    """Lowest eigenpair by Lanczos with an unbounded basis (tnkit/eigensolver.py:374-504):
    the three-term recurrence, reorthogonalisation against the whole basis ('full'), every
    `reorth_period` steps ('periodic') or never ('none'), and tnkit's stopping rules: the Ritz
    residual |beta y_last| < tol ('residual') or |e - e_prev| < tol ('energy'); breakdown at
    beta < 1e-12 max(1, |alpha|). Returns (e, x, {'iters', 'converged'}).

    The basis lives on the device in buffers of 64 rows, allocated as it grows. A full
    reorthogonalisation is two classical Gram-Schmidt passes per buffer (`kernels.krylov_orth`),
    older buffers first; the call on the current buffer also subtracts beta_prev v_prev and
    alpha v, and returns (alpha, beta) in one host read. tnkit subtracts one vector at a time
    (modified Gram-Schmidt); the two agree to rounding."""
    if reorth is True:
        reorth = "full"
    elif reorth is False:
        reorth = "none"
    if reorth not in ("full", "periodic", "none"):
        raise ValueError(f"reorth must be 'full', 'periodic' or 'none', got {reorth!r}")
    if stop not in ("residual", "energy"):
        raise ValueError(f"stop must be 'residual' or 'energy', got {stop!r}")
    check_reorth_periodic(reorth, stop)
    if v0.dtype != torch.float64:
        raise TypeError("lanczos_ground (hip): float64 vectors only")
    _t = time.perf_counter if timers is not None else None
    dev = v0.device
    v0 = v0.reshape(-1).contiguous()
    n = v0.numel()
    kw = kernels.KrylovWorkspace(n, dev)
    n0 = float(kernels.norm(v0, kw).item())
    if not np.isfinite(n0) or n0 < 1e-300:
        return 0.0, v0, {"iters": 0, "converged": False}
    chunks = []

    def row(i):
        c = i // _CHUNK
        while len(chunks) <= c:
            chunks.append(torch.empty((_CHUNK, n), dtype=torch.float64, device=dev))
        return chunks[c][i % _CHUNK]

    kernels.scale_by(v0, None, 1.0 / n0, row(0))
    alphas, betas = [], []
    beta_prev = 0.0
    prev_e = None
    converged = False
    e = 0.0
    it = 0
    for it in range(max_iter):
        if _t:
            t0 = _t()
        w = matvec(row(it))
        if w.dtype != torch.float64 or not w.is_contiguous():
            w = w.to(torch.float64).contiguous()
        w = w.reshape(-1)
        c, jj = divmod(it, _CHUNK)
        do_reorth = reorth == "full" or (reorth == "periodic" and it % reorth_period == 0)
        if _t:
            timers["matvec"] += _t() - t0
            t0 = _t()
        if do_reorth:
            if jj == 0 and it > 0:   # v_prev closes the previous buffer
                kernels.lincomb(chunks[c - 1][_CHUNK - 1:], [-beta_prev], w, beta=1.0)
            for cc in range(c):
                kernels.krylov_orth(chunks[cc], _CHUNK, w, 0, [], 0, kw)
            sub, base = ([beta_prev], jj - 1) if jj >= 1 else ([], 0)
            kernels.krylov_orth(chunks[c], jj + 1, w, jj, sub, base, kw)
            alpha, beta = (float(x) for x in kw.scal.cpu().tolist())
        else:
            if it > 0:
                kernels.lincomb(row(it - 1).unsqueeze(0), [-beta_prev], w, beta=1.0)
            alpha = float(kernels.dot(row(it), w, kw).item())
            kernels.lincomb(row(it).unsqueeze(0), [-alpha], w, beta=1.0)
            beta = float(kernels.norm(w, kw).item())
        alphas.append(alpha)
        if _t:
            timers["reorth"] += _t() - t0
            t0 = _t()
        k = len(alphas)
        e, y = _tridiag_ground(alphas, betas, k, vector=(stop == "residual"))
        if _t:
            timers["tridiag"] += _t() - t0
        breakdown = (not np.isfinite(beta) or not np.isfinite(alpha)
                     or beta < 1e-12 * max(1.0, abs(alpha)))
        if stop == "residual":
            hit = abs(beta) * abs(float(y[-1])) < tol
        else:
            hit = prev_e is not None and abs(e - prev_e) < tol
        if breakdown or hit:
            converged = True
            it += 1
            break
        prev_e = e
        betas.append(beta)
        beta_prev = beta
        kernels.scale_by(w, None, 1.0 / beta, row(it + 1))
    else:
        it = max_iter
    _, ground = _tridiag_ground(alphas, betas, len(alphas))
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    for c in range(-(-len(ground) // _CHUNK)):
        coef = [float(g) for g in ground[c * _CHUNK:(c + 1) * _CHUNK]]
        kernels.lincomb(chunks[c][:len(coef)], coef, x, beta=1.0)
    nx = float(kernels.norm(x, kw).item())
    kernels.scale_by(x, None, 1.0 / nx, x)
    return e, x, {"iters": it, "converged": converged}


def lanczos_ground_restart(matvec, v0, ops=None, tol=1e-8, max_iter=200, m=40, k=10,
                           timers=None, reorth_alg="cgs2"):
    """Lowest eigenpair of the Hermitian operator `matvec` by thick-restart Lanczos with CGS2
    reorthogonalisation inside a window of m vectors, keeping k Ritz vectors on restart.
    `v0` is a float64 device tensor. Returns (theta, x, info) with info = {'iters',
    'converged', 'restarts'} (tnkit/eigensolver.py:507-674)."""
    if reorth_alg not in ("cgs2", "cgs2v2"):
        raise ValueError(f"reorth_alg {reorth_alg!r}: the HIP basis kernels implement CGS2")
    if v0.dtype != torch.float64:
        raise TypeError("lanczos_ground_restart (hip): float64 vectors only")
    if m + 1 > 64:
        raise ValueError("lanczos_ground_restart (hip): restart_m must be at most 63")
    if k >= m:
        k = max(1, m // 2)
    dev = v0.device
    n = v0.numel()
    kw = kernels.KrylovWorkspace(n, dev)
    v0 = v0.reshape(-1).contiguous()
    n0 = float(kernels.norm(v0, kw).item())
    if not np.isfinite(n0) or n0 < 1e-300:
        return 0.0, v0, {"iters": 0, "converged": False, "restarts": 0}
    V = torch.zeros((m + 1, n), dtype=torch.float64, device=dev)
    kernels.scale_by(v0, None, 1.0 / n0, V[0])
    live = 1
    T = np.zeros((m + 1, m + 1))
    kept = 0
    border = None
    beta_prev = 0.0
    theta = 0.0
    y = np.array([1.0])
    converged = False
    restarts = 0
    iters = 0
    j = 0
    while iters < max_iter:
        w = matvec(V[j])
        if w.dtype != torch.float64 or not w.is_contiguous():
            w = w.to(torch.float64).contiguous()
        w = w.reshape(-1)
        if kept and j == kept:                      # first step after a restart: arrowhead
            sub, base = list(border), 0
        elif j > 0:
            sub, base = [beta_prev], j - 1
        else:
            sub, base = [], 0
        kernels.krylov_orth(V, live, w, j, sub, base, kw)
        alpha, beta = (float(x) for x in kw.scal.cpu().tolist())
        T[j, j] = alpha
        cur = j + 1
        if cur == 1:
            th, y = alpha, np.array([1.0])
        else:
            ev, evec = np.linalg.eigh(T[:cur, :cur])
            th, y = float(ev[0]), evec[:, 0]
        iters += 1
        res = abs(beta) * abs(float(y[-1]))
        breakdown = (not np.isfinite(beta) or not np.isfinite(alpha)
                     or beta < 1e-12 * max(1.0, abs(alpha)))
        theta = th
        if breakdown or res < tol:
            converged = True
            break
        if cur >= m:                                # thick restart
            ev, evec = np.linalg.eigh(T[:cur, :cur])
            kk = min(k, cur - 1)
            Yk = torch.as_tensor(np.ascontiguousarray(evec[:, :kk])).to(dev)   # (cur, kk)
            U = kernels.gemm_tn(Yk, V[:cur])                                   # (kk, n)
            V[:kk].copy_(U)
            kernels.scale_by(w, None, 1.0 / beta, V[kk])
            live = kk + 1
            T[:] = 0.0
            border = beta * np.asarray(evec[cur - 1, :kk], dtype=np.float64)
            for r in range(kk):
                T[r, r] = float(ev[r])
                T[r, kk] = T[kk, r] = border[r]
            kept = kk
            j = kk
            beta_prev = 0.0
            y = np.zeros(kk + 1)
            y[0] = 1.0
            restarts += 1
            continue
        T[j, j + 1] = T[j + 1, j] = beta
        beta_prev = beta
        kernels.scale_by(w, None, 1.0 / beta, V[j + 1])
        live = j + 2
        j += 1
    ny = min(len(y), live)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    kernels.lincomb(V, [float(c) for c in y[:ny]], x)
    nx = float(kernels.norm(x, kw).item())
    kernels.scale_by(x, None, 1.0 / nx, x)
    return theta, x, {"iters": iters, "converged": converged, "restarts": restarts}


def lanczos_ground_mixed(matvec_lo, matvec_hi, v0, ops=None, tol=1e-8, max_iter=200, m=40, k=10,
                         tol_lo=1e-4, timers=None, counts=None):
    """Two-stage mixed-precision thick restart (tnkit/eigensolver.py:677-725): stage 1 with the
    low-precision matvec to max(tol_lo, tol), stage 2 in float64 warm-started from stage 1.
    Both matvecs take and return float64 vectors; every projection is float64."""
    tol_lo = max(tol_lo, tol)
    c1 = {"n": 0}
    c2 = {"n": 0}

    def mv1(x):
        c1["n"] += 1
        return matvec_lo(x)

    def mv2(x):
        c2["n"] += 1
        return matvec_hi(x)

    e_lo, x_lo, i_lo = lanczos_ground_restart(mv1, v0, ops, tol=tol_lo, max_iter=max_iter, m=m,
                                              k=k, timers=timers)
    e_hi, x_hi, i_hi = lanczos_ground_restart(mv2, x_lo, ops, tol=tol, max_iter=max_iter, m=m,
                                              k=k, timers=timers)
    if counts is not None:
        counts["lo"] = counts.get("lo", 0) + c1["n"]
        counts["hi"] = counts.get("hi", 0) + c2["n"]
    return e_hi, x_hi, {"iters": i_lo["iters"] + i_hi["iters"], "converged": i_hi["converged"],
                        "restarts": i_lo["restarts"] + i_hi["restarts"], "n_lo": c1["n"],
                        "n_hi": c2["n"], "e_lo": e_lo}


# ---------------------------------------------------------------------------------------------
# adaptive Krylov exponential
# ---------------------------------------------------------------------------------------------

KRYLOV_MAX = 32  # basis rows the device step keeps in registers (tnh_expm_step_c)


def _tridiag_matrix(alphas, betas):
    j = len(alphas)
    T = np.diag(np.asarray(alphas, dtype=np.float64))
    if j > 1:
        off = np.asarray(betas[:j - 1], dtype=np.float64)
        T += np.diag(off, 1) + np.diag(off, -1)
    return T


def _expm_tridiag(tau, alphas, betas) -> np.ndarray:
    from scipy.linalg import expm
    return expm(tau * _tridiag_matrix(alphas, betas))


class ExpmMonitor:
    """Stop rule and diagnostics of the adaptive Krylov exponential (tnkit/eigensolver.py:230-287).

    After the j-th Lanczos vector (alphas a_1..a_j, betas b_1..b_{j-1}, and the norm b_j of the
    residual that would become vector j + 1):
      * b_j < tol means the Krylov space is invariant: stop without adding the residual;
      * Saad's first-neglected-term estimate e_j = |b_j| |tau| |exp(tau T_j)[j, 1]| beta0 ends an
        adaptive solve once e_j < tol, from j = 2 on;
      * for `info`, the same estimate is frozen at the first j >= est_ref_j (0.0 when the space
        breaks down before), so it keeps measuring the step size.
    The estimate is computed only when one of the last two needs it. `adaptive=None` reads
    TNKIT_EXPM_ADAPTIVE (any value but '0' means adaptive). The final coefficients are
    beta0 exp(tau T) e_1."""

    def __init__(self, tau, beta0, tol, adaptive, info, est_ref_j):
        if adaptive is None:
            adaptive = os.environ.get("TNKIT_EXPM_ADAPTIVE", "1") != "0"
        self.tau, self.beta0, self.tol, self.adaptive = tau, beta0, tol, bool(adaptive)
        self.info, self.ref_dim = info, est_ref_j
        self.record = {"err_est": float("inf"), "beta1": float("nan"),
                       "est_ref": float("nan"), "est_ref_j": 0}

    def check(self, k, alphas, betas, beta):
        """After step k: 'breakdown' (stop, beta unused), 'converged' (stop after this vector)
        or None."""
        rec, j = self.record, len(alphas)
        if k == 0:
            rec["beta1"] = beta
        tracking = self.info is not None and rec["est_ref_j"] == 0
        if beta < self.tol:
            if tracking:
                rec["est_ref"], rec["est_ref_j"] = 0.0, j
            return "breakdown"
        freeze = tracking and j >= self.ref_dim
        may_stop = self.adaptive and j >= 2
        if not (freeze or may_stop):
            return None
        corner = _expm_tridiag(self.tau, alphas, betas)[j - 1, 0]
        est = abs(beta) * abs(self.tau) * abs(corner) * self.beta0
        rec["err_est"] = est
        if freeze:
            rec["est_ref"], rec["est_ref_j"] = est, j
        return "converged" if may_stop and est < self.tol else None

    def finish(self, alphas, betas, capped):
        if self.info is not None:
            self.info.update(self.record, krylov_dim=len(alphas), capped=capped)
        return self.beta0 * _expm_tridiag(self.tau, alphas, betas)[:, 0]


def lanczos_expm_flat(matvec, x: torch.Tensor, tau, m=30, tol=1e-12, adaptive=None, info=None,
                      est_ref_j=2) -> torch.Tensor:
    """exp(tau H) x for a flat complex128 device vector x (any float/complex dtype is promoted),
    `matvec` mapping flat complex128 device vectors to the same. Each step is one
    `tnh_expm_step_c` (three-term subtraction, two Gram-Schmidt passes, alpha, beta and the next
    basis vector on the device) and one host read of (alpha, beta) for the monitor."""
    if m > KRYLOV_MAX:
        raise ValueError(f"lanczos_expm: m = {m} above the device Krylov window {KRYLOV_MAX}")
    x = x.reshape(-1).to(torch.complex128)
    n = x.numel()
    beta0 = float(torch.linalg.vector_norm(x).item())
    if not np.isfinite(beta0) or beta0 < 1e-300:
        return x
    mon = ExpmMonitor(tau, beta0, tol, adaptive, info, est_ref_j)
    dev = x.device
    V = torch.zeros((1, m, n), dtype=torch.complex128, device=dev)
    ab = torch.zeros((1, 2 * m), dtype=torch.float64, device=dev)
    ws = kernels.expm_ws(1, dev)
    V[0, 0] = x / beta0
    alphas, betas = [], []
    capped = False
    for k in range(m):
        w = matvec(V[0, k]).reshape(1, n).to(torch.complex128).contiguous()
        kernels.expm_step_c(V, w, ab, k, ws)
        alpha, beta = ab[0, [k, m + k]].tolist()
        alphas.append(alpha)
        stop = mon.check(k, alphas, betas, beta)
        if stop == "breakdown":
            break
        betas.append(beta)
        if stop:
            break
    else:
        capped = bool(mon.adaptive)
    c = mon.finish(alphas, betas, capped)
    out = torch.empty((1, n), dtype=torch.complex128, device=dev)
    kernels.lincomb_c(V, torch.as_tensor(c[None, :]).to(dev), out, len(alphas))
    return out.reshape(-1)


def lanczos_expm(matvec, v, tau, ops=None, m=30, tol=1e-12, adaptive=None, info=None,
                 est_ref_j=2):
    """Krylov approximation of exp(tau H) v with full reorthogonalisation and tnkit's adaptive
    stop and diagnostics (tnkit/eigensolver.py:290-349; `info` keys krylov_dim, err_est, beta1,
    est_ref, est_ref_j, capped). `v` is a device tensor of any shape (matvec sees that shape) or
    a BlockTensor, which goes to `block_lanczos.lanczos_expm` as in tnkit. m <= 32."""
    from .block_ops import BlockTensor
    if isinstance(v, BlockTensor):
        from . import block_lanczos
        return block_lanczos.lanczos_expm(matvec, v, tau, ops, m=m, tol=tol, adaptive=adaptive,
                                          info=info, est_ref_j=est_ref_j)
    shape = tuple(v.shape)
    out = lanczos_expm_flat(lambda x: matvec(x.reshape(shape)).reshape(-1), v, tau, m=m, tol=tol,
                            adaptive=adaptive, info=info, est_ref_j=est_ref_j)
    return out.reshape(shape)
