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

import numpy as np
import torch

from . import kernels


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
