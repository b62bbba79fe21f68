"""numpy restatement of tnkit's excited-state DMRG by orthogonality penalty -- ORACLE (test
infrastructure only; the product never imports it).

`dmrg_excited` follows tnkit/dmrg_excited.py:603-836 on dense MPOs: two-site DMRG on
H + w sum_i |psi_i><psi_i| with the penalty applied locally through the overlap environments of
each previous state (tnkit/dmrg_excited.py:338-377), or, for deflate='project', on
Q H Q + sigma (1 - Q) with Q the projector off the block vectors (tnkit/dmrg_excited.py:384-418).
`dmrg_tower` is the sequential tower of tnkit/dmrg_excited.py:1062-1128 (no pair_solve).

Pinned by tnkit's own golden runs (tests/golden/excited_golden.json: per-sweep penalised
energies to 1e-12, sweeps exactly, overlaps and discarded weight, matvec counts) and by exact
diagonalisation of the chain's low spectrum.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .dense import (lanczos_ground, lanczos_ground_restart, random_mps, right_canonicalize,
                    truncate_gram, truncate_svd, update_left_env, update_right_env)

PROJ_RTOL = 1e-5     # tnkit/dmrg_excited.py:380
PROJ_KEEP = 0.5      # tnkit/dmrg_excited.py:381


def mps_inner(A, B, dtype=np.complex128):
    """<A|B> of two dense MPSs (tnkit/dmrg_excited.py:140-175, dense branch)."""
    E = np.ones((1, 1), dtype=dtype)
    for a, b in zip(A, B):
        E = np.einsum("ax,asb->xsb", E, np.conj(a), optimize=True)
        E = np.einsum("xsb,xsy->by", E, b, optimize=True)
    return complex(E.reshape(-1)[0])


def mpo_expectation(M, mpo, dtype=np.complex128):
    """<psi|W|psi> / <psi|psi> by the left-environment chain (tnkit/dmrg_excited.py:185-211)."""
    F = np.ones((1, 1, 1), dtype=dtype)
    for A, W in zip(M, mpo):
        F = update_left_env(F, A, W)
    return complex(F.reshape(-1)[0]) / mps_inner(M, M, dtype)


def overlap_envs(M, P, dtype):
    """EL[0] and every ER[i] between the current MPS and a fixed P (tnkit/dmrg_excited.py:338-356):
    ER[i][a, x] contracts sites i..N-1, conjugated on the M side."""
    N = len(M)
    EL: List[Optional[np.ndarray]] = [None] * (N + 1)
    ER: List[Optional[np.ndarray]] = [None] * (N + 1)
    EL[0] = np.ones((1, 1), dtype=dtype)
    ER[N] = np.ones((1, 1), dtype=dtype)
    for i in range(N - 1, -1, -1):
        ER[i] = grow_ER(ER[i + 1], M[i], P[i])
    return EL, ER


def grow_EL(EL_i, Mi, Pi):
    """tnkit/dmrg_excited.py:359-361"""
    t = np.einsum("ax,asb->xsb", EL_i, np.conj(Mi), optimize=True)
    return np.einsum("xsb,xsy->by", t, Pi, optimize=True)


def grow_ER(ER_i1, Mi, Pi):
    """tnkit/dmrg_excited.py:364-366"""
    t = np.einsum("asb,by->asy", np.conj(Mi), ER_i1, optimize=True)
    return np.einsum("asy,xsy->ax", t, Pi, optimize=True)


def block_vector(EL_i, Pi, Pi1, ER_i2):
    """V[a,s1,s2,b] = EL[a,x] P_i[x,s1,z] P_{i+1}[z,s2,y] ER[b,y] (tnkit/dmrg_excited.py:369-374)."""
    t = np.einsum("ax,xsz->asz", EL_i, Pi, optimize=True)
    t = np.einsum("asz,zty->asty", t, Pi1, optimize=True)
    return np.einsum("asty,by->astb", t, ER_i2, optimize=True)


def projector(Vs):
    """(Q, rank): Q x = x - B^T (conj(B) x), B the right singular vectors of the stacked block
    vectors above PROJ_RTOL of the largest singular value (tnkit/dmrg_excited.py:384-397)."""
    if not Vs:
        return (lambda x: x), 0
    _, s, Vh = np.linalg.svd(np.stack(Vs), full_matrices=False)
    r = int(np.sum(s > PROJ_RTOL * s[0])) if s.size and s[0] > 0 else 0
    B = Vh[:r]
    Bc = np.conj(B)
    return (lambda x: x - (Bc @ x) @ B), r


def lifted(h, proj, starts, margin):
    """Q H Q + sigma (1 - Q), sigma = margin + the largest Rayleigh quotient of Q H Q over
    `starts` (tnkit/dmrg_excited.py:400-418)."""
    def hq(x):
        return proj(h(proj(x)))

    sigma = margin + max(float(np.real(np.vdot(v, hq(v)))) / float(np.linalg.norm(v)) ** 2
                         for v in starts)

    def matvec(x):
        px = proj(x)
        return proj(h(px)) + sigma * (x - px)
    return matvec


def _truncate(mat, chi, trunc, direction):
    if trunc == "svd":
        return truncate_svd(mat, chi)
    if trunc == "gram":
        return truncate_gram(mat, chi, "left" if direction == "right" else "right")
    raise ValueError(f"oracle trunc {trunc!r} not restated for dmrg_excited")


def dmrg_excited(mpo, chi, prev_states=(), penalty=1.0, n_sweeps=12, tol=1e-11, eig_tol=1e-9,
                 trunc="svd", seed=0, init_mps=None, dtype=np.float64, eig_max_iter=None,
                 reorth="full", eig="lanczos-restart", restart_m=40, restart_k=10,
                 deflate="penalty", start_noise=0.0) -> Tuple[float, List[np.ndarray], Dict]:
    """Lowest eigenstate of H + penalty sum_i |psi_i><psi_i| by two-site DMRG
    (tnkit/dmrg_excited.py:603-836, dense branch). `dtype` is the run's arithmetic (tnkit's
    'auto' backend takes float64 for a real MPO; backend='numpy' is complex128).

    Returns (E_phys, M, info) with tnkit's info keys; info['matvecs'] lists the H_eff applies
    per sweep (what tnkit's golden runs count)."""
    real = not np.issubdtype(np.dtype(dtype), np.complexfloating)
    N = len(mpo)
    d = mpo[0].shape[2]
    W = [np.asarray(np.real(w) if real else w, dtype=dtype) for w in mpo]
    prev = [[np.asarray(x, dtype=dtype) for x in P] for P in prev_states]
    if init_mps is not None:
        M = right_canonicalize([np.asarray(m, dtype=dtype) for m in init_mps])
    else:
        M = right_canonicalize(random_mps(N, chi, d, np.random.default_rng(seed), dtype))
    if eig_max_iter is None:
        eig_max_iter = 100
    L: List[Optional[np.ndarray]] = [None] * (N + 1)
    R: List[Optional[np.ndarray]] = [None] * (N + 1)
    L[0] = np.ones((1, 1, 1), dtype=dtype)
    R[N] = np.ones((1, 1, 1), dtype=dtype)
    for i in range(N - 1, 0, -1):
        R[i] = update_right_env(R[i + 1], M[i], W[i])
    ELs, ERs = [], []
    for P in prev:
        el, er = overlap_envs(M, P, dtype)
        ELs.append(el)
        ERs.append(er)
    rng = np.random.default_rng(seed + 1)
    noise_rng = np.random.default_rng(seed + 2) if start_noise else None
    count = [0]
    state = {"disc": 0.0, "fallbacks": 0}

    def nrm(x):
        return float(np.linalg.norm(x))

    def solve_block(i, direction, noise):
        from .dense import apply_two_site
        W1, W2 = W[i], W[i + 1]
        Li, Ri = L[i], R[i + 2]
        theta0 = np.einsum("asc,ctb->astb", M[i], M[i + 1], optimize=True)
        a, s1, s2, b = theta0.shape
        shape = theta0.shape
        dim = int(np.prod(shape))
        v0 = theta0.reshape(-1)
        Vs = [block_vector(ELs[j][i], P[i], P[i + 1], ERs[j][i + 2]).reshape(-1)
              for j, P in enumerate(prev)]

        def h(x):
            count[0] += 1
            return apply_two_site(x.reshape(shape), Li, W1, W2, Ri).reshape(-1)

        def penalised(x):
            y = h(x)
            for v in Vs:
                c = complex(np.vdot(v, x))
                if real:
                    c = c.real
                y = y + (penalty * c) * v
            return y

        def eigsolve(matvec, v):
            if eig == "lanczos-restart":
                e, x, _ = lanczos_ground_restart(matvec, v, tol=eig_tol,
                                                 max_iter=min(dim, eig_max_iter), m=restart_m,
                                                 k=restart_k, dtype=dtype)
            else:
                e, x, _ = lanczos_ground(matvec, v, tol=eig_tol, max_iter=min(dim, eig_max_iter),
                                         reorth=reorth)
            return float(np.real(e)), x

        if noise:
            r = rng_noise_draw(noise_rng, dim, dtype)
            v0 = v0 + (start_noise * nrm(v0) / nrm(r)) * r
        x = None
        if deflate == "project":
            proj, rank = projector(Vs)
            if rank < dim:
                v = proj(v0)
                if nrm(v) <= 1e-6 * max(nrm(v0), 1e-300):
                    v = proj(np.asarray(rng.standard_normal(dim), dtype=dtype))
                e, x = eigsolve(lifted(h, proj, [v], penalty), v)
                x = proj(x)
                nx = nrm(x)
                x = x / nx if nx > PROJ_KEEP else None
            state["fallbacks"] += x is None
        if x is None:
            e, x = eigsolve(penalised, v0)
        U, S, Vh, disc = _truncate(x.reshape(a * s1, s2 * b), chi, trunc, direction)
        S = S / np.linalg.norm(S)
        state["disc"] += disc
        if direction == "right":
            M[i] = U.reshape(a, s1, -1)
            M[i + 1] = (S[:, None] * Vh).reshape(-1, s2, b)
            L[i + 1] = update_left_env(L[i], M[i], W1)
            for j, P in enumerate(prev):
                ELs[j][i + 1] = grow_EL(ELs[j][i], M[i], P[i])
        else:
            M[i] = (U * S).reshape(a, s1, -1)
            M[i + 1] = Vh.reshape(-1, s2, b)
            R[i + 1] = update_right_env(R[i + 2], M[i + 1], W2)
            for j, P in enumerate(prev):
                ERs[j][i + 1] = grow_ER(ERs[j][i + 2], M[i + 1], P[i + 1])
        return e

    energies, per_sweep = [], []
    prev_E = None
    converged = False
    E = None
    for sweep in range(n_sweeps):
        noise = bool(start_noise) and sweep == 0
        c0 = count[0]
        for i in range(0, N - 1):
            E = solve_block(i, "right", noise)
        for i in range(N - 2, -1, -1):
            E = solve_block(i, "left", noise)
        energies.append(E)
        per_sweep.append(count[0] - c0)
        if prev_E is not None and abs(E - prev_E) < tol:
            converged = True
            break
        prev_E = E
    ov = [abs(mps_inner(M, [np.asarray(x, dtype=dtype) for x in P], dtype)) for P in prev_states]
    E_phys = float(np.real(mpo_expectation(M, W, dtype)))
    info = {"sweeps": len(energies), "energies": energies, "E_penalised": E,
            "penalty_overlaps": ov, "discarded_weight": state["disc"],
            "project_fallbacks": state["fallbacks"], "converged": converged,
            "matvecs": per_sweep, "expand": 0}
    return E_phys, M, info


def rng_noise_draw(rng, n, dtype):
    """A start-noise vector drawn as tnkit's backend randn draws it (tnkit/backend.py:258-263)."""
    r = rng.standard_normal(n)
    if np.issubdtype(np.dtype(dtype), np.complexfloating):
        r = r + 1j * rng.standard_normal(n)
    return np.asarray(r, dtype=dtype)


def tower_penalty(energies, penalty):
    """3x the spread of the levels found plus `penalty` (tnkit/dmrg_excited.py:1055-1059)."""
    if not energies:
        return penalty
    return max(penalty, 3.0 * (max(energies) - min(energies)) + penalty)


def dmrg_tower(mpo, chi, nlevels, penalty=1.0, seed=0, ov_tol=1e-5, max_retries=2,
               warm_chain=True, deflate="penalty", **kw):
    """The lowest `nlevels` eigenstates by repeated penalised DMRG with the adaptive penalty and
    the overlap retry of tnkit/dmrg_excited.py:1062-1128 (sequential tower, no pair_solve).
    Returns (energies, states, infos); each info gains 'penalty' and 'attempt'."""
    states, energies, infos = [], [], []
    for n in range(nlevels):
        w = tower_penalty(energies, penalty)
        for attempt in range(max_retries):
            warm = warm_chain and deflate == "penalty" and states and attempt == 0
            init = states[-1] if warm else None
            E, M, info = dmrg_excited(mpo, chi, prev_states=states, penalty=w,
                                      seed=seed + 17 * n + 101 * attempt, init_mps=init,
                                      deflate=deflate, **kw)
            ov = max(info["penalty_overlaps"], default=0.0)
            info["penalty"] = w
            info["attempt"] = attempt
            if ov < ov_tol:
                break
            w *= 4.0
        states.append([np.asarray(m) for m in M])
        energies.append(E)
        infos.append(info)
    return np.array(energies), states, infos
