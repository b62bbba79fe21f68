"""numpy restatement of tnkit's dense two-site DMRG path -- the ORACLE (test infrastructure).

Only `tests/`, `__graft_entry__.smoke()` and `bench.py`'s `cpu_baseline` leg use this module.

Conventions follow tnkit (`tnkit/_dmrg.py:7-11`, `tnkit/mps.py:32-41`):
    MPS  M[i] : (a, s, b)                    left bond, physical, right bond
    MPO  W[i] : (wl, wr, s_out, s_in)
    env  L[i], R[i] : (bra, mpo, ket)
All contractions go through `np.einsum(..., optimize=True)`, which is what tnkit's numpy
backend reproduces bit for bit (`tnkit/backend.py:112-156`: a compiled plan is installed only if
it equals `np.einsum(..., optimize=True)` element for element).

`dtype` of every routine follows its inputs, so the same code gives the float64 reference and
the float32 low-precision matvec used by the mixed eigensolver.
"""
from __future__ import annotations

import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np

# --------------------------------------------------------------------------------------------
# contractions
# --------------------------------------------------------------------------------------------


def apply_two_site(theta, L, W1, W2, R):
    """H_eff theta by four pairwise contractions.  tnkit/_dmrg.py:162-169"""
    x = np.einsum("pwq,qijb->pwijb", L, theta, optimize=True)
    x = np.einsum("wmsi,pwijb->pmsjb", W1, x, optimize=True)
    x = np.einsum("mrtj,pmsjb->pstrb", W2, x, optimize=True)
    return np.einsum("drb,pstrb->pstd", R, x, optimize=True)


def fused_two_site(L, W1, W2, R):
    """(Lm, S, Rm): L as (p wl, q), the MPO pair as S[(s1 s2 wr), (wl s1' s2')], R as
    (b', wr b).  tnkit/_dmrg.py:172-180"""
    p, wl, q = L.shape
    wr, d1, d2 = W2.shape[1], W1.shape[2], W2.shape[2]
    S = np.einsum("wmsi,mrtj->strwij", W1, W2, optimize=True).reshape(d1 * d2 * wr, wl * d1 * d2)
    return L.reshape(p * wl, q), S, R.reshape(R.shape[0], wr * R.shape[2])


def apply_two_site_fused(theta, Lm, S, Rm):
    """H_eff theta from the fused matrices: GEMM, batched 'S' GEMM, GEMM.
    tnkit/_dmrg.py:183-189"""
    q, d1, d2, b = theta.shape
    x = (Lm @ theta.reshape(q, -1)).reshape(-1, S.shape[1], b)
    x = S @ x
    return (x.reshape(-1, Rm.shape[1]) @ Rm.T).reshape(-1, d1, d2, Rm.shape[0])


def apply_one_site(M, L, W, R):
    """One-site H_eff M.  tnkit/_dmrg.py:192-199"""
    x = np.einsum("aky,ysb->aksb", L, M, optimize=True)
    x = np.einsum("klst,aktb->alsb", W, x, optimize=True)
    return np.einsum("bly,alsy->asb", R, x, optimize=True)


def update_left_env(L, A, W):
    """L' = sum L A* W A past a left-canonical site.  tnkit/_dmrg.py:216-222"""
    t = np.einsum("pwq,psd->wqsd", L, A.conj(), optimize=True)
    t = np.einsum("wqsd,wrsy->qdry", t, W, optimize=True)
    return np.einsum("qdry,qye->dre", t, A, optimize=True)


def update_right_env(R, B, W):
    """R' = sum R B* W B past a right-canonical site.  tnkit/_dmrg.py:225-231"""
    t = np.einsum("xry,usx->ryus", R, B.conj(), optimize=True)
    t = np.einsum("ryus,wrsp->yuwp", t, W, optimize=True)
    return np.einsum("yuwp,vpy->uwv", t, B, optimize=True)


def two_site_flops(a, b, d1, d2, wl, wm, wr, complex_=False):
    """Algorithmic flop count of one fused two-site matvec (2 per real MAC, x4 complex):
    GEMM1 2*a*wl*a*(d1 d2 b) + S-mix 2*(d1 d2 wr)*(wl d1 d2)*a*b + GEMM3 2*(a d1 d2)*(wr b)*b.
    At a = b = chi, wl = wr = D, d1 = d2 = d this is 4 d^2 D chi^3 + 2 d^4 D^2 chi^2, the count of
    tnkit/bench/profile/gpu_matvec_prof.py:92."""
    f = (2 * a * wl * a * d1 * d2 * b + 2 * (d1 * d2 * wr) * (wl * d1 * d2) * a * b
         + 2 * a * d1 * d2 * wr * b * b)
    return 4 * f if complex_ else f


def env_flops(a, d, c, wl, wr, complex_=False):
    """Algorithmic flops of one left-environment update L(a, wl, a) -> L'(c, wr, c)."""
    f = 2 * wl * a * d * c * a + 2 * (wl * d) * (wr * d) * a * c + 2 * (c * wr) * (a * d) * c
    return 4 * f if complex_ else f


# --------------------------------------------------------------------------------------------
# states
# --------------------------------------------------------------------------------------------


def random_mps(N, chi, d=2, rng=None, dtype=np.float64):
    """Gaussian MPS with bonds min(chi, d^min(i, N-i)), drawn from one numpy Generator in site
    order (complex draws take the imaginary part right after the real part).
    tnkit/_dmrg.py:253-261, tnkit/backend.py:258-263"""
    rng = np.random.default_rng(0) if rng is None else rng
    dims = [1] + [min(chi, d ** min(i, N - i)) for i in range(1, N)] + [1]
    cplx = np.issubdtype(np.dtype(dtype), np.complexfloating)
    out = []
    for i in range(N):
        shape = (dims[i], d, dims[i + 1])
        r = rng.standard_normal(shape)
        if cplx:
            r = r + 1j * rng.standard_normal(shape)
        out.append(np.asarray(r, dtype=dtype))
    return out


def svd(A):
    """Economy SVD with tnkit's gesdd -> gesvd fallback.  tnkit/backend.py:441-447"""
    try:
        return np.linalg.svd(A, full_matrices=False)
    except np.linalg.LinAlgError:
        from scipy.linalg import svd as ssvd
        return ssvd(A, full_matrices=False, lapack_driver="gesvd")


def right_canonicalize_lq(M):
    """Right-canonical form by an LQ factorisation per site (QR of the transpose), the form
    tnkit uses for block tensors (tnkit/_dmrg.py:295-306); equal to `right_canonicalize` up to
    an orthogonal gauge on every bond."""
    M = list(M)
    for i in range(len(M) - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        Q, Rr = np.linalg.qr(M[i].reshape(Dl, d * Dr).T)
        M[i] = Q.T.reshape(-1, d, Dr)
        M[i - 1] = np.einsum("asb,bc->asc", M[i - 1], Rr.T, optimize=True)
        M[i - 1] = M[i - 1] / np.linalg.norm(M[i - 1])
    M[0] = M[0] / np.linalg.norm(M[0])
    return M


def right_canonicalize(M):
    """Right-canonical form by one SVD per site from the right, renormalising the running
    centre at every step; the centre ends at site 0.  tnkit/_dmrg.py:264-292"""
    M = list(M)
    for i in range(len(M) - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        U, S, Vh = svd(M[i].reshape(Dl, d * Dr))
        M[i] = Vh.reshape(-1, d, Dr)
        M[i - 1] = np.einsum("asb,bc->asc", M[i - 1], U * S, optimize=True)
        M[i - 1] = M[i - 1] / np.linalg.norm(M[i - 1])
    M[0] = M[0] / np.linalg.norm(M[0])
    return M


# --------------------------------------------------------------------------------------------
# Krylov solvers
# --------------------------------------------------------------------------------------------


def _tridiag_lowest(alphas, betas, k):
    """Lowest eigenpair of the k x k symmetric tridiagonal (diag alphas[:k], off-diag
    betas[:k-1]) by dense eigh (tnkit/eigensolver.py:352-371 uses LAPACK's tridiagonal
    routine; the pair is the same to rounding)."""
    d = np.asarray(alphas[:k], dtype=np.float64)
    if k == 1:
        return float(d[0]), np.array([1.0])
    e = np.asarray(betas[:k - 1], dtype=np.float64)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    vals, vecs = np.linalg.eigh(T)
    return float(vals[0]), vecs[:, 0]


def lanczos_ground(matvec, v0, tol=1e-8, max_iter=200, reorth="full", reorth_period=5,
                   stop="residual"):
    """Lowest eigenpair by Lanczos with an unbounded basis. tnkit/eigensolver.py:374-504:
    w = H v - beta_prev v_prev; alpha = <v, w>; w -= alpha v; modified Gram-Schmidt against
    every stored vector when reorth == 'full' (or every `reorth_period` steps for 'periodic');
    stop on |beta y_last| < tol ('residual') or |e - e_prev| < tol ('energy'); breakdown at
    beta < 1e-12 max(1, |alpha|); the vector is assembled once at the end."""
    n0 = float(np.linalg.norm(v0))
    if not np.isfinite(n0) or n0 < 1e-300:
        return 0.0, v0, {"iters": 0, "converged": False}
    v = v0 / n0
    V = [v]
    alphas, betas = [], []
    beta_prev = 0.0
    v_prev = np.zeros_like(v)
    prev_e = None
    converged = False
    e = 0.0
    it = 0
    for it in range(max_iter):
        w = matvec(v) - beta_prev * v_prev
        alpha = float(np.vdot(v, w).real)
        alphas.append(alpha)
        w = w - alpha * v
        if reorth == "full" or (reorth == "periodic" and it % reorth_period == 0):
            for u in V:
                w = w - np.vdot(u, w) * u
        e, y = _tridiag_lowest(alphas, betas, len(alphas))
        beta = float(np.linalg.norm(w))
        breakdown = (not np.isfinite(beta) or not np.isfinite(alpha)
                     or beta < 1e-12 * max(1.0, abs(alpha)))
        hit = (abs(beta) * abs(float(y[-1])) < tol) if stop == "residual" else \
            (prev_e is not None and abs(e - prev_e) < tol)
        if breakdown or hit:
            converged = True
            it += 1
            break
        prev_e = e
        betas.append(beta)
        v_prev, beta_prev = v, beta
        v = w / beta
        V.append(v)
    else:
        it = max_iter
    _, ground = _tridiag_lowest(alphas, betas, len(alphas))
    x = np.zeros_like(V[0])
    for c, vj in zip(ground[:len(V)], V):
        x = x + float(c) * vj
    return e, x / np.linalg.norm(x), {"iters": it, "converged": converged}


def lanczos_ground_restart(matvec, v0, tol=1e-8, max_iter=200, m=40, k=10,
                           dtype=np.float64):
    """Lowest eigenpair by thick-restart Lanczos (Wu & Simon 2000) with CGS2 reorthogonalisation
    against the live window and the Ritz residual |beta y_last| checked every step.
    tnkit/eigensolver.py:507-674 (reorth_alg='cgs2').

    Returns (theta, x, info) with info = {'iters', 'converged', 'restarts'}."""
    if k >= m:
        k = max(1, m // 2)
    n0 = float(np.linalg.norm(v0))
    if not np.isfinite(n0) or n0 < 1e-300:
        return 0.0, v0, {"iters": 0, "converged": False, "restarts": 0}
    n = v0.size
    V = np.zeros((m + 1, n), dtype=dtype)
    V[0] = v0.reshape(-1) / n0
    live = 1
    T = np.zeros((m + 1, m + 1))
    kept = 0
    border = None
    beta_prev = 0.0
    v_prev = np.zeros(n, dtype=dtype)
    theta = 0.0
    y = np.array([1.0])
    converged = False
    restarts = 0
    iters = 0
    j = 0
    while iters < max_iter:
        v = V[j]
        w = matvec(v)
        if kept and j == kept:                      # first step after a restart: arrowhead
            w = w - np.einsum("r,rn->n", border.astype(dtype), V[:kept], optimize=True)
        elif j > 0:
            w = w - beta_prev * v_prev
        alpha = complex(np.vdot(v, w)).real
        w = w - alpha * v
        Vl = V[:live]
        Vc = np.conj(Vl)
        for _ in range(2):                          # CGS2
            c = np.einsum("kn,n->k", Vc, w, optimize=True)
            w = w - np.einsum("k,kn->n", c, Vl, optimize=True)
        T[j, j] = alpha
        beta = float(np.linalg.norm(w))
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
            Yk = evec[:, :kk]
            vnew = w / beta
            U = np.einsum("rc,rn->cn", Yk.astype(dtype), V[:cur], optimize=True)
            V[:kk] = U
            V[kk] = vnew
            live = kk + 1
            T[:] = 0.0
            border = beta * np.asarray(Yk[cur - 1, :], dtype=np.float64)
            for r in range(kk):
                T[r, r] = float(ev[r])
                T[r, kk] = border[r]
                T[kk, r] = border[r]
            kept = kk
            j = kk
            beta_prev = 0.0
            v_prev = np.zeros(n, dtype=dtype)
            y = np.zeros(kk + 1)
            y[0] = 1.0
            restarts += 1
            continue
        T[j, j + 1] = T[j + 1, j] = beta
        v_prev = v
        beta_prev = beta
        V[j + 1] = w / beta
        live = j + 2
        j += 1
    ny = min(len(y), live)
    x = np.einsum("r,rn->n", y[:ny].astype(dtype), V[:ny], optimize=True)
    x = x / np.linalg.norm(x)
    return theta, x, {"iters": iters, "converged": converged, "restarts": restarts}


def lanczos_ground_mixed(matvec_lo, matvec_hi, v0, tol=1e-8, max_iter=200, m=40, k=10,
                         tol_lo=1e-4):
    """Stage 1: thick restart with the low-precision matvec to max(tol_lo, tol); stage 2: the
    float64 thick restart warm-started from the stage-1 vector.  tnkit/eigensolver.py:677-725.
    Returns (e, x, info) with info['n_lo'], info['n_hi'], info['e_lo']."""
    tol_lo = max(tol_lo, tol)
    c = {"lo": 0, "hi": 0}

    def lo(x):
        c["lo"] += 1
        return matvec_lo(x)

    def hi(x):
        c["hi"] += 1
        return matvec_hi(x)

    e_lo, x_lo, i_lo = lanczos_ground_restart(lo, v0, tol=tol_lo, max_iter=max_iter, m=m, k=k)
    e, x, i_hi = lanczos_ground_restart(hi, x_lo, tol=tol, max_iter=max_iter, m=m, k=k)
    return e, x, {"iters": i_lo["iters"] + i_hi["iters"], "converged": i_hi["converged"],
                  "restarts": i_lo["restarts"] + i_hi["restarts"],
                  "n_lo": c["lo"], "n_hi": c["hi"], "e_lo": e_lo}


def _tridiag_matrix(alphas, betas):
    j = len(alphas)
    T = np.diag(np.asarray(alphas, dtype=np.float64))
    if j > 1:
        off = np.asarray(betas[:j - 1], dtype=np.float64)
        T += np.diag(off, 1) + np.diag(off, -1)
    return T


def lanczos_expm(matvec, v, tau, m=30, tol=1e-12, adaptive=True, info=None, est_ref_j=2):
    """exp(tau H) v by an adaptive Krylov space with full reorthogonalisation
    (tnkit/eigensolver.py:230-349): stop at breakdown (beta < tol) or when Saad's first
    neglected term |beta| |tau| |exp(tau T)[j-1, 0]| beta0 drops below tol (j >= 2); `info`
    gets krylov_dim, err_est, beta1, est_ref (the same estimate frozen at j = est_ref_j),
    est_ref_j and capped."""
    import scipy.linalg
    beta0 = float(np.linalg.norm(v))
    if not np.isfinite(beta0) or beta0 < 1e-300:
        return v
    V = [v / beta0]
    alphas, betas = [], []
    beta = 0.0
    v_prev = np.zeros_like(v)
    est, est_ref, est_ref_at, beta1 = float("inf"), float("nan"), 0, float("nan")
    capped = False
    for k in range(m):
        w = matvec(V[-1]) - beta * v_prev
        alpha = float(np.vdot(V[-1], w).real)
        alphas.append(alpha)
        w = w - alpha * V[-1]
        for u in V:
            w = w - np.vdot(u, w) * u
        beta = float(np.linalg.norm(w))
        if k == 0:
            beta1 = beta
        if beta < tol:
            if est_ref_at == 0:
                est_ref, est_ref_at = 0.0, k + 1
            break
        jj = len(alphas)
        want_ref = info is not None and est_ref_at == 0 and jj >= est_ref_j
        stop = False
        if (adaptive and jj >= 2) or want_ref:
            Et = scipy.linalg.expm(tau * _tridiag_matrix(alphas, betas))
            est = abs(beta) * abs(tau) * abs(Et[jj - 1, 0]) * beta0
            if want_ref:
                est_ref, est_ref_at = est, jj
            stop = bool(adaptive and jj >= 2 and est < tol)
        betas.append(beta)
        v_prev = V[-1]
        V.append(w / beta)
        if stop:
            break
    else:
        capped = bool(adaptive)
    c = beta0 * scipy.linalg.expm(tau * _tridiag_matrix(alphas, betas))[:, 0]
    if info is not None:
        info.update(krylov_dim=len(alphas), err_est=est, beta1=beta1, est_ref=est_ref,
                    est_ref_j=est_ref_at, capped=capped)
    out = np.zeros_like(V[0], dtype=np.complex128)
    for coeff, vj in zip(c, V):
        out = out + complex(coeff) * vj
    return out


def lanczos_expm_fixed(matvec, v, tau, m):
    """exp(tau H) v from m Krylov steps with CGS2, alpha read from the first projection and a
    clamped division instead of a breakdown branch; exponential of the tridiagonal by its
    eigendecomposition.  tnkit/eigensolver.py:95-217 (expm_where='eigh'),
    tnkit/backend.py:372-381 (guard_denom), :364-371 (expm_sym_tridiag_eigh)."""
    n = v.size
    m = max(1, min(int(m), n))
    V = np.zeros((m, n), dtype=np.result_type(v.dtype, np.complex64)
                 if np.iscomplexobj(tau) or np.iscomplexobj(v) else v.dtype)
    beta0 = float(np.linalg.norm(v))
    V[0] = v.reshape(-1) / max(beta0, 1e-200)
    alphas, betas = [], []
    for j in range(m):
        w = matvec(V[j])
        if j > 0:
            w = w - betas[j - 1] * V[j - 1]
        Vl = V[:j + 1]
        c = np.conj(Vl) @ w
        alphas.append(float(c[j].real))
        w = w - c @ Vl
        c = np.conj(Vl) @ w
        w = w - c @ Vl
        b = float(np.linalg.norm(w))
        if j < m - 1:
            betas.append(b)
            V[j + 1] = w / max(b, 1e-200)
    Tm = np.diag(np.asarray(alphas))
    if betas:
        Tm = Tm + np.diag(betas, 1) + np.diag(betas, -1)
    lam, Z = np.linalg.eigh(Tm)
    E = (Z * np.exp(tau * lam)) @ Z.T
    return ((beta0 * E[:, 0]) @ V).reshape(v.shape)


# --------------------------------------------------------------------------------------------
# truncation
# --------------------------------------------------------------------------------------------


def truncate_svd(A, chi):
    """Rank-chi SVD cut with the ABSOLUTE discarded weight.  tnkit/_dmrg.py:132-137"""
    U, S, Vh = svd(A)
    keep = min(chi, S.shape[0])
    return U[:, :keep], S[:keep], Vh[:keep], float(np.sum(S[keep:] ** 2))


def randomized_svd(A, chi, rng, oversample=8, power_iters=2, mixed=False):
    """Rank-chi factorisation by the Halko-Martinsson-Tropp range finder
    (tnkit/linalg.py:81-135): a Gaussian sketch of ell = chi + oversample columns drawn from
    `rng` on the host (real for real A), `power_iters` rounds of QR-orthonormalised subspace
    iteration, then the SVD of the ell x n projection. `mixed` runs the range finder in single
    precision and the projection and its SVD in double. Below min(m, n) <= ell it is the full
    SVD. Returns (U, S, Vh, absolute discarded weight)."""
    m, n = A.shape
    if min(m, n) <= chi + oversample:
        U, S, Vh = np.linalg.svd(A, full_matrices=False)
        keep = min(chi, S.shape[0])
        return U[:, :keep], S[:keep], Vh[:keep], float(np.sum(S[keep:] ** 2))
    ell = int(min(chi + oversample, n))
    Om = rng.standard_normal((n, ell))
    if np.iscomplexobj(A):
        Om = Om + 1j * rng.standard_normal((n, ell))
    if mixed:
        low = np.complex64 if np.iscomplexobj(A) else np.float32
        As = A.astype(low)
        Q, _ = np.linalg.qr(As @ Om.astype(low))
        for _ in range(power_iters):
            Q, _ = np.linalg.qr(As.conj().T @ Q)
            Q, _ = np.linalg.qr(As @ Q)
        Q = Q.astype(A.dtype)
    else:
        Q, _ = np.linalg.qr(A @ Om)
        for _ in range(power_iters):
            Q, _ = np.linalg.qr(A.conj().T @ Q)
            Q, _ = np.linalg.qr(A @ Q)
    Ub, S, Vh = np.linalg.svd(Q.conj().T @ A, full_matrices=False)
    U = Q @ Ub
    keep = min(chi, S.shape[0])
    return U[:, :keep], S[:keep], Vh[:keep], float(np.sum(S[keep:] ** 2))


def _gram_left(A, chi):
    """Top-chi left factorisation A ~= Q C from the smaller Gram matrix.  tnkit/linalg.py:138-159"""
    m, n = A.shape
    Ah = A.conj().T
    w, Z = np.linalg.eigh(A @ Ah if m <= n else Ah @ A)
    r = min(m, n)
    k = min(int(chi), r)
    Z = Z[:, list(range(r - 1, r - k - 1, -1))]
    if m <= n:
        return Z, Z.conj().T @ A, w[::-1], k
    Q, Rr = np.linalg.qr(A @ Z)
    return Q, Rr @ Z.conj().T, w[::-1], k


def truncate_gram(A, chi, iso="left"):
    """Gram-matrix truncation with an exact isometry on the `iso` side; S = norms of the
    carried factor; discarded weight = sum of the dropped eigenvalues clipped at 0.
    tnkit/linalg.py:162-204"""
    if np.finfo(A.dtype).eps > np.finfo(np.float64).eps:
        raise ValueError("trunc='gram' needs double precision")
    if iso == "right":
        Q, C, w, k = _gram_left(A.conj().T, chi)
        Vh, US = Q.conj().T, C.conj().T
        S = np.linalg.norm(US, axis=0)
        Sg = np.where(S > 0, S, 1.0)
        return US / Sg[None, :], S, Vh, float(np.sum(np.clip(w[k:], 0.0, None)))
    if iso != "left":
        raise ValueError(f"iso must be 'left' or 'right', got {iso!r}")
    U, C, w, k = _gram_left(A, chi)
    S = np.linalg.norm(C, axis=1)
    Sg = np.where(S > 0, S, 1.0)
    return U, S, C / Sg[:, None], float(np.sum(np.clip(w[k:], 0.0, None)))


# --------------------------------------------------------------------------------------------
# the sweep
# --------------------------------------------------------------------------------------------


def dmrg(mpo: Sequence[np.ndarray], chi: int, n_sweeps: int = 8, tol: float = 1e-10,
         rtol: float = 1e-10, eig_tol: float = 1e-9, eig: str = "lanczos-restart",
         trunc: str = "svd", seed: int = 0, max_iter: Optional[int] = None,
         restart_m: int = 40, restart_k: int = 10, mixed_tol_lo: float = 1e-4,
         init_mps: Optional[List[np.ndarray]] = None, dtype=np.float64,
         max_steps: Optional[int] = None, canonical: str = "svd", reorth: str = "full",
         eig_stop: str = "residual", expand: int = 0, expand_alpha: float = 3e-3,
         expand_decay: float = 0.3) -> Tuple[float, List[np.ndarray], Dict]:
    """Dense two-site DMRG, the sweep of tnkit/_dmrg.py:560-799 restricted to the dense
    path: warm-started thick-restart Lanczos (or the mixed two-stage form with a float32
    matvec, or the unbounded-basis Lanczos of eig='lanczos'), SVD or Gram truncation, S renormalised, environments updated after each cut.

    `expand` > 0 runs that many opening sweeps with local subspace expansion at
    alpha = expand_alpha * expand_decay**sweep (tnkit/_dmrg.py:700-705, :780; the split is
    `oracle.expansion.enriched_split`), and a sweep with expansion never ends the run.
    `max_steps` (oracle-only) stops after that many local steps, for timing samples.
    `canonical` 'lq' starts from the LQ right-canonical form (the same state in another gauge).
    Returns (E, M, info) with tnkit's info keys 'sweeps', 'energies', 'discarded_weight',
    'matvec_calls', 'sweep_s', 'mixed_counts'."""
    N = len(mpo)
    d = mpo[0].shape[2]
    real = not np.issubdtype(np.dtype(dtype), np.complexfloating)
    if real and any(np.abs(np.imag(W)).max(initial=0.0) > 1e-12 for W in mpo):
        raise TypeError("real dtype requested for an MPO with a significant imaginary part")
    mpo = [np.asarray(np.real(W) if real else W, dtype=dtype) for W in mpo]
    canon = right_canonicalize if canonical == "svd" else right_canonicalize_lq
    if init_mps is not None:
        M = canon([np.asarray(x, dtype=dtype) for x in init_mps])
    else:
        M = canon(random_mps(N, chi, d, np.random.default_rng(seed), dtype))
    L = [None] * (N + 1)
    R = [None] * (N + 1)
    L[0] = np.ones((1, 1, 1), dtype=dtype)
    R[N] = np.ones((1, 1, 1), dtype=dtype)
    for i in range(N - 1, 0, -1):
        R[i] = update_right_env(R[i + 1], M[i], mpo[i])
    energies, sweep_s = [], []
    total_disc = 0.0
    rng_trunc = np.random.default_rng(seed + 1)     # tnkit/_dmrg.py:613
    counts = {"matvec": 0}
    mixed_counts = {"lo": 0, "hi": 0}
    steps = [0]

    def solve(i, direction, alpha=0.0):
        nonlocal total_disc
        W1, W2 = mpo[i], mpo[i + 1]
        theta0 = np.einsum("asc,ctb->astb", M[i], M[i + 1], optimize=True)
        shape = theta0.shape
        Lm, S, Rm = fused_two_site(L[i], W1, W2, R[i + 2])

        def mv(x):
            counts["matvec"] += 1
            return apply_two_site_fused(x.reshape(shape), Lm, S, Rm).reshape(-1)

        n = int(np.prod(shape))
        mi = max_iter if max_iter is not None else min(n, max(200, 4 * chi))
        if eig == "lanczos-mixed":
            lo_ops = (Lm.astype(np.float32), S.astype(np.float32), Rm.astype(np.float32))

            def mv_lo(x):
                counts["matvec"] += 1
                y = apply_two_site_fused(x.reshape(shape).astype(np.float32), *lo_ops)
                return y.reshape(-1).astype(np.float64)

            e, x, inf = lanczos_ground_mixed(mv_lo, mv, theta0.reshape(-1), tol=eig_tol,
                                             max_iter=mi, m=restart_m, k=restart_k,
                                             tol_lo=mixed_tol_lo)
            mixed_counts["lo"] += inf["n_lo"]
            mixed_counts["hi"] += inf["n_hi"]
        elif eig == "lanczos-restart":
            e, x, _ = lanczos_ground_restart(mv, theta0.reshape(-1), tol=eig_tol, max_iter=mi,
                                             m=restart_m, k=restart_k, dtype=dtype)
        elif eig == "lanczos":
            e, x, _ = lanczos_ground(mv, theta0.reshape(-1), tol=eig_tol, max_iter=mi,
                                     reorth=reorth, stop=eig_stop)
        else:
            raise ValueError(f"oracle eig {eig!r} not restated")
        x = x.reshape(shape)
        a, s1, s2, b = shape
        if alpha:
            from .expansion import enriched_split, expansion_term
            P = expansion_term(x, L[i], W1, W2, R[i + 2], direction == "right")
            A, C, disc = enriched_split(x, P, alpha, chi, direction == "right")
            if direction == "right":
                M[i], M[i + 1] = A, C / np.linalg.norm(C)
                L[i + 1] = update_left_env(L[i], M[i], W1)
            else:
                M[i + 1], M[i] = A, C / np.linalg.norm(C)
                R[i + 1] = update_right_env(R[i + 2], M[i + 1], W2)
            total_disc += disc
            steps[0] += 1
            return e
        mat = x.reshape(a * s1, s2 * b)
        if trunc == "svd":
            U, Sv, Vh, disc = truncate_svd(mat, chi)
        elif trunc == "gram":
            U, Sv, Vh, disc = truncate_gram(mat, chi, "left" if direction == "right" else "right")
        elif trunc in ("randomized", "randomized-mixed"):
            U, Sv, Vh, disc = randomized_svd(mat, chi, rng_trunc,
                                             mixed=trunc == "randomized-mixed")
        else:
            raise ValueError(f"oracle trunc {trunc!r} not restated")
        Sv = Sv / np.linalg.norm(Sv)
        if direction == "right":
            M[i] = U.reshape(a, s1, -1)
            M[i + 1] = (Sv[:, None] * Vh).reshape(-1, s2, b)
            L[i + 1] = update_left_env(L[i], M[i], W1)
        else:
            M[i] = (U * Sv).reshape(a, s1, -1)
            M[i + 1] = Vh.reshape(-1, s2, b)
            R[i + 1] = update_right_env(R[i + 2], M[i + 1], W2)
        total_disc += disc
        steps[0] += 1
        return e

    prev = None
    E = None
    sweep = -1
    alphas = []
    for sweep in range(n_sweeps):
        t0 = time.perf_counter()
        alpha = expand_alpha * expand_decay ** sweep if sweep < expand else 0.0
        alphas.append(alpha)
        for i in range(0, N - 1):
            E = solve(i, "right", alpha)
            if max_steps is not None and steps[0] >= max_steps:
                break
        else:
            for i in range(N - 2, -1, -1):
                E = solve(i, "left", alpha)
                if max_steps is not None and steps[0] >= max_steps:
                    break
        energies.append(E)
        sweep_s.append(time.perf_counter() - t0)
        if max_steps is not None and steps[0] >= max_steps:
            break
        if prev is not None and abs(E - prev) < max(tol, rtol * abs(E)):
            break
        prev = None if alpha else E
    info = {"sweeps": sweep + 1, "energies": energies, "discarded_weight": total_disc,
            "matvec_calls": counts["matvec"], "sweep_s": sweep_s, "mixed_counts": mixed_counts,
            "steps": steps[0], "expand": expand, "expand_alphas": alphas}
    return E, M, info


# --------------------------------------------------------------------------------------------
# two-site TDVP
# --------------------------------------------------------------------------------------------


def tdvp2_evolve(mps, mpo, chi, dt, n_steps, krylov_m=20, measure=None, svd_tol=0.0):
    """Two-site TDVP (tnkit/tdvp.py:555-800): each local step evolves theta by exp(-i H2 dt/2)
    with the fixed-m Krylov exponential, cuts it with an SVD to at most `chi` values (and the
    `svd_tol` relative budget), renormalises S, updates the environment and back-evolves the
    one-site centre by exp(+i H1 dt/2) unless it is at the chain end; right then left sweeps.
    Returns (M, records) with records [(t, measure(M))]."""
    M = right_canonicalize_lq([np.asarray(t, dtype=np.complex128) for t in mps])
    W = [np.asarray(w, dtype=np.complex128) for w in mpo]
    N = len(M)
    L = [None] * (N + 1)
    R = [None] * (N + 1)
    L[0] = np.ones((1, 1, 1), dtype=np.complex128)
    R[N] = np.ones((1, 1, 1), dtype=np.complex128)
    for i in range(N - 1, 0, -1):
        R[i] = update_right_env(R[i + 1], M[i], W[i])
    rec = [(0.0, measure(M))] if measure else []

    def cut(mat):
        U, S, Vh = svd(mat)
        tot = float(np.sum(S ** 2)) or 1.0
        keep = min(chi, S.shape[0])
        if svd_tol > 0:
            tail = np.cumsum((S ** 2)[::-1])[::-1] / tot
            drop = np.append(tail, 0.0)
            ok = np.nonzero(drop[1:keep + 1] <= svd_tol)[0]
            keep = int(ok[0]) + 1 if len(ok) else keep
        S = S[:keep] / np.linalg.norm(S[:keep])
        return U[:, :keep], S, Vh[:keep]

    def step(i, moving):
        th = np.einsum("asc,ctb->astb", M[i], M[i + 1])
        a, s1, s2, b = th.shape
        op = fused_two_site(L[i], W[i], W[i + 1], R[i + 2])
        th = lanczos_expm_fixed(lambda x: apply_two_site_fused(x.reshape(th.shape), *op)
                                .reshape(-1), th.reshape(-1), -0.5j * dt, krylov_m)
        U, S, Vh = cut(th.reshape(a * s1, s2 * b))
        nk = S.shape[0]
        if moving == "right":
            M[i] = U.reshape(a, s1, nk)
            L[i + 1] = update_left_env(L[i], M[i], W[i])
            C = (S[:, None] * Vh).reshape(nk, s2, b)
            if i < N - 2:
                C = lanczos_expm_fixed(lambda x: apply_one_site(x.reshape(C.shape), L[i + 1],
                                                                W[i + 1], R[i + 2]).reshape(-1),
                                       C.reshape(-1), 0.5j * dt, krylov_m).reshape(C.shape)
            M[i + 1] = C
        else:
            M[i + 1] = Vh.reshape(nk, s2, b)
            R[i + 1] = update_right_env(R[i + 2], M[i + 1], W[i + 1])
            C = (U * S).reshape(a, s1, nk)
            if i > 0:
                C = lanczos_expm_fixed(lambda x: apply_one_site(x.reshape(C.shape), L[i], W[i],
                                                                R[i + 1]).reshape(-1),
                                       C.reshape(-1), 0.5j * dt, krylov_m).reshape(C.shape)
            M[i] = C

    for n in range(n_steps):
        for i in range(N - 1):
            step(i, "right")
        for i in range(N - 2, -1, -1):
            step(i, "left")
        if measure:
            rec.append(((n + 1) * dt, measure(M)))
    return M, rec


def apply_zero_site(C, L, R):
    """Bond H_eff C (tnkit/_dmrg.py:202-213): C (r, b) -> sum L[p,w,q] C[q,b'] R[b,w,b']."""
    return np.einsum("pwq,qc,bwc->pb", L, C, R, optimize=True)


def tdvp1_evolve(mps, mpo, dt, n_steps, fixed_m=None, krylov_m=30, measure=None,
                 meas_every=1):
    """One-site TDVP (tnkit/tdvp.py:804-852 and the sweep of :1188-1360): right-canonicalise,
    then per step a rightward sweep (evolve site i by exp(-i H1 dt/2), QR it, evolve the bond
    by exp(+i H0 dt/2), absorb it into site i+1) and the mirrored leftward sweep (LQ via the
    QR of the transpose). The Krylov solver is the adaptive `lanczos_expm` capped at
    `krylov_m`, or `lanczos_expm_fixed` with `fixed_m` steps. Bond dimensions never change.
    Returns (M, records) with records (t, measure(M)) at t = 0 and every `meas_every` steps."""
    N = len(mps)
    M = right_canonicalize([np.asarray(m, dtype=np.complex128) for m in mps])
    W = [np.asarray(w, dtype=np.complex128) for w in mpo]
    L, R = [None] * (N + 1), [None] * (N + 1)
    L[0] = np.ones((1, 1, 1), dtype=np.complex128)
    R[N] = np.ones((1, 1, 1), dtype=np.complex128)
    for i in range(N - 1, 0, -1):
        R[i] = update_right_env(R[i + 1], M[i], W[i])

    def ex(mv, x, tau):
        if fixed_m:
            return lanczos_expm_fixed(mv, x.reshape(-1), tau, fixed_m).reshape(x.shape)
        return lanczos_expm(mv, x.reshape(-1), tau, m=krylov_m).reshape(x.shape)

    records = [(0.0, measure(M))] if measure is not None else []
    for step in range(n_steps):
        for i in range(N):
            a, s, b = M[i].shape
            mv1 = lambda x: apply_one_site(x.reshape(a, s, b), L[i], W[i], R[i + 1]).reshape(-1)  # noqa: E731,E501
            M[i] = ex(mv1, M[i], -0.5j * dt)
            if i == N - 1:
                continue
            Q, C = np.linalg.qr(M[i].reshape(a * s, b))
            r = C.shape[0]
            M[i] = Q.reshape(a, s, r)
            L[i + 1] = update_left_env(L[i], M[i], W[i])
            mv0 = lambda x: apply_zero_site(x.reshape(r, b), L[i + 1], R[i + 1]).reshape(-1)  # noqa: E731,E501
            C = ex(mv0, C, 0.5j * dt)
            M[i + 1] = np.einsum("rb,bsc->rsc", C, M[i + 1], optimize=True)
        for i in range(N - 1, -1, -1):
            a, s, b = M[i].shape
            mv1 = lambda x: apply_one_site(x.reshape(a, s, b), L[i], W[i], R[i + 1]).reshape(-1)  # noqa: E731,E501
            M[i] = ex(mv1, M[i], -0.5j * dt)
            if i == 0:
                continue
            Qt, Ct = np.linalg.qr(np.einsum("asb->sba", M[i]).reshape(s * b, a))
            r = Ct.shape[0]
            M[i] = np.einsum("sbr->rsb", Qt.reshape(s, b, r))
            R[i] = update_right_env(R[i + 1], M[i], W[i])
            C = Ct.T
            mv0 = lambda x: apply_zero_site(x.reshape(a, r), L[i], R[i]).reshape(-1)  # noqa: E731,E501
            C = ex(mv0, C, 0.5j * dt)
            M[i - 1] = np.einsum("csa,ar->csr", M[i - 1], C, optimize=True)
        if measure is not None and (step + 1) % meas_every == 0:
            records.append(((step + 1) * dt, measure(M)))
    return M, records


def mps_dense(M):
    """The state vector of an MPS (site 0 slowest)."""
    v = np.ones((1, 1), dtype=np.result_type(*[np.asarray(t).dtype for t in M]))
    for t in M:
        t = np.asarray(t)
        v = np.einsum("xa,asb->xsb", v, t).reshape(-1, t.shape[2])
    return v.reshape(-1)


def local_expectations(M, op):
    """<op_i> of a dense MPS: right-canonicalise a copy, sweep the norm environment from the
    left (tnkit/observables.py:112-147)."""
    M = right_canonicalize_lq([np.asarray(t, dtype=np.complex128) for t in M])
    Lenv = np.ones((1, 1), dtype=np.complex128)
    out = []
    for A in M:
        out.append(np.einsum("ab,asc,st,btc->", Lenv, A, op, A.conj()).real)
        Lenv = np.einsum("ab,asc,bsd->cd", Lenv, A, A.conj())
    return np.array(out)


# --------------------------------------------------------------------------------------------
# energy and its derivative with respect to the MPO (Hellmann-Feynman checks)
# --------------------------------------------------------------------------------------------


def mpo_energy(M, mpo):
    """<psi|H|psi> / <psi|psi> by the left environment chain (norm by the same chain with the
    identity MPO)."""
    L = np.ones((1, 1, 1))
    Ln = np.ones((1, 1))
    for A, W in zip(M, mpo):
        L = update_left_env(L, A, W)
        Ln = np.einsum("pq,psd,qse->de", Ln, A.conj(), A, optimize=True)
    return float(np.real(L.reshape(-1)[0] / Ln.reshape(-1)[0]))


def mpo_gradient(L, A, W, R):
    """d<psi|H|psi>/dW at one site for fixed environments:
    dW[w,r,s,y] = sum L[p,w,q] conj(A[p,s,d]) A[q,y,e] R[d,r,e] (the site's term of the
    environment chain, differentiated in W)."""
    return np.einsum("pwq,psd,qye,dre->wrsy", L, A.conj(), A, R, optimize=True)


def field_gradient(M, mpo_fields_fn, g):
    """dE/dg_i for the per-site-field TFIM, E = <psi|H(g)|psi> / <psi|psi> at a right-canonical
    state with its centre at site 0 (so <psi|psi> = |M[0]|^2): the MPO gradient at site i
    contracted with dW_i/dg_i = -Z in the (0, last) channel."""
    mpo = mpo_fields_fn(g)
    N = len(mpo)
    R = [None] * (N + 1)
    R[N] = np.ones((1, 1, 1))
    for i in range(N - 1, 0, -1):
        R[i] = update_right_env(R[i + 1], M[i], mpo[i])
    Z = np.diag([1.0, -1.0])
    out = np.zeros(N)
    L = np.ones((1, 1, 1))
    norm2 = float(np.real(np.vdot(M[0], M[0])))
    for i in range(N):
        dW = mpo_gradient(L, M[i], mpo[i], R[i + 1])
        wl, wr = mpo[i].shape[:2]
        out[i] = float(np.real(np.sum(dW[0, wr - 1] * (-Z)))) / norm2
        L = update_left_env(L, M[i], mpo[i])
    return out
