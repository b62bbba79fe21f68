"""numpy restatement of tnkit's local subspace expansion for two-site DMRG -- ORACLE (test
infrastructure only).

`enriched_split` follows tnkit/expansion.py:173-213 on dense tensors (Hubig, McCulloch,
Schollwoeck and Wolf, Phys. Rev. B 91, 155115 (2015)): the optimised centre x is split at the
bond the sweep crosses, with the new basis the `chi` leading eigenvectors of
x x^H + alpha^2 P P^H on the kept legs, P the expansion term less its part in the span of x.
`expansion_term` is the P of tnkit/_dmrg.py:700-705 (H_eff applied to x on the kept side only).

Pinned by tnkit's own outputs (tests/golden/golden_expand.json: A, C and the discarded weight on
fixed inputs, both sides, three (alpha, chi)), and by what the mathematics fixes: with room for
every state A C = x exactly; with alpha = 0 the split is the truncated SVD of x.
"""
from __future__ import annotations

import numpy as np

from .dense import svd


def top(mat, k, rel_tol):
    """U, S, Vh of the k largest singular values above rel_tol times the largest (at least
    one) (tnkit/expansion.py:35-50, dense branch)."""
    U, S, Vh = svd(mat)
    n = min(int(k), max(1, int((S > rel_tol * S[0]).sum())))
    return U[:, :n], S[:n], Vh[:n]


def enriched_split(x, extra, alpha, chi, right, rel_tol=1e-14):
    """(A, C, disc) with x ~ A C: right=True keeps the first two legs of x (A is (a, s1, k), C is
    (k, s2, b)); right=False keeps the last two (A is (k, s2, b), C is (a, s1, k)). `extra` has
    the same kept legs as x. disc = |x|^2 - |C|^2, clipped at 0."""
    a, s1, s2, b = x.shape
    if right:
        X = x.reshape(a * s1, s2 * b)
        U, _, _ = top(X, min(X.shape), rel_tol)
        E = extra.reshape(a * s1, -1)
        E = E - U @ (U.conj().T @ E)
        rho = X @ X.conj().T + alpha ** 2 * (E @ E.conj().T)
        Ur, _, _ = top(rho, chi, rel_tol ** 2)
        A = Ur.reshape(a, s1, -1)
        C = (Ur.conj().T @ X).reshape(-1, s2, b)
    else:
        X = x.reshape(a * s1, s2 * b)
        _, _, Vh = top(X, min(X.shape), rel_tol)
        E = extra.reshape(-1, s2 * b)
        E = E - (E @ Vh.conj().T) @ Vh
        rho = X.conj().T @ X + alpha ** 2 * (E.conj().T @ E)
        _, _, Vr = top(rho, chi, rel_tol ** 2)
        A = Vr.reshape(-1, s2, b)
        C = (X @ Vr.conj().T).reshape(a, s1, -1)
    disc = max(float(np.linalg.norm(x)) ** 2 - float(np.linalg.norm(C)) ** 2, 0.0)
    return A, C, disc


def expansion_term(x, L, W1, W2, R, right):
    """P of tnkit/_dmrg.py:700-705: H_eff applied to x on the side that keeps the basis,
    (a, s1, w, s2, b) with w the MPO channel between the two sites."""
    if right:
        t = np.einsum("pwq,qijb->pwijb", L, x, optimize=True)
        return np.einsum("wmsi,pwijb->psmjb", W1, t, optimize=True)
    t = np.einsum("drb,astb->astdr", R, x, optimize=True)
    return np.einsum("wrjt,astdr->aswjd", W2, t, optimize=True)
