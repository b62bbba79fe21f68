"""numpy restatement of tnkit's stateful TDVP on its per-bond schedules -- ORACLE (test
infrastructure only).

`Evolver` follows tnkit/tdvp_evolver.py:65-364 without expansion or graphs: '1site' is the
one-site step of tnkit/tdvp.py:804-850 on every site, '2site' the two-site step of
tnkit/tdvp.py:555-650 on every bond, and '12site' picks per bond at the start of every step:
two-site when the bond is below the cap and was not pinned by degenerate_tol at its last
two-site cut, one-site otherwise; every `two_site_every`-th step is two-site everywhere
(tnkit/tdvp_evolver.py:240-276, :313-325). The truncation is tnkit's `_cut` rule
(tnkit/_dmrg.py:77-113) on an SVD; Krylov exponentials are the adaptive `lanczos_expm` or the
fixed-m `lanczos_expm_fixed` of oracle.dense.

Pinned by tnkit's recorded runs (tests/golden/golden_evolver.json: <Z_i> per step, bonds and the
two-site bond counts of every schedule) and, at a cap above every exact bond, by exact evolution
of the state vector.
"""
from __future__ import annotations

import numpy as np

from .dense import (apply_one_site, apply_two_site, apply_zero_site, lanczos_expm,
                    lanczos_expm_fixed, right_canonicalize, svd, update_left_env,
                    update_right_env)


def cut(s, chi, svd_tol=0.0, degenerate_tol=0.0, disc0=0.0, tot=None):
    """(k, pinned, relative dropped weight) of tnkit/_dmrg.py:77-113."""
    n = len(s)
    tot = tot or float(np.sum(s ** 2)) or 1.0
    k = min(chi, n)
    if degenerate_tol > 0 and k < n and s[k] > 0:
        j = k
        while j > 0 and s[j - 1] - s[k] <= degenerate_tol * s[j - 1]:
            j -= 1
        k = j or k
    pinned = k < min(chi, n)
    disc = disc0 + float(np.sum(s[k:] ** 2)) / tot
    if svd_tol > 0.0 and k > 1:
        tail = np.cumsum((s[:k] ** 2)[::-1])[::-1]
        drop = tail / tot + disc
        ok = np.nonzero(drop[1:] <= svd_tol)[0]
        if len(ok):
            c = j = int(ok[0]) + 1
            if degenerate_tol > 0:
                while j < k and s[c - 1] - s[j] <= degenerate_tol * s[c - 1]:
                    j += 1
            if j < k:
                k, pinned, disc = j, False, float(drop[j])
    return k, pinned, disc


class Evolver:
    """State, environments and schedule of one TDVP trajectory (complex128)."""

    def __init__(self, mps, mpo, dt, chi, fast_expm=None, krylov_m=30, svd_tol=0.0,
                 degenerate_tol=0.0, renorm=True):
        self.W = [np.asarray(w, dtype=np.complex128) for w in mpo]
        self.M = right_canonicalize([np.asarray(m, dtype=np.complex128) for m in mps])
        self.dt, self.cap, self.fm, self.km = dt, chi, fast_expm, krylov_m
        self.svd_tol, self.dtol, self.renorm = svd_tol, degenerate_tol, renorm
        N = len(self.M)
        self.pinned = [False] * (N - 1)
        self.L, self.R = [None] * (N + 1), [None] * (N + 1)
        self.L[0] = np.ones((1, 1, 1), dtype=np.complex128)
        self.R[N] = np.ones((1, 1, 1), dtype=np.complex128)
        for i in range(N - 1, 0, -1):
            self.R[i] = update_right_env(self.R[i + 1], self.M[i], self.W[i])
        self.step, self.t = 0, 0.0

    def _ex(self, mv, x, tau):
        if self.fm:
            return lanczos_expm_fixed(mv, x.reshape(-1), tau, self.fm).reshape(x.shape)
        return lanczos_expm(mv, x.reshape(-1), tau, m=self.km).reshape(x.shape)

    def one_site(self, i, direction):
        """tnkit/tdvp.py:804-850."""
        M, L, R, W, h = self.M, self.L, self.R, self.W, self.dt
        N = len(M)
        a, s, b = M[i].shape
        M[i] = self._ex(lambda x: apply_one_site(x.reshape(a, s, b), L[i], W[i], R[i + 1])
                        .reshape(-1), M[i], -0.5j * h)
        if direction == "right":
            if i == N - 1:
                return
            Q, C = np.linalg.qr(M[i].reshape(a * s, b))
            r = C.shape[0]
            M[i] = Q.reshape(a, s, r)
            L[i + 1] = update_left_env(L[i], M[i], W[i])
            C = self._ex(lambda x: apply_zero_site(x.reshape(r, b), L[i + 1], R[i + 1])
                         .reshape(-1), C, 0.5j * h)
            M[i + 1] = np.einsum("rb,bsc->rsc", C, M[i + 1])
        else:
            if i == 0:
                return
            Qt, Ct = np.linalg.qr(np.einsum("asb->sba", M[i]).reshape(s * b, a))
            r = Ct.shape[0]
            M[i] = np.einsum("sbr->rsb", Qt.reshape(s, b, r))
            R[i] = update_right_env(R[i + 1], M[i], W[i])
            C = self._ex(lambda x: apply_zero_site(x.reshape(a, r), L[i], R[i]).reshape(-1),
                         Ct.T, 0.5j * h)
            M[i - 1] = np.einsum("csa,ar->csr", M[i - 1], C)

    def two_site(self, i, direction):
        """tnkit/tdvp.py:555-650 with the `_svd_cut` of :527-552; returns the pin flag."""
        M, L, R, W, h = self.M, self.L, self.R, self.W, self.dt
        N = len(M)
        th = np.einsum("asc,ctb->astb", M[i], M[i + 1])
        a, s1, s2, b = th.shape
        th = self._ex(lambda x: apply_two_site(x.reshape(a, s1, s2, b), L[i], W[i], W[i + 1],
                                               R[i + 2]).reshape(-1), th, -0.5j * h)
        U, S, Vh = svd(th.reshape(a * s1, s2 * b))
        tot = float(np.linalg.norm(th)) ** 2 or 1.0
        k, pinned, _ = cut(S, self.cap, self.svd_tol, self.dtol, 0.0, tot)
        U, S, Vh = U[:, :k], S[:k], Vh[:k]
        if self.renorm:
            S = S / np.linalg.norm(S)
        if direction == "right":
            M[i] = U.reshape(a, s1, k)
            L[i + 1] = update_left_env(L[i], M[i], W[i])
            C = (S[:, None] * Vh).reshape(k, s2, b)
            if i < N - 2:
                C = self._ex(lambda x: apply_one_site(x.reshape(k, s2, b), L[i + 1], W[i + 1],
                                                      R[i + 2]).reshape(-1), C, 0.5j * h)
            M[i + 1] = C
        else:
            M[i + 1] = Vh.reshape(k, s2, b)
            R[i + 1] = update_right_env(R[i + 2], M[i + 1], W[i + 1])
            C = (U * S).reshape(a, s1, k)
            if i > 0:
                C = self._ex(lambda x: apply_one_site(x.reshape(a, s1, k), L[i], W[i],
                                                      R[i + 1]).reshape(-1), C, 0.5j * h)
            M[i] = C
        return pinned

    def sweep12(self, two):
        """One step on the per-bond schedule (tnkit/tdvp_evolver.py:240-276)."""
        N, pin = len(self.M), self.pinned
        for i in range(N):
            if i < N - 1 and two[i]:
                pin[i] = self.two_site(i, "right")
            elif not (i == N - 1 and N > 1 and two[N - 2]):
                self.one_site(i, "right")
        for i in range(N - 1, -1, -1):
            if i > 0 and two[i - 1]:
                pin[i - 1] = self.two_site(i - 1, "left")
            elif not (i == 0 and N > 1 and two[0]):
                self.one_site(i, "left")

    def run(self, n_steps, method="2site", two_site_every=0, measure=None):
        """Take n_steps steps; returns ([(t, measure(M))] after every step, [n_two per step])."""
        recs, n_two = [], []
        N = len(self.M)
        for _ in range(n_steps):
            if method == "12site":
                every = two_site_every and (self.step + 1) % two_site_every == 0
                two = [bool(every) or (self.M[i].shape[2] != self.cap and not self.pinned[i])
                       for i in range(N - 1)]
                n_two.append(sum(two))
                self.sweep12(two)
            elif method == "2site":
                for i in range(N - 1):
                    self.pinned[i] = self.two_site(i, "right")
                for i in range(N - 2, -1, -1):
                    self.pinned[i] = self.two_site(i, "left")
                n_two.append(None)
            else:
                for i in range(N):
                    self.one_site(i, "right")
                for i in range(N - 1, -1, -1):
                    self.one_site(i, "left")
                n_two.append(None)
            self.step += 1
            self.t += self.dt
            if measure is not None:
                recs.append((self.t, measure(self.M)))
        return recs, n_two
