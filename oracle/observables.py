"""numpy restatement of tnkit's entanglement observables -- ORACLE (test infrastructure only).

`all_entropies`, `entanglement_entropy` and `two_site_entropy` follow
tnkit/observables.py:313-414: a right-canonical copy (SVD form), the centre carried left to
right with the Schmidt spectrum of each cut read off the SVD that moves it; entropies in nats
with the spectrum normalised and probabilities below 1e-16 dropped.

Pinned by exact diagonalisation (Schmidt values of the state vector) and, for the transverse-field
Ising chain, by the free-fermion block entropy of `oracle.exact.tfim_block_entropies`.
"""
from __future__ import annotations

import numpy as np

from .dense import right_canonicalize, svd


def schmidt_spectra(M, upto=None):
    """The Schmidt values of bonds 1..upto (default N-1) of the state M (any gauge)."""
    M = right_canonicalize([np.asarray(t) for t in M])
    upto = len(M) - 1 if upto is None else upto
    out, carry = [], None
    for i in range(upto):
        A = M[i] if carry is None else np.einsum("ab,bsc->asc", carry, M[i])
        Dl, d, Dr = A.shape
        U, S, Vh = svd(A.reshape(Dl * d, Dr))
        out.append(S)
        carry = np.diag(S) @ Vh
    return out


def entropy(s):
    """-sum p ln p over p = s^2 / |s|^2 > 1e-16 (tnkit/observables.py:333-338)."""
    s = np.asarray(s) / np.linalg.norm(s)
    p = s ** 2
    p = p[p > 1e-16]
    return float(-np.sum(p * np.log(p)))


def all_entropies(M):
    """Von Neumann entropy of every bond (tnkit/observables.py:373-387)."""
    return np.array([entropy(s) for s in schmidt_spectra(M)])


def entanglement_entropy(M, bond):
    """Entropy of the cut between sites bond-1 and bond (tnkit/observables.py:341-370)."""
    return entropy(schmidt_spectra(M, bond)[-1])


def two_site_entropy(M):
    """-Tr rho ln rho of the reduced density matrix of sites (n, n+1) (tnkit/observables.py:
    390-414)."""
    M = right_canonicalize([np.asarray(t) for t in M])
    Lenv = np.ones((1, 1), dtype=np.result_type(*M))
    out = []
    for i in range(len(M) - 1):
        th = np.einsum("asc,cte->aste", M[i], M[i + 1])
        r = np.einsum("ab,aste,bxye->stxy", Lenv, th, th.conj())
        Lenv = np.einsum("ab,asc,bsd->cd", Lenv, M[i], M[i].conj())
        d2 = r.shape[0] * r.shape[1]
        p = np.linalg.eigvalsh(r.reshape(d2, d2))
        p = p / p.sum()
        p = p[p > 1e-16]
        out.append(float(-np.sum(p * np.log(p))))
    return np.array(out)
