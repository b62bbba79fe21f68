"""Exact free-fermion results for the open transverse-field Ising chain (host, numpy).

H = -sum_j X_j X_{j+1} - sum_j g_j Z_j maps under Jordan-Wigner to H = (i/4) a^T A a in the
Majoranas a_{2j} = (prod_{k<j} Z_k) X_j, a_{2j+1} = (prod_{k<j} Z_k) Y_j. These are the
reference curves of the TFIM runs: tnkit's tfim_free_fermion_energy (tnkit/free_fermion.py:19-31)
and its quench dynamics (tnkit/free_fermion_dynamics.py:31-115), with per-site fields allowed.

Conventions: Gamma_kl = i <a_k a_l> (k != l), Z_j = -i a_{2j} a_{2j+1}, so <Z_j> = -Gamma_{2j,2j+1}.
The ground state is Gamma = i sign(iA); a quench evolves Gamma(t) = R Gamma R^T with
R = exp(A t), built from the eigen-decomposition of the Hermitian iA so that R is orthogonal to
rounding at every t.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np


def _fields(N: int, g) -> np.ndarray:
    return np.broadcast_to(np.asarray(g, dtype=np.float64), (N,)).copy()


def tfim_ground_energy(N: int, g) -> float:
    """E0 = -1/2 sum_k Lambda_k, Lambda^2 = eig((A - B)(A + B)) with A = diag(2 g_i) minus the
    nearest-neighbour ones and B_{i,i+1} = -1 = -B_{i+1,i} (tnkit/free_fermion.py:19-31)."""
    g = _fields(N, g)
    A = np.diag(2.0 * g)
    B = np.zeros((N, N))
    idx = np.arange(N - 1)
    A[idx, idx + 1] = A[idx + 1, idx] = -1.0
    B[idx, idx + 1], B[idx + 1, idx] = -1.0, 1.0
    M = (A - B) @ (A + B)
    return -0.5 * float(np.sqrt(np.clip(np.linalg.eigvalsh(0.5 * (M + M.T)), 0.0, None)).sum())


def majorana_matrix(N: int, g) -> np.ndarray:
    """The real antisymmetric A of H = (i/4) a^T A a: 2 g_j at (2j, 2j+1), 2 at (2j+1, 2j+2)."""
    g = _fields(N, g)
    A = np.zeros((2 * N, 2 * N))
    j = np.arange(N)
    A[2 * j, 2 * j + 1] = 2.0 * g
    A[2 * j[:-1] + 1, 2 * j[:-1] + 2] = 2.0
    return A - A.T


def ground_covariance(N: int, g) -> np.ndarray:
    w, V = np.linalg.eigh(1j * majorana_matrix(N, g))
    return np.real(1j * (V * np.sign(w)) @ V.conj().T)


def quench_z(N: int, g0, g1, times: Sequence[float]) -> np.ndarray:
    """<Z_j>(t), shape (len(times), N), after a quench from the ground state at fields g0 to
    fields g1."""
    G0 = ground_covariance(N, g0)
    w, V = np.linalg.eigh(1j * majorana_matrix(N, g1))      # A = -i V diag(w) V^H
    out = np.empty((len(times), N))
    j = np.arange(N)
    for n, t in enumerate(times):
        R = np.real((V * np.exp(-1j * w * t)) @ V.conj().T)
        G = R @ G0 @ R.T
        out[n] = -G[2 * j, 2 * j + 1]
    return out
