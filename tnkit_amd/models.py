"""MPO builders of the models on this path.

`tfim_mpo` restates tnkit/models.py:50-95 (D = 3); `tfim_mpo_fields` is the same chain with
per-site fields g_i (the W[0, 2] = -g_i Z entry of tnkit/models.py:77-84 made site-dependent);
`xxz_mpo` is the D = 5 XXZ chain of tnkit's benchmark runner
(bench/suite/runners/tnkit_runner.py:32-37), with spin-1/2 operators S = sigma / 2.
"""
from __future__ import annotations

import numpy as np

I2 = np.eye(2)
SX = np.array([[0.0, 1.0], [1.0, 0.0]])
SY = np.array([[0.0, -1j], [1j, 0.0]])
SZ = np.array([[1.0, 0.0], [0.0, -1.0]])
SP = np.array([[0.0, 1.0], [0.0, 0.0]])     # S+ (spin-1/2)
SM = SP.T.copy()                             # S-
SZH = np.diag([0.5, -0.5])                   # S^z = sigma^z / 2


def tfim_mpo_fields(g, h=0.0):
    """H = -sum X_i X_{i+1} - sum g_i Z_i - h sum X_i, open chain, real float64 tensors
    (1,3,2,2), (3,3,2,2) ..., (3,1,2,2)."""
    g = np.asarray(g, dtype=np.float64)
    N = g.size
    if N < 2:
        raise ValueError("tfim_mpo needs N >= 2")
    mpo = []
    for i in range(N):
        W = np.zeros((3, 3, 2, 2))
        W[0, 0] = I2
        W[0, 1] = SX
        W[0, 2] = -g[i] * SZ - h * SX
        W[1, 2] = -SX
        W[2, 2] = I2
        if i == 0:
            W = W[0:1]
        elif i == N - 1:
            W = W[:, 2:3]
        mpo.append(np.ascontiguousarray(W))
    return mpo


def tfim_mpo(N, g, h=0.0):
    """Open transverse-field Ising chain H = -sum X X - g sum Z - h sum X (tnkit/models.py:50-95),
    as real float64 tensors (tnkit returns complex128 with zero imaginary part)."""
    return tfim_mpo_fields(np.full(int(N), float(g)), h)


def xxz_mpo(N, J=1.0, Jz=1.0):
    """H = J sum (S+ S- + S- S+)/2 + Jz sum Sz Sz, open chain, D = 5, real
    (bench/suite/runners/tnkit_runner.py:32-37; the runner's prune() removes nothing here)."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0], W[0, 1], W[0, 2], W[0, 3] = I2, SP, SM, SZH
    W[1, 4], W[2, 4], W[3, 4], W[4, 4] = 0.5 * J * SM, 0.5 * J * SP, Jz * SZH, I2
    return [np.ascontiguousarray(W[:1])] + [W.copy() for _ in range(int(N) - 2)] + \
        [np.ascontiguousarray(W[:, 4:])]
