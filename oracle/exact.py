"""Closed forms and brute force that pin the oracle (TEST INFRASTRUCTURE ONLY).

Only `tests/`, `__graft_entry__.smoke()` and `bench.py`'s `cpu_baseline` leg use this module.
"""
from __future__ import annotations

import numpy as np


def tfim_free_fermion_energy(N, g):
    """Exact ground energy of the open chain H = -sum X_i X_{i+1} - sum g_i Z_i (g scalar or
    length-N) by Jordan-Wigner: E0 = -1/2 sum_k Lambda_k with Lambda^2 = eig((A-B)(A+B)),
    A = diag(2 g_i) - (nearest-neighbour 1), B_{i,i+1} = -1 = -B_{i+1,i}.
    tnkit/free_fermion.py:19-31 (uniform g); the per-site form only changes diag(A)."""
    g = np.broadcast_to(np.asarray(g, dtype=np.float64), (N,))
    A = np.diag(2.0 * g)
    B = np.zeros((N, N))
    for i in range(N - 1):
        A[i, i + 1] = A[i + 1, i] = -1.0
        B[i, i + 1] = -1.0
        B[i + 1, i] = 1.0
    M = (A - B) @ (A + B)
    w = np.linalg.eigvalsh(0.5 * (M + M.T))
    return -0.5 * float(np.sum(np.sqrt(np.clip(w, 0.0, None))))


def mpo_to_dense(mpo):
    """The full 2^N x 2^N operator of an open MPO (W: wl, wr, s, s'), contracted left to right
    with the boundary channels 0 (left) and 0 (right, the MPO's last tensor has wr = 1)."""
    T = mpo[0][0]                                         # (wr, s, s')
    T = T.reshape(T.shape[0], T.shape[1], T.shape[2])
    for W in mpo[1:]:
        # T: (w, S, S') ; W: (w, r, s, s') -> (r, S s, S' s')
        T = np.einsum("wab,wrst->rasbt", T, W)
        r, a, s, b, t = T.shape
        T = T.reshape(r, a * s, b * t)
    assert T.shape[0] == 1
    return T[0]


def ground_energy_dense(mpo):
    """Lowest eigenvalue of `mpo_to_dense(mpo)` (exact diagonalisation, small N only)."""
    H = mpo_to_dense(mpo)
    H = 0.5 * (H + H.conj().T)
    return float(np.linalg.eigvalsh(H)[0])


def heff_dense(L, W1, W2, R):
    """The two-site effective Hamiltonian as an explicit (a d1 d2 b)^2 matrix, by contracting
    the network with identity inputs: H[(p s t d), (q i j b)] = sum L[p,w,q] W1[w,m,s,i]
    W2[m,r,t,j] R[d,r,b].  A single einsum over the whole network (no pairwise order)."""
    H = np.einsum("pwq,wmsi,mrtj,drb->pstdqijb", L, W1, W2, R)
    a, d1, d2, b = L.shape[0], W1.shape[2], W2.shape[2], R.shape[0]
    n = a * d1 * d2 * b
    return H.reshape(n, L.shape[2] * W1.shape[3] * W2.shape[3] * R.shape[2])


def mps_to_vector(M):
    """The full state vector of an MPS with outer bonds 1."""
    psi = np.asarray(M[0])
    for t in M[1:]:
        psi = np.tensordot(psi, np.asarray(t), axes=([psi.ndim - 1], [0]))
    return psi.reshape(-1)


def site_expectations(psi, op, N):
    """<psi| op_i |psi> / <psi|psi> for every site of an N-site state vector of qubits."""
    psi = psi.reshape((2,) * N)
    nrm = np.vdot(psi, psi).real
    out = []
    for i in range(N):
        phi = np.moveaxis(np.tensordot(op, psi, axes=([1], [i])), 0, i)
        out.append(np.vdot(psi, phi).real / nrm)
    return np.array(out)


def evolve_dense(mpo, psi0, t):
    """exp(-i H t) psi0 for the dense H of an MPO (small N only)."""
    import scipy.linalg
    H = mpo_to_dense(mpo)
    return scipy.linalg.expm(-1j * t * H) @ psi0


def tfim_majorana_matrix(g):
    """A (2N x 2N, real antisymmetric) of H = -sum X_j X_{j+1} - sum g_j Z_j = (i/4) g^T A g in
    the Majoranas g_{2j} = (prod_{k<j} Z_k) X_j, g_{2j+1} = (prod_{k<j} Z_k) Y_j, for which
    Z_j = -i g_{2j} g_{2j+1} and X_j X_{j+1} = -i g_{2j+1} g_{2j+2}: the field puts 2 g_j at
    (2j, 2j+1), the bond 2 at (2j+1, 2j+2). The Jordan-Wigner construction behind
    tnkit/free_fermion_dynamics.py:31-51, with per-site fields."""
    g = np.asarray(g, dtype=np.float64)
    N = g.size
    A = np.zeros((2 * N, 2 * N))
    for j in range(N):
        A[2 * j, 2 * j + 1] = 2.0 * g[j]
        if j + 1 < N:
            A[2 * j + 1, 2 * j + 2] = 2.0
    return A - A.T


def tfim_ground_covariance(g):
    """Majorana covariance Gamma_kl = i <g_k g_l> (k != l) of the ground state of
    `tfim_majorana_matrix(g)`: Gamma = i sign(iA)."""
    A = tfim_majorana_matrix(g)
    w, V = np.linalg.eigh(1j * A)
    return np.real(1j * (V * np.sign(w)) @ V.conj().T)


def tfim_quench_z(g0, g1, times):
    """<Z_j>(t) after the quench from the ground state of fields g0 to fields g1: Heisenberg
    evolution g(t) = exp(A1 t) g(0), Gamma(t) = R Gamma(0) R^T, <Z_j> = -Gamma_{2j, 2j+1}
    (tnkit/free_fermion_dynamics.py:54-115). Returns an array (len(times), N)."""
    import scipy.linalg
    G0 = tfim_ground_covariance(g0)
    A1 = tfim_majorana_matrix(g1)
    N = A1.shape[0] // 2
    out = []
    for t in times:
        R = scipy.linalg.expm(A1 * t)
        G = R @ G0 @ R.T
        out.append(-np.array([G[2 * j, 2 * j + 1] for j in range(N)]))
    return np.array(out)


def tfim_block_entropies(g):
    """Entanglement entropy (nats) of the first l sites, l = 1..N-1, of the free-fermion ground
    state of `tfim_majorana_matrix(g)`: the reduced state is Gaussian with covariance
    Gamma[:2l, :2l], whose i Gamma has eigenvalues +-nu_k, and S = sum_k h((1 + nu_k) / 2) with
    h(p) = -p ln p - (1 - p) ln(1 - p) (Peschel's formula for a block at the open end)."""
    G = tfim_ground_covariance(g)
    N = G.shape[0] // 2
    out = []
    for l in range(1, N):
        nu = np.linalg.eigvalsh(1j * G[:2 * l, :2 * l])
        nu = np.clip(nu[nu > 0], 0.0, 1.0)
        p = (1 + nu) / 2
        q = 1 - p
        s = 0.0
        for a in np.concatenate([p, q]):
            if a > 1e-16:
                s -= a * np.log(a)
        out.append(s)
    return np.array(out)
