"""Observables of a dense MPS on the device (tnkit/observables.py:112-147, :313-414).

Every routine works on a right-canonical copy (an LQ per site, the same state as tnkit's SVD
form in another bond gauge) and leaves the input untouched:

* `local_expectations(mps, op)`: the norm environment Lenv swept from the left,
  <op_i> = sum Lenv[a,b] A[a,s,c] op[s,t] conj(A[b,t,c]). Per site the two chi^3 products
  T = Lenv^T A and Lenv' = T^T conj(A) are K-major MFMA GEMMs of libtnkit_hip (`gemm_tn`); the
  d x d operator mix and the final reduction are vector work.
* `all_entropies(mps)` / `entanglement_entropy(mps, bond)`: the centre carried left to right,
  each cut's Schmidt values from the SVD that moves it (rocSOLVER), all spectra brought to the
  host in one transfer; the entropy is tnkit's arithmetic (nats, spectrum normalised,
  probabilities below 1e-16 dropped).
* `two_site_entropy(mps)`: the reduced density matrix of each pair of sites, one transfer for all
  of them, eigenvalues on the host.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np
import torch

from . import kernels


def _device_mps(mps) -> List[torch.Tensor]:
    dev = torch.device("cuda", torch.cuda.current_device())
    M = [t if isinstance(t, torch.Tensor) else torch.as_tensor(np.asarray(t)) for t in mps]
    if getattr(mps[0], "is_block", False):
        raise TypeError("observables need dense tensors: convert a block MPS to dense first")
    dt = torch.complex128 if any(t.is_complex() for t in M) else torch.float64
    return [t.to(dev, dt).contiguous() for t in M]


def right_canonical(mps) -> List[torch.Tensor]:
    """A right-canonical, normalised copy of a dense MPS on the device (LQ per site, on the
    library's QR)."""
    M = [t.clone() for t in _device_mps(mps)]
    for i in range(len(M) - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        Q, Rr = kernels.qr_any(M[i].reshape(Dl, d * Dr).mT)
        M[i] = Q.mT.reshape(-1, d, Dr).contiguous()
        M[i - 1] = torch.tensordot(M[i - 1], Rr.mT, dims=([2], [0])).contiguous()
    M[0] = (M[0] / torch.linalg.norm(M[0])).contiguous()
    return M


def to_numpy(mps) -> List[np.ndarray]:
    """Host copies of the site tensors (tnkit/observables.py:76-84)."""
    return [np.ascontiguousarray(t.detach().cpu().numpy() if isinstance(t, torch.Tensor)
                                 else np.asarray(t)) for t in mps]


def _env_step(Lenv: torch.Tensor, A: torch.Tensor):
    """(T, Lenv') with T[b, s, c] = sum_a Lenv[a, b] A[a, s, c] and
    Lenv'[c, d] = sum_{b, s} T[b, s, c] conj(A)[b, s, d]: two K-major GEMMs."""
    a, s, c = A.shape
    T = kernels.gemm_tn(Lenv.contiguous(), A.reshape(a, s * c))
    Ac = A.conj().resolve_conj().contiguous() if A.is_complex() else A
    Ln = kernels.gemm_tn(T.reshape(a * s, c), Ac.reshape(a * s, c))
    return T.reshape(a, s, c), Ln, Ac


def local_expectations(mps, op, canon: Optional[List[torch.Tensor]] = None) -> np.ndarray:
    """Real part of <op_i> at every site of a dense MPS (device or host tensors; it need not be
    normalised). `canon` is a right-canonical copy from `right_canonical`, to skip that step."""
    M = right_canonical(mps) if canon is None else canon
    dt = M[0].dtype
    if np.iscomplexobj(op) and dt != torch.complex128:
        M = [t.to(torch.complex128) for t in M]
        dt = torch.complex128
    opd = torch.as_tensor(np.asarray(op)).to(M[0].device, dt)
    vals = []
    Lenv = torch.ones((1, 1), dtype=dt, device=M[0].device)
    for A in M:
        T, Ln, Ac = _env_step(Lenv, A)
        U = torch.einsum("bsc,st->btc", T, opd).contiguous()       # the d x d operator mix
        if dt == torch.float64:
            vals.append(kernels.dot(U.reshape(-1), Ac.reshape(-1)))
        else:
            vals.append(torch.sum(U * Ac))
        Lenv = Ln
    return np.ascontiguousarray(np.real(torch.stack(vals).cpu().numpy()))


def _spectra(M: List[torch.Tensor], upto: int) -> List[np.ndarray]:
    out, carry = [], None
    for i in range(upto):
        A = M[i] if carry is None else torch.tensordot(carry, M[i], dims=([1], [0]))
        Dl, d, Dr = A.shape
        _, S, Vh = kernels.thin_svd(A.reshape(Dl * d, Dr))
        out.append(S)
        carry = S.to(Vh.dtype)[:, None] * Vh
    sizes = [int(s.numel()) for s in out]
    host = torch.cat(out).cpu().numpy() if out else np.zeros(0)
    return np.split(host, np.cumsum(sizes)[:-1]) if out else []


def _entropy(s: np.ndarray) -> float:
    s = s / np.linalg.norm(s)
    p = s ** 2
    p = p[p > 1e-16]
    return float(-np.sum(p * np.log(p)))


def all_entropies(mps, canon: Optional[List[torch.Tensor]] = None) -> np.ndarray:
    """Von Neumann entropy (nats) of every bond in one sweep; entry i is the bond between sites
    i and i + 1 (tnkit/observables.py:373-387)."""
    M = right_canonical(mps) if canon is None else canon
    return np.array([_entropy(s) for s in _spectra(M, len(M) - 1)])


def entanglement_entropy(mps, bond: int, canon: Optional[List[torch.Tensor]] = None) -> float:
    """Entropy of the cut between sites bond - 1 and bond, 1 <= bond <= N - 1
    (tnkit/observables.py:341-370)."""
    M = right_canonical(mps) if canon is None else canon
    if not 1 <= int(bond) <= len(M) - 1:
        raise ValueError(f"bond must be in 1..{len(M) - 1}, got {bond}")
    return _entropy(_spectra(M, int(bond))[-1])


def two_site_entropy(mps, canon: Optional[List[torch.Tensor]] = None) -> np.ndarray:
    """-Tr rho ln rho of the reduced density matrix of sites (n, n+1), n = 0..N-2, in nats
    (tnkit/observables.py:390-414). Finite open chains only (outer bonds 1)."""
    if int(mps[0].shape[0]) != 1 or int(mps[-1].shape[2]) != 1:
        raise ValueError("two_site_entropy needs a finite open chain (outer bonds 1)")
    M = right_canonical(mps) if canon is None else canon
    rhos = []
    Lenv = torch.ones((1, 1), dtype=M[0].dtype, device=M[0].device)
    for i in range(len(M) - 1):
        th = torch.tensordot(M[i], M[i + 1], dims=([2], [0]))
        rhos.append(torch.einsum("ab,aste,bxye->stxy", Lenv, th, th.conj()))
        Lenv = _env_step(Lenv, M[i])[1]
    out = []
    for r in torch.stack(rhos).cpu().numpy():
        d2 = r.shape[0] * r.shape[1]
        p = np.linalg.eigvalsh(r.reshape(d2, d2))
        p = p / p.sum()
        p = p[p > 1e-16]
        out.append(float(-np.sum(p * np.log(p))))
    return np.array(out)
