"""On-site expectation values (tnkit/observables.py:112-147), on the device.

`local_expectations(mps, op)` brings a copy of the state to right-canonical form (LQ per site),
then sweeps the norm environment from the left: <op_i> = sum Lenv[a,b] A[a,s,c] op[s,t]
conj(A[b,t,c]). Returns the real parts on the host.
"""
from __future__ import annotations

import numpy as np
import torch


def _right_canonical(M):
    M = [t.clone() for t in M]
    for i in range(len(M) - 1, 0, -1):
        Dl, d, Dr = M[i].shape
        Q, Rr = torch.linalg.qr(M[i].reshape(Dl, d * Dr).mT)
        M[i] = Q.mT.reshape(-1, d, Dr)
        M[i - 1] = torch.tensordot(M[i - 1], Rr.mT, dims=([2], [0]))
    M[0] = M[0] / torch.linalg.norm(M[0])
    return M


def local_expectations(mps, op) -> np.ndarray:
    """<op_i> for every site of a dense MPS (device or host tensors)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    M = [torch.as_tensor(np.asarray(t) if not isinstance(t, torch.Tensor) else t).to(dev)
         for t in mps]
    dt = torch.complex128 if any(t.is_complex() for t in M) or np.iscomplexobj(op) \
        else torch.float64
    M = _right_canonical([t.to(dt) for t in M])
    opd = torch.as_tensor(np.asarray(op)).to(dev, dt)
    out = []
    Lenv = torch.ones((1, 1), dtype=dt, device=dev)
    for A in M:
        out.append(torch.einsum("ab,asc,st,btc->", Lenv, A, opd, A.conj()))
        Lenv = torch.einsum("ab,asc,bsd->cd", Lenv, A, A.conj())
    return np.array([complex(x.item()).real for x in out])
