"""tnkit_amd: the MI355X (gfx950) two-site effective-Hamiltonian path of tnkit, as
`backend="hip"`.

    from tnkit_amd import dmrg, get_backend, tfim_mpo
    E, mps, info = dmrg(tfim_mpo(32, 1.0), chi=32, backend="hip-real")

Importing the package does not import torch or load the HIP library; both happen when a HIP
backend is first constructed (tnkit imports torch lazily in the same way, tnkit/backend.py).
"""
from __future__ import annotations

__version__ = "0.1.0"

from .models import tfim_mpo, tfim_mpo_fields, xxz_mpo  # noqa: F401
from .mps import MPS  # noqa: F401


def __getattr__(name):
    if name in ("dmrg", "DMRGOptions", "random_mps", "right_canonicalize"):
        from . import _dmrg
        return getattr(_dmrg, name)
    if name in ("get_backend", "HipBackend"):
        from . import backend
        return getattr(backend, name)
    raise AttributeError(name)
