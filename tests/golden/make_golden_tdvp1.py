"""Write tests/golden/golden_tdvp1.json by running tnkit's one-site TDVP (`tdvp1_evolve`).

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_tdvp1.py

The problem is the one of tnkit's graph-fallback tests: a random complex MPS (seed 2026, N =
10, bonds up to 6) under the TFIM with g = 1, h = 0.1. Recorded: <sigma^z_i> after every step
and the final state as a dense vector, for the adaptive Krylov solver and for the fixed
Krylov dimension 8.
"""
from __future__ import annotations

import json
import os
import sys

os.environ["TNKIT_EINSUM_PLAN"] = "0"
for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.environ.get("TNKIT_REF", "/root/reference"))

from tnkit.models import tfim_mpo  # noqa: E402
from tnkit.observables import local_expectations  # noqa: E402
from tnkit.tdvp import tdvp1_evolve  # noqa: E402

SZ = np.diag([1.0, -1.0]).astype(np.complex128)


def problem(n=10, chi=6, seed=2026):
    rng = np.random.default_rng(seed)
    dims = [1] + [min(chi, 2 ** min(i + 1, n - i - 1)) for i in range(n - 1)] + [1]
    return [(rng.normal(size=(dims[i], 2, dims[i + 1]))
             + 1j * rng.normal(size=(dims[i], 2, dims[i + 1]))) / 10.0 for i in range(n)]


def dense(mps):
    v = np.ones((1, 1), dtype=np.complex128)
    for t in mps:
        t = np.asarray(t)
        v = np.einsum("xa,asb->xsb", v, t).reshape(-1, t.shape[2])
    return v.reshape(-1)


def main():
    mpo = tfim_mpo(10, 1.0, 0.10)
    runs = []
    for name, kw in (("adaptive", {}), ("fixed8", {"fast_expm": 8})):
        st, rec = tdvp1_evolve([m.copy() for m in problem()], mpo, 0.05, 4, backend="numpy",
                               graph="off", measure=lambda s: local_expectations(s, SZ).real,
                               **kw)
        v = dense(list(st))
        runs.append({"name": name, "dt": 0.05, "n_steps": 4, "kwargs": kw,
                     "z": [[float(x).hex() for x in z] for _, z in rec],
                     "times": [float(t) for t, _ in rec],
                     "psi_re": [float(x).hex() for x in v.real],
                     "psi_im": [float(x).hex() for x in v.imag]})
    with open(os.path.join(HERE, "golden_tdvp1.json"), "w") as f:
        json.dump({"problem": {"n": 10, "chi": 6, "seed": 2026, "g": 1.0, "h": 0.1},
                   "runs": runs}, f)
    print("wrote golden_tdvp1.json")


if __name__ == "__main__":
    main()
