"""Write tests/golden/golden_expm.json by running tnkit's adaptive Krylov exponential
(`eigensolver.lanczos_expm` with its `_ExpmMonitor`) itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_expm.py

The fixture holds, per case, the inputs' seeds and tnkit's `info` (krylov_dim, err_est, beta1,
est_ref, est_ref_j, capped) and the output vector (real and imaginary parts as hex floats).
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

from tnkit import eigensolver as E  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402


def hexs(xs):
    return [float(x).hex() for x in xs]


def main():
    be = get_backend("numpy")
    cases = []
    for (seed, n, dt, m, tol, adaptive) in [(3, 60, 0.05, 30, 1e-12, True),
                                            (3, 60, 0.5, 30, 1e-12, True),
                                            (4, 40, 0.1, 6, 1e-12, True),
                                            (5, 50, 0.2, 12, 1e-12, False)]:
        rng = np.random.default_rng(seed)
        A = rng.standard_normal((n, n))
        A = 0.5 * (A + A.T)
        v0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        info = {}
        out = E.lanczos_expm(lambda v: A @ v, v0.astype(np.complex128), -1j * dt, be, m=m,
                             tol=tol, adaptive=adaptive, info=info)
        cases.append({"seed": seed, "n": n, "dt": dt, "m": m, "tol": tol, "adaptive": adaptive,
                      "krylov_dim": int(info["krylov_dim"]), "err_est": float(info["err_est"]).hex(),
                      "beta1": float(info["beta1"]).hex(), "est_ref": float(info["est_ref"]).hex(),
                      "est_ref_j": int(info["est_ref_j"]), "capped": bool(info["capped"]),
                      "out_re": hexs(out.real), "out_im": hexs(out.imag)})
    # an exact breakdown: a start vector inside a 3-dimensional invariant subspace
    lam = np.arange(1.0, 41.0)
    v0 = np.zeros(40, dtype=np.complex128)
    v0[[2, 9, 30]] = [1.0, 2.0, -1.0]
    info = {}
    out = E.lanczos_expm(lambda v: lam * v, v0, -1j * 0.3, be, m=30, tol=1e-12, info=info)
    cases.append({"diag": True, "dt": 0.3, "m": 30, "tol": 1e-12, "adaptive": True,
                  "krylov_dim": int(info["krylov_dim"]), "est_ref": float(info["est_ref"]).hex(),
                  "est_ref_j": int(info["est_ref_j"]), "capped": bool(info["capped"]),
                  "out_re": hexs(out.real), "out_im": hexs(out.imag)})
    with open(os.path.join(HERE, "golden_expm.json"), "w") as f:
        json.dump({"lanczos_expm": cases}, f, indent=1)
    print("wrote golden_expm.json")


if __name__ == "__main__":
    main()
