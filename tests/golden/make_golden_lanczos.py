"""Write tests/golden/golden_lanczos.json by running tnkit's unbounded-basis Lanczos
(`eigensolver.lanczos_ground`, eig='lanczos') itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_lanczos.py

The tests read only the JSON written here (hex floats). BLAS runs on one thread and the einsum
plan cache is off, as in make_golden.py.
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
sys.path.insert(0, HERE)

from tnkit.api import dmrg  # noqa: E402
from tnkit import eigensolver as E  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402
from tnkit.models import tfim_mpo  # noqa: E402
from make_golden import run_dmrg, xxz_mpo  # noqa: E402,F401


def main():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((120, 120))
    A = 0.5 * (A + A.T)
    v0 = rng.standard_normal(120)
    be = get_backend("numpy-real")
    runs = []
    for (reorth, stop, tol, max_iter) in [("full", "residual", 1e-10, 200),
                                          ("full", "energy", 1e-12, 200),
                                          ("none", "energy", 1e-10, 200),
                                          ("full", "residual", 1e-10, 25)]:
        e, x, info = E.lanczos_ground(lambda v: A @ v, v0.copy(), be, tol=tol,
                                      max_iter=max_iter, reorth=reorth, stop=stop)
        runs.append({"reorth": reorth, "stop": stop, "tol": tol, "max_iter": max_iter,
                     "e": float(e).hex(), "iters": int(info["iters"]),
                     "converged": bool(info["converged"]),
                     "x_dot_v0": float(abs(x @ v0)).hex()})
    out = {"lanczos_ground": {"seed": 7, "n": 120, "runs": runs}}
    d = run_dmrg("xxz12_chi16_lanczos", xxz_mpo(12), 16, n_sweeps=3, tol=0.0, rtol=0.0, seed=1,
                 eig="lanczos")
    d["model"] = {"kind": "xxz", "N": 12, "J": 1.0, "Jz": 1.0}
    t = run_dmrg("tfim10_chi8_lanczos_energy", tfim_mpo(10, 1.2), 8, n_sweeps=3, tol=0.0,
                 rtol=0.0, seed=3, eig="lanczos", eig_stop="energy")
    t["model"] = {"kind": "tfim", "N": 10, "g": 1.2}
    out["dmrg"] = [d, t]
    with open(os.path.join(HERE, "golden_lanczos.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote golden_lanczos.json")


if __name__ == "__main__":
    main()
