"""Write tests/golden/golden_rsvd.json by running tnkit's randomized truncations
(`linalg.randomized_svd_device`, trunc='randomized' and 'randomized-mixed') itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_rsvd.py
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

from tnkit import linalg as TL  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402
from make_golden import hexs, run_dmrg, xxz_mpo  # noqa: E402


def main():
    be = get_backend("numpy-real")
    mats = []
    for (seed, m, n, chi, mixed) in [(1, 60, 50, 12, False), (2, 40, 90, 20, True),
                                     (3, 20, 24, 10, False)]:
        rng = np.random.default_rng(seed)
        # a matrix with a decaying spectrum, as two-site wavefunctions have
        U, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
        V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
        s = np.exp(-0.3 * np.arange(min(m, n)))
        A = (U * s) @ V.T
        Ut, St, Vht, disc = TL.randomized_svd_device(A, chi, be, rng=np.random.default_rng(seed + 10),
                                                     mixed=mixed)
        mats.append({"seed": seed, "m": m, "n": n, "chi": chi, "mixed": mixed,
                     "S": hexs(St), "disc": float(disc).hex(),
                     "recon_err": float(np.abs((Ut * St) @ Vht - A).max()).hex()})
    runs = []
    for trunc in ("randomized", "randomized-mixed"):
        r = run_dmrg(f"xxz16_chi12_{trunc}", xxz_mpo(16), 12, n_sweeps=3, tol=0.0, rtol=0.0,
                     seed=2, trunc=trunc)
        r["model"] = {"kind": "xxz", "N": 16, "J": 1.0, "Jz": 1.0}
        runs.append(r)
    with open(os.path.join(HERE, "golden_rsvd.json"), "w") as f:
        json.dump({"randomized_svd": mats, "dmrg": runs}, f, indent=1)
    print("wrote golden_rsvd.json")


if __name__ == "__main__":
    main()
