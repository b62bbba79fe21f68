"""Write tests/golden/golden_expand.json: tnkit's DMRG with subspace expansion, recorded by running
tnkit itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden_expand.py

Recorded (numbers as float.hex, one BLAS thread):
  * `enriched_split` (tnkit/expansion.py:173-213) on fixed dense inputs, both sides: the new
    isometry, the carried factor and the discarded weight;
  * dense two-site DMRG with `expand` > 0 (tnkit/_dmrg.py:700-705, :780): per-sweep energies,
    the alphas, matvec count, discarded weight and final bonds;
  * U(1) block DMRG with the default `expand=None` (which is min(4, n_sweeps - 2) expansion
    sweeps on block MPOs, tnkit/_dmrg.py:583-584) and with an explicit `expand`: per-sweep
    energies, alphas, matvec count and the final bond legs (charges and sizes).
"""
from __future__ import annotations

import json
import os
import sys

for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.environ.get("TNKIT_REF", "/root/reference"))

from tnkit.api import dmrg  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402
from tnkit.block_backend import to_block_mpo  # noqa: E402
from tnkit.expansion import enriched_split  # noqa: E402
from tnkit.models import tfim_mpo  # noqa: E402
from tnkit.symmetry import U1  # noqa: E402

SP = np.array([[0, 1], [0, 0]], dtype=float)
SM = SP.T.copy()
SZ = np.diag([0.5, -0.5])
I2 = np.eye(2)


def xxz_mpo(N, J=1.0, Jz=1.0):
    """The D = 5 XXZ chain of tnkit's benchmark runner (bench/suite/runners/tnkit_runner.py)."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0], W[0, 1], W[0, 2], W[0, 3] = I2, SP, SM, SZ
    W[1, 4], W[2, 4], W[3, 4], W[4, 4] = 0.5 * J * SM, 0.5 * J * SP, Jz * SZ, I2
    return [W[:1]] + [W.copy() for _ in range(N - 2)] + [W[:, 4:]]


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64).reshape(-1)]


def split_cases():
    be = get_backend("numpy-real")
    rng = np.random.default_rng(11)
    out = []
    for right in (True, False):
        x = rng.standard_normal((3, 2, 2, 4))
        extra = rng.standard_normal((3, 2, 5, 2, 4) if right else (3, 2, 5, 2, 4))
        for alpha, chi in ((0.1, 5), (0.05, 3), (0.0, 4)):
            A, C, disc = enriched_split(x, extra, alpha, chi, be, right)
            out.append({"right": right, "alpha": alpha, "chi": chi, "x": hexs(x),
                        "x_shape": list(x.shape), "extra": hexs(extra),
                        "extra_shape": list(extra.shape), "A": hexs(A), "A_shape": list(A.shape),
                        "C": hexs(C), "C_shape": list(C.shape), "disc": float(disc).hex()})
    return out


def dense_runs():
    out = []
    cases = (("tfim", 10, 0.9, 8, 5, 2, 3e-3, 0), ("xxz", 12, 1.0, 16, 6, 3, 3e-3, 1),
             ("xxz", 10, 1.0, 6, 5, 2, 1e-2, 2))
    for kind, N, g, chi, sweeps, expand, alpha, seed in cases:
        mpo = tfim_mpo(N, g) if kind == "tfim" else xxz_mpo(N)
        E, mps, info = dmrg(mpo, chi, n_sweeps=sweeps, tol=0.0, rtol=0.0, expand=expand,
                            expand_alpha=alpha, seed=seed)
        out.append({"kind": kind, "N": N, "g": g, "chi": chi, "n_sweeps": sweeps,
                    "expand": expand, "expand_alpha": alpha, "seed": seed,
                    "energies": [float(e).hex() for e in info["energies"]],
                    "expand_alphas": [float(a).hex() for a in info["expand_alphas"]],
                    "matvec_calls": int(info["matvec_calls"]),
                    "discarded_weight": float(info["discarded_weight"]).hex(),
                    "bonds": [int(t.shape[2]) for t in mps]})
        print("dense", kind, N, chi, expand, E, info["matvec_calls"])
    return out


def block_runs():
    be = get_backend("numpy-real")
    out = []
    for N, chi, charge, seed, sweeps, expand in ((8, 8, 0, 1, 6, None), (10, 16, 0, 2, 5, None),
                                                 (10, 12, 2, 3, 5, 2)):
        mpo, _ = to_block_mpo(xxz_mpo(N), [[1, -1]] * N, be, sym=U1())
        kw = {} if expand is None else {"expand": expand}
        E, mps, info = dmrg(mpo, chi, charge=charge, n_sweeps=sweeps, tol=0.0, rtol=0.0,
                            seed=seed, backend="numpy-real", **kw)
        out.append({"N": N, "chi": chi, "charge": charge, "seed": seed, "n_sweeps": sweeps,
                    "expand": expand, "info_expand": int(info["expand"]),
                    "energies": [float(e).hex() for e in info["energies"]],
                    "expand_alphas": [float(a).hex() for a in info["expand_alphas"]],
                    "matvec_calls": int(info["matvec_calls"]),
                    "discarded_weight": float(info["discarded_weight"]).hex(),
                    "bonds": [[list(t.legs[2].charges), list(t.legs[2].dims)] for t in mps]})
        print("block", N, chi, charge, expand, E, info["matvec_calls"], info["expand"])
    return out


def main():
    res = {"split": split_cases(), "dense": dense_runs(), "block": block_runs()}
    with open(os.path.join(HERE, "golden_expand.json"), "w") as f:
        json.dump(res, f, indent=1)
    print("wrote golden_expand.json")


if __name__ == "__main__":
    main()
