"""Write the golden fixtures of tests/golden/ by running tnkit itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_golden.py

Nothing else in this repository imports tnkit; the tests read only the fixtures written here
(JSON with hex floats, and .npz arrays loaded with allow_pickle=False). The einsum plan cache is
switched off (TNKIT_EINSUM_PLAN=0) so that every tnkit contraction is exactly
`np.einsum(..., optimize=True)`, and BLAS runs on one thread so the float sequence is
reproducible.
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

from tnkit.api import dmrg  # noqa: E402
from tnkit import _dmrg as D  # noqa: E402
from tnkit import eigensolver as E  # noqa: E402
from tnkit import linalg as TL  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402
from tnkit.models import tfim_mpo  # noqa: E402
from tnkit.free_fermion import tfim_free_fermion_energy  # noqa: E402

SP = np.array([[0, 1], [0, 0]], dtype=float)
SM = SP.T.copy()
SZ = np.diag([0.5, -0.5])
I2 = np.eye(2)


def xxz_mpo(N, J=1.0, Jz=1.0):
    """The D=5 XXZ chain of the benchmark runner (S = sigma / 2), real."""
    W = np.zeros((5, 5, 2, 2))
    W[0, 0], W[0, 1], W[0, 2], W[0, 3] = I2, SP, SM, SZ
    W[1, 4], W[2, 4], W[3, 4], W[4, 4] = 0.5 * J * SM, 0.5 * J * SP, Jz * SZ, I2
    return [W[:1]] + [W.copy() for _ in range(N - 2)] + [W[:, 4:]]


def hexs(xs):
    return [float(x).hex() for x in xs]


def run_dmrg(name, mpo, chi, **kw):
    E0, mps, info = dmrg(mpo, chi=chi, backend="numpy-real", **kw)
    return {"name": name, "chi": chi, "kwargs": kw, "energy": float(E0).hex(),
            "energies": hexs(info["energies"]), "sweeps": int(info["sweeps"]),
            "matvec_calls": int(info["matvec_calls"]),
            "discarded_weight": float(info["discarded_weight"]).hex(),
            "bonds": [int(t.shape[2]) for t in mps]}


def main():
    out = {}
    # --- DMRG trajectories -------------------------------------------------------------
    runs = []
    runs.append(run_dmrg("tfim32_chi32", tfim_mpo(32, 1.0), 32, n_sweeps=10, tol=0.0, rtol=0.0,
                         seed=0))
    runs[-1]["model"] = {"kind": "tfim", "N": 32, "g": 1.0}
    runs[-1]["free_fermion"] = float(tfim_free_fermion_energy(32, 1.0)).hex()
    runs.append(run_dmrg("tfim10_chi8_svd", tfim_mpo(10, 1.2), 8, n_sweeps=4, tol=0.0,
                         rtol=0.0, seed=3))
    runs[-1]["model"] = {"kind": "tfim", "N": 10, "g": 1.2}
    runs.append(run_dmrg("tfim10_chi8_gram", tfim_mpo(10, 1.2), 8, n_sweeps=4, tol=0.0,
                         rtol=0.0, seed=3, trunc="gram"))
    runs[-1]["model"] = {"kind": "tfim", "N": 10, "g": 1.2}
    runs.append(run_dmrg("xxz12_chi16", xxz_mpo(12), 16, n_sweeps=4, tol=0.0, rtol=0.0, seed=1))
    runs[-1]["model"] = {"kind": "xxz", "N": 12, "J": 1.0, "Jz": 1.0}
    runs.append(run_dmrg("xxz16_chi12_restart", xxz_mpo(16), 12, n_sweeps=3, tol=0.0, rtol=0.0,
                         seed=2, restart_m=8, restart_k=3))
    runs[-1]["model"] = {"kind": "xxz", "N": 16, "J": 1.0, "Jz": 1.0}
    out["dmrg"] = runs

    # --- thick-restart Lanczos on a fixed symmetric matrix -----------------------------
    rng = np.random.default_rng(7)
    A = rng.standard_normal((120, 120))
    A = 0.5 * (A + A.T)
    v0 = rng.standard_normal(120)
    be = get_backend("numpy-real")
    lz = []
    for (m, k, tol) in [(40, 10, 1e-10), (12, 4, 1e-10), (6, 2, 1e-8)]:
        e, x, info = E.lanczos_ground_restart(lambda v: A @ v, v0.copy(), be, tol=tol,
                                              max_iter=2000, m=m, k=k)
        lz.append({"m": m, "k": k, "tol": tol, "e": float(e).hex(), "iters": info["iters"],
                   "restarts": info["restarts"], "converged": bool(info["converged"])})
    out["lanczos"] = {"seed": 7, "n": 120, "runs": lz}

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)

    # --- contraction vectors -------------------------------------------------------------
    rng = np.random.default_rng(11)
    a, wl, wm, wr, d1, d2, b = 7, 5, 5, 4, 2, 3, 9
    L = rng.standard_normal((a, wl, a))
    R = rng.standard_normal((b, wr, b))
    W1 = rng.standard_normal((wl, wm, d1, d1))
    W2 = rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    Lm, S, Rm = D.fused_two_site(L, W1, W2, R)
    y2 = D.apply_two_site_fused(th, Lm, S, Rm)
    y2u = D.apply_two_site(th, L, W1, W2, R)
    A3 = rng.standard_normal((a, d1, 6))
    B3 = rng.standard_normal((6, d2, b))
    Wl = rng.standard_normal((wl, wm, d1, d1))
    Wr = rng.standard_normal((wm, wr, d2, d2))
    Lp = D.update_left_env(L, A3, Wl)
    Rp = D.update_right_env(R, B3, Wr)
    M1 = rng.standard_normal((a, d1, b))
    W1s = rng.standard_normal((wl, wr, d1, d1))
    y1 = D.apply_one_site(M1, L, W1s, R)
    # random_mps + right_canonicalize
    mps = D.right_canonicalize(D.random_mps(6, 4, 2, np.random.default_rng(0), be), be)
    # truncations
    X = rng.standard_normal((10, 14))
    U, Sv, Vh = be.svd(X)
    gl = TL.gram_truncation(X, 6, be, "left")
    gr = TL.gram_truncation(X, 6, be, "right")
    arrays = dict(L=L, R=R, W1=W1, W2=W2, theta=th, S=S, y_fused=y2, y_unfused=y2u,
                  A=A3, B=B3, Wl=Wl, Wr=Wr, Lp=Lp, Rp=Rp, M1=M1, W1s=W1s, y_one=y1,
                  X=X, X_S=Sv, gl_U=gl[0], gl_S=gl[1], gl_Vh=gl[2], gl_disc=np.array(gl[3]),
                  gr_U=gr[0], gr_S=gr[1], gr_Vh=gr[2], gr_disc=np.array(gr[3]))
    for i, t in enumerate(mps):
        arrays[f"mps{i}"] = np.asarray(t)
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **arrays)
    print("wrote golden.json and vectors.npz")


if __name__ == "__main__":
    main()
