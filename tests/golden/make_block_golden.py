"""Write tests/golden/blocks.npz: tnkit's U(1) block-sparse layouts, contraction plans, gather
indices and one two-site matvec, recorded by running tnkit itself.

Run in the build container only, where the tnkit sources are mounted read-only:

    TNKIT_REF=/root/reference python tests/golden/make_block_golden.py

Every array is numeric (np.load with allow_pickle=False reads it back). For each case, the legs
of theta (a, s1, s2, b), of the environments and of the bulk XXZ MPO (from tnkit's
`to_block_mpo`) are stored, then for each of the four contractions of tnkit's
`_dmrg.apply_two_site` (block_backend.py einsum -> blocks.tensordot):
  * the plan's GEMM list (pa, pb, po, R, K, C), its macs and nmm (blocks.py:1089-1091),
  * both gather indices expanded on the host (blocks.py:450-456), or an empty array when the
    operand needs none,
  * the output layout (blocks, offsets, strides, shapes, size) before and after the einsum's
    leg permutation,
and finally the gather of the result into the canonical layout and the result's data.

tests/golden/block_dmrg.json holds tnkit's block DMRG runs (subspace expansion off) on small U(1)
XXZ chains: per-sweep energies, matvec counts, discarded weight, the final bond legs, and the
bond legs and norms of its random start state.
"""
from __future__ import annotations

import os
import sys

for _v in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS", "MKL_NUM_THREADS"):
    os.environ[_v] = "1"

import numpy as np  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.environ.get("TNKIT_REF", "/root/reference"))

from tnkit import blocks as B  # noqa: E402
from tnkit.backend import get_backend  # noqa: E402
from tnkit.block_backend import BlockBackend, _parse, to_block_mpo  # noqa: E402
from tnkit.symmetry import U1  # noqa: E402

SP = np.array([[0, 1], [0, 0]], dtype=float)
SM = SP.T.copy()
SZ = np.diag([0.5, -0.5])
I2 = np.eye(2)

SUBS = ("pwq,qijb->pwijb", "wmsi,pwijb->pmsjb", "mrtj,pmsjb->pstrb", "drb,pstrb->pstd")


def xxz_mpo(N, J=1.0, Jz=1.0):
    W = np.zeros((5, 5, 2, 2))
    W[0, 0], W[0, 1], W[0, 2], W[0, 3] = I2, SP, SM, SZ
    W[1, 4], W[2, 4], W[3, 4], W[4, 4] = 0.5 * J * SM, 0.5 * J * SP, Jz * SZ, I2
    return [W[:1]] + [W.copy() for _ in range(N - 2)] + [W[:, 4:]]


def struct_arrays(prefix, s, out):
    off, st, shp = s.arrays
    out[prefix + "blocks"] = np.array(s.blocks, dtype=np.int64).reshape(len(s.blocks), len(s.legs))
    out[prefix + "off"] = np.asarray(off, np.int64)
    out[prefix + "st"] = np.asarray(st, np.int64).reshape(len(s.blocks), len(s.legs))
    out[prefix + "shp"] = np.asarray(shp, np.int64).reshape(len(s.blocks), len(s.legs))
    out[prefix + "size"] = np.array(s.size, np.int64)
    out[prefix + "charge"] = np.array(s.charge, np.int64)


def leg_arrays(prefix, leg, out):
    out[prefix + "charges"] = np.array(leg.charges, np.int64)
    out[prefix + "dims"] = np.array(leg.dims, np.int64)
    out[prefix + "sign"] = np.array(leg.sign, np.int64)


def index_array(ix):
    return np.zeros(0, np.int64) if ix is None else np.asarray(ix.host, np.int64)


def case(name, a_leg, b_leg, seed, out):
    be = get_backend("numpy-real")
    bb = BlockBackend(be)
    mpo, _ = to_block_mpo(xxz_mpo(6), [[1, -1]] * 6, be, sym=U1())
    W = mpo[2]                                           # a bulk tensor: legs (wl+, wr-, s+, s'-)
    wl, wr, p = W.legs[0], W.legs[1], W.legs[2]
    rng = np.random.default_rng(seed)
    theta = B.random((a_leg, p, p, b_leg.flip()), 0, be, rng)
    L = B.random((a_leg, wl.flip(), a_leg.flip()), 0, be, rng)
    R = B.random((b_leg.flip(), wr.flip(), b_leg), 0, be, rng)
    k = name + "/"
    for nm, leg in (("a", a_leg), ("b", b_leg), ("wl", wl), ("wr", wr), ("p", p)):
        leg_arrays(k + "leg_" + nm + "/", leg, out)
    for nm, t in (("theta", theta), ("L", L), ("R", R), ("W", W)):
        struct_arrays(k + nm + "/", t.struct, out)
        out[k + nm + "/data"] = np.asarray(t.data)
    ops = (L, W, W, R)
    x = theta
    for c, (sub, A) in enumerate(zip(SUBS, ops)):
        ax_a, ax_b, perm = _parse(sub, 2)
        plan = B._plan(A.struct, x.struct, ax_a, ax_b)
        pre = k + f"c{c + 1}/"
        out[pre + "gemms"] = np.array(plan.gemms, np.int64).reshape(len(plan.gemms), 6)
        out[pre + "macs"] = np.array(plan.macs, np.int64)
        out[pre + "nmm"] = np.array(plan.nmm, np.int64)
        out[pre + "ia"] = index_array(plan.ia)
        out[pre + "ib"] = index_array(plan.ib)
        out[pre + "spec"] = np.array(list(ax_a) + [-1] + list(ax_b) + [-1] + list(perm), np.int64)
        struct_arrays(pre + "out/", plan.out, out)
        y = bb.einsum(sub, A, x)
        struct_arrays(pre + "perm/", y.struct, out)
        x = y
    canon = B.canonical(x.legs, x.charge)
    out[k + "final/ix"] = index_array(B._to_canonical(x.struct))
    struct_arrays(k + "final/", canon, out)
    y = x.canonical()
    out[k + "final/data"] = np.asarray(y.data)
    # the same product from tnkit's dense einsum on the densified operands
    from tnkit import _dmrg as D
    dense = D.apply_two_site(theta.to_dense(), L.to_dense(), W.to_dense(), W.to_dense(),
                             R.to_dense())
    out[k + "final/dense"] = np.asarray(dense)
    # truncation maps: fuse theta to (a s1, s2 b), truncated SVD, split both factors back
    from tnkit.legs import fuse as lfuse
    mat, rec = lfuse(theta, ((0, 1), (2, 3)), be)
    out[k + "fuse/ix"] = index_array(B._fuse_index(theta.struct, rec)[1])
    struct_arrays(k + "fuse/", mat.struct, out)
    for g, fl in enumerate(rec.fused):
        leg_arrays(k + f"fuse/leg{g}/", fl, out)
    chi = max(2, theta.struct.size // 40)
    U, S, Vh, disc = B.truncated_svd(mat, chi)
    out[k + "tsvd/S"] = np.asarray(S)
    out[k + "tsvd/disc"] = np.array(disc)
    out[k + "tsvd/chi"] = np.array(chi)
    leg_arrays(k + "tsvd/bond/", Vh.legs[0], out)
    struct_arrays(k + "tsvd/U/", U.struct, out)
    struct_arrays(k + "tsvd/Vh/", Vh.struct, out)
    su, iu = B._split_index(U.struct, rec, (0,))
    sv, iv = B._split_index(Vh.struct, rec, (1,))
    out[k + "split_u/ix"] = index_array(iu)
    struct_arrays(k + "split_u/", su, out)
    out[k + "split_v/ix"] = index_array(iv)
    struct_arrays(k + "split_v/", sv, out)
    out[k + "scale_u/ix"] = index_array(B._leg_positions(U.struct, 1))
    out[k + "scale_v/ix"] = index_array(B._leg_positions(Vh.struct, 0))
    print(name, "theta", theta.struct.size, "blocks", len(theta.struct.blocks),
          "gemms", [len(out[k + f"c{c + 1}/gemms"]) for c in range(4)])


def gaussian_leg(chi, sigma, parity, sign=1):
    """Sector sizes proportional to exp(-(q/2)^2 / (2 sigma^2)) over charges of one parity,
    rounded, at least 1, summing to chi."""
    qs = np.arange(-40 + parity, 41, 2)
    w = np.exp(-(qs / 2.0) ** 2 / (2 * sigma ** 2))
    dims = np.maximum(1, np.floor(chi * w / w.sum())).astype(int)
    keep = w * chi / w.sum() >= 0.5
    qs, dims = qs[keep], dims[keep]
    while dims.sum() > chi:
        dims[np.argmax(dims)] -= 1
    while dims.sum() < chi:
        dims[np.argmax(w[keep] * chi / w.sum() - dims)] += 1
    return B.Leg(tuple(int(q) for q in qs), tuple(int(d) for d in dims), sign, U1())


def main():
    out = {}
    case("small", B.Leg((-3, -1, 1, 3), (1, 3, 2, 1)), B.Leg((-3, -1, 1, 3), (2, 1, 3, 1)), 5,
         out)
    case("medium", gaussian_leg(48, 1.6, 0), gaussian_leg(40, 1.4, 0), 6, out)
    case("odd", gaussian_leg(30, 1.2, 1), gaussian_leg(36, 1.5, 1), 7, out)
    np.savez_compressed(os.path.join(HERE, "blocks.npz"), **out)
    print("wrote blocks.npz with", len(out), "arrays")


def block_dmrg_runs():
    """tnkit's block DMRG (expand=0) on small U(1) XXZ chains: per-sweep energies, matvec counts,
    discarded weight and the final bond legs; plus the legs of random_block_mps."""
    import json
    from tnkit.api import dmrg
    from tnkit.block_backend import random_block_mps
    be = get_backend("numpy-real")
    runs = []
    for N, chi, charge, seed, sweeps in ((10, 16, 0, 1, 3), (12, 24, 0, 2, 3), (10, 12, 2, 3, 3)):
        mpo, _ = to_block_mpo(xxz_mpo(N), [[1, -1]] * N, be, sym=U1())
        E, mps, info = dmrg(mpo, chi, charge=charge, expand=0, n_sweeps=sweeps, tol=0.0,
                            rtol=0.0, seed=seed, backend="numpy-real")
        start = random_block_mps(mpo, chi, be, charge, np.random.default_rng(seed))
        runs.append({"N": N, "chi": chi, "charge": charge, "seed": seed, "n_sweeps": sweeps,
                     "energy": float(E).hex(),
                     "energies": [float(e).hex() for e in info["energies"]],
                     "matvec_calls": int(info["matvec_calls"]),
                     "discarded_weight": float(info["discarded_weight"]).hex(),
                     "bonds": [[list(t.legs[2].charges), list(t.legs[2].dims)] for t in mps],
                     "start_bonds": [[list(t.legs[2].charges), list(t.legs[2].dims)]
                                     for t in start],
                     "start_norms": [float(np.linalg.norm(t.data)).hex() for t in start]})
        print("block dmrg", N, chi, charge, E, info["matvec_calls"])
    with open(os.path.join(HERE, "block_dmrg.json"), "w") as f:
        json.dump(runs, f, indent=1)


if __name__ == "__main__":
    main()
    block_dmrg_runs()
