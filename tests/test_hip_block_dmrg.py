"""Block truncation and the U(1) block DMRG sweep on the device against tnkit's recorded runs and
exact diagonalisation in a charge sector."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import exact as X
from tnkit_amd import block_linalg as BL
from tnkit_amd import block_mps as BM
from tnkit_amd import blocks as B
from tnkit_amd._dmrg import dmrg
from tnkit_amd.block_ops import BlockTensor
from tnkit_amd.models import xxz_mpo

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
GOLD = np.load(os.path.join(HERE, "golden", "blocks.npz"))
RUNS = json.load(open(os.path.join(HERE, "golden", "block_dmrg.json")))


def _leg(k):
    return B.Leg(tuple(GOLD[k + "charges"].tolist()), tuple(GOLD[k + "dims"].tolist()),
                 int(GOLD[k + "sign"]))


@pytest.mark.parametrize("case", ["small", "medium", "odd"])
def test_truncated_svd_matches_tnkit(hip, case):
    """Fused theta, cut across sectors at tnkit's chi: singular values, discarded weight, the new
    bond leg and the factor layouts equal tnkit's; U S Vh reproduces the kept part."""
    a, b, p = _leg(f"{case}/leg_a/"), _leg(f"{case}/leg_b/"), _leg(f"{case}/leg_p/")
    th = BlockTensor(B.canonical((a, p, p, b.flip()), 0),
                     torch.from_numpy(GOLD[f"{case}/theta/data"]).to(hip))
    mat, rec = BL.fuse(th, ((0, 1), (2, 3)))
    chi = int(GOLD[f"{case}/tsvd/chi"])
    U, S, Vh, disc = BL.truncated_svd(mat, chi)
    gS = GOLD[f"{case}/tsvd/S"]
    assert np.abs(S.cpu().numpy() - gS).max() <= 1e-12 * gS.max()
    assert abs(disc - float(GOLD[f"{case}/tsvd/disc"])) <= 1e-12 * float((gS ** 2).sum() + disc)
    assert Vh.legs[0] == _leg(f"{case}/tsvd/bond/")
    assert U.struct is B.canonical(U.legs, 0)
    # the split factors: U S Vh over the kept sectors equals the rank-chi product of theta
    A = BL.split(U, rec, legs=(0,))
    Bv = BL.split(BL.scale(Vh, S, 0), rec, legs=(1,))
    from tnkit_amd.block_ops import einsum
    approx = einsum("asc,ctb->astb", A, Bv).canonical()
    resid = float(torch.linalg.norm(approx.data - th.data) ** 2)
    assert abs(resid - disc) <= 1e-9 * float((gS ** 2).sum() + disc)


@pytest.mark.parametrize("run", RUNS, ids=lambda r: f"N{r['N']}_chi{r['chi']}_q{r['charge']}")
def test_block_dmrg_matches_tnkit(hip, run):
    """tnkit's block DMRG (expand=0), same seed and options: per-sweep energies to 1e-9, the
    final bond sectors exactly, matvec count within 5%."""
    print(f"block dmrg {run['N']} {run['chi']} {run['charge']}", flush=True)
    N = run["N"]
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    E, mps, info = dmrg(mpo, run["chi"], charge=run["charge"], expand=0,
                        n_sweeps=run["n_sweeps"], tol=0.0, rtol=0.0, seed=run["seed"])
    ref = [float.fromhex(h) for h in run["energies"]]
    assert info["blocks"] and len(info["energies"]) == len(ref)
    assert max(abs(e - r) for e, r in zip(info["energies"], ref)) < 1e-9
    assert [[list(t.legs[2].charges), list(t.legs[2].dims)] for t in mps] == run["bonds"]
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.05 * run["matvec_calls"]
    print("matvecs", info["matvec_calls"], "tnkit", run["matvec_calls"])


def _sector_ground(N, q):
    H = X.mpo_to_dense(xxz_mpo(N))
    bits = (np.arange(2 ** N)[:, None] >> np.arange(N - 1, -1, -1)) & 1
    sz2 = (1 - 2 * bits).sum(1)
    sel = sz2 == q
    Hs = H[np.ix_(sel, sel)]
    return float(np.linalg.eigvalsh(0.5 * (Hs + Hs.T))[0])


@pytest.mark.parametrize("q", [0, 2, -4])
def test_block_dmrg_exact_in_sector(hip, q):
    """N = 10 at chi = 32 (exact bond): the lowest energy with 2 S^z = q equals exact
    diagonalisation restricted to that sector; the mixed-precision solver agrees to 1e-9."""
    N = 10
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    e0 = _sector_ground(N, q)
    E, mps, _ = dmrg(mpo, 32, charge=q, expand=0, n_sweeps=4, tol=0.0, rtol=0.0, seed=5)
    assert mps[-1].legs[2].charges == (q,)
    assert abs(E - e0) < 1e-9
    Em, _, info = dmrg(mpo, 32, charge=q, expand=0, n_sweeps=4, tol=0.0, rtol=0.0, seed=5,
                       eig="lanczos-mixed")
    assert abs(Em - e0) < 1e-9 and info["mixed_counts"]["lo"] > 0


@pytest.mark.parametrize("case", ["medium", "odd"])
@pytest.mark.parametrize("iso", ["left", "right"])
def test_truncated_gram_matches_svd(hip, case, iso):
    """The batched Gram truncation keeps the same sectors and singular values as the SVD one,
    gives an exact isometry on its `iso` side, and the same rank-chi product."""
    from tnkit_amd.block_ops import einsum
    a, b, p = _leg(f"{case}/leg_a/"), _leg(f"{case}/leg_b/"), _leg(f"{case}/leg_p/")
    th = BlockTensor(B.canonical((a, p, p, b.flip()), 0),
                     torch.from_numpy(GOLD[f"{case}/theta/data"]).to(hip))
    mat, rec = BL.fuse(th, ((0, 1), (2, 3)))
    chi = int(GOLD[f"{case}/tsvd/chi"])
    U1, S1, V1, d1 = BL.truncated_svd(mat, chi)
    U2, S2, V2, d2 = BL.truncated_gram(mat, chi, iso)
    assert V1.legs[0] == V2.legs[0]
    assert float((S1 - S2).abs().max()) <= 1e-10 * float(S1.max())
    assert abs(d1 - d2) <= 1e-10 * float((S1 ** 2).sum() + d1)
    iso_t = U2 if iso == "left" else V2
    gram = einsum("ab,ac->bc", iso_t, iso_t.conj()) if iso == "left" else \
        einsum("ba,ca->bc", iso_t, iso_t.conj())
    for r, c, G in BL._matrix_blocks(gram.canonical()):
        assert float((G - torch.eye(G.shape[0], dtype=G.dtype, device=hip)).abs().max()) < 1e-12
    prod = lambda U, S, V: einsum("asc,ctb->astb", BL.split(U, rec, legs=(0,)),
                                  BL.split(BL.scale(V, S, 0), rec, legs=(1,))).canonical().data
    assert float((prod(U1, S1, V1) - prod(U2, S2, V2)).abs().max()) \
        <= 1e-9 * float(th.data.abs().max())


@pytest.mark.parametrize("q", [0, 2])
def test_block_dmrg_gram_truncation_exact_in_sector(hip, q):
    N = 10
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    E, _, _ = dmrg(mpo, 32, charge=q, expand=0, n_sweeps=4, tol=0.0, rtol=0.0, seed=6,
                   trunc="gram")
    assert abs(E - _sector_ground(N, q)) < 1e-9
    run = RUNS[1]
    mpo, _ = BM.to_block_mpo(xxz_mpo(run["N"]), [[1, -1]] * run["N"], hip)
    E, mps, info = dmrg(mpo, run["chi"], charge=run["charge"], expand=0,
                        n_sweeps=run["n_sweeps"], tol=0.0, rtol=0.0, seed=run["seed"],
                        trunc="gram")
    assert abs(E - float.fromhex(run["energy"])) < 1e-8


@pytest.mark.parametrize("stop", ["residual", "energy"])
def test_block_dmrg_lanczos_exact_in_sector(hip, stop):
    """eig='lanczos' (unbounded basis, full reorthogonalisation) on the flat block buffer:
    N = 10 at the exact bond reaches the sector's exact ground energy."""
    N = 10
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    E, _, info = dmrg(mpo, 32, charge=2, expand=0, n_sweeps=4, tol=0.0, rtol=0.0, seed=5,
                      eig="lanczos", eig_stop=stop, eig_tol=1e-11 if stop == "energy" else 1e-9)
    assert abs(E - _sector_ground(N, 2)) < 1e-9
