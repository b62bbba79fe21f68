"""Subspace expansion on the HIP device: the dense and block enriched_split against tnkit's
recorded outputs and the oracle, and dense / U(1) block DMRG with expansion against tnkit's
recorded runs."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import blocks as OB
from oracle import expansion as OX
from tnkit_amd import block_mps as BM
from tnkit_amd import blocks as B
from tnkit_amd import expansion as X
from tnkit_amd._dmrg import dmrg
from tnkit_amd.block_ops import BlockTensor, einsum, random
from tnkit_amd.models import tfim_mpo, xxz_mpo

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_expand.json")) as f:
    GOLD = json.load(f)


def _h(v):
    return np.array([float.fromhex(t) for t in v])


def _legs(x):
    return tuple((g.charges, g.dims, g.sign) for g in x.legs)


def _dense(x):
    return OB.canonical_to_dense(_legs(x), x.canonical().data.cpu().numpy(), x.charge)


@pytest.mark.parametrize("case", GOLD["split"], ids=lambda c: f"right{c['right']}_a{c['alpha']}_chi{c['chi']}")
def test_dense_enriched_split_matches_tnkit(hip, case):
    """The kept basis is fixed up to a sign per vector, so compare its projector, the product
    A C and the discarded weight."""
    x = _h(case["x"]).reshape(case["x_shape"])
    e = _h(case["extra"]).reshape(case["extra_shape"])
    A, C, disc = X.enriched_split(torch.as_tensor(x, device=hip), torch.as_tensor(e, device=hip),
                                  case["alpha"], case["chi"], None, case["right"])
    A, C = A.cpu().numpy(), C.cpu().numpy()
    Ar = _h(case["A"]).reshape(case["A_shape"])
    Cr = _h(case["C"]).reshape(case["C_shape"])
    assert A.shape == Ar.shape
    if case["right"]:
        Am, Amr = A.reshape(-1, A.shape[-1]), Ar.reshape(-1, Ar.shape[-1])
        y, yr = np.einsum("abk,kcd->abcd", A, C), np.einsum("abk,kcd->abcd", Ar, Cr)
    else:
        Am, Amr = A.reshape(A.shape[0], -1).T, Ar.reshape(Ar.shape[0], -1).T
        y, yr = np.einsum("abk,kcd->abcd", C, A), np.einsum("abk,kcd->abcd", Cr, Ar)
    assert np.abs(Am @ Am.T - Amr @ Amr.T).max() < 1e-10
    assert np.abs(y - yr).max() < 1e-10
    assert abs(disc - float.fromhex(case["disc"])) < 1e-12


@pytest.mark.parametrize("right", [True, False])
def test_block_enriched_split_matches_dense(hip, right):
    """On block tensors (charge-conserving x and expansion term): the same projector, product
    and discarded weight as the dense oracle on the dense images; with room for every state
    A C = x and the basis gains the term's sectors (charges +-3, which x lacks)."""
    rng = np.random.default_rng(1)
    phys = B.Leg((-1, 1), (1, 1))
    a = B.Leg((-2, 0, 2), (1, 2, 1))
    b = B.Leg((0,), (3,), -1)
    w = B.Leg((-2, 0, 2), (1, 1, 1))
    if right:
        x = random((a, phys, phys, b), 0, rng, hip)
        e = random((a, phys, w, phys, b), 0, rng, hip)
    else:
        x = random((b.flip(), phys, phys, a.flip()), 0, rng, hip)
        e = random((b.flip(), phys, w, phys, a.flip()), 0, rng, hip)
    for alpha, chi in ((0.1, 100), (0.3, 3)):
        A, C, disc = X.enriched_split(x, e, alpha, chi, None, right)
        Ad, Cd = _dense(A), _dense(C)
        Ao, Co, do = OX.enriched_split(_dense(x), _dense(e), alpha, chi, right)
        if right:
            Am, Amo = Ad.reshape(-1, Ad.shape[-1]), Ao.reshape(-1, Ao.shape[-1])
            y = np.einsum("abk,kcd->abcd", Ad, Cd)
            yo = np.einsum("abk,kcd->abcd", Ao, Co)
        else:
            Am, Amo = Ad.reshape(Ad.shape[0], -1).T, Ao.reshape(Ao.shape[0], -1).T
            y = np.einsum("abk,kcd->abcd", Cd, Ad)
            yo = np.einsum("abk,kcd->abcd", Co, Ao)
        assert np.abs(Am @ Am.T - Amo @ Amo.T).max() < 1e-10
        assert np.abs(y - yo).max() < 1e-10 and abs(disc - do) < 1e-10
        if chi == 100:
            assert np.abs(y - _dense(x)).max() < 1e-12 and disc < 1e-12
        if chi == 100 and right:
            plain = X.enriched_split(x, e, 0.0, chi, None, right)[0]
            assert set(A.legs[2].charges) > set(plain.legs[2].charges)


def _check_energies(got, ref, expand):
    """Sweeps with expansion end on states that are not yet eigenstates, whose energies move at
    first order with the local eigenvectors (converged to a 1e-9 residual, so the iteration
    where the Krylov solve stops can differ between two machines): 1e-8 there, 1e-9 after."""
    assert len(got) == len(ref)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert abs(g - r) < (1e-8 if k < expand else 1e-9), (k, g, r)


@pytest.mark.parametrize("run", GOLD["dense"], ids=lambda r: f"{r['kind']}{r['N']}_exp{r['expand']}")
def test_dense_dmrg_with_expansion_matches_tnkit(hip, run):
    print(f"dense expand {run['kind']} N={run['N']}", flush=True)
    mpo = tfim_mpo(run["N"], run["g"]) if run["kind"] == "tfim" else xxz_mpo(run["N"])
    E, M, info = dmrg(mpo, run["chi"], n_sweeps=run["n_sweeps"], tol=0.0, rtol=0.0,
                      expand=run["expand"], expand_alpha=run["expand_alpha"], seed=run["seed"],
                      backend="hip-real")
    _check_energies(info["energies"], _h(run["energies"]), run["expand"])
    assert info["expand"] == run["expand"]
    assert info["expand_alphas"] == pytest.approx(list(_h(run["expand_alphas"])), rel=1e-15)
    # The expanded split keeps min(chi, #eigenvalues above 1e-28 of the largest) states of the
    # density, so where its rank is below chi it also keeps rounding-level directions, which
    # differ between LAPACK and rocSOLVER. The expansion sweeps then take another (equally valid)
    # path: their energies agree to 1e-8, the converged sweeps to 1e-9, and the matvec count and
    # discarded weight are path quantities (the oracle matches tnkit's exactly on the host).
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.15 * run["matvec_calls"]
    assert [int(t.shape[2]) for t in M] == run["bonds"]


@pytest.mark.parametrize("run", GOLD["block"], ids=lambda r: f"N{r['N']}_chi{r['chi']}_q{r['charge']}_exp{r['expand']}")
def test_block_dmrg_with_expansion_matches_tnkit(hip, run):
    """The default expand on block MPOs (None: min(4, n_sweeps - 2) sweeps) and an explicit one:
    per-sweep energies, the alphas, the final bond sectors, the matvec count within 5%."""
    print(f"block expand N={run['N']} chi={run['chi']} q={run['charge']}", flush=True)
    N = run["N"]
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    kw = {} if run["expand"] is None else {"expand": run["expand"]}
    E, mps, info = dmrg(mpo, run["chi"], charge=run["charge"], n_sweeps=run["n_sweeps"],
                        tol=0.0, rtol=0.0, seed=run["seed"], **kw)
    assert info["expand"] == run["info_expand"]
    assert info["expand_alphas"] == pytest.approx(list(_h(run["expand_alphas"])), rel=1e-15)
    _check_energies(info["energies"], _h(run["energies"]), run["info_expand"])
    assert [[list(t.legs[2].charges), list(t.legs[2].dims)] for t in mps] == run["bonds"]
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.15 * run["matvec_calls"]
