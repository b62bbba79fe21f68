"""Excited-state DMRG on the HIP device (tnkit_amd.dmrg_excited) against tnkit's recorded
runs, the oracle and exact diagonalisation; the rank-update kernel against numpy."""
import json
import os
import warnings

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import exact as EX
from oracle import excited as OX
from tnkit_amd import _dmrg as H
from tnkit_amd import dmrg_excited as DX
from tnkit_amd import kernels
from tnkit_amd.models import tfim_mpo, xxz_mpo

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "excited_golden.json")) as f:
    GOLD = json.load(f)

W = tfim_mpo(GOLD["model"]["N"], GOLD["model"]["g"])


def _hex(s):
    return float.fromhex(s)


@pytest.mark.parametrize("n,k", [(1, 1), (1000, 3), (4097, 17), (70001, 64)])
def test_rank_update_matches_numpy(hip, n, k):
    rng = np.random.default_rng(n + k)
    V = rng.standard_normal((k, n))
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    ref = y + 0.37 * V.T @ (V @ x)
    Vd, xd, yd = (torch.as_tensor(a, device=hip) for a in (V, x, y))
    kernels.rank_update(Vd, xd, 0.37, yd)
    assert np.abs(yd.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()
    # in place (x is y): the projector form with an orthonormal basis
    Q, _ = np.linalg.qr(V.T)
    Qd = torch.as_tensor(np.ascontiguousarray(Q.T), device=hip)
    z = torch.as_tensor(x, device=hip)
    kernels.rank_update(Qd, z, -1.0, z)
    pz = x - Q @ (Q.T @ x)
    assert np.abs(z.cpu().numpy() - pz).max() <= 1e-12 * np.abs(x).max()


def test_rank_update_refuses_bad_shapes(hip):
    V = torch.zeros((65, 8), dtype=torch.float64, device=hip)
    x = torch.zeros(8, dtype=torch.float64, device=hip)
    with pytest.raises(ValueError):
        kernels.rank_update(V, x, 1.0, x)
    with pytest.raises(ValueError):
        kernels.rank_update(V[:2], x[:7].contiguous(), 1.0, x)


@pytest.fixture(scope="module")
def gs(hip):
    _, M, _ = H.dmrg(W, GOLD["ground_state"]["chi"], n_sweeps=GOLD["ground_state"]["n_sweeps"],
                     backend="hip-real")
    return M


@pytest.mark.parametrize("name", ["excited-penalty", "excited-project", "excited-lanczos",
                                  "excited-init"])
def test_excited_matches_tnkit_runs(hip, gs, name):
    """Sweeps exactly; energies per sweep within 1e-10; the overlap with the ground state at
    rounding level; discarded weight within 1e-13; H_eff applies per sweep within 5%."""
    rec = GOLD["runs"][name]["record"]
    kw = dict(GOLD["runs"][name]["call"]["kw"])
    if kw.pop("init_gs", False):
        kw["init_mps"] = gs
    ex = GOLD["excited"]
    E, M, info = DX.dmrg_excited(W, ex["chi"], prev_states=[gs], n_sweeps=ex["n_sweeps"],
                                 seed=ex["seed"], backend="hip-real", **kw)
    assert info["sweeps"] == rec["sweeps"]
    assert info["project_fallbacks"] == rec["project_fallbacks"]
    for got, ref in zip(info["energies"], rec["energies"]):
        assert abs(got - _hex(ref)) < 1e-10
    assert abs(E - _hex(rec["E"])) < 1e-10
    assert info["penalty_overlaps"][0] < 1e-9
    assert abs(info["discarded_weight"] - _hex(rec["discarded_weight"])) < 1e-13
    for got, ref in zip(info["matvecs_per_sweep"], rec["matvecs"]):
        assert abs(got - ref) <= max(2, 0.05 * ref)
    assert M.center == 0 and all(t.device.type == "cuda" for t in M)


@pytest.mark.parametrize("name", ["tower-penalty", "tower-project"])
def test_tower_matches_tnkit_runs_and_ed(hip, name):
    rec = GOLD["runs"][name]["record"]
    tw = GOLD["tower"]
    E, states, infos = DX.dmrg_tower(W, tw["chi"], tw["nlevels"], n_sweeps=tw["n_sweeps"],
                                     backend="hip-real", **GOLD["runs"][name]["call"]["kw"])
    ref = np.array([_hex(e) for e in rec["energies"]])
    assert np.abs(E - ref).max() < 1e-10
    spec = np.linalg.eigvalsh(EX.mpo_to_dense(W))[:tw["nlevels"]]
    assert np.abs(E - spec).max() < 1e-9
    for info, lv in zip(infos, rec["levels"]):
        assert info["attempt"] == lv["attempt"]
        assert info["sweeps"] == lv["sweeps"]
        assert abs(info["penalty"] - _hex(lv["penalty"])) <= 1e-9 * _hex(lv["penalty"])
    calls = sum((i["matvecs_all_attempts"] for i in infos), [])
    assert len(calls) == len(rec["matvecs"])
    for got, want in zip(calls, rec["matvecs"]):
        assert abs(got - want) <= max(2, 0.05 * want)
    assert all(isinstance(t, np.ndarray) for t in states[0])


def test_excited_xxz_against_oracle_and_ed(hip):
    """A second model (XXZ N=10, D=5, truncated at chi=12): the device run follows the oracle's
    trajectory and lands on the first excited level of ED within the truncation error."""
    mpo = xxz_mpo(10)
    _, g_ref, _ = D.dmrg(mpo, 16, n_sweeps=4, seed=1)
    E_ref, _, inf_ref = OX.dmrg_excited(mpo, 12, prev_states=[g_ref], penalty=3.0, n_sweeps=4,
                                        seed=5)
    E, M, info = DX.dmrg_excited(mpo, 12, prev_states=[g_ref], penalty=3.0, n_sweeps=4, seed=5,
                                 backend="hip-real")
    assert info["sweeps"] == inf_ref["sweeps"]
    # the first excited level is a triplet: the state is one mixture inside it, fixed by the
    # start vector, so the trajectories agree to rounding amplified by the truncation only
    assert np.abs(np.array(info["energies"]) - np.array(inf_ref["energies"])).max() < 1e-8
    assert abs(E - E_ref) < 1e-8
    spec = np.linalg.eigvalsh(EX.mpo_to_dense(mpo))
    assert abs(E - spec[1]) < 5e-6


def test_observables_match_oracle(hip, gs):
    Mh = [t.cpu().numpy() for t in gs]
    rng = np.random.default_rng(4)
    B = D.random_mps(len(Mh), 6, 2, rng)
    assert abs(DX.mps_inner(gs, B) - OX.mps_inner(Mh, B)) < 1e-12
    assert abs(DX.mpo_expectation(B, W) - OX.mpo_expectation(B, W)) < 1e-11


def test_excited_option_errors(hip, gs):
    with pytest.raises(ValueError):
        DX.dmrg_excited(W, 8, prev_states=[gs], eig="lanczos-mixed", backend="hip-real")
    with pytest.raises(ValueError):
        DX.dmrg_excited(W, 8, prev_states=[gs], deflate="lift", backend="hip-real")
    with pytest.raises(ValueError):
        DX.dmrg_excited(W, 8, prev_states=[gs], start_noise=-1.0, backend="hip-real")
    with pytest.raises(TypeError):
        DX.dmrg_excited(W, 8, prev_states=[gs], n_sweps=2)
    with pytest.raises(ValueError):
        DX.dmrg_tower(W, 8, 2, pair_solve=True)


def test_low_penalty_warns(hip, gs):
    """A penalty far below the gap returns the ground state again, with tnkit's warning."""
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        _, _, info = DX.dmrg_excited(W, 8, prev_states=[gs], penalty=1e-3, n_sweeps=3, seed=3,
                                     backend="hip-real")
    assert info["penalty_overlaps"][0] ** 2 > DX.OVERLAP_WARN
    assert any(issubclass(w.category, DX.DeflationWarning) for w in rec)
