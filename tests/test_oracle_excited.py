"""The excited-state oracle (oracle/excited.py) against tnkit's recorded dmrg_excited /
dmrg_tower runs and against exact diagonalisation of the same chain."""
import json
import os

import numpy as np
import pytest

from oracle import dense as D
from oracle import excited as X
from oracle import exact as EX
from tnkit_amd.models import tfim_mpo

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "excited_golden.json")) as f:
    GOLD = json.load(f)

W = tfim_mpo(GOLD["model"]["N"], GOLD["model"]["g"])


def _hex(s):
    return float.fromhex(s)


@pytest.fixture(scope="module")
def gs():
    _, M, _ = D.dmrg(W, GOLD["ground_state"]["chi"], n_sweeps=GOLD["ground_state"]["n_sweeps"])
    return M


@pytest.fixture(scope="module")
def spectrum():
    return np.linalg.eigvalsh(EX.mpo_to_dense(W))


def _run_excited(name, gs):
    call = GOLD["runs"][name]["call"]
    kw = dict(call["kw"])
    if kw.pop("init_gs", False):
        kw["init_mps"] = gs
    ex = GOLD["excited"]
    return X.dmrg_excited(W, ex["chi"], prev_states=[gs], n_sweeps=ex["n_sweeps"],
                          seed=ex["seed"], dtype=np.dtype(call["dtype"]), **kw)


@pytest.mark.parametrize("name", ["excited-penalty", "excited-complex", "excited-project",
                                  "excited-lanczos", "excited-init"])
def test_excited_matches_tnkit_runs(name, gs):
    """tnkit's comparison rule for its own golden runs off the recording machine
    (tests/test_excited_symmetric.py:120-153): sweeps and fallbacks exactly, energies within
    1e-12 relative, overlaps and discarded weight within 1e-12 absolute; matvec counts may move
    by a step near a convergence threshold, so they are held to 1%."""
    rec = GOLD["runs"][name]["record"]
    E, M, info = _run_excited(name, gs)
    assert info["sweeps"] == rec["sweeps"]
    assert info["project_fallbacks"] == rec["project_fallbacks"]
    for got, ref in zip(info["energies"], rec["energies"]):
        assert abs(got - _hex(ref)) <= 1e-12 * max(1.0, abs(_hex(ref)))
    assert abs(E - _hex(rec["E"])) <= 1e-12 * abs(_hex(rec["E"]))
    assert abs(info["E_penalised"] - _hex(rec["E_penalised"])) <= 1e-12 * abs(E)
    for got, ref in zip(info["penalty_overlaps"], rec["penalty_overlaps"]):
        assert abs(got - _hex(ref)) <= 1e-12
    assert abs(info["discarded_weight"] - _hex(rec["discarded_weight"])) <= 1e-12
    for got, ref in zip(info["matvecs"], rec["matvecs"]):
        assert abs(got - ref) <= max(1, 0.01 * ref)


def test_excited_is_first_excited_level(gs, spectrum):
    """Closed check: the state found is the chain's first excited eigenstate (ED), orthogonal to
    the ground state, and its energy is <M|H|M> of the returned MPS."""
    E, M, info = _run_excited("excited-penalty", gs)
    assert abs(E - spectrum[1]) < 1e-9
    psi = EX.mps_to_vector(M)
    H = EX.mpo_to_dense(W)
    assert abs(psi @ H @ psi / (psi @ psi) - E) < 1e-10
    assert abs(psi @ EX.mps_to_vector(gs)) < 1e-8


@pytest.mark.parametrize("name", ["tower-penalty", "tower-project"])
def test_tower_matches_tnkit_runs_and_ed(name, spectrum):
    rec = GOLD["runs"][name]["record"]
    tw = GOLD["tower"]
    E, states, infos = X.dmrg_tower(W, tw["chi"], tw["nlevels"], n_sweeps=tw["n_sweeps"],
                                    **GOLD["runs"][name]["call"]["kw"])
    ref = np.array([_hex(e) for e in rec["energies"]])
    assert np.abs(E - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(E - spectrum[:tw["nlevels"]]).max() < 1e-9
    for info, lv in zip(infos, rec["levels"]):
        assert info["attempt"] == lv["attempt"]
        assert info["sweeps"] == lv["sweeps"]
        assert abs(info["penalty"] - _hex(lv["penalty"])) <= 1e-12 * _hex(lv["penalty"])


def test_planted_sign_error_is_caught(gs):
    """A penalty of the wrong sign (an attraction) returns the ground state again: overlap ~1,
    energy the ground energy. The golden comparison above would fail on it."""
    E, _, info = X.dmrg_excited(W, 8, prev_states=[gs], penalty=-1.0, n_sweeps=4, seed=3)
    assert info["penalty_overlaps"][0] > 0.99
    rec = GOLD["runs"]["excited-penalty"]["record"]
    assert abs(E - _hex(rec["E"])) > 0.1


def test_projector_and_lift_algebra():
    """The projector removes exactly the span of the block vectors; the lifted operator equals
    Q H Q on the range of Q and sigma on the deflated span."""
    rng = np.random.default_rng(0)
    n = 30
    Vs = [rng.standard_normal(n) for _ in range(3)]
    Vs.append(Vs[0] + 2 * Vs[1])                     # a dependent vector: rank stays 3
    proj, r = X.projector(Vs)
    assert r == 3
    x = rng.standard_normal(n)
    px = proj(x)
    for v in Vs:
        assert abs(v @ px) < 1e-12 * np.linalg.norm(x) * np.linalg.norm(v)
    assert np.allclose(proj(px), px, atol=1e-13)
    A = rng.standard_normal((n, n))
    H = A + A.T
    mv = X.lifted(lambda z: H @ z, proj, [px], 1.0)
    Q = np.stack([proj(e) for e in np.eye(n)], axis=1)
    sigma = 1.0 + (px @ Q @ H @ Q @ px) / (px @ px)
    Hl = np.stack([mv(e) for e in np.eye(n)], axis=1)
    assert np.allclose(Hl, Q @ H @ Q + sigma * (np.eye(n) - Q), atol=1e-11)
