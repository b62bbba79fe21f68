"""The stateful-TDVP oracle (oracle/evolver.py) against tnkit's recorded TDVPEvolver runs, exact
evolution at full bond dimension, and the cut rule against hand-worked spectra."""
import json
import os

import numpy as np
import pytest

from oracle import dense as D
from oracle import evolver as OE
from oracle import exact as X
from tnkit_amd.models import tfim_mpo

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "golden_evolver.json")) as f:
    GOLD = json.load(f)

Z = np.diag([1.0, -1.0])


def _run(r):
    _, gs, _ = D.dmrg(tfim_mpo(r["N"], r["g0"]), 4, n_sweeps=4, tol=0.0, rtol=0.0)
    ev = OE.Evolver(gs, tfim_mpo(r["N"], r["g1"]), r["dt"], r["chi"], fast_expm=r["fast_expm"],
                    svd_tol=r["svd_tol"], degenerate_tol=r["degenerate_tol"])
    recs, n_two = [], []
    for part in r["split"]:
        a, b = ev.run(part, r["method"], r["two_site_every"],
                      measure=lambda m: D.local_expectations(m, Z))
        recs += a
        n_two += b
    return ev, recs, n_two


@pytest.mark.parametrize("run", GOLD["runs"], ids=lambda r: r["name"])
def test_evolver_matches_tnkit(run):
    ev, recs, n_two = _run(run)
    z = np.array([z for _, z in recs])
    zr = np.array([[float.fromhex(v) for v in row] for row in run["z"]])
    assert z.shape == zr.shape and np.abs(z - zr).max() < 1e-12
    if run["method"] == "12site":
        assert n_two == run["n_two"]
    assert [int(t.shape[2]) for t in ev.M[:-1]] == run["final_bonds"]


@pytest.mark.parametrize("method", ["2site", "12site"])
def test_evolver_exact_at_full_bond(method):
    """With a cap above every exact bond (2^4 at N = 8) the schedules evolve exactly: <Z_i>(t)
    equals exact evolution of the state vector."""
    N, g0, g1, dt = 8, 1.5, 0.8, 0.05
    _, gs, _ = D.dmrg(tfim_mpo(N, g0), 16, n_sweeps=4, tol=0.0, rtol=0.0)
    psi0 = X.mps_to_vector(gs)
    ev = OE.Evolver(gs, tfim_mpo(N, g1), dt, 16)
    recs, _ = ev.run(10, method, measure=lambda m: D.local_expectations(m, Z))
    for t, z in recs:
        ref = X.site_expectations(X.evolve_dense(tfim_mpo(N, g1), psi0, t), Z, N)
        assert np.abs(z - ref).max() < 1e-9


def test_cut_rule():
    s = np.array([1.0, 0.5, 0.5 * (1 - 1e-9), 0.1, 0.0])
    tot = float(np.sum(s ** 2))
    # the cap at 2 falls inside the degenerate pair (0.5, 0.5(1 - 1e-9)): it moves down to 1
    assert OE.cut(s, 2, degenerate_tol=1e-6)[:2] == (1, True)
    assert OE.cut(s, 2)[:2] == (2, False)
    k, pinned, disc = OE.cut(s, 4, svd_tol=0.011 / tot)
    assert (k, pinned) == (3, False) and abs(disc - 0.01 / tot) < 1e-15
    # a cap among zero values stays put; a group starting at s[0] is split
    assert OE.cut(s, 4, degenerate_tol=0.5)[0] == 4
    assert OE.cut(np.array([1.0, 1.0, 1.0]), 2, degenerate_tol=1e-6)[0] == 2


def test_product_cut_is_the_oracle_rule():
    """The evolver's host-side cut (tnkit_amd.tdvp_evolver.cut) against the oracle on random
    spectra with groups of near-equal values (no device)."""
    from tnkit_amd.tdvp_evolver import cut
    rng = np.random.default_rng(7)
    for _ in range(300):
        n = int(rng.integers(1, 12))
        s = np.sort(rng.random(n))[::-1]
        if n > 3:
            s[2] = s[1] * (1 - 1e-10)
        chi = int(rng.integers(1, 14))
        st = float(rng.choice([0.0, 1e-3, 0.05]))
        dt = float(rng.choice([0.0, 1e-6, 0.2]))
        assert cut(s, chi, st, dt) == OE.cut(s, chi, st, dt)


def test_product_cut_groups_zeros_and_prior_loss():
    """The same with long runs of near-equal values, zero tails, a prior loss disc0 and an
    explicit total."""
    from tnkit_amd.tdvp_evolver import cut
    rng = np.random.default_rng(11)
    for _ in range(500):
        n = int(rng.integers(1, 16))
        s = np.sort(rng.random(n) ** 3)[::-1]
        for _g in range(int(rng.integers(0, 3))):
            a = int(rng.integers(0, n))
            b = int(rng.integers(a, n))
            s[a:b + 1] = s[a] * (1 - 1e-9 * np.arange(b - a + 1))
        if rng.random() < 0.3:
            s[int(rng.integers(0, n)):] = 0.0
        chi = int(rng.integers(1, 18))
        st = float(rng.choice([0.0, 1e-4, 1e-2, 0.3]))
        dt = float(rng.choice([0.0, 1e-7, 1e-3, 0.5]))
        d0 = float(rng.choice([0.0, 1e-5]))
        tot = None if rng.random() < 0.5 else float(np.sum(s ** 2)) * 1.5
        assert cut(s, chi, st, dt, d0, tot) == OE.cut(s, chi, st, dt, d0, tot)
