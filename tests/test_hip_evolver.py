"""The stateful TDVP on the device (tnkit_amd.tdvp_evolver.TDVPEvolver) against tnkit's recorded
runs on the '1site', '2site' and '12site' schedules, and the evolver's state handling."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import evolver as OE
from tnkit_amd.models import tfim_mpo
from tnkit_amd.tdvp_evolver import TDVPEvolver

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_evolver.json")) as f:
    GOLD = json.load(f)

Z = np.diag([1.0, -1.0])


def _gs(r):
    _, gs, _ = D.dmrg(tfim_mpo(r["N"], r["g0"]), 4, n_sweeps=4, tol=0.0, rtol=0.0)
    return gs


@pytest.mark.parametrize("run", GOLD["runs"], ids=lambda r: r["name"])
def test_evolver_matches_tnkit(hip, run):
    print(f"evolver {run['name']}", flush=True)
    steps = []
    ev = TDVPEvolver(_gs(run), tfim_mpo(run["N"], run["g1"]), run["dt"], chi=run["chi"],
                     fast_expm=run["fast_expm"], svd_tol=run["svd_tol"],
                     degenerate_tol=run["degenerate_tol"],
                     measure=lambda m: D.local_expectations(m, Z),
                     step_callback=lambda d, rec: steps.append(d))
    recs = []
    for part in run["split"]:
        recs += ev.run(part, method=run["method"], two_site_every=run["two_site_every"])
    z = np.array([z for _, z in recs])
    zr = np.array([[float.fromhex(v) for v in row] for row in run["z"]])
    assert z.shape == zr.shape and np.abs(z - zr).max() < 1e-9
    assert [float.fromhex(t) for t in run["t"]] == pytest.approx([t for t, _ in recs], rel=1e-14)
    if run["method"] == "12site":
        assert [d["n_two"] for d in steps] == run["n_two"]
    assert [int(t.shape[2]) for t in ev.mps[:-1]] == run["final_bonds"]
    assert ev.center == 0 and ev.step == run["n_steps"]


def test_evolver_state_handling(hip):
    """A run split across calls equals one call; reset takes a new state and rebuilds;
    a replaced tensor forces a rebuild; the options tnkit refuses, refused."""
    r = GOLD["runs"][0]
    gs = _gs(r)
    H = tfim_mpo(r["N"], r["g1"])
    a = TDVPEvolver(gs, H, 0.05, chi=8)
    a.run(4, method="12site")
    b = TDVPEvolver(gs, H, 0.05, chi=8)
    b.run(1, method="12site")
    b.run(3, method="12site")
    assert all(torch.allclose(x, y, atol=1e-12) for x, y in zip(a.mps, b.mps))
    b.reset(gs)
    assert b.center is None
    b.run(2, method="2site")
    ref = OE.Evolver(gs, H, 0.05, 8)
    ref.run(2, "2site")
    za = D.local_expectations([t.cpu().numpy() for t in b.mps], Z)
    zb = D.local_expectations(ref.M, Z)
    assert np.abs(za - zb).max() < 1e-9 and b.step == 6
    with pytest.raises(ValueError):
        TDVPEvolver(gs, H, 0.05, chi=8, expand=2)
    with pytest.raises(ValueError):
        TDVPEvolver(gs, H, 0.05, chi=8, adaptive_dt={"tol": 1e-6})
    with pytest.raises(ValueError):
        TDVPEvolver(gs, H, 0.05).run(1, method="2site")
    with pytest.raises(ValueError):
        TDVPEvolver(gs, H, 0.05, chi=8).run(1, method="2site", two_site_every=2)
    with pytest.raises(TypeError):
        TDVPEvolver(gs, H, 0.05, chi=8, krylov=3)
