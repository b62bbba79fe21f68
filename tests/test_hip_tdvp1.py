"""One-site TDVP on the HIP device against tnkit's recorded runs and exact evolution, and its
captured site steps: replays equal the eager fixed-schedule steps bit for bit, and capture or
replay failures fall back to those eager steps."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import exact as X
from tnkit_amd import tdvp1 as T1
from tnkit_amd.models import tfim_mpo

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_tdvp1.json")) as f:
    GOLD = json.load(f)
SZ = np.diag([1.0, -1.0])
DT = 0.05


def problem(n=10, chi=6, seed=2026):
    rng = np.random.default_rng(seed)
    dims = [1] + [min(chi, 2 ** min(i + 1, n - i - 1)) for i in range(n - 1)] + [1]
    return [(rng.normal(size=(dims[i], 2, dims[i + 1]))
             + 1j * rng.normal(size=(dims[i], 2, dims[i + 1]))) / 10.0 for i in range(n)]


def zs(M):
    return D.local_expectations([np.asarray(t) for t in M], SZ).real


@pytest.mark.parametrize("run", GOLD["runs"], ids=[r["name"] for r in GOLD["runs"]])
def test_tdvp1_matches_tnkit(hip, run):
    """<sigma^z_i> after every step and the final state vector of tnkit's run (adaptive Krylov
    to 1e-12; fixed dimension 8 with the Taylor exponential on the device) to 1e-9."""
    print("tdvp1", run["name"], flush=True)
    st, rec = T1.tdvp1_evolve(problem(), tfim_mpo(10, 1.0, 0.1), run["dt"], run["n_steps"],
                              measure=zs, **run["kwargs"])
    ref_z = np.array([[float.fromhex(x) for x in z] for z in run["z"]])
    assert np.abs(np.array([z for _, z in rec]) - ref_z).max() < 1e-9
    psi = np.array([float.fromhex(x) for x in run["psi_re"]]) + \
        1j * np.array([float.fromhex(x) for x in run["psi_im"]])
    assert np.abs(D.mps_dense([t.cpu().numpy() for t in st]) - psi).max() < 1e-9
    assert st.center == 0


def test_tdvp1_exact_at_full_bond(hip):
    """With every bond at its full dimension the one-site projection is exact: <sigma^z_i>(t)
    equals exact evolution of the dense state (N = 8, 10 steps)."""
    N = 8
    rng = np.random.default_rng(4)
    dims = [1] + [2 ** min(i + 1, N - i - 1) for i in range(N - 1)] + [1]
    mps = [rng.normal(size=(dims[i], 2, dims[i + 1])) + 1j * rng.normal(
        size=(dims[i], 2, dims[i + 1])) for i in range(N)]
    mpo = tfim_mpo(N, 0.8, 0.2)
    st, rec = T1.tdvp1_evolve(mps, mpo, 0.05, 10, measure=zs, meas_every=5)
    psi0 = X.mps_to_vector(mps)
    psi0 = psi0 / np.linalg.norm(psi0)
    for t, z in rec:
        ref = X.site_expectations(X.evolve_dense(mpo, psi0, t), SZ, N)
        assert np.abs(z - ref).max() < 1e-8


def _same(a, b):
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert torch.equal(x, y), f"site {i} differs"


def test_site_step_graphs_replay_eager_bits(hip):
    """cuda_graph=SiteStepGraphCache(dt, 8): the bulk site steps replay captured graphs (many
    replays per signature, with eager edge steps in between) and the final state equals the
    eager fixed-schedule run bit for bit; graph='on' builds the same cache."""
    mps, mpo = problem(), tfim_mpo(10, 1.0, 0.1)
    ref, _ = T1.tdvp1_evolve(mps, mpo, DT, 4, fast_expm=8)
    gc = T1.SiteStepGraphCache(DT, 8)
    got, _ = T1.tdvp1_evolve(mps, mpo, DT, 4, fast_expm=8, cuda_graph=gc)
    _same(got, ref)
    st = gc.stats
    print("stats", {k: v for k, v in st.items() if k != "signatures"}, flush=True)
    assert st["replayed"] > 0 and st["eager"] > 0 and st["captures"] >= 2
    assert st["capture_failed"] == st["replay_failed"] == 0
    got2, _ = T1.tdvp1_evolve(mps, mpo, DT, 4, fast_expm=8, cuda_graph=gc)   # reuse the cache
    _same(got2, ref)
    got3, _ = T1.tdvp1_evolve(mps, mpo, DT, 4, fast_expm=8, graph="on")
    _same(got3, ref)


class _FailingCache(T1.SiteStepGraphCache):
    def __init__(self, *a, mode="capture", **kw):
        super().__init__(*a, **kw)
        self.mode = mode

    def _make(self, sig, direction):
        if self.mode == "capture":
            raise RuntimeError("capture refused (injected)")
        cache = self

        class Stub:
            pool = None

            def run(self, *args):
                raise RuntimeError("replay refused (injected)")
        del cache
        return Stub()


@pytest.mark.parametrize("mode", ["capture", "replay"])
def test_site_step_graph_failures_fall_back(hip, mode, capsys):
    """Injected capture or replay failures: every step runs eagerly with the eager result, the
    failures are counted and printed, and the breaker stops new captures after the budget."""
    mps, mpo = problem(), tfim_mpo(10, 1.0, 0.1)
    ref, _ = T1.tdvp1_evolve(mps, mpo, DT, 2, fast_expm=8)
    gc = _FailingCache(DT, 8, mode=mode, fault_budget=0)
    got, _ = T1.tdvp1_evolve(mps, mpo, DT, 2, fast_expm=8, cuda_graph=gc)
    _same(got, ref)
    assert gc.stats[f"{mode}_failed"] > 0 and gc.stats["replayed"] == 0
    assert "TNKIT_GRAPH_FALLBACK" in capsys.readouterr().out
    gc2 = _FailingCache(DT, 8, mode="capture", fault_budget=2)
    T1.tdvp1_evolve(mps, mpo, DT, 2, fast_expm=8, cuda_graph=gc2)
    assert gc2.stats["capture_failed"] == 2 and gc2.stats["breaker_tripped"] == 2


def test_tdvp1_option_errors(hip):
    mps, mpo = problem(), tfim_mpo(10, 1.0, 0.1)
    with pytest.raises(ValueError):
        T1.tdvp1_evolve(mps, mpo, DT, 1, cuda_graph=T1.SiteStepGraphCache(DT, 8))
    with pytest.raises(ValueError):
        T1.tdvp1_evolve(mps, mpo, DT, 1, graph="on")
    with pytest.raises(ValueError):
        T1.tdvp1_evolve(mps, mpo, DT, 1, adaptive_dt={"tol": 1e-6})
