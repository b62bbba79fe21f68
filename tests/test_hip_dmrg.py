"""Two-site DMRG on the HIP backend against tnkit's golden trajectories, the oracle, exact
free fermions and exact diagonalisation."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import exact as X
from tnkit_amd import _dmrg as H
from tnkit_amd import eigensolver as E
from tnkit_amd import kernels
from tnkit_amd.backend import get_backend
from tnkit_amd.models import tfim_mpo, xxz_mpo

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden.json")) as f:
    GOLD = json.load(f)
with open(os.path.join(GOLDEN, "golden_lanczos.json")) as f:
    GOLD_LZ = json.load(f)
with open(os.path.join(GOLDEN, "golden_rsvd.json")) as f:
    GOLD_RS = json.load(f)


def _model(m):
    return tfim_mpo(m["N"], m["g"]) if m["kind"] == "tfim" else xxz_mpo(m["N"], m["J"], m["Jz"])


def test_lanczos_restart_hip_vs_eigh(hip):
    """The device thick restart on an explicit symmetric matrix: eigenpair against eigh, and the
    (iters, restarts) of tnkit's run on the same matrix and start."""
    run = GOLD["lanczos"]
    rng = np.random.default_rng(run["seed"])
    n = run["n"]
    A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    v0 = rng.standard_normal(n)
    Ad = torch.as_tensor(A).to(hip)
    for case in run["runs"]:
        print("lanczos", case, flush=True)
        e, x, info = E.lanczos_ground_restart(lambda v: Ad @ v, torch.as_tensor(v0).to(hip),
                                              tol=case["tol"], max_iter=2000, m=case["m"],
                                              k=case["k"])
        w, V = np.linalg.eigh(A)
        # the Ritz value is within the residual tolerance of the eigenvalue, as tnkit's is
        assert abs(e - w[0]) <= max(case["tol"], 1e-11)
        assert abs(float.fromhex(case["e"]) - w[0]) <= max(case["tol"], 1e-11)
        assert abs(info["iters"] - case["iters"]) <= 2
        assert abs(info["restarts"] - case["restarts"]) <= 1
        assert abs(abs(float(x.cpu().numpy() @ V[:, 0])) - 1) < 1e-6


@pytest.mark.parametrize("run", GOLD["dmrg"], ids=[r["name"] for r in GOLD["dmrg"]])
def test_dmrg_matches_tnkit_trajectory(hip, run):
    """Per-sweep energies within 1e-10 of tnkit's numpy run (same seed, options); matvec counts
    within 2%."""
    print("dmrg", run["name"], flush=True)
    Ehip, M, info = H.dmrg(_model(run["model"]), run["chi"], backend="hip-real", **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    assert len(info["energies"]) == len(ref)
    for a, b in zip(info["energies"], ref):
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b))
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.02 * run["matvec_calls"]
    assert [int(t.shape[2]) for t in M] == run["bonds"]


def test_config1_tfim32_free_fermion(hip):
    """TFIM N=32, g=1, chi=32, 10 sweeps, tol=rtol=0, seed 0: the exact free-fermion energy."""
    E0, M, info = H.dmrg(tfim_mpo(32, 1.0), 32, n_sweeps=10, tol=0.0, rtol=0.0, seed=0,
                         backend="hip-real")
    ff = X.tfim_free_fermion_energy(32, 1.0)
    assert abs(E0 - ff) < 1e-9
    assert info["sweeps"] == 10 and info["backend"] == "hip-real"


@pytest.mark.parametrize("trunc", ["svd", "gram"])
def test_mixed_precision_dmrg_matches_f64(hip, trunc):
    """'lanczos-mixed' (float32 matrix-core matvec, float64 Krylov and refinement) reaches the
    float64 energy to 1e-9 on XXZ N=20, chi=32."""
    mpo = xxz_mpo(20)
    kw = dict(n_sweeps=4, tol=0.0, rtol=0.0, seed=1, backend="hip-real", trunc=trunc)
    e64, _, i64 = H.dmrg(mpo, 32, **kw)
    emx, _, imx = H.dmrg(mpo, 32, eig="lanczos-mixed", **kw)
    assert abs(emx - e64) < 1e-9
    assert imx["mixed_counts"]["lo"] > 0 and imx["mixed_counts"]["hi"] > 0
    eo, _, _ = D.dmrg(mpo, 32, n_sweeps=4, tol=0.0, rtol=0.0, seed=1, trunc=trunc)
    assert abs(e64 - eo) < 1e-10


def test_dmrg_vs_ed(hip):
    mpo = xxz_mpo(10, 1.0, 0.8)
    E0, _, _ = H.dmrg(mpo, 32, n_sweeps=4, tol=0.0, rtol=0.0, seed=0, backend="hip-real")
    assert abs(E0 - X.ground_energy_dense(mpo)) < 1e-9


def test_backend_contract(hip):
    be = get_backend("hip-real")
    assert be.name == "hip-real" and be.dtype == torch.float64
    assert get_backend(be) is be
    assert get_backend("auto", mpo=tfim_mpo(4, 1.0)).name == "hip-real"
    with pytest.raises(ValueError):
        get_backend("cuda")
    with pytest.raises(TypeError):
        be.array(np.array([1.0 + 1.0j]))
    x = be.randn(np.random.default_rng(0), (3, 4))
    np.testing.assert_array_equal(be.to_numpy(x), np.random.default_rng(0).standard_normal((3, 4)))
    v = be.array(np.arange(5.0))
    assert float(be.norm(v)) == pytest.approx(np.sqrt(30.0))
    assert float(be.vdot(v, v)) == pytest.approx(30.0)
    with pytest.raises(ValueError):
        H.dmrg(tfim_mpo(4, 1.0), 4, backend="hip-real", trunc="cutn")


@pytest.mark.parametrize("case", GOLD_LZ["lanczos_ground"]["runs"],
                         ids=lambda c: f"{c['reorth']}-{c['stop']}-{c['max_iter']}")
def test_lanczos_ground_hip_vs_tnkit(hip, case):
    """The device unbounded-basis Lanczos (basis in 64-row buffers; the 66-step run crosses a
    buffer) against tnkit's run on the same matrix and start: Ritz value to 1e-10, iteration
    count within 1 (two Gram-Schmidt passes per buffer instead of tnkit's one-by-one
    subtraction), the converged flag, and the vector against eigh."""
    run = GOLD_LZ["lanczos_ground"]
    rng = np.random.default_rng(run["seed"])
    n = run["n"]
    A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    v0 = rng.standard_normal(n)
    Ad = torch.as_tensor(A).to(hip)
    print("lanczos_ground", case, flush=True)
    e, x, info = E.lanczos_ground(lambda v: Ad @ v, torch.as_tensor(v0).to(hip),
                                  tol=case["tol"], max_iter=case["max_iter"],
                                  reorth=case["reorth"], stop=case["stop"])
    assert abs(e - float.fromhex(case["e"])) < 1e-10
    assert abs(info["iters"] - case["iters"]) <= 1
    assert info["converged"] == case["converged"]
    if case["converged"] and case["stop"] == "residual":
        w, V = np.linalg.eigh(A)
        assert abs(abs(float(x.cpu().numpy() @ V[:, 0])) - 1) < 1e-8


@pytest.mark.parametrize("run", GOLD_LZ["dmrg"], ids=[r["name"] for r in GOLD_LZ["dmrg"]])
def test_dmrg_lanczos_matches_tnkit(hip, run):
    """eig='lanczos' on the HIP backend against tnkit's trajectory: energies to 1e-10, matvec
    counts within 2%, the same bonds."""
    print("dmrg", run["name"], flush=True)
    Ehip, M, info = H.dmrg(_model(run["model"]), run["chi"], backend="hip-real", **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    for a, b in zip(info["energies"], ref):
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b))
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.02 * run["matvec_calls"]
    assert [int(t.shape[2]) for t in M] == run["bonds"]


def test_lanczos_ground_long_basis(hip):
    """200 steps without convergence on a 3000-dimensional diagonal operator with a dense
    spectrum: four basis buffers, the energy against the oracle's run to 1e-10 and the basis
    still orthogonal enough that the Ritz value is above the true minimum."""
    lam = np.sort(np.random.default_rng(5).uniform(-1.0, 1.0, 3000))
    v0 = np.random.default_rng(6).standard_normal(3000)
    ld = torch.as_tensor(lam).to(hip)
    e, x, info = E.lanczos_ground(lambda v: ld * v, torch.as_tensor(v0).to(hip), tol=1e-14,
                                  max_iter=200)
    e_ref, _, info_ref = D.lanczos_ground(lambda v: lam * v, v0.copy(), tol=1e-14, max_iter=200)
    assert info["iters"] == info_ref["iters"] == 200
    assert abs(e - e_ref) < 1e-10 and e >= lam[0] - 1e-12


@pytest.mark.parametrize("r", GOLD_RS["randomized_svd"], ids=lambda r: f"{r['m']}x{r['n']}")
def test_randomized_svd_device_vs_tnkit(hip, r):
    """The device range finder with tnkit's generator: tnkit's singular values and discarded
    weight (1e-12; 1e-6 relative for the float32 range finder), an orthonormal U, and the
    reconstruction error of tnkit's factorisation."""
    from tnkit_amd.linalg import randomized_svd_device
    rng = np.random.default_rng(r["seed"])
    m, n = r["m"], r["n"]
    U0, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    V0, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    A = (U0 * np.exp(-0.3 * np.arange(min(m, n)))) @ V0.T
    U, S, Vh, disc = randomized_svd_device(torch.as_tensor(A).to(hip), r["chi"],
                                           rng=np.random.default_rng(r["seed"] + 10),
                                           mixed=r["mixed"])
    Sref = np.array([float.fromhex(x) for x in r["S"]])
    tol = 1e-6 if r["mixed"] else 1e-12
    assert np.abs(S.cpu().numpy() - Sref).max() < tol * Sref[0]
    assert abs(disc - float.fromhex(r["disc"])) < tol * max(Sref[0] ** 2, 1.0)
    Un = U.cpu().numpy()
    assert np.abs(Un.T @ Un - np.eye(Un.shape[1])).max() < (1e-6 if r["mixed"] else 1e-12)
    rec = np.abs((Un * S.cpu().numpy()) @ Vh.cpu().numpy() - A).max()
    assert abs(rec - float.fromhex(r["recon_err"])) < (1e-4 if r["mixed"] else 1e-6) * rec + 1e-12


@pytest.mark.parametrize("run", GOLD_RS["dmrg"], ids=[r["name"] for r in GOLD_RS["dmrg"]])
def test_dmrg_randomized_matches_tnkit(hip, run):
    """trunc='randomized' / 'randomized-mixed' on the HIP backend against tnkit's trajectory
    with the same seed (the sketches are drawn from the same host generator)."""
    print("dmrg", run["name"], flush=True)
    Ehip, M, info = H.dmrg(_model(run["model"]), run["chi"], backend="hip-real", **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    for a, b in zip(info["energies"], ref):
        assert abs(a - b) <= 1e-9 * max(1.0, abs(b))
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= 0.02 * run["matvec_calls"]
    assert [int(t.shape[2]) for t in M] == run["bonds"]
