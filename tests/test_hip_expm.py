"""The adaptive Krylov exponential on the device (eig twin of TDVP's Krylov step): against
tnkit's recorded runs on fixed matrices, and on U(1) block tensors against scipy's expm of the
dense sector Hamiltonian built by the oracle."""
import json
import os

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import blocks as OB
from oracle import exact as X
from tnkit_amd import block_lanczos as BL
from tnkit_amd import block_ops as BO
from tnkit_amd import eigensolver as E

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "golden_expm.json")) as f:
    GOLD_EX = json.load(f)


def _case(c):
    if c.get("diag"):
        lam = np.arange(1.0, 41.0)
        v0 = np.zeros(40, dtype=np.complex128)
        v0[[2, 9, 30]] = [1.0, 2.0, -1.0]
        return np.diag(lam), v0
    rng = np.random.default_rng(c["seed"])
    A = rng.standard_normal((c["n"], c["n"]))
    A = 0.5 * (A + A.T)
    return A, rng.standard_normal(c["n"]) + 1j * rng.standard_normal(c["n"])


@pytest.mark.parametrize("c", GOLD_EX["lanczos_expm"], ids=lambda c: f"dt{c['dt']}-m{c['m']}")
def test_lanczos_expm_hip_vs_tnkit(hip, c):
    """Same Krylov dimension, cap flag and frozen estimate (1e-8 relative) as tnkit's run, and
    its output vector to 1e-10."""
    print("lanczos_expm", {k: v for k, v in c.items() if k not in ("out_re", "out_im")},
          flush=True)
    A, v0 = _case(c)
    Ad = torch.as_tensor(A).to(hip).to(torch.complex128)
    info = {}
    out = E.lanczos_expm(lambda v: Ad @ v, torch.as_tensor(v0).to(hip), -1j * c["dt"],
                         m=c["m"], tol=c["tol"], adaptive=c["adaptive"], info=info)
    ref = np.array([float.fromhex(x) for x in c["out_re"]]) + \
        1j * np.array([float.fromhex(x) for x in c["out_im"]])
    assert info["krylov_dim"] == c["krylov_dim"] and info["capped"] == c["capped"]
    assert info["est_ref_j"] == c["est_ref_j"]
    er = float.fromhex(c["est_ref"])
    assert abs(info["est_ref"] - er) <= 1e-8 * abs(er) + 1e-14
    assert np.abs(out.cpu().numpy() - ref).max() < 1e-10


def test_lanczos_expm_window_limit(hip):
    with pytest.raises(ValueError):
        E.lanczos_expm(lambda v: v, torch.ones(8, dtype=torch.complex128, device=hip), -0.1j,
                       m=33)


@pytest.mark.parametrize("dt", [0.05, 0.3])
def test_block_lanczos_expm_vs_dense_expm(hip, dt):
    """A U(1) two-site problem whose environments come from a random block MPS (so H_eff is
    symmetric): exp(-i dt H) theta from the block Krylov exponential (real block operator on the
    real and imaginary parts) equals scipy's expm of the oracle's dense sector matrix."""
    import bench
    prob = bench.build_block_problem(chi=16, sigma=1.2, half_width=3, seed=2)
    L, R, W, th = prob["L"], prob["R"], prob["W"], prob["theta"]
    op = BO.BlockTwoSite(L, W, W, R, theta_legs=th.legs)
    T = lambda t: tuple((g.charges, g.dims, g.sign) for g in t.legs)  # noqa: E731
    dn = lambda t: OB.canonical_to_dense(T(t), t.data.cpu().numpy())  # noqa: E731
    H = X.heff_dense(dn(L), dn(W), dn(W), dn(R))
    n = op.size
    pos = OB.canonical_to_dense(T(th), np.arange(1, n + 1, dtype=np.float64)).reshape(-1)
    nz = np.nonzero(pos)[0]
    P = np.empty(n, dtype=np.int64)
    P[pos[nz].astype(np.int64) - 1] = nz
    Hs = H[np.ix_(P, P)]
    assert np.abs(Hs - Hs.T).max() < 1e-10 * np.abs(Hs).max()
    v = th.data.cpu().numpy()
    ref = scipy.linalg.expm(-1j * dt * Hs) @ v
    mv = lambda t: BO.BlockTensor(op.in_struct, op.apply(t.data))  # noqa: E731
    info = {}
    got = E.lanczos_expm(mv, th, -1j * dt, m=30, tol=1e-12, info=info)
    assert got.struct is op.in_struct and got.data.dtype == torch.complex128
    assert np.abs(got.data.cpu().numpy() - ref).max() < 1e-9 * np.abs(ref).max()
    assert info["krylov_dim"] <= 30 and not info["capped"]
    got2 = BL.lanczos_expm(mv, th, -1j * dt, m=30, tol=1e-12)
    assert torch.equal(got2.data, got.data)
