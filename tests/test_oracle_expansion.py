"""Subspace expansion: the oracle's enriched_split and DMRG with expansion against tnkit's
recorded outputs and against what the mathematics fixes; the HIP drivers' expand rules (host
logic, no device)."""
import json
import os

import numpy as np
import pytest

from oracle import dense as D
from oracle import expansion as X
from tnkit_amd.models import tfim_mpo, xxz_mpo

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "golden_expand.json")) as f:
    GOLD = json.load(f)


def _h(v):
    return np.array([float.fromhex(t) for t in v])


@pytest.mark.parametrize("case", GOLD["split"], ids=lambda c: f"right{c['right']}_a{c['alpha']}_chi{c['chi']}")
def test_enriched_split_matches_tnkit(case):
    x = _h(case["x"]).reshape(case["x_shape"])
    e = _h(case["extra"]).reshape(case["extra_shape"])
    A, C, disc = X.enriched_split(x, e, case["alpha"], case["chi"], case["right"])
    Ar = _h(case["A"]).reshape(case["A_shape"])
    Cr = _h(case["C"]).reshape(case["C_shape"])
    assert A.shape == Ar.shape and C.shape == Cr.shape
    # singular vectors are fixed up to a sign per column by the same LAPACK call
    assert np.abs(A - Ar).max() < 1e-12 and np.abs(C - Cr).max() < 1e-12
    assert abs(disc - float.fromhex(case["disc"])) < 1e-12


@pytest.mark.parametrize("right", [True, False])
def test_enriched_split_closed_forms(right):
    """With room for every state A C = x exactly and A is an isometry; with alpha = 0 the kept
    basis is the leading singular subspace of x (the plain truncated SVD)."""
    rng = np.random.default_rng(3)
    x = rng.standard_normal((4, 2, 2, 6))
    e = rng.standard_normal((4, 2, 3, 2, 6))
    A, C, disc = X.enriched_split(x, e, 0.2, 100, right)
    y = np.einsum("abk,kcd->abcd", A, C) if right else np.einsum("abk,kcd->abcd", C, A)
    assert np.abs(y - x).max() < 1e-12 and disc < 1e-12
    Am = A.reshape(-1, A.shape[-1]) if right else A.reshape(A.shape[0], -1).T
    assert np.abs(Am.T @ Am - np.eye(Am.shape[1])).max() < 1e-12
    A0, _, d0 = X.enriched_split(x, e, 0.0, 3, right)
    Xm = x.reshape(8, 12)
    U, S, Vh = np.linalg.svd(Xm)
    P = U[:, :3] @ U[:, :3].T if right else Vh[:3].T @ Vh[:3]
    A0m = A0.reshape(-1, 3) if right else A0.reshape(3, -1).T
    assert np.abs(A0m @ A0m.T - P).max() < 1e-10
    assert abs(d0 - (S[3:] ** 2).sum()) < 1e-10


@pytest.mark.parametrize("run", GOLD["dense"], ids=lambda r: f"{r['kind']}{r['N']}_exp{r['expand']}")
def test_dense_dmrg_with_expansion_matches_tnkit(run):
    mpo = tfim_mpo(run["N"], run["g"]) if run["kind"] == "tfim" else xxz_mpo(run["N"])
    E, M, info = D.dmrg(mpo, run["chi"], n_sweeps=run["n_sweeps"], tol=0.0, rtol=0.0,
                        expand=run["expand"], expand_alpha=run["expand_alpha"], seed=run["seed"])
    ref = _h(run["energies"])
    assert np.abs(np.array(info["energies"]) - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.allclose(info["expand_alphas"], _h(run["expand_alphas"]), rtol=0, atol=0)
    assert info["matvec_calls"] == run["matvec_calls"]
    # |x|^2 - |C|^2 per expanded step: a difference of two numbers near 1, so absolute
    dw = float.fromhex(run["discarded_weight"])
    assert abs(info["discarded_weight"] - dw) <= 1e-12 + 1e-9 * dw
    assert [int(t.shape[2]) for t in M] == run["bonds"]


def test_expand_rules_and_errors():
    """tnkit's expand rules (host logic of the HIP drivers): None is min(4, n_sweeps - 2) on
    block MPOs and 0 on dense ones; non-integers, a last biased sweep and alpha <= 0 raise."""
    from tnkit_amd._dmrg import DMRGOptions, expand_sweeps
    for n in (1, 2, 3, 4, 5, 6, 8):
        assert expand_sweeps(DMRGOptions(n_sweeps=n), True) == min(4, max(0, n - 2))
        assert expand_sweeps(DMRGOptions(n_sweeps=n), False) == 0
    for bad in (2.5, True, "2"):
        with pytest.raises(TypeError, match="expand must be an int"):
            expand_sweeps(DMRGOptions(expand=bad), False)
    with pytest.raises(ValueError, match="unbiased"):
        expand_sweeps(DMRGOptions(n_sweeps=4, expand=4), False)
    with pytest.raises(ValueError, match="pass expand=0"):
        expand_sweeps(DMRGOptions(expand=2, expand_alpha=0.0), False)
