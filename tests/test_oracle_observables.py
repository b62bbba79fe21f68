"""Entanglement oracle against exact diagonalisation and the free-fermion block entropy; MPS /
MPO archives (host logic, no device)."""
import pickle

import numpy as np
import pytest

from oracle import dense as D
from oracle import exact as X
from oracle import observables as O
from tnkit_amd import io
from tnkit_amd.models import tfim_mpo_fields, xxz_mpo
from tnkit_amd.mps import MPS


@pytest.fixture(scope="module")
def tfim_state():
    g = 1.0 + 0.3 * np.random.default_rng(0).standard_normal(10)
    _, M, _ = D.dmrg(tfim_mpo_fields(g), 32, n_sweeps=6)
    return g, M


def _ed_entropies(psi, N):
    return np.array([O.entropy(np.linalg.svd(psi.reshape(2 ** l, -1), compute_uv=False))
                     for l in range(1, N)])


def test_entropies_match_free_fermions_and_ed(tfim_state):
    g, M = tfim_state
    S = O.all_entropies(M)
    assert np.abs(S - X.tfim_block_entropies(g)).max() < 1e-9
    assert np.abs(S - _ed_entropies(X.mps_to_vector(M), len(M))).max() < 1e-12
    for bond in (1, 4, 9):
        assert abs(O.entanglement_entropy(M, bond) - S[bond - 1]) < 1e-14


def test_two_site_entropy_matches_ed(tfim_state):
    _, M = tfim_state
    psi = X.mps_to_vector(M)
    psi = psi / np.linalg.norm(psi)
    N = len(M)
    ref = []
    for n in range(N - 1):
        t = psi.reshape(2 ** n, 4, -1)
        r = np.einsum("aib,ajb->ij", t, t.conj())
        p = np.linalg.eigvalsh(r)
        p = p[p > 1e-16]
        ref.append(-np.sum(p * np.log(p)))
    assert np.abs(O.two_site_entropy(M) - np.array(ref)).max() < 1e-12


def test_planted_wrong_cut_is_caught(tfim_state):
    """Reading the spectrum of the wrong bond (an index slip) misses the free-fermion value."""
    g, M = tfim_state
    ff = X.tfim_block_entropies(g)
    assert np.abs(np.array([O.entanglement_entropy(M, b + 2) for b in range(7)]) - ff[:7]).max() > 1e-3


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32])
def test_dense_archives_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(1)
    M = MPS([np.asarray(t, dtype=dtype) for t in D.random_mps(6, 4, 2, rng,
                                                                np.complex128 if dtype == np.complex128 else np.float64)], center=0)
    p = tmp_path / "state.npz"
    io.save_mps(p, M)
    R = io.load_mps(p)
    assert R.center == 0 and len(R) == 6
    assert all(a.dtype == b.dtype and np.array_equal(a, b) for a, b in zip(M, R))
    W = xxz_mpo(5)
    io.save_mpo(tmp_path / "h.npz", W)
    W2 = io.load_mpo(tmp_path / "h.npz")
    assert all(np.array_equal(a, b) for a, b in zip(W, W2))
    with pytest.raises(ValueError, match="MPS archive"):
        io.load_mps(tmp_path / "h.npz")
    M.center = None
    io.save_mps(p, M)
    assert io.load_mps(p).center is None


def test_mps_pickles_with_its_centre():
    M = MPS([np.ones((1, 2, 1)), np.ones((1, 2, 1))], center=1)
    R = pickle.loads(pickle.dumps(M))
    assert isinstance(R, MPS) and R.center == 1
