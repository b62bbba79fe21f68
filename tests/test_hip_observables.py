"""Device observables (local expectations, entanglement entropies) against the oracle and the
free-fermion block entropy; MPS archives of device and block tensors."""
import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import exact as X
from oracle import observables as O
from tnkit_amd import block_mps as BM
from tnkit_amd import io
from tnkit_amd import observables as H
from tnkit_amd._dmrg import dmrg
from tnkit_amd.models import tfim_mpo_fields, xxz_mpo

pytestmark = pytest.mark.gpu

Z = np.diag([1.0, -1.0])
Y = np.array([[0, -1j], [1j, 0]])


@pytest.fixture(scope="module")
def tfim(hip):
    g = 1.0 + 0.3 * np.random.default_rng(0).standard_normal(12)
    _, M, _ = dmrg(tfim_mpo_fields(g), 32, n_sweeps=6, backend="hip-real")
    return g, M


def test_local_expectations_match_oracle(hip, tfim):
    _, M = tfim
    Mh = [t.cpu().numpy() for t in M]
    X_ = np.array([[0.0, 1.0], [1.0, 0.0]])
    for op in (Z, X_, Y):
        got = H.local_expectations(M, op)
        ref = D.local_expectations(Mh, op)
        assert np.abs(got - ref).max() < 1e-12
    rng = np.random.default_rng(3)
    Mc = D.random_mps(9, 8, 2, rng, np.complex128)          # complex, unnormalised, any gauge
    assert np.abs(H.local_expectations(Mc, Z) - D.local_expectations(Mc, Z)).max() < 1e-12


def test_entropies_match_oracle_and_free_fermions(hip, tfim):
    g, M = tfim
    S = H.all_entropies(M)
    assert np.abs(S - O.all_entropies([t.cpu().numpy() for t in M])).max() < 1e-12
    assert np.abs(S - X.tfim_block_entropies(g)).max() < 1e-9
    canon = H.right_canonical(M)
    for bond in (1, 6, 11):
        assert abs(H.entanglement_entropy(M, bond, canon=canon) - S[bond - 1]) < 1e-13
    S2 = H.two_site_entropy(M)
    assert np.abs(S2 - O.two_site_entropy([t.cpu().numpy() for t in M])).max() < 1e-12
    with pytest.raises(ValueError):
        H.entanglement_entropy(M, 0)


def test_archives_of_device_and_block_states(hip, tmp_path, tfim):
    _, M = tfim
    io.save_mps(tmp_path / "d.npz", M)
    R = io.load_mps(tmp_path / "d.npz", device=hip)
    assert R.center == M.center and all(torch.equal(a, b) for a, b in zip(M, R))
    N = 8
    mpo, _ = BM.to_block_mpo(xxz_mpo(N), [[1, -1]] * N, hip)
    E, mb, _ = dmrg(mpo, 16, charge=0, expand=0, n_sweeps=2, tol=0.0, rtol=0.0, seed=1)
    io.save_mps(tmp_path / "b.npz", mb)
    io.save_mpo(tmp_path / "bw.npz", mpo)
    Rb = io.load_mps(tmp_path / "b.npz", device=hip)
    Wb = io.load_mpo(tmp_path / "bw.npz", device=hip)
    for a, b in zip(mb, Rb):
        assert a.legs == b.legs and a.charge == b.charge
        assert torch.equal(a.canonical().data, b.data)
    # the loaded state and MPO restart the sweep at the same energy
    E2, _, _ = dmrg(Wb, 16, init_mps=list(Rb), expand=0, n_sweeps=1, tol=0.0, rtol=0.0)
    assert abs(E2 - E) < 1e-8
