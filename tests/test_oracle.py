"""The oracle against tnkit's own golden vectors and against mathematics that fixes it.

CPU only. Every oracle function is checked by something other than a retyping of its formula:
tnkit's outputs (tests/golden, written by tests/golden/make_golden.py), exact diagonalisation,
the free-fermion closed form, numpy/scipy linear algebra, or an invariant.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

from oracle import dense as D
from oracle import exact as X
from tnkit_amd.models import tfim_mpo, tfim_mpo_fields, xxz_mpo

from conftest import GOLDEN

with open(os.path.join(GOLDEN, "golden.json")) as f:
    GOLD = json.load(f)
with open(os.path.join(GOLDEN, "golden_lanczos.json")) as f:
    GOLD_LZ = json.load(f)
with open(os.path.join(GOLDEN, "golden_expm.json")) as f:
    GOLD_EX = json.load(f)
with open(os.path.join(GOLDEN, "golden_rsvd.json")) as f:
    GOLD_RS = json.load(f)
VEC = np.load(os.path.join(GOLDEN, "vectors.npz"), allow_pickle=False)


def _model(m):
    if m["kind"] == "tfim":
        return tfim_mpo(m["N"], m["g"])
    return xxz_mpo(m["N"], m["J"], m["Jz"])


# --------------------------------------------------------------------------------------------
# golden vectors from tnkit
# --------------------------------------------------------------------------------------------

@pytest.mark.parametrize("run", GOLD["dmrg"], ids=[r["name"] for r in GOLD["dmrg"]])
def test_dmrg_trajectory_matches_tnkit(run):
    """Per-sweep energies to 1e-12 relative, the exact matvec count, the discarded weight and
    the final bonds of tnkit's numpy-real run (same seed, same options)."""
    E, M, info = D.dmrg(_model(run["model"]), run["chi"], **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    assert len(info["energies"]) == len(ref)
    for a, b in zip(info["energies"], ref):
        assert abs(a - b) <= 1e-12 * max(1.0, abs(b))
    assert info["matvec_calls"] == run["matvec_calls"]
    dw = float.fromhex(run["discarded_weight"])
    assert abs(info["discarded_weight"] - dw) <= 1e-9 * max(dw, 1e-12) + 1e-16
    assert [int(t.shape[2]) for t in M] == run["bonds"]


def test_config1_reaches_free_fermion():
    run = GOLD["dmrg"][0]
    assert run["name"] == "tfim32_chi32"
    ff = X.tfim_free_fermion_energy(32, 1.0)
    assert float.fromhex(run["free_fermion"]) == pytest.approx(ff, abs=1e-12)
    assert abs(float.fromhex(run["energy"]) - ff) < 1e-9


@pytest.mark.parametrize("case", GOLD["lanczos"]["runs"], ids=lambda c: f"m{c['m']}k{c['k']}")
def test_lanczos_restart_matches_tnkit(case):
    """(eigenvalue, iterations, restarts) of tnkit's thick restart on a fixed matrix: pins the
    restart bookkeeping (arrowhead, kept vectors, residual stop)."""
    rng = np.random.default_rng(GOLD["lanczos"]["seed"])
    n = GOLD["lanczos"]["n"]
    A = rng.standard_normal((n, n))
    A = 0.5 * (A + A.T)
    v0 = rng.standard_normal(n)
    e, x, info = D.lanczos_ground_restart(lambda v: A @ v, v0, tol=case["tol"], max_iter=2000,
                                          m=case["m"], k=case["k"])
    assert info["iters"] == case["iters"]
    assert info["restarts"] == case["restarts"]
    assert abs(e - float.fromhex(case["e"])) < 1e-11
    assert abs(e - np.linalg.eigvalsh(A)[0]) < 1e-9


def test_contractions_match_tnkit_vectors():
    v = VEC
    Lm, S, Rm = D.fused_two_site(v["L"], v["W1"], v["W2"], v["R"])
    np.testing.assert_allclose(S, v["S"], rtol=0, atol=1e-13)
    y = D.apply_two_site_fused(v["theta"], Lm, S, Rm)
    np.testing.assert_allclose(y, v["y_fused"], rtol=0, atol=1e-12 * np.abs(v["y_fused"]).max())
    y2 = D.apply_two_site(v["theta"], v["L"], v["W1"], v["W2"], v["R"])
    np.testing.assert_allclose(y2, v["y_unfused"], rtol=0, atol=1e-12 * np.abs(y2).max())
    Lp = D.update_left_env(v["L"], v["A"], v["Wl"])
    np.testing.assert_allclose(Lp, v["Lp"], rtol=0, atol=1e-12 * np.abs(Lp).max())
    Rp = D.update_right_env(v["R"], v["B"], v["Wr"])
    np.testing.assert_allclose(Rp, v["Rp"], rtol=0, atol=1e-12 * np.abs(Rp).max())
    y1 = D.apply_one_site(v["M1"], v["L"], v["W1s"], v["R"])
    np.testing.assert_allclose(y1, v["y_one"], rtol=0, atol=1e-12 * np.abs(y1).max())


def test_random_mps_and_canonical_form_match_tnkit():
    M = D.right_canonicalize(D.random_mps(6, 4, 2, np.random.default_rng(0)))
    for i, t in enumerate(M):
        np.testing.assert_allclose(t, VEC[f"mps{i}"], rtol=0, atol=1e-13)


def test_gram_truncation_matches_tnkit_vectors():
    v = VEC
    for iso, p in (("left", "gl"), ("right", "gr")):
        U, S, Vh, disc = D.truncate_gram(v["X"], 6, iso)
        np.testing.assert_allclose(S, v[f"{p}_S"], rtol=1e-10)
        np.testing.assert_allclose(U @ np.diag(S) @ Vh,
                                   v[f"{p}_U"] @ np.diag(v[f"{p}_S"]) @ v[f"{p}_Vh"], atol=1e-10)
        assert disc == pytest.approx(float(v[f"{p}_disc"]), rel=1e-9)


# --------------------------------------------------------------------------------------------
# mathematics
# --------------------------------------------------------------------------------------------

SHAPES = [(1, 1, 5, 5, 2, 2, 2), (4, 5, 5, 1, 2, 2, 1), (7, 20, 20, 20, 2, 3, 9),
          (3, 3, 3, 3, 2, 2, 5)]


@pytest.mark.parametrize("a,wl,wm,wr,d1,d2,b", SHAPES)
def test_two_site_forms_agree_with_explicit_heff(a, wl, wm, wr, d1, d2, b):
    """Fused, unfused and the explicit (a d1 d2 b)^2 matrix from ONE einsum over the network
    agree (tnkit/tests/test_two_site_fused.py shapes plus one more)."""
    rng = np.random.default_rng(1)
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d1, d1)), rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    H = X.heff_dense(L, W1, W2, R)
    ref = (H @ th.reshape(-1)).reshape(th.shape)
    fused = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    unf = D.apply_two_site(th, L, W1, W2, R)
    scale = np.abs(ref).max()
    assert np.abs(fused - ref).max() <= 1e-12 * scale
    assert np.abs(unf - ref).max() <= 1e-12 * scale


def test_heff_hermitian_for_hermitian_envs():
    """<x, H y> = <H x, y> when L, R are symmetric in (bra, ket) and the MPO is Hermitian."""
    rng = np.random.default_rng(2)
    N, chi = 8, 6
    mpo = xxz_mpo(N)
    M = D.right_canonicalize(D.random_mps(N, chi, 2, rng))
    Rs = [None] * (N + 1)
    Rs[N] = np.ones((1, 1, 1))
    for i in range(N - 1, 1, -1):
        Rs[i] = D.update_right_env(Rs[i + 1], M[i], mpo[i])
    L0 = np.ones((1, 1, 1))
    A = M[0]
    L1 = D.update_left_env(L0, A, mpo[0])
    op = D.fused_two_site(L1, mpo[1], mpo[2], Rs[3])
    shape = (L1.shape[0], 2, 2, Rs[3].shape[0])
    x, y = rng.standard_normal(shape), rng.standard_normal(shape)
    lhs = np.vdot(x, D.apply_two_site_fused(y, *op))
    rhs = np.vdot(D.apply_two_site_fused(x, *op), y)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_environment_chain_gives_mpo_expectation():
    """Contracting L through every site of a normalised MPS gives <psi|H|psi>, which equals the
    expectation from the dense Hamiltonian (N = 8)."""
    rng = np.random.default_rng(3)
    N = 8
    mpo = xxz_mpo(N, 1.0, 0.7)
    M = D.right_canonicalize(D.random_mps(N, 8, 2, rng))
    psi = M[0]
    for t in M[1:]:
        psi = np.tensordot(psi, t, axes=([psi.ndim - 1], [0]))
    psi = psi.reshape(-1)
    H = X.mpo_to_dense(mpo)
    ref = float(psi @ H @ psi) / float(psi @ psi)
    L = np.ones((1, 1, 1))
    for i in range(N):
        L = D.update_left_env(L, M[i], mpo[i])
    Rr = np.ones((1, 1, 1))
    for i in range(N - 1, -1, -1):
        Rr = D.update_right_env(Rr, M[i], mpo[i])
    assert abs(L.reshape(-1)[0] - ref) < 1e-12 * max(1, abs(ref))
    assert abs(Rr.reshape(-1)[0] - ref) < 1e-12 * max(1, abs(ref))


@pytest.mark.parametrize("N", [2, 5, 8])
def test_free_fermion_energy_equals_ed(N):
    """The closed form against exact diagonalisation, uniform and per-site fields."""
    g = 1.0 + 0.3 * np.random.default_rng(N).standard_normal(N)
    assert X.tfim_free_fermion_energy(N, 0.7) == pytest.approx(
        X.ground_energy_dense(tfim_mpo(N, 0.7)), abs=1e-10)
    assert X.tfim_free_fermion_energy(N, g) == pytest.approx(
        X.ground_energy_dense(tfim_mpo_fields(g)), abs=1e-10)


def test_xxz_mpo_is_the_heisenberg_chain():
    """The D = 5 MPO equals sum_i (S+S- + S-S+)/2 + Jz SzSz built from Kronecker products."""
    N, J, Jz = 6, 1.0, 0.6
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])
    Sz = np.diag([0.5, -0.5])
    H = np.zeros((2 ** N, 2 ** N))
    for i in range(N - 1):
        def op(a, b):
            mats = [np.eye(2)] * N
            mats = list(mats)
            mats[i], mats[i + 1] = a, b
            out = mats[0]
            for m in mats[1:]:
                out = np.kron(out, m)
            return out
        H += 0.5 * J * (op(Sp, Sp.T) + op(Sp.T, Sp)) + Jz * op(Sz, Sz)
    np.testing.assert_allclose(X.mpo_to_dense(xxz_mpo(N, J, Jz)), H, atol=1e-14)


def test_lanczos_restart_vs_eigh_and_breakdown():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((200, 200))
    A = A + A.T
    w, V = np.linalg.eigh(A)
    e, x, info = D.lanczos_ground_restart(lambda v: A @ v, rng.standard_normal(200), tol=1e-10,
                                          max_iter=2000, m=20, k=5)
    assert info["converged"] and info["restarts"] > 0
    assert abs(e - w[0]) < 1e-9
    assert abs(abs(x @ V[:, 0]) - 1) < 1e-9
    # an exact eigenvector stops at the first step (breakdown)
    e2, x2, info2 = D.lanczos_ground_restart(lambda v: A @ v, V[:, 3].copy(), tol=1e-10)
    assert info2["iters"] == 1 and abs(e2 - w[3]) < 1e-10


def test_mixed_lanczos_reaches_float64_accuracy():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((300, 300))
    A = A + A.T
    A32 = A.astype(np.float32)
    e, x, info = D.lanczos_ground_mixed(lambda v: (A32 @ v.astype(np.float32)).astype(np.float64),
                                        lambda v: A @ v, rng.standard_normal(300), tol=1e-10,
                                        max_iter=3000, m=30, k=8)
    assert abs(e - np.linalg.eigvalsh(A)[0]) < 1e-9
    assert info["n_lo"] > 0 and info["n_hi"] > 0


def test_expm_fixed_vs_scipy():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((40, 40))
    A = 0.5 * (A + A.T)
    v = rng.standard_normal(40) + 0j
    tau = -0.05j
    got = D.lanczos_expm_fixed(lambda x: A @ x, v, tau, 20)
    ref = scipy.linalg.expm(tau * A) @ v
    assert np.abs(got - ref).max() < 1e-11


@pytest.mark.parametrize("shape", [(12, 20), (20, 12), (16, 16)])
def test_truncations_vs_svd(shape):
    """SVD cut: Eckart-Young (error^2 = discarded weight); Gram cut: exact isometry on the
    requested side, the same discarded weight, the same reconstruction."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal(shape)
    s = np.linalg.svd(A, compute_uv=False)
    chi = 7
    U, S, Vh, disc = D.truncate_svd(A, chi)
    assert disc == pytest.approx(np.sum(s[chi:] ** 2), rel=1e-12)
    assert np.linalg.norm(A - (U * S) @ Vh) ** 2 == pytest.approx(disc, rel=1e-9)
    for iso in ("left", "right"):
        Ug, Sg, Vhg, dg = D.truncate_gram(A, chi, iso)
        assert dg == pytest.approx(disc, rel=1e-8)
        np.testing.assert_allclose((Ug * Sg) @ Vhg, (U * S) @ Vh, atol=1e-10)
        iso_m = Ug.T @ Ug if iso == "left" else Vhg @ Vhg.T
        np.testing.assert_allclose(iso_m, np.eye(chi), atol=1e-13)
    with pytest.raises(ValueError):
        D.truncate_gram(A.astype(np.float32), chi)


def test_dmrg_vs_exact_diagonalisation():
    """A small XXZ chain reaches the ED ground energy (chi large enough to be exact)."""
    mpo = xxz_mpo(10, 1.0, 0.8)
    E, M, info = D.dmrg(mpo, 32, n_sweeps=4, tol=0.0, rtol=0.0, seed=0)
    assert abs(E - X.ground_energy_dense(mpo)) < 1e-9


def test_oracle_plants_are_caught():
    """A dropped MPO term or a transposed environment changes the matvec by O(1): the explicit
    H_eff test above would fail on such an oracle mistake."""
    rng = np.random.default_rng(8)
    a, wl, wm, wr, d1, d2, b = 3, 3, 3, 3, 2, 2, 4
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d1, d1)), rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    ref = (X.heff_dense(L, W1, W2, R) @ th.reshape(-1)).reshape(th.shape)
    bad1 = D.apply_two_site_fused(th, *D.fused_two_site(L.transpose(2, 1, 0), W1, W2, R))
    W1b = W1.copy()
    W1b[0, 0] = 0
    bad2 = D.apply_two_site_fused(th, *D.fused_two_site(L, W1b, W2, R))
    for bad in (bad1, bad2):
        assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()


def test_lq_canonical_start_gives_the_same_trajectory():
    """Two-site DMRG is covariant under the bond gauge: an LQ right-canonical start gives the
    per-sweep energies of the SVD one to rounding (the HIP engine canonicalises by LQ)."""
    run = GOLD["dmrg"][3]
    E1, _, i1 = D.dmrg(_model(run["model"]), run["chi"], **run["kwargs"])
    E2, _, i2 = D.dmrg(_model(run["model"]), run["chi"], canonical="lq", **run["kwargs"])
    for a, b in zip(i1["energies"], i2["energies"]):
        assert abs(a - b) < 1e-11 * max(1, abs(a))
    assert abs(i1["matvec_calls"] - i2["matvec_calls"]) <= 0.02 * i1["matvec_calls"]
    M = D.right_canonicalize_lq(D.random_mps(6, 4, 2, np.random.default_rng(0)))
    for t in M[1:]:
        a, d, b = t.shape
        np.testing.assert_allclose(t.reshape(a, d * b) @ t.reshape(a, d * b).T, np.eye(a),
                                   atol=1e-13)


def test_tdvp2_oracle_is_exact_at_full_bond_dimension():
    """With bonds at 2^(N/2) the two-site projector is the identity, so TDVP2 reproduces the
    exact evolution exp(-iHt)|psi0> (dense matrix exponential), N = 8, quench g 1.5 -> 1.0."""
    N, dt, n = 8, 0.05, 6
    E0, M0, _ = D.dmrg(tfim_mpo(N, 1.5), 16, n_sweeps=4, tol=0.0, rtol=0.0, seed=0)
    mpo = tfim_mpo(N, 1.0)
    Z = np.diag([1.0, -1.0])
    M, rec = D.tdvp2_evolve(M0, mpo, 16, dt, n, krylov_m=20,
                            measure=lambda m: D.local_expectations(m, Z))
    psi0 = X.mps_to_vector(M0)
    for t, z in rec:
        ref = X.site_expectations(X.evolve_dense(mpo, psi0, t), Z, N)
        assert np.abs(z - ref).max() < 1e-9, (t, np.abs(z - ref).max())
    # and the local expectations themselves against the state vector
    np.testing.assert_allclose(D.local_expectations(M0, Z), X.site_expectations(psi0, Z, N),
                               atol=1e-12)


def test_mpo_gradient_vs_finite_differences_and_hellmann_feynman():
    """The MPO gradient of <psi|H|psi> at a DMRG ground state: central differences of the energy
    functional (step 1e-5) and dE0/dg_i of the exact free-fermion energy with per-site fields,
    and -<Z_i> (Hellmann-Feynman), all to 1e-7."""
    N = 12
    g = 1.0 + 0.1 * np.random.default_rng(2).standard_normal(N)
    mpo = tfim_mpo_fields(g)
    E, M, _ = D.dmrg(mpo, 32, n_sweeps=6, tol=0.0, rtol=0.0, seed=0)
    assert abs(E - X.tfim_free_fermion_energy(N, g)) < 1e-9
    grad = D.field_gradient(M, tfim_mpo_fields, g)
    h = 1e-5
    for i in (0, 3, N - 1):
        e = np.zeros(N)
        e[i] = h
        fd = (D.mpo_energy(M, tfim_mpo_fields(g + e)) - D.mpo_energy(M, tfim_mpo_fields(g - e))) / (2 * h)
        ff = (X.tfim_free_fermion_energy(N, g + e) - X.tfim_free_fermion_energy(N, g - e)) / (2 * h)
        assert abs(grad[i] - fd) < 1e-7
        assert abs(grad[i] - ff) < 1e-7
    z = D.local_expectations(M, np.diag([1.0, -1.0]))
    np.testing.assert_allclose(grad, -z, atol=1e-7)


def test_free_fermion_quench_matches_exact_evolution():
    """The Majorana covariance dynamics (TFIM quench g0 -> g1, per-site fields) against dense
    exact evolution of the exact ground state, N = 7."""
    from oracle import exact as X
    from tnkit_amd.models import tfim_mpo_fields
    N = 7
    rng = np.random.default_rng(4)
    g0 = 1.3 + 0.2 * rng.standard_normal(N)
    g1 = np.full(N, 0.9)
    times = [0.0, 0.3, 1.1]
    ff = X.tfim_quench_z(g0, g1, times)
    H0 = X.mpo_to_dense(tfim_mpo_fields(g0))
    w, V = np.linalg.eigh(0.5 * (H0 + H0.conj().T))
    psi0 = V[:, 0]
    mpo1 = tfim_mpo_fields(g1)
    Z = np.diag([1.0, -1.0])
    for t, z in zip(times, ff):
        ref = X.site_expectations(X.evolve_dense(mpo1, psi0, t), Z, N)
        assert np.abs(z - ref).max() < 1e-10
    # the ground state energy from the covariance: E = (i/4) sum A_kl <g_k g_l> = -tr(A Gamma)/4
    A = X.tfim_majorana_matrix(g0)
    G = X.tfim_ground_covariance(g0)
    assert abs(-0.25 * np.trace(A @ G) - w[0]) < 1e-10
    assert abs(X.tfim_free_fermion_energy(N, g0) - w[0]) < 1e-10


def test_product_free_fermion_matches_oracle():
    """tnkit_amd.free_fermion (host reference curves used by the benchmarks) against the oracle's
    covariance dynamics and free-fermion energy, N = 40 with per-site fields."""
    from oracle import exact as X
    from tnkit_amd import free_fermion as F
    rng = np.random.default_rng(9)
    N = 40
    g0 = rng.uniform(1.2, 2.0, N)
    times = [0.0, 0.5, 2.0]
    assert np.abs(F.quench_z(N, g0, 1.0, times) - X.tfim_quench_z(g0, np.ones(N), times)).max() \
        < 1e-10
    assert abs(F.tfim_ground_energy(N, g0) - X.tfim_free_fermion_energy(N, g0)) < 1e-10
    assert abs(F.tfim_ground_energy(32, 1.0) - X.tfim_free_fermion_energy(32, 1.0)) < 1e-12


# --------------------------------------------------------------------------------------------
# unbounded-basis Lanczos (eig='lanczos')
# --------------------------------------------------------------------------------------------

def _lz_matrix():
    run = GOLD_LZ["lanczos_ground"]
    rng = np.random.default_rng(run["seed"])
    A = rng.standard_normal((run["n"], run["n"]))
    return 0.5 * (A + A.T), rng.standard_normal(run["n"])


@pytest.mark.parametrize("case", GOLD_LZ["lanczos_ground"]["runs"],
                         ids=lambda c: f"{c['reorth']}-{c['stop']}-{c['max_iter']}")
def test_lanczos_ground_matches_tnkit(case):
    """tnkit's lanczos_ground on a fixed symmetric matrix: the same iteration count, converged
    flag, Ritz value (1e-12) and overlap with the start vector; a converged Ritz value is also
    within the residual tolerance of eigh's lowest eigenvalue."""
    A, v0 = _lz_matrix()
    e, x, info = D.lanczos_ground(lambda v: A @ v, v0.copy(), tol=case["tol"],
                                  max_iter=case["max_iter"], reorth=case["reorth"],
                                  stop=case["stop"])
    assert info["iters"] == case["iters"] and info["converged"] == case["converged"]
    assert abs(e - float.fromhex(case["e"])) < 1e-12
    assert abs(abs(x @ v0) - float.fromhex(case["x_dot_v0"])) < 1e-10
    if case["converged"] and case["stop"] == "residual":
        assert abs(e - np.linalg.eigvalsh(A)[0]) < 1e-9


def test_lanczos_ground_invariants():
    """On a diagonal operator with known spectrum the full-reorth Lanczos finds the lowest
    eigenvalue and its unit vector; a start vector that is an eigenvector breaks down after one
    step with that eigenvalue."""
    lam = np.linspace(-3.0, 5.0, 300)
    rng = np.random.default_rng(1)
    e, x, info = D.lanczos_ground(lambda v: lam * v, rng.standard_normal(300), tol=1e-10,
                                  max_iter=300)
    assert info["converged"] and abs(e + 3.0) < 1e-9 and abs(abs(x[0]) - 1) < 1e-8
    v = np.zeros(300)
    v[7] = 2.0
    e, x, info = D.lanczos_ground(lambda u: lam * u, v, tol=1e-10, max_iter=50)
    assert info["iters"] == 1 and abs(e - lam[7]) < 1e-15


@pytest.mark.parametrize("run", GOLD_LZ["dmrg"], ids=[r["name"] for r in GOLD_LZ["dmrg"]])
def test_dmrg_lanczos_matches_tnkit(run):
    """eig='lanczos' DMRG (with eig_stop='energy' on the TFIM run): per-sweep energies to
    1e-12, the exact matvec count and the final bonds of tnkit's run."""
    E, M, info = D.dmrg(_model(run["model"]), run["chi"], **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    assert max(abs(a - b) for a, b in zip(info["energies"], ref)) < 1e-12
    assert info["matvec_calls"] == run["matvec_calls"]
    assert [int(t.shape[2]) for t in M] == run["bonds"]


def test_eig_option_rules():
    """The product's option rules (tnkit/_dmrg.py:603-611, tnkit/eigensolver.py:27-58):
    reorth='periodic' raises with eig_stop='residual' and warns with 'energy'; an unknown
    eig_stop raises."""
    from types import SimpleNamespace
    from tnkit_amd.eigensolver import check_eig_options
    with pytest.raises(ValueError):
        check_eig_options(SimpleNamespace(eig="lanczos-restart", reorth="periodic",
                                          eig_stop="residual"))
    with pytest.warns(RuntimeWarning):
        check_eig_options(SimpleNamespace(eig="lanczos-restart", reorth="periodic",
                                          eig_stop="energy"))
    with pytest.raises(ValueError):
        check_eig_options(SimpleNamespace(eig="lanczos", reorth="full", eig_stop="value"))
    check_eig_options(SimpleNamespace(eig="lanczos", reorth="periodic", eig_stop="residual"))


# --------------------------------------------------------------------------------------------
# adaptive Krylov exponential
# --------------------------------------------------------------------------------------------

def expm_case(c):
    """(matvec, start vector, dense matrix) of one recorded lanczos_expm case."""
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
def test_lanczos_expm_matches_tnkit(c):
    """tnkit's adaptive exponential: the same Krylov dimension, cap flag, frozen estimate and
    output vector (1e-13); a converged run is also exp(tau A) v from scipy to 1e-9."""
    A, v0 = expm_case(c)
    info = {}
    out = D.lanczos_expm(lambda v: A @ v, v0, -1j * c["dt"], m=c["m"], tol=c["tol"],
                         adaptive=c["adaptive"], info=info)
    ref = np.array([float.fromhex(x) for x in c["out_re"]]) + \
        1j * np.array([float.fromhex(x) for x in c["out_im"]])
    assert info["krylov_dim"] == c["krylov_dim"] and info["capped"] == c["capped"]
    assert info["est_ref_j"] == c["est_ref_j"]
    assert abs(info["est_ref"] - float.fromhex(c["est_ref"])) <= 1e-12 * abs(
        float.fromhex(c["est_ref"])) + 1e-300
    assert np.abs(out - ref).max() < 1e-13
    if not c["capped"] and c["adaptive"]:
        exact = scipy.linalg.expm(-1j * c["dt"] * A) @ v0
        assert np.abs(out - exact).max() < 1e-9


# --------------------------------------------------------------------------------------------
# randomized range-finder truncation
# --------------------------------------------------------------------------------------------

def rsvd_matrix(r):
    rng = np.random.default_rng(r["seed"])
    m, n = r["m"], r["n"]
    U, _ = np.linalg.qr(rng.standard_normal((m, min(m, n))))
    V, _ = np.linalg.qr(rng.standard_normal((n, min(m, n))))
    return (U * np.exp(-0.3 * np.arange(min(m, n)))) @ V.T


@pytest.mark.parametrize("r", GOLD_RS["randomized_svd"], ids=lambda r: f"{r['m']}x{r['n']}")
def test_randomized_svd_matches_tnkit(r):
    """tnkit's randomized_svd_device with the same generator: singular values, discarded weight
    and reconstruction error to 1e-13; and, against the exact SVD, the kept values are the
    leading ones to the range finder's accuracy."""
    A = rsvd_matrix(r)
    U, S, Vh, disc = D.randomized_svd(A, r["chi"], np.random.default_rng(r["seed"] + 10),
                                      mixed=r["mixed"])
    Sref = np.array([float.fromhex(x) for x in r["S"]])
    assert np.abs(S - Sref).max() < 1e-13
    assert abs(disc - float.fromhex(r["disc"])) < 1e-13
    assert abs(np.abs((U * S) @ Vh - A).max() - float.fromhex(r["recon_err"])) < 1e-13
    Sx = np.linalg.svd(A, compute_uv=False)
    assert np.abs(S - Sx[:r["chi"]]).max() < 1e-4 * Sx[0]
    # the float32 range finder leaves Q orthonormal to float32 rounding only
    assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < (1e-6 if r["mixed"] else 1e-12)


@pytest.mark.parametrize("run", GOLD_RS["dmrg"], ids=[r["name"] for r in GOLD_RS["dmrg"]])
def test_dmrg_randomized_matches_tnkit(run):
    """trunc='randomized' (and the float32 range finder of 'randomized-mixed') DMRG: per-sweep
    energies, matvec counts, bonds and discarded weight of tnkit's run with the same seed."""
    E, M, info = D.dmrg(_model(run["model"]), run["chi"], **run["kwargs"])
    ref = [float.fromhex(h) for h in run["energies"]]
    mixed = run["kwargs"]["trunc"] == "randomized-mixed"
    assert max(abs(a - b) for a, b in zip(info["energies"], ref)) < (1e-10 if mixed else 1e-12)
    assert abs(info["matvec_calls"] - run["matvec_calls"]) <= (2 if mixed else 0)
    assert [int(t.shape[2]) for t in M] == run["bonds"]
    dw = float.fromhex(run["discarded_weight"])
    assert abs(info["discarded_weight"] - dw) < 1e-4 * dw


# --------------------------------------------------------------------------------------------
# one-site TDVP
# --------------------------------------------------------------------------------------------

with open(os.path.join(GOLDEN, "golden_tdvp1.json")) as f:
    GOLD_T1 = json.load(f)


def _tdvp1_problem(n=10, chi=6, seed=2026):
    rng = np.random.default_rng(seed)
    dims = [1] + [min(chi, 2 ** min(i + 1, n - i - 1)) for i in range(n - 1)] + [1]
    return [(rng.normal(size=(dims[i], 2, dims[i + 1]))
             + 1j * rng.normal(size=(dims[i], 2, dims[i + 1]))) / 10.0 for i in range(n)]


@pytest.mark.parametrize("run", GOLD_T1["runs"], ids=[r["name"] for r in GOLD_T1["runs"]])
def test_tdvp1_matches_tnkit(run):
    """tnkit's tdvp1_evolve (adaptive and fixed Krylov dimension 8): <sigma^z_i> per step and
    the final state vector to 1e-13."""
    SZ = np.diag([1.0, -1.0])
    M, rec = D.tdvp1_evolve(_tdvp1_problem(), tfim_mpo(10, 1.0, 0.1), run["dt"],
                            run["n_steps"], fixed_m=run["kwargs"].get("fast_expm"),
                            measure=lambda M: D.local_expectations(M, SZ).real)
    ref_z = np.array([[float.fromhex(x) for x in z] for z in run["z"]])
    assert np.abs(np.array([z for _, z in rec]) - ref_z).max() < 1e-13
    psi = np.array([float.fromhex(x) for x in run["psi_re"]]) + \
        1j * np.array([float.fromhex(x) for x in run["psi_im"]])
    assert np.abs(D.mps_dense(M) - psi).max() < 1e-13
