"""Complex and batched kernels, the device Krylov exponential, TDVP2 and the batched evolution
against the oracle, scipy and exact evolution."""
import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import dense as D
from oracle import exact as X
from tnkit_amd import kernels
from tnkit_amd import tdvp as T
from tnkit_amd.models import tfim_mpo
from tnkit_amd.observables import local_expectations

pytestmark = pytest.mark.gpu

Z = np.diag([1.0, -1.0])


def _c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _t(x, dev, dtype=torch.complex128):
    return torch.as_tensor(np.ascontiguousarray(x)).to(dev, dtype)


def _rel(got, ref):
    got = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else got
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("M,N,K", [(7, 9, 5), (64, 64, 64), (130, 70, 33), (10, 300, 41)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_gemm_tn_complex(hip, M, N, K, dtype):
    print(f"gemm_tn complex {M} {N} {K} {dtype}", flush=True)
    rng = np.random.default_rng(M + N)
    A, B = _c(rng, K, M), _c(rng, K, N)
    got = kernels.gemm_tn(_t(A, hip, dtype), _t(B, hip, dtype))
    tol = 1e-13 if dtype == torch.complex128 else 5e-6
    assert _rel(got, A.T @ B) < tol


@pytest.mark.parametrize("a,wl,wm,wr,d1,d2,b", [(1, 1, 3, 3, 2, 2, 2), (6, 3, 3, 3, 2, 2, 7),
                                                (40, 5, 5, 5, 2, 2, 33), (5, 3, 3, 1, 2, 1, 4)])
def test_complex_matvec_and_envs(hip, a, wl, wm, wr, d1, d2, b):
    print(f"complex two-site {(a, wl, wm, wr, d1, d2, b)}", flush=True)
    rng = np.random.default_rng(a * b)
    L, R = _c(rng, a, wl, a), _c(rng, b, wr, b)
    W1, W2 = _c(rng, wl, wm, d1, d1), _c(rng, wm, wr, d2, d2)
    th = _c(rng, a, d1, d2, b)
    ref = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    op = kernels.fused_two_site(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip))
    assert _rel(kernels.apply_two_site_fused(_t(th, hip), *op), ref) < 1e-12
    op64 = kernels.fused_two_site(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip),
                                  dtype=torch.complex64)
    got = kernels.apply_two_site_fused(_t(th, hip), *op64)
    assert got.dtype == torch.complex128 and _rel(got, ref) < 5e-5
    A = _c(rng, a, d1, 6)
    W = _c(rng, wl, wr, d1, d1)
    assert _rel(kernels.update_left_env(_t(L, hip), _t(A, hip), _t(W, hip)),
                D.update_left_env(L, A, W)) < 1e-12
    Bm = _c(rng, 6, d1, b)
    W = _c(rng, 3, wr, d1, d1)
    assert _rel(kernels.update_right_env(_t(R, hip), _t(Bm, hip), _t(W, hip)),
                D.update_right_env(R, Bm, W)) < 1e-12


def test_batched_kernels_match_per_leg(hip):
    rng = np.random.default_rng(3)
    k, a, w, d, b = 3, 9, 3, 2, 11
    L, R = _c(rng, k, a, w, a), _c(rng, k, b, w, b)
    W1, W2 = _c(rng, k, w, w, d, d), _c(rng, k, w, w, d, d)
    th = _c(rng, k, a, d, d, b)
    op = kernels.fused_two_site_b(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip))
    got = kernels.apply_two_site_b(_t(th, hip), *op).cpu().numpy()
    A = _c(rng, k, a, d, 5)
    gl = kernels.update_left_env_b(_t(L, hip), _t(A, hip), _t(W1, hip)).cpu().numpy()
    Bm = _c(rng, k, 5, d, b)
    gr = kernels.update_right_env_b(_t(R, hip), _t(Bm, hip), _t(W2, hip)).cpu().numpy()
    for j in range(k):
        ref = D.apply_two_site_fused(th[j], *D.fused_two_site(L[j], W1[j], W2[j], R[j]))
        assert _rel(got[j], ref) < 1e-12
        assert _rel(gl[j], D.update_left_env(L[j], A[j], W1[j])) < 1e-12
        assert _rel(gr[j], D.update_right_env(R[j], Bm[j], W2[j])) < 1e-12


@pytest.mark.parametrize("n,m", [(1, 5), (6, 10), (200, 20), (5000, 30)])
def test_expm_batched_vs_scipy(hip, n, m):
    print(f"expm n={n} m={m}", flush=True)
    rng = np.random.default_rng(n)
    k = 3
    H = _c(rng, k, n, n)
    H = 0.5 * (H + np.conj(np.swapaxes(H, 1, 2))) / np.sqrt(n)
    v = _c(rng, k, n)
    Hd = _t(H, hip)
    tau = -0.05j
    got = T.lanczos_expm_fixed_batched(lambda x: torch.bmm(Hd, x.reshape(k, n, 1)).reshape(k, n),
                                       _t(v, hip), tau, m).cpu().numpy()
    for j in range(k):
        ref = scipy.linalg.expm(tau * H[j]) @ v[j]
        assert np.abs(got[j] - ref).max() < 1e-11 * max(1, np.abs(ref).max())
    ref0 = D.lanczos_expm_fixed(lambda x: H[0] @ x, v[0], tau, m)
    assert np.abs(got[0] - ref0).max() < 1e-12 * np.abs(ref0).max()


def _ground(N, g, chi, seed=0):
    E, M, _ = D.dmrg(tfim_mpo(N, g), chi, n_sweeps=4, tol=0.0, rtol=0.0, seed=seed)
    return M


def test_tdvp2_exact_at_full_bond(hip):
    """N = 8 at chi = 16: the HIP TDVP2 equals exact evolution to 1e-9 (quench g 1.5 -> 1)."""
    N, dt, n = 8, 0.05, 6
    M0 = _ground(N, 1.5, 16)
    mpo = tfim_mpo(N, 1.0)
    M, rec = T.tdvp2_evolve(M0, mpo, 16, dt, n, krylov_m=20,
                            measure=lambda m: local_expectations(m, Z))
    psi0 = X.mps_to_vector(M0)
    assert len(rec) == n + 1
    for t, z in rec:
        ref = X.site_expectations(X.evolve_dense(mpo, psi0, t), Z, N)
        assert np.abs(z - ref).max() < 1e-9, (t, np.abs(z - ref).max())


def test_tdvp2_truncated_matches_oracle(hip):
    """N = 10 with the cap chi = 6 < 2^5: HIP (Gram cut) and oracle (SVD cut) trajectories agree."""
    N, dt, n = 10, 0.05, 5
    M0 = _ground(N, 1.3, 6)
    mpo = tfim_mpo(N, 1.0)
    _, rh = T.tdvp2_evolve(M0, mpo, 6, dt, n, krylov_m=20,
                           measure=lambda m: local_expectations(m, Z))
    _, ro = D.tdvp2_evolve(M0, mpo, 6, dt, n, krylov_m=20,
                           measure=lambda m: D.local_expectations(m, Z))
    for (t1, z1), (t2, z2) in zip(rh, ro):
        assert abs(t1 - t2) < 1e-12 and np.abs(z1 - z2).max() < 1e-7


def test_evolve_batch_uneven_caps_match_single_runs(hip):
    """Three quenches (g0 = 1.2, 1.6, 2.0 -> 1.0) with per-leg caps 4, 6, 8 in one batch equal
    three single-trajectory runs (and the oracle), leg by leg."""
    N, dt, n = 8, 0.05, 4
    g0s, caps = [1.2, 1.6, 2.0], [4, 6, 8]
    states = [_ground(N, g, 4, seed=j) for j, g in enumerate(g0s)]
    mpo = tfim_mpo(N, 1.0)
    stats = {}
    legs, rec = T.evolve_batch(states, mpo, dt, n, chi_max=caps, krylov_m=20, keep_device=True,
                               measure=lambda ss: [local_expectations(s, Z) for s in ss],
                               stats=stats)
    assert len(rec) == n + 1
    for j in range(3):
        _, r1 = T.tdvp2_evolve(states[j], mpo, caps[j], dt, n, krylov_m=20,
                               measure=lambda m: local_expectations(m, Z))
        _, ro = D.tdvp2_evolve(states[j], mpo, caps[j], dt, n, krylov_m=20,
                               measure=lambda m: D.local_expectations(m, Z))
        for (t, zs), (_, z1), (_, zo) in zip(rec, r1, ro):
            assert np.abs(zs[j] - z1).max() < 1e-10
            assert np.abs(zs[j] - zo).max() < 1e-7
        assert max(int(t.shape[2]) for t in legs[j][:-1]) <= caps[j]
    assert stats["krylov_m"] >= 20


def test_batch_geometry_errors(hip):
    with pytest.raises(T.BatchGeometryError):
        T.evolve_batch([_ground(6, 1.2, 4), _ground(8, 1.2, 4)], tfim_mpo(6, 1.0), 0.05, 1,
                       chi_max=4)
    with pytest.raises(ValueError):
        T.evolve_batch([_ground(6, 1.2, 4)], tfim_mpo(6, 1.0), 0.05, 1)


def test_evolve_batch_wide_caps_small_chain(hip):
    """Four quenches on N = 16 with caps 64..256, far above the bonds the chain can hold: padded
    legs of low Schmidt rank give Gram matrices with large null spaces (the case where the
    batched Hermitian eigensolver needs its SVD fallback); every leg stays finite and matches the
    exact free-fermion <Z_j>(t)."""
    from tnkit_amd import free_fermion as F
    from tnkit_amd._dmrg import dmrg as hdmrg
    N, legs = 16, 4
    g0 = np.random.default_rng(1).uniform(1.2, 2.0, legs)
    caps = [int(c) for c in np.linspace(64, 256, legs).round()]
    states = []
    for j, g in enumerate(g0):
        _, mps, _ = hdmrg(tfim_mpo(N, float(g)), 16, n_sweeps=4, tol=1e-12, rtol=0.0, seed=j,
                          backend="hip-real")
        states.append([t.to(torch.complex128) for t in mps])
    _, rec = T.evolve_batch(states, tfim_mpo(N, 1.0), 0.05, 4, chi_max=caps, krylov_m=20,
                            keep_device=True, meas_every=2,
                            measure=lambda ss: [local_expectations(s, Z) for s in ss])
    for t, zs in rec:
        for j in range(legs):
            ref = F.quench_z(N, float(g0[j]), 1.0, [t])[0]
            assert np.isfinite(zs[j]).all() and np.abs(zs[j] - ref).max() < 1e-6, (t, j)


def test_graphed_krylov_equals_eager(hip, monkeypatch):
    """The batched evolution with the Krylov loops captured in hipGraphs (each signature eager
    once, then captured and replayed, with eager work between replays) equals the eager run."""
    N, legs = 10, 3
    states = [_ground(N, g, 6, seed=j) for j, g in enumerate((1.2, 1.5, 1.9))]
    mpo = tfim_mpo(N, 1.0)
    kw = dict(chi_max=[6, 8, 10], krylov_m=12, keep_device=True)
    monkeypatch.setenv("TNKIT_AMD_GRAPHS", "0")
    ref, _ = T.evolve_batch(states, mpo, 0.05, 4, **kw)
    monkeypatch.setenv("TNKIT_AMD_GRAPHS", "1")
    T.clear_graphs()
    got, _ = T.evolve_batch(states, mpo, 0.05, 4, **kw)
    assert len(T._GRAPHS) > 0
    for a, b in zip(ref, got):
        for x, y in zip(a, b):
            assert x.shape == y.shape
            assert float((x - y).abs().max()) <= 1e-12 * max(1.0, float(x.abs().max()))
    T.clear_graphs()


def test_ragged_modes_same_trajectory(hip):
    """tnkit's three ragged strategies on legs of different entry bonds (ground states at chi 3,
    5, 5, 8) with caps 6, 10, 12, 16 and a discarded-weight budget: 'pad' (fixed widest shapes),
    'tight' and 'group' give the same <Z_j>(t) to 1e-10, equal single runs, and 'group' makes
    one batch per entry profile with its records per group."""
    N, dt, n = 10, 0.05, 3
    g0s, chis, caps = [1.3, 1.5, 1.7, 1.9], [3, 5, 5, 8], [6, 10, 12, 16]
    states = [_ground(N, g, x, seed=j) for j, (g, x) in enumerate(zip(g0s, chis))]
    mpo = tfim_mpo(N, 1.0)
    meas = lambda ss: [local_expectations(s, Z) for s in ss]  # noqa: E731
    out = {}
    for mode in ("pad", "tight", "group"):
        stats = {}
        print("ragged", mode, flush=True)
        legs, rec = T.evolve_batch(states, mpo, dt, n, chi_max=caps, krylov_m=20,
                                   keep_device=True, measure=meas, stats=stats, ragged=mode,
                                   svd_tol=1e-10)
        out[mode] = (legs, rec, stats)
    _, rec_t, st_t = out["tight"]
    _, rec_p, st_p = out["pad"]
    _, rec_g, st_g = out["group"]
    profiles = [tuple([1] + [int(t.shape[2]) for t in s[:-1]] + [1]) for s in states]
    assert st_g["n_groups"] == len(set(profiles)) and len(rec_g) == st_g["n_groups"]
    assert st_p["n_groups"] == 1 and st_p["fill"] < st_t["fill"]
    groups = T._partition(states, "group")
    for j in range(len(states)):
        gi = next(q for q, g in enumerate(groups) if j in g)
        slot = groups[gi].index(j)
        for k in range(n + 1):
            zt = rec_t[k][1][j]
            assert np.abs(rec_p[k][1][j] - zt).max() < 1e-10
            assert np.abs(rec_g[gi][k][1][slot] - zt).max() < 1e-10
        _, r1 = T.tdvp2_evolve(states[j], mpo, caps[j], dt, n, krylov_m=20, svd_tol=1e-10,
                               measure=lambda m: local_expectations(m, Z))
        assert max(np.abs(rec_t[k][1][j] - r1[k][1]).max() for k in range(n + 1)) < 1e-10
    assert st_t["bonds"] == st_p["bonds"]


def test_ragged_mode_refused():
    with pytest.raises(ValueError, match="ragged"):
        T.evolve_batch([[np.ones((1, 2, 1))]], [np.ones((1, 1, 2, 2))], 0.05, 1, chi_max=4,
                       ragged="fold")


@pytest.mark.parametrize("low", [False, True])
@pytest.mark.parametrize("a,b", [(256, 256), (256, 128), (128, 256), (96, 80)])
def test_ragged_matvec_equals_the_padded_one(hip, a, b, low):
    """The ragged batched matvec (each leg's live bonds given) against the padded one on a batch
    whose legs are zero outside their own bonds: equal on every leg's live block, zero outside it,
    complex128 and complex64 operators. (96, 80) takes the staged GEMMs, which compute in full."""
    from tnkit_amd import kernels
    rng = np.random.default_rng(a + b + low)
    k, D, d = 5, 3, 2
    ra = np.array([a, a // 2, 17, a - 3, 1])
    rb = np.array([b, 33, b // 3, b - 1, 1])
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
    L = np.zeros((k, a, D, a), complex)
    R = np.zeros((k, b, D, b), complex)
    th = np.zeros((k, a, d, d, b), complex)
    for z in range(k):
        L[z, :ra[z], :, :ra[z]] = c(ra[z], D, ra[z])
        R[z, :rb[z], :, :rb[z]] = c(rb[z], D, rb[z])
        th[z, :ra[z], :, :, :rb[z]] = c(ra[z], d, d, rb[z])
    W1, W2 = c(k, D, D, d, d), c(k, D, D, d, d)
    t = lambda x: torch.as_tensor(x).to(hip)  # noqa: E731
    op = kernels.fused_two_site_b(t(L), t(W1), t(W2), t(R),
                                  dtype=torch.complex64 if low else None)
    full = kernels.apply_two_site_b(t(th), *op).cpu().numpy()
    live = torch.as_tensor(np.stack([ra, rb], 1).astype(np.int32), device=hip)
    rag = kernels.apply_two_site_b(t(th), *op, live=live).cpu().numpy()
    for z in range(k):
        blk = np.zeros((a, d, d, b), bool)
        blk[:ra[z], :, :, :rb[z]] = True
        assert np.abs(rag[z][blk] - full[z][blk]).max() <= 1e-14 * max(np.abs(full[z]).max(), 1e-300)
        assert np.abs(rag[z][~blk]).max(initial=0.0) == 0.0
        assert np.abs(full[z][~blk]).max(initial=0.0) == 0.0


def test_ragged_krylov_step_equals_the_padded_one(hip):
    """tnh_expm_step_c_ragged and tnh_lincomb_c_ragged on a batch whose legs are zero outside
    their own bonds give the basis, coefficients and combination of the padded forms, bit for bit
    (the skipped entries are zeros)."""
    from tnkit_amd import kernels
    rng = np.random.default_rng(11)
    k, m, a, d, b = 4, 6, 48, 2, 40
    ra, rb = np.array([48, 20, 7, 33]), np.array([40, 9, 40, 1])
    n = a * d * d * b
    V = np.zeros((k, m, a, d, d, b), complex)
    w = np.zeros((k, a, d, d, b), complex)
    for z in range(k):
        sh = (ra[z], d, d, rb[z])
        V[z, :, :ra[z], :, :, :rb[z]] = rng.standard_normal((m,) + sh) + 1j * rng.standard_normal((m,) + sh)
        w[z, :ra[z], :, :, :rb[z]] = rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    live = torch.as_tensor(np.stack([ra, rb], 1).astype(np.int32), device=hip)
    cf = torch.as_tensor(rng.standard_normal((k, m)) + 1j * rng.standard_normal((k, m))).to(hip)
    outs = []
    for rag in (None, (live, d * d * b, b)):
        Vt = torch.as_tensor(V.reshape(k, m, n)).to(hip)
        Vt[:, 0] /= torch.linalg.vector_norm(Vt[:, 0], dim=1, keepdim=True)
        ab = torch.zeros((k, 2 * m), dtype=torch.float64, device=hip)
        ws = kernels.expm_ws(k, hip)
        for j in range(3):
            wt = torch.as_tensor(w.reshape(k, n) * (j + 1)).to(hip)
            kernels.expm_step_c(Vt, wt, ab, j, ws, ragged=rag)
        out = torch.full((k, n), 7.0 + 0j, dtype=torch.complex128, device=hip)
        kernels.lincomb_c(Vt, cf, out, 4, ragged=rag)
        outs.append((Vt[:, :4].cpu().numpy(), ab.cpu().numpy(), out.cpu().numpy()))
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][2], outs[1][2])
