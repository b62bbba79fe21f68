"""The batched small Hermitian eigensolver (`tnh_eigh_jacobi`, one workgroup per matrix, Jacobi in
LDS) against numpy's eigh: eigenvalues to 1e-13 of the spectral radius, A Z = Z diag(w) and
Z^H Z = 1 to 1e-12, on Gram matrices of the shapes the truncation meets (low rank, zero rows
with the negative diagonal shifts of the padded legs, degenerate spectra, eigenvalues spread
over 24 decades), complex128 and float64. The small thin SVD (`tnh_svd_jacobi`, one-sided
Jacobi) against numpy's SVD on random, rank-deficient, zero, spread and degenerate matrices of both
orientations."""
import numpy as np
import pytest
import torch

from tnkit_amd import kernels, tdvp

pytestmark = pytest.mark.gpu


def _cases(n, cplx, rng):
    def rnd(*s):
        x = rng.standard_normal(s)
        return x + 1j * rng.standard_normal(s) if cplx else x

    out = []
    X = rnd(n, 2 * n)
    out.append(("psd", X @ X.conj().T, None))
    r = max(1, n // 5)
    X = rnd(n, r)
    G = X @ X.conj().T
    add = np.zeros(n)
    z = max(0, n - 2 * r)
    if z:
        G[n - z:, :] = 0
        G[:, n - z:] = 0
        add[n - z:] = -(1.0 + np.arange(n - z, n) / n)
    out.append(("lowrank+pad", G, add))
    Q, _ = np.linalg.qr(rnd(n, n))
    vals = np.repeat(np.arange(1.0, 5.0), (n + 3) // 4)[:n]
    out.append(("degenerate", (Q * vals) @ Q.conj().T, None))
    out.append(("spread", (Q * np.logspace(0, -24, n)) @ Q.conj().T, None))
    out.append(("zero+shift", np.zeros((n, n)), -(1.0 + np.arange(n) / n)))
    out.append(("diagonal", np.diag(rng.standard_normal(n)), None))
    return [(name, 0.5 * (G + G.conj().T), add) for name, G, add in out]


@pytest.mark.parametrize("n", [1, 2, 3, 8, 17, 50, 63, 64])
@pytest.mark.parametrize("cplx", [True, False])
def test_eigh_jacobi_against_numpy(hip, n, cplx):
    print(f"eigh_jacobi n={n} complex={cplx}", flush=True)
    rng = np.random.default_rng(100 * n + cplx)
    cases = _cases(n, cplx, rng)
    dt = torch.complex128 if cplx else torch.float64
    G = torch.as_tensor(np.stack([c[1] for c in cases])).to(hip, dt)
    add = torch.as_tensor(np.stack([c[2] if c[2] is not None else np.zeros(n) for c in cases]),
                          dtype=torch.float64, device=hip)
    w, Z, sw = kernels.eigh_jacobi(G, add)
    w, Z, sw = w.cpu().numpy(), Z.cpu().numpy(), sw.cpu().numpy()
    for b, (name, A, a) in enumerate(cases):
        A = A + np.diag(a if a is not None else np.zeros(n))
        ref = np.linalg.eigvalsh(A)
        scale = max(np.abs(ref).max(), 1e-300)
        assert sw[b] > 0, (name, sw[b])
        assert np.all(np.diff(w[b]) >= 0), name
        assert np.abs(w[b] - ref).max() <= 1e-13 * scale, name
        assert np.abs(A @ Z[b] - Z[b] * w[b]).max() <= 1e-12 * scale, name
        assert np.abs(Z[b].conj().T @ Z[b] - np.eye(n)).max() <= 1e-12, name


def test_eigh_jacobi_reports_unconverged(hip):
    rng = np.random.default_rng(5)
    X = rng.standard_normal((2, 30, 30))
    G = torch.as_tensor(X @ X.transpose(0, 2, 1), device=hip)
    _, _, sw = kernels.eigh_jacobi(G, max_sweeps=1)
    assert sw.tolist() == [-1, -1]
    w, _, sw = kernels.eigh_jacobi(G)
    assert bool((sw > 1).all())
    assert np.abs(w.cpu().numpy() - np.linalg.eigvalsh(X @ X.transpose(0, 2, 1))).max() < 1e-11


def test_eigh_jacobi_refuses_large_and_wrong_dtype(hip):
    with pytest.raises(ValueError):
        kernels.eigh_jacobi(torch.zeros((1, 65, 65), dtype=torch.float64, device=hip))
    with pytest.raises(TypeError):
        kernels.eigh_jacobi(torch.zeros((1, 4, 4), dtype=torch.float32, device=hip))


@pytest.mark.parametrize("iso", ["left", "right"])
def test_gram_cut_same_with_jacobi_and_vendor(hip, iso, monkeypatch):
    """The truncation of padded legs keeps the same product U S Vh and Schmidt values whichever
    eigensolver runs (the legs' spectra are non-degenerate at the cut)."""
    rng = np.random.default_rng(7)
    k, m, n = 6, 40, 24
    mat = rng.standard_normal((k, m, n)) + 1j * rng.standard_normal((k, m, n))
    mat[1, :, 10:] = 0                      # a leg padded to the widest bond
    mat[2, 30:, :] = 0
    t = torch.as_tensor(mat, device=hip)
    caps = np.array([8, 12, 5, 24, 3, 16])
    ceil = np.full(k, 24)
    out = {}
    for flag in (True, False):
        monkeypatch.setattr(tdvp, "_JACOBI", flag)
        U, S, Vh, ranks, disc = tdvp._gram_cut(t, caps, ceil, 0.0, iso)
        out[flag] = (torch.matmul(U * S[:, None, :].to(U.dtype), Vh).cpu().numpy(),
                     S.cpu().numpy(), ranks, disc)
    a, b = out[True], out[False]
    assert np.array_equal(a[2], b[2])
    assert np.abs(a[1] - b[1]).max() < 1e-11 * np.abs(b[1]).max()
    assert np.abs(a[0] - b[0]).max() < 1e-10 * np.abs(b[0]).max()
    assert np.abs(np.asarray(a[3]) - np.asarray(b[3])).max() < 1e-12


def _svd_cases(m, n, cplx, rng):
    def rnd(*s):
        x = rng.standard_normal(s)
        return x + 1j * rng.standard_normal(s) if cplx else x

    out = [("random", rnd(m, n))]
    A = rnd(m, n)
    if min(m, n) >= 3:                       # exact rank deficiency on the short side
        if m >= n:
            A[:, 0] = 0
            A[:, 2] = 0
        else:
            A[0, :] = 0
            A[2, :] = 0
    out.append(("zero lines", A))
    out.append(("rank one", np.outer(rnd(m), rnd(n))))
    out.append(("zero", np.zeros((m, n)) * (1j if cplx else 1)))
    r = min(m, n)
    Q1, _ = np.linalg.qr(rnd(m, r))
    Q2, _ = np.linalg.qr(rnd(n, r))
    out.append(("spread", (Q1 * np.logspace(0, -24, r)) @ Q2.conj().T))
    out.append(("degenerate", (Q1 * np.repeat([1.0, 0.5], (r + 1) // 2)[:r]) @ Q2.conj().T))
    return out


@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (5, 1), (7, 5), (5, 7), (32, 32), (64, 64),
                                 (80, 64), (64, 80), (200, 16), (16, 200), (64, 33)])
@pytest.mark.parametrize("cplx", [True, False])
def test_svd_jacobi_against_numpy(hip, m, n, cplx):
    """One-sided Jacobi SVD against numpy's: singular values to 1e-13 of the largest, U S Vh = A
    to 1e-13, U and Vh isometries to 1e-12 (also with exactly zero singular values), one batched
    launch over all cases."""
    print(f"svd_jacobi {m}x{n} complex={cplx}", flush=True)
    rng = np.random.default_rng(7 * m + n + cplx)
    cases = _svd_cases(m, n, cplx, rng)
    dt = torch.complex128 if cplx else torch.float64
    A = torch.as_tensor(np.stack([c[1] for c in cases])).to(hip, dt)
    U, S, Vh, sw = kernels.svd_jacobi(A)
    U, S, Vh, sw = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy(), sw.cpu().numpy()
    r = min(m, n)
    for b, (name, X) in enumerate(cases):
        ref = np.linalg.svd(X, compute_uv=False)
        scale = max(ref[0], 1e-300)
        assert sw[b] > 0, (name, sw[b])
        assert np.all(np.diff(S[b]) <= 0), name
        assert np.abs(S[b] - ref).max() <= 1e-13 * scale, name
        assert np.abs((U[b] * S[b]) @ Vh[b] - X).max() <= 1e-13 * scale, name
        assert np.abs(U[b].conj().T @ U[b] - np.eye(r)).max() <= 1e-12, name
        assert np.abs(Vh[b] @ Vh[b].conj().T - np.eye(r)).max() <= 1e-12, name


def test_thin_svd_routes_and_falls_back(hip):
    """thin_svd equals torch's SVD on a matrix the kernel takes and on one it does not."""
    rng = np.random.default_rng(9)
    for m, n in ((40, 30), (300, 100)):
        X = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        U, S, Vh = kernels.thin_svd(torch.as_tensor(X, device=hip))
        assert kernels.svd_jacobi_fits(m, n, torch.complex128) == (m == 40)
        ref = np.linalg.svd(X, compute_uv=False)
        assert np.abs(S.cpu().numpy() - ref).max() < 1e-12 * ref[0]
        assert np.abs(((U * S.to(U.dtype)) @ Vh).cpu().numpy() - X).max() < 1e-12 * ref[0]
    with pytest.raises(ValueError):
        kernels.svd_jacobi(torch.zeros((100, 100), dtype=torch.complex128, device=hip))


def test_jacobi_kernels_keep_good_matrices_and_flag_nan(hip):
    """A NaN in one matrix of a batch leaves the others exact and shows up in that matrix's
    outputs (never finite garbage), which is what thin_svd and the truncation check before
    deferring to the vendor routines."""
    rng = np.random.default_rng(11)
    X = rng.standard_normal((2, 6, 6))
    G = X @ X.transpose(0, 2, 1)
    G[1, 2, 3] = G[1, 3, 2] = np.nan
    w, Z, sw = kernels.eigh_jacobi(torch.as_tensor(G, device=hip))
    w, Z, sw = w.cpu().numpy(), Z.cpu().numpy(), sw.cpu().numpy()
    assert np.abs(w[0] - np.linalg.eigvalsh(G[0])).max() < 1e-12 * np.abs(w[0]).max()
    assert sw[0] > 0
    assert sw[1] < 0                           # reported, so the caller redoes it
    A = rng.standard_normal((2, 7, 5))
    A[1, 0, 0] = np.nan
    U, S, Vh, sw = kernels.svd_jacobi(torch.as_tensor(A, device=hip))
    S = S.cpu().numpy()
    assert np.abs(S[0] - np.linalg.svd(A[0], compute_uv=False)).max() < 1e-13 * S[0, 0]
    assert np.isnan(S[1]).any() and sw.cpu().numpy()[1] < 0
