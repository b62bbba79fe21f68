"""The library's own symmetric eigensolver for orders 65..2048 (`tnh_sytrd` + `tnh_syev_top`:
Householder tridiagonalisation in one cooperative launch, divide and conquer, WY
back-transformation) against numpy's eigh: eigenvalues to 1e-13 of the spectral radius,
A Z = Z diag(w) to 1e-13 and Z^T Z = 1 to 1e-12 for the k leading vectors, on the spectra the Gram
truncation meets (Gram matrices with decaying columns, low rank with a large zero cluster,
degenerate, 30 decades, random symmetric, diagonal, zero). The tridiagonalisation alone is
checked by rebuilding Q T Q^T from its reflectors; the Gram truncation that runs on it against
the oracle's (tnkit/linalg.py:138-204) and against Eckart-Young."""
import numpy as np
import pytest
import torch

from oracle import dense as D
from tnkit_amd import _lib, kernels, linalg

pytestmark = pytest.mark.gpu


def _matrix(kind, n, rng):
    if kind == "gram":
        X = rng.standard_normal((n, n)) * np.exp(-np.arange(n) / max(1.0, n / 20.0))[None, :]
        A = X @ X.T
    elif kind == "lowrank":
        X = rng.standard_normal((n, max(1, n // 7)))
        A = X @ X.T
    elif kind == "degenerate":
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (Q * np.repeat(np.arange(1.0, 6.0), (n + 4) // 5)[:n]) @ Q.T
    elif kind == "graded":
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        A = (Q * np.logspace(0, -30, n)) @ Q.T
    elif kind == "random":
        X = rng.standard_normal((n, n))
        A = X + X.T
    elif kind == "diagonal":
        A = np.diag(rng.standard_normal(n))
    else:
        A = np.zeros((n, n))
    return 0.5 * (A + A.T)


def _check(A, w, Z, k):
    wr = np.linalg.eigvalsh(A)
    sc = max(np.abs(wr).max(), 1e-300)
    assert np.abs(w - wr).max() <= 1e-13 * sc
    lam = w[::-1][:k]
    assert np.abs(A @ Z - Z * lam[None, :]).max() <= 1e-13 * sc
    assert np.abs(Z.T @ Z - np.eye(k)).max() <= 1e-12


@pytest.mark.parametrize("n,kind", [(65, "random"), (100, "gram"), (128, "lowrank"),
                                    (200, "degenerate"), (257, "graded"), (300, "diagonal"),
                                    (96, "zero"), (500, "gram"), (777, "lowrank"),
                                    (1024, "gram"), (2048, "gram"), (2048, "lowrank")])
def test_syev_top_against_numpy(hip, n, kind):
    rng = np.random.default_rng(n)
    A = _matrix(kind, n, rng)
    At = torch.as_tensor(A, device=hip)
    for k in sorted({1, n // 2, n}):
        w, Z = kernels.syev_top(At, k)
        _check(A, w.cpu().numpy(), Z.cpu().numpy(), k)


def test_syev_top_repeat_calls_are_identical(hip):
    """Repeated solves of one matrix give bit-identical results (every workgroup computes the
    shared scalars in one fixed order)."""
    A = torch.as_tensor(_matrix("gram", 700, np.random.default_rng(7)), device=hip)
    w1, Z1 = kernels.syev_top(A, 350)
    w2, Z2 = kernels.syev_top(A, 350)
    assert torch.equal(w1, w2) and torch.equal(Z1, Z2)


@pytest.mark.parametrize("n", [2, 3, 65, 300, 1500])
def test_sytrd_rebuilds_the_matrix(hip, n):
    """Q T Q^T = A with Q = H_0 ... H_{n-2} rebuilt from the reflector rows, Q orthogonal."""
    rng = np.random.default_rng(n + 1)
    A = _matrix("random" if n < 100 else "gram", n, rng)
    lib = _lib.load()
    At = torch.as_tensor(A, device=hip)
    d = torch.zeros(n, dtype=torch.float64, device=hip)
    e = torch.zeros(n, dtype=torch.float64, device=hip)
    tau = torch.zeros(n, dtype=torch.float64, device=hip)
    Vr = torch.zeros((n, n), dtype=torch.float64, device=hip)
    st = torch.ones(1, dtype=torch.int32, device=hip)
    ws = torch.empty(lib.tnh_sytrd_ws_bytes(n), dtype=torch.uint8, device=hip)
    _lib.check(lib.tnh_sytrd(n, At.data_ptr(), n, d.data_ptr(), e.data_ptr(), tau.data_ptr(),
                             Vr.data_ptr(), st.data_ptr(), ws.data_ptr(), ws.numel(),
                             kernels.stream(hip)), "sytrd")
    torch.cuda.synchronize()
    assert int(st.item()) == 0
    d, e, tau, V = (x.cpu().numpy() for x in (d, e, tau, Vr))
    Q = np.eye(n)
    for j in range(n - 2, -1, -1):
        Q = Q - tau[j] * np.outer(V[j], V[j] @ Q)
    T = np.diag(d) + np.diag(e[:n - 1], 1) + np.diag(e[:n - 1], -1)
    assert np.abs(Q @ T @ Q.T - A).max() <= 1e-13 * np.abs(A).max()
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-13


@pytest.mark.parametrize("m,n,chi,iso", [(200, 300, 120, "left"), (512, 256, 100, "right"),
                                         (1024, 1024, 512, "left")])
def test_gram_truncation_on_the_solver(hip, m, n, chi, iso):
    """linalg.gram_truncation (Gram matrix on the MFMA GEMM, eigh on the library's solver) equals
    the oracle's Gram truncation: singular values to 1e-10 of the largest, the exact isometry to
    1e-12, the discarded weight, and the truncation error is Eckart-Young's. U S Vh agrees to
    1e-8 of the largest singular value: kept directions whose Gram eigenvalues sit within
    ~n eps of their neighbours across the cut are not ranked by any float64 Gram eigensolver
    (tnkit/linalg.py:175-185 states the same floor)."""
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) * np.exp(-np.arange(n) / 40.0)[None, :]
    U, S, Vh, disc = linalg.gram_truncation(torch.as_tensor(A, device=hip), chi, iso)
    U, S, Vh = (x.cpu().numpy() for x in (U, S, Vh))
    Ur, Sr, Vr, dr = D.truncate_gram(A, chi, iso)
    assert np.abs(S - Sr).max() <= 1e-10 * Sr[0]
    assert np.abs((U * S) @ Vh - (Ur * Sr) @ Vr).max() <= 1e-8 * Sr[0]
    iso_m = U if iso == "left" else Vh.T
    assert np.abs(iso_m.T @ iso_m - np.eye(iso_m.shape[1])).max() <= 1e-12
    assert abs(disc - dr) <= 1e-10 * Sr[0] ** 2
    sv = np.linalg.svd(A, compute_uv=False)
    err = np.linalg.norm(A - (U * S) @ Vh) ** 2
    assert abs(err - np.sum(sv[chi:] ** 2)) <= 1e-9 * sv[0] ** 2


def _hermitian(kind, n, rng):
    """Complex Hermitian test matrices: the complex counterparts of _matrix, plus a tridiagonal
    with complex couplings (every reflector has nothing below its subdiagonal entry but a phase)
    and a padded Gram matrix (rows and columns beyond the live extent exactly zero)."""
    def cn(*s):
        return rng.standard_normal(s) + 1j * rng.standard_normal(s)
    if kind == "gram":
        X = cn(n, n) * np.exp(-np.arange(n) / max(1.0, n / 20.0))[None, :]
        A = X @ X.conj().T
    elif kind == "lowrank":
        X = cn(n, max(1, n // 7))
        A = X @ X.conj().T
    elif kind == "degenerate":
        Q, _ = np.linalg.qr(cn(n, n))
        A = (Q * np.repeat(np.arange(1.0, 6.0), (n + 4) // 5)[:n]) @ Q.conj().T
    elif kind == "graded":
        Q, _ = np.linalg.qr(cn(n, n))
        A = (Q * np.logspace(0, -30, n)) @ Q.conj().T
    elif kind == "random":
        X = cn(n, n)
        A = X + X.conj().T
    elif kind == "tridiagonal":
        A = np.diag(rng.standard_normal(n)).astype(complex)
        c = cn(n - 1)
        A += np.diag(c, -1) + np.diag(c.conj(), 1)
    elif kind == "padded":
        m = max(1, (2 * n) // 3)
        X = np.zeros((n, n), complex)
        X[:m, :m] = cn(m, m)
        A = X @ X.conj().T
    elif kind == "real":
        X = rng.standard_normal((n, n))
        A = (X @ X.T).astype(complex)
    else:
        A = np.zeros((n, n), complex)
    return 0.5 * (A + A.conj().T)


def _check_c(A, w, Z, k):
    wr = np.linalg.eigvalsh(A)
    sc = max(np.abs(wr).max(), 1e-300)
    assert np.abs(w - wr).max() <= 1e-13 * sc
    lam = w[::-1][:k]
    assert np.abs(A @ Z - Z * lam[None, :]).max() <= 1e-13 * sc
    assert np.abs(Z.conj().T @ Z - np.eye(k)).max() <= 1e-12


@pytest.mark.parametrize("n,kind", [(65, "random"), (100, "gram"), (128, "lowrank"),
                                    (200, "degenerate"), (257, "graded"), (300, "tridiagonal"),
                                    (96, "zero"), (384, "padded"), (500, "gram"), (512, "real"),
                                    (777, "lowrank"), (1024, "gram"), (2, "random"),
                                    (3, "tridiagonal")])
def test_heev_top_against_numpy(hip, n, kind):
    """The batched complex solver on one matrix: eigenvalues to 1e-13 of the spectral radius,
    A Z = Z diag(w) to 1e-13, Z^H Z = 1 to 1e-12 for k = 1, n/2, n."""
    rng = np.random.default_rng(n + 11)
    A = _hermitian(kind, n, rng)
    At = torch.as_tensor(A, device=hip).unsqueeze(0).contiguous()
    for k in sorted({1, max(1, n // 2), n}):
        w, Z = kernels.heev_top(At, k)
        _check_c(A, w[0].cpu().numpy(), Z[0].cpu().numpy(), k)


@pytest.mark.parametrize("n,batch", [(128, 5), (512, 9), (1024, 3)])
def test_heev_top_batch_matches_single_solves(hip, n, batch):
    """A batch that needs several cooperative launches (more matrices than one launch holds at
    n = 512 and 1024): every matrix to the same bar, and bit-identical to its solve alone (each
    matrix is reduced by the same workgroup layout whatever the batch)."""
    rng = np.random.default_rng(n * batch)
    kinds = ["gram", "padded", "random", "lowrank", "tridiagonal"]
    As = np.stack([_hermitian(kinds[i % len(kinds)], n, rng) for i in range(batch)])
    At = torch.as_tensor(As, device=hip).contiguous()
    k = n // 4
    w, Z = kernels.heev_top(At, k)
    wn, Zn = w.cpu().numpy(), Z.cpu().numpy()
    for i in range(batch):
        _check_c(As[i], wn[i], Zn[i], k)
    for i in (0, batch - 1):
        w1, Z1 = kernels.heev_top(At[i:i + 1].contiguous(), k)
        assert torch.equal(w1[0], w[i]) and torch.equal(Z1[0], Z[i])


def test_heev_top_rejects_bad_shapes(hip):
    A = torch.zeros((1, 1025, 1025), dtype=torch.complex128, device=hip)
    with pytest.raises(ValueError):
        kernels.heev_top(A, 4)
    with pytest.raises(ValueError):
        kernels.heev_top(torch.zeros((2, 8, 8), dtype=torch.complex128, device=hip), 9)
