"""The library's thin QR (`tnh_qr`: LAPACK geqrf's reflectors in one cooperative launch per group
of matrices, Q by WY blocks) against numpy's QR, which takes the same reflectors (geqrf + ungqr):
Q and R elementwise on well-conditioned matrices, and Q^H Q = 1, Q R = A, R upper with a real
diagonal on rank-deficient ones (zero and repeated columns, the projected factors of a
truncation whose kept spectrum reaches rounding level)."""
import numpy as np
import pytest
import torch

from tnkit_amd import kernels

pytestmark = pytest.mark.gpu


def _rand(shape, cplx, rng):
    X = rng.standard_normal(shape)
    return X + 1j * rng.standard_normal(shape) if cplx else X


def _check_factors(A, Q, R, tol=1e-13):
    n = A.shape[1]
    sc = max(np.abs(A).max(), 1e-300)
    assert np.abs(Q.conj().T @ Q - np.eye(n)).max() <= tol * 10
    assert np.abs(Q @ R - A).max() <= tol * sc * 10
    assert np.abs(np.tril(R, -1)).max() == 0.0
    assert np.abs(np.diag(R).imag).max() == 0.0


@pytest.mark.parametrize("m,n,cplx", [(2048, 1024, False), (1500, 700, False), (300, 300, False),
                                      (65, 1, False), (100, 37, False), (1024, 512, True),
                                      (512, 256, True), (200, 200, True), (7, 3, True),
                                      (513, 64, True), (1, 1, False)])
def test_qr_against_numpy(hip, m, n, cplx):
    rng = np.random.default_rng(m * 7 + n)
    A = _rand((m, n), cplx, rng)
    Q, R = kernels.qr(torch.as_tensor(A, device=hip))
    Q, R = Q.cpu().numpy(), R.cpu().numpy()
    Qn, Rn = np.linalg.qr(A)
    _check_factors(A, Q, R)
    assert np.abs(R - Rn).max() <= 1e-12 * np.abs(A).max() * np.sqrt(m)
    assert np.abs(Q - Qn).max() <= 1e-11


@pytest.mark.parametrize("kind", ["zero_columns", "repeated", "graded"])
@pytest.mark.parametrize("cplx", [False, True])
def test_qr_rank_deficient(hip, kind, cplx):
    rng = np.random.default_rng(3)
    m, n = 400, 160
    A = _rand((m, n), cplx, rng)
    if kind == "zero_columns":
        A[:, 100:] = 0.0
        A[:, 3] = 0.0
    elif kind == "repeated":
        A[:, 80:] = A[:, :80]
    else:                                   # columns spanning 30 decades, as A Z at a deep cut
        A = A * np.logspace(0, -30, n)[None, :]
    Q, R = kernels.qr(torch.as_tensor(A, device=hip))
    _check_factors(A, Q.cpu().numpy(), R.cpu().numpy())


@pytest.mark.parametrize("B,m,n,cplx", [(32, 512, 128, True), (9, 256, 256, True),
                                         (10, 1024, 512, True), (3, 2048, 1024, False)])
def test_qr_batch_matches_single(hip, B, m, n, cplx):
    """Batches in one launch and in several (the last two hold 4 and 2 matrices per launch on
    256 compute units): every matrix against numpy, and bit-identical to its factorisation
    alone."""
    rng = np.random.default_rng(B + m + n)
    A = _rand((B, m, n), cplx, rng)
    At = torch.as_tensor(A, device=hip)
    Q, R = kernels.qr(At)
    Qh, Rh = Q.cpu().numpy(), R.cpu().numpy()
    for i in range(B):
        _check_factors(A[i], Qh[i], Rh[i])
    for i in (0, B - 1):
        Q1, R1 = kernels.qr(At[i])
        assert torch.equal(Q1, Q[i]) and torch.equal(R1, R[i])


def test_qr_refuses_what_it_does_not_serve(hip):
    assert not kernels.qr_supported(torch.zeros((3, 5), dtype=torch.float64, device=hip))
    assert not kernels.qr_supported(torch.zeros((1100, 5), dtype=torch.complex128, device=hip))
    assert not kernels.qr_supported(torch.zeros((8, 5), dtype=torch.float32, device=hip))
    with pytest.raises(ValueError):
        kernels.qr(torch.zeros((3, 5), dtype=torch.float64, device=hip))
