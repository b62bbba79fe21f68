"""The randomized range-finder SVD on the library's kernels (tnkit/linalg.py:81-135): the
orthonormalisation it runs on the K-major GEMM and the small eigensolvers, against numpy's QR
span on sketches from perfectly to exactly singular conditioning, and the whole factorisation
against the oracle's (numpy QR and SVD, the same host sketch) at sizes where the core SVD goes
through the Gram truncation (ell > 64) as well as through the Jacobi kernel."""
import numpy as np
import pytest
import torch

from oracle import dense as D
from tnkit_amd import linalg

pytestmark = pytest.mark.gpu


def _sketch(m, l, kind, rng, cplx=False):
    X = rng.standard_normal((m, l)) + (1j * rng.standard_normal((m, l)) if cplx else 0.0)
    U, _ = np.linalg.qr(X)
    Y = rng.standard_normal((l, l)) + (1j * rng.standard_normal((l, l)) if cplx else 0.0)
    V, _ = np.linalg.qr(Y)
    if kind == "flat":
        s = np.ones(l)
    elif kind == "graded":
        s = np.logspace(0, -8, l)
    elif kind == "steep":
        s = np.logspace(0, -14, l)
    else:                                   # exactly rank deficient: rank l // 2
        s = np.where(np.arange(l) < l // 2, np.logspace(0, -3, l), 0.0)
    return (U * s) @ V.conj().T


@pytest.mark.parametrize("m,l,kind,cplx", [(500, 100, "flat", False), (500, 100, "graded", False),
                                           (800, 120, "steep", False), (600, 90, "deficient", False),
                                           (300, 48, "graded", True), (300, 40, "steep", True),
                                           (64, 20, "flat", False)])
def test_orthonormal_spans_the_sketch(hip, m, l, kind, cplx):
    """Q^H Q = 1 to 1e-13, and Q Q^H Y = Y to 1e-12 of |Y| (the span of numpy's QR factor)."""
    rng = np.random.default_rng(m + l)
    Y = _sketch(m, l, kind, rng, cplx)
    Q = linalg._orthonormal(torch.as_tensor(Y).to(hip)).cpu().numpy()
    assert Q.shape == (m, l)
    assert np.abs(Q.conj().T @ Q - np.eye(l)).max() <= 1e-13
    assert np.abs(Q @ (Q.conj().T @ Y) - Y).max() <= 1e-12 * np.abs(Y).max()


@pytest.mark.parametrize("m,n,chi,mixed", [(900, 700, 100, False), (700, 1100, 150, False),
                                           (600, 500, 40, False), (900, 700, 100, True)])
def test_randomized_svd_against_the_oracle(hip, m, n, chi, mixed):
    """The same host sketch through the device range finder and the oracle's: singular values to
    1e-10 of the largest (1e-6 with the float32 range finder), U S Vh to 1e-9 (1e-5), U an
    isometry to 1e-12, the discarded weight to 1e-9 of |A|^2."""
    rng = np.random.default_rng(m * n + chi)
    r = min(m, n)
    U0, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V0, _ = np.linalg.qr(rng.standard_normal((n, r)))
    A = (U0 * np.exp(-np.arange(r) / 25.0)) @ V0.T
    U, S, Vh, disc = linalg.randomized_svd_device(torch.as_tensor(A).to(hip), chi,
                                                  rng=np.random.default_rng(7), mixed=mixed)
    U, S, Vh = (x.cpu().numpy() for x in (U, S, Vh))
    Ur, Sr, Vr, dr = D.randomized_svd(A, chi, np.random.default_rng(7), mixed=mixed)
    tol = 1e-6 if mixed else 1e-10
    assert U.shape == Ur.shape and Vh.shape == Vr.shape
    assert np.abs(S - Sr).max() <= tol * Sr[0]
    assert np.abs((U * S) @ Vh - (Ur * Sr) @ Vr).max() <= (1e-5 if mixed else 1e-9) * Sr[0]
    assert np.abs(U.T @ U - np.eye(U.shape[1])).max() <= 1e-12
    assert abs(disc - dr) <= 1e-9 * Sr[0] ** 2
