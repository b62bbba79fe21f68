"""Backward kernels: the H_eff matvec's VJP, the MPO gradient and dE/dg_i of the per-site-field
TFIM against the oracle, finite differences, the exact free-fermion derivative and -<Z_i>."""
import numpy as np
import pytest
import torch

from oracle import dense as D
from oracle import exact as X
from tnkit_amd import _dmrg as H
from tnkit_amd import autodiff as AD
from tnkit_amd.models import tfim_mpo_fields
from tnkit_amd.observables import local_expectations

pytestmark = pytest.mark.gpu


def _t(x, dev):
    return torch.as_tensor(np.ascontiguousarray(x)).to(dev)


def test_heff_matvec_backward_is_the_transpose(hip):
    rng = np.random.default_rng(0)
    a, wl, wm, wr, d, b = 5, 3, 4, 3, 2, 6
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d, d)), rng.standard_normal((wm, wr, d, d))
    th = rng.standard_normal((a, d, d, b))
    ybar = rng.standard_normal((a, d, d, b))
    Hm = X.heff_dense(L, W1, W2, R)
    x = _t(th, hip).requires_grad_(True)
    y = AD.HeffMatvec.apply(x, _t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip))
    np.testing.assert_allclose(y.detach().cpu().numpy().reshape(-1), Hm @ th.reshape(-1),
                               atol=1e-12 * np.abs(Hm).max() * 10)
    (gx,) = torch.autograd.grad(y, x, _t(ybar, hip))
    ref = Hm.T @ ybar.reshape(-1)
    assert np.abs(gx.cpu().numpy().reshape(-1) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("a,d,c,wl,wr", [(1, 2, 2, 1, 3), (6, 2, 5, 3, 3), (40, 2, 33, 5, 5),
                                         (7, 2, 3, 3, 1)])
def test_mpo_gradient_vs_oracle(hip, a, d, c, wl, wr):
    rng = np.random.default_rng(a + c)
    L, A = rng.standard_normal((a, wl, a)), rng.standard_normal((a, d, c))
    W, R = rng.standard_normal((wl, wr, d, d)), rng.standard_normal((c, wr, c))
    got = AD.mpo_gradient(_t(L, hip), _t(A, hip), _t(W, hip), _t(R, hip)).cpu().numpy()
    ref = D.mpo_gradient(L, A, W, R)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_field_gradient_hellmann_feynman(hip):
    """TFIM N=16 with g_i = 1 + 0.1 N(0,1) (seed 2), chi=32: dE/dg_i from the backward kernels
    equals central differences of the energy functional (h = 1e-5), the derivative of the exact
    free-fermion energy and -<Z_i>, to 1e-7."""
    N = 16
    g = 1.0 + 0.1 * np.random.default_rng(2).standard_normal(N)
    E, M, _ = H.dmrg(tfim_mpo_fields(g), 32, n_sweeps=6, tol=0.0, rtol=0.0, backend="hip-real")
    assert abs(E - X.tfim_free_fermion_energy(N, g)) < 1e-9
    dg = AD.field_gradient(list(M), torch.as_tensor(g)).cpu().numpy()
    Mh = [t.cpu().numpy() for t in M]
    z = local_expectations(M, np.diag([1.0, -1.0]))
    np.testing.assert_allclose(dg, -z, atol=1e-7)
    np.testing.assert_allclose(dg, D.field_gradient(Mh, tfim_mpo_fields, g), atol=1e-10)
    h = 1e-5
    for i in (0, 7, N - 1):
        e = np.zeros(N)
        e[i] = h
        fd = (D.mpo_energy(Mh, tfim_mpo_fields(g + e)) - D.mpo_energy(Mh, tfim_mpo_fields(g - e))) / (2 * h)
        ff = (X.tfim_free_fermion_energy(N, g + e) - X.tfim_free_fermion_energy(N, g - e)) / (2 * h)
        assert abs(dg[i] - fd) < 1e-7 and abs(dg[i] - ff) < 1e-7


def _heff_ref(th, L, W1, W2, R):
    return torch.einsum("pwq,wmsu,mrtv,brc,quvc->pstb", L, W1, W2, R, th)


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128])
def test_heff_matvec_gradients_of_every_input(hip, dt):
    """The gradients of theta, L, W1, W2 and R (torch's convention, so H_eff^H y_bar for theta)
    equal torch autograd through a CPU einsum of the same network, real and complex."""
    g = torch.Generator().manual_seed(3)
    a, wl, wm, wr, d, b = 6, 3, 4, 2, 2, 5
    r = lambda *s: torch.randn(*s, dtype=dt, generator=g)  # noqa: E731
    cpu = [r(a, d, d, b), r(a, wl, a), r(wl, wm, d, d), r(wm, wr, d, d), r(b, wr, b)]
    ybar = r(a, d, d, b)
    dev = [t.to(hip).requires_grad_(True) for t in cpu]
    y = AD.HeffMatvec.apply(*dev)
    got = torch.autograd.grad(y, dev, ybar.to(hip))
    ref_in = [t.clone().requires_grad_(True) for t in cpu]
    ref = torch.autograd.grad(_heff_ref(*ref_in), ref_in, ybar)
    for name, x, z in zip(("theta", "L", "W1", "W2", "R"), got, ref):
        err = float((x.cpu() - z).abs().max() / z.abs().max())
        assert err < 1e-12, (name, err)


def test_heff_matvec_theta_only_skips_operand_gradients(hip):
    """Only theta requires grad: the other gradients are None and theta's is H_eff^H y_bar."""
    g = torch.Generator().manual_seed(4)
    a, wl, d, b = 4, 3, 2, 4
    r = lambda *s: torch.randn(*s, dtype=torch.complex128, generator=g)  # noqa: E731
    th, L, W1, W2, R = r(a, d, d, b), r(a, wl, a), r(wl, wl, d, d), r(wl, wl, d, d), r(b, wl, b)
    x = th.to(hip).requires_grad_(True)
    y = AD.HeffMatvec.apply(x, *(t.to(hip) for t in (L, W1, W2, R)))
    ybar = r(a, d, d, b)
    (gx,) = torch.autograd.grad(y, x, ybar.to(hip))
    Hm = X.heff_dense(L.numpy(), W1.numpy(), W2.numpy(), R.numpy())
    ref = Hm.conj().T @ ybar.numpy().reshape(-1)
    assert np.abs(gx.cpu().numpy().reshape(-1) - ref).max() < 1e-12 * np.abs(ref).max()


def _energy_ref(M, W):
    L = torch.ones((1, 1, 1), dtype=M[0].dtype)
    n = torch.ones((1, 1), dtype=M[0].dtype)
    for A, w in zip(M, W):
        L = torch.einsum("pwq,psd,wrsy,qye->dre", L, A.conj(), w, A)
        n = torch.einsum("pq,psd,qse->de", n, A.conj(), A)
    return (L.reshape(-1)[0] / n.reshape(-1)[0]).real


def test_mpo_expectation_gradient_complex(hip):
    """dE/dW of E = Re <psi|H|psi> / <psi|psi> for a complex MPS and complex MPO tensors equals
    torch autograd through a CPU einsum chain."""
    g = torch.Generator().manual_seed(5)
    dims = [1, 2, 4, 3, 1]
    D_ = [1, 3, 3, 3, 1]
    r = lambda *s: torch.randn(*s, dtype=torch.complex128, generator=g)  # noqa: E731
    M = [r(dims[i], 2, dims[i + 1]) for i in range(4)]
    W = [r(D_[i], D_[i + 1], 2, 2) for i in range(4)]
    Wd = [w.to(hip).requires_grad_(True) for w in W]
    E = AD.MPOExpectation.apply([m.to(hip) for m in M], *Wd)
    got = torch.autograd.grad(E, Wd)
    Wr = [w.clone().requires_grad_(True) for w in W]
    Er = _energy_ref(M, Wr)
    assert abs(float(E) - float(Er)) < 1e-12 * abs(float(Er))
    ref = torch.autograd.grad(Er, Wr)
    for x, z in zip(got, ref):
        assert float((x.cpu() - z).abs().max() / z.abs().max()) < 1e-12
