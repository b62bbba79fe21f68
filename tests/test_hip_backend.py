"""The backend boundary on the device: tnkit's own einsum equations through HipBackend.einsum
(library GEMM and row kernels) against numpy, tnkit's unfused apply_two_site and environment
chains restated on be.einsum against the oracle, the duck-type helpers, the backward entries
(matvec2_bwd, the environment VJPs) against CPU autograd, and a hipGraph recorded through the C
ABI replaying the eager result."""
import ctypes

import numpy as np
import pytest
import scipy.linalg
import torch

from oracle import dense as D
from oracle import exact as X
from tnkit_amd import _lib, contract, kernels
from tnkit_amd import autodiff as AD
from tnkit_amd.backend import get_backend

from test_contract_plans import TNKIT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("eq,shapes,route", TNKIT, ids=[t[0] for t in TNKIT])
@pytest.mark.parametrize("dt", [np.float64, np.complex128, np.float32])
def test_tnkit_equations_on_device(hip, eq, shapes, route, dt):
    """Every equation tnkit sends through be.einsum, on the library route, equals numpy's
    einsum (float64/complex128 to 1e-13, float32 to 1e-5 of the largest entry)."""
    be = get_backend("hip")
    rng = np.random.default_rng(len(eq))
    ops = []
    for s in shapes:
        x = rng.standard_normal(s)
        if dt == np.complex128:
            x = x + 1j * rng.standard_normal(s)
        ops.append(x.astype(dt))
    ref = np.einsum(eq, *ops)
    got = be.einsum(eq, *(torch.as_tensor(o).to(hip) for o in ops))
    assert contract.plan_of(eq, tuple(len(s) for s in shapes)) == route
    g = got.cpu().numpy()
    assert g.shape == ref.shape and g.dtype == ref.dtype
    tol = 1e-5 if dt == np.float32 else 1e-13
    assert np.abs(g - ref).max() <= tol * max(1.0, np.abs(ref).max())


def _be_apply_two_site(be, theta, L, W1, W2, R):
    # the four equations of tnkit/_dmrg.py:162-169, in the same order
    x1 = be.einsum("pwq,qijb->pwijb", L, theta)
    x2 = be.einsum("wmsi,pwijb->pmsjb", W1, x1)
    x3 = be.einsum("mrtj,pmsjb->pstrb", W2, x2)
    return be.einsum("drb,pstrb->pstd", R, x3)


def _be_left_env(be, L, A, W):
    t1 = be.einsum("pwq,psd->wqsd", L, be.conj(A))         # tnkit/_dmrg.py:216-222
    t2 = be.einsum("wqsd,wrsy->qdry", t1, W)
    return be.einsum("qdry,qye->dre", t2, A)


def _be_right_env(be, R, B, W):
    t1 = be.einsum("xry,usx->ryus", R, be.conj(B))          # tnkit/_dmrg.py:225-231
    t2 = be.einsum("ryus,wrsp->yuwp", t1, W)
    return be.einsum("yuwp,vpy->uwv", t2, B)


@pytest.mark.parametrize("a,b,D_", [(64, 48, 5), (256, 256, 5)])
def test_tnkit_chains_on_backend_vs_oracle(hip, a, b, D_):
    """tnkit's unfused two-site matvec and both environment updates, run equation by equation
    on be.einsum, equal the oracle to 1e-12."""
    be = get_backend("hip-real")
    rng = np.random.default_rng(a + b)
    L, R = rng.standard_normal((a, D_, a)), rng.standard_normal((b, D_, b))
    W = rng.standard_normal((D_, D_, 2, 2))
    th = rng.standard_normal((a, 2, 2, b))
    t = lambda x: torch.as_tensor(x).to(hip)  # noqa: E731
    got = _be_apply_two_site(be, t(th), t(L), t(W), t(W), t(R)).cpu().numpy()
    ref = D.apply_two_site(th, L, W, W, R)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    A = rng.standard_normal((a, 2, b))
    gl = _be_left_env(be, t(L), t(A), t(W)).cpu().numpy()
    rl = D.update_left_env(L, A, W)
    assert np.abs(gl - rl).max() < 1e-12 * np.abs(rl).max()
    Bt = rng.standard_normal((a, 2, b))
    gr = _be_right_env(be, t(R), t(Bt), t(W)).cpu().numpy()
    rr = D.update_right_env(R, Bt, W)
    assert np.abs(gr - rr).max() < 1e-12 * np.abs(rr).max()


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128, torch.float32, torch.complex64])
def test_row_kernels_at_krylov_size(hip, dt):
    """rows_dot / rows_combine over 41 rows of n = 2^22 (the chi=1024 Krylov basis) and over 70
    rows (two launches of the dot) against float64/complex128 references."""
    g = torch.Generator(device=hip)
    g.manual_seed(0)
    for k, n in ((41, 1 << 22), (70, 3000)):
        V = torch.randn((k, n), dtype=dt, device=hip, generator=g)
        x = torch.randn(n, dtype=dt, device=hip, generator=g)
        c = torch.randn(k, dtype=dt, device=hip, generator=g)
        hi = torch.complex128 if dt.is_complex else torch.float64
        ref_d = (V.to(hi) @ x.to(hi))
        ref_c = (c.to(hi) @ V.to(hi))
        got_d = kernels.rows_dot(V, x).to(hi)
        got_c = kernels.rows_combine(c, V).to(hi)
        tol = 1e-12 if dt in (torch.float64, torch.complex128) else 1e-5
        assert float((got_d - ref_d).abs().max()) <= tol * float(ref_d.abs().max()) * 10
        assert float((got_c - ref_c).abs().max()) <= tol * float(ref_c.abs().max()) * 10


def test_duck_type_helpers(hip):
    """tridiag/stack/guard_denom/mv/vm and the device exponentials of tnkit/backend.py:592-751:
    the eigh route (Jacobi kernel) and the Taylor route against scipy's expm."""
    be = get_backend("hip")
    rng = np.random.default_rng(3)
    k, m = 5, 12
    d = rng.standard_normal((k, m))
    e = np.abs(rng.standard_normal((k, m - 1)))
    tau = -0.3j
    got = be.expm_sym_tridiag_eigh_batched(torch.as_tensor(d).to(hip), torch.as_tensor(e).to(hip),
                                           tau).cpu().numpy()
    for j in range(k):
        T = np.diag(d[j]) + np.diag(e[j], 1) + np.diag(e[j], -1)
        assert np.abs(got[j] - scipy.linalg.expm(tau * T)).max() < 1e-12
        T1 = be.tridiag(torch.as_tensor(d[j]).to(hip), torch.as_tensor(e[j]).to(hip))
        ex = be.expm_taylor(tau * T1, 8).cpu().numpy()
        assert np.abs(ex - scipy.linalg.expm(tau * T)).max() < 1e-12
    one = be.expm_sym_tridiag_eigh(torch.as_tensor(d[0]).to(hip), torch.as_tensor(e[0]).to(hip),
                                   tau).cpu().numpy()
    assert np.abs(one - got[0]).max() < 1e-13
    assert float(be.guard_denom(torch.tensor(0.0, dtype=torch.float64, device=hip))) == 1e-200
    s = be.stack([torch.tensor(1.0, device=hip), torch.tensor(2.0, device=hip)])
    assert s.cpu().tolist() == [1.0, 2.0]
    A = torch.as_tensor(rng.standard_normal((6, 40))).to(hip).to(torch.complex128)
    x = torch.as_tensor(rng.standard_normal(40)).to(hip).to(torch.complex128)
    c = torch.as_tensor(rng.standard_normal(6)).to(hip).to(torch.complex128)
    assert torch.allclose(be.mv(A, x), A @ x, atol=1e-13)
    assert torch.allclose(be.vm(c, A), c @ A, atol=1e-13)
    G = rng.standard_normal((30, 30)) + 1j * rng.standard_normal((30, 30))
    G = G + G.conj().T
    w, Z = be.eigh(torch.as_tensor(G).to(hip))
    assert np.abs(w.cpu().numpy() - np.linalg.eigvalsh(G)).max() < 1e-12 * np.abs(G).max()
    Zh = Z.cpu().numpy()
    assert np.abs(G @ Zh - Zh * w.cpu().numpy()).max() < 1e-12 * np.abs(G).max()


@pytest.mark.parametrize("shape,cplx", [((150, 150), True), ((3, 200, 200), True),
                                        ((300, 300), False)])
def test_backend_eigh_above_the_jacobi_orders(hip, shape, cplx):
    """be.eigh past order 64 on the library's solvers (complex batched, real single): ascending
    eigenvalues to 1e-12 of the radius, A Z = Z diag(w), Z^H Z = 1."""
    rng = np.random.default_rng(sum(shape))
    X = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    G = X + np.swapaxes(X, -1, -2).conj()
    be = get_backend("hip")
    w, Z = be.eigh(torch.as_tensor(G).to(hip))
    w, Z = w.cpu().numpy(), Z.cpu().numpy()
    wr = np.linalg.eigvalsh(G)
    sc = np.abs(wr).max()
    assert np.abs(w - wr).max() < 1e-12 * sc
    assert np.abs(G @ Z - Z * w[..., None, :]).max() < 1e-12 * sc
    n = shape[-1]
    assert np.abs(np.swapaxes(Z, -1, -2).conj() @ Z - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("shape,cplx", [((300, 120), False), ((4, 256, 64), True),
                                        ((20, 40), True)])
def test_backend_qr(hip, shape, cplx):
    """be.qr: the library's QR for tall matrices (numpy's Q and R to 1e-12), the vendor's for
    wide ones; Q R = A either way."""
    rng = np.random.default_rng(sum(shape) + 1)
    A = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    Q, R = get_backend("hip").qr(torch.as_tensor(A).to(hip))
    Q, R = Q.cpu().numpy(), R.cpu().numpy()
    Qn, Rn = np.linalg.qr(A)
    assert np.abs(Q @ R - A).max() < 1e-12 * np.abs(A).max()
    assert np.abs(R - Rn).max() < 1e-12 * np.abs(A).max() * 10
    assert np.abs(Q - Qn).max() < 1e-11


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128])
def test_matvec2_bwd_is_the_adjoint(hip, dt):
    """tnh_matvec2_bwd(ybar) = H_eff^H ybar of the dense H_eff (the oracle's explicit matrix)."""
    rng = np.random.default_rng(6)
    a, wl, wm, wr, b = 7, 3, 4, 2, 5
    cx = dt == torch.complex128
    r = lambda *s: rng.standard_normal(s) + (1j * rng.standard_normal(s) if cx else 0)  # noqa
    L, W1, W2, R, yb = r(a, wl, a), r(wl, wm, 2, 2), r(wm, wr, 2, 2), r(b, wr, b), r(a, 2, 2, b)
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x)).to(hip).to(dt)  # noqa: E731
    got = kernels.matvec2_bwd(t(L), t(W1), t(W2), t(R), t(yb)).cpu().numpy().reshape(-1)
    ref = X.heff_dense(L, W1, W2, R).conj().T @ yb.reshape(-1)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def _lenv_ref(L, A, W):
    return torch.einsum("pwq,psd,wrsy,qye->dre", L, A.conj(), W, A)


def _renv_ref(R, B, W):
    return torch.einsum("xry,usx,wrsp,vpy->uwv", R, B.conj(), W, B)


@pytest.mark.parametrize("dt", [torch.float64, torch.complex128])
def test_environment_vjps_vs_cpu_autograd(hip, dt):
    """LeftEnvUpdate / RightEnvUpdate gradients of the environment and of W (tnh_env_*_bwd and
    the MPO gradient) equal torch autograd through CPU einsums."""
    g = torch.Generator().manual_seed(9)
    a, d, c, wl, wr = 6, 2, 5, 3, 4
    r = lambda *s: torch.randn(*s, dtype=dt, generator=g)  # noqa: E731
    for fn, ref_fn, shapes in (
            (AD.LeftEnvUpdate, _lenv_ref, [(a, wl, a), (a, d, c), (wl, wr, d, d)]),
            (AD.RightEnvUpdate, _renv_ref, [(c, wr, c), (a, d, c), (wl, wr, d, d)])):
        env, A, W = (r(*s) for s in shapes)
        out_shape = ref_fn(env, A, W).shape
        cot = r(*out_shape)
        de, dW = env.to(hip).requires_grad_(True), W.to(hip).requires_grad_(True)
        y = fn.apply(de, A.to(hip), dW)
        assert float((y.cpu() - ref_fn(env, A, W)).abs().max()) < 1e-12 * float(y.abs().max())
        got = torch.autograd.grad(y, (de, dW), cot.to(hip))
        re, rW = env.clone().requires_grad_(True), W.clone().requires_grad_(True)
        ref = torch.autograd.grad(ref_fn(re, A, rW), (re, rW), cot)
        for x, z in zip(got, ref):
            assert float((x.cpu() - z).abs().max() / z.abs().max()) < 1e-12


def test_graph_through_the_c_abi_replays_eager(hip):
    """A two-site matvec recorded with tnh_graph_begin/end on a side stream (buffers allocated
    before the capture) replays the eager result bit for bit, three times, with other device
    work between the replays."""
    lib = _lib.load()
    rng = np.random.default_rng(12)
    a = b = 64
    t = lambda x: torch.as_tensor(x).to(hip)  # noqa: E731
    L, R = t(rng.standard_normal((a, 5, a))), t(rng.standard_normal((b, 5, b)))
    W = t(rng.standard_normal((5, 5, 2, 2)))
    th = t(rng.standard_normal((a, 2, 2, b)))
    LT, S, RT = kernels.fused_two_site(L, W, W, R)
    nbytes = lib.tnh_two_site_ws_bytes(_lib.F64, a, 2, 2, b, 5, 5)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=hip)
    out = torch.empty_like(th)
    eager = kernels.apply_two_site_fused(th, LT, S, RT)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    graph = kernels.Graph(side)
    with graph.capture():
        _lib.check(lib.tnh_apply_two_site_fused(_lib.F64, _lib.F64, a, 2, 2, b, 5, 5,
                                                LT.data_ptr(), S.data_ptr(), RT.data_ptr(),
                                                th.data_ptr(), out.data_ptr(), ws.data_ptr(),
                                                nbytes, 0, 0, ctypes.c_void_p(side.cuda_stream)),
                   "captured matvec")
    try:
        for k in range(3):
            out.fill_(float("nan"))
            torch.cuda.synchronize()
            graph.launch()
            side.synchronize()
            assert torch.equal(out, eager), k
            _ = (th * 2.0).sum().item()           # other work between replays
    finally:
        graph.close()
