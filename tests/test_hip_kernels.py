"""libtnkit_hip kernels against the oracle (float64 to 1e-12 relative; float32 operator to a
bound set by float32 rounding), through the C-ABI on the HIP device."""
import numpy as np
import pytest
import torch

from oracle import dense as D
from tnkit_amd import _lib, kernels
from tnkit_amd.models import xxz_mpo

pytestmark = pytest.mark.gpu


def _t(x, dev, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(x)).to(dev, dtype)


def _rel(got, ref):
    got = got.detach().cpu().double().numpy() if isinstance(got, torch.Tensor) else got
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


@pytest.mark.parametrize("M,N,K,tile", [(1, 1, 1, 0), (17, 33, 9, 0), (128, 128, 128, 1),
                                        (130, 200, 77, 1), (64, 96, 40, 2), (50, 70, 33, 3),
                                        (10, 300, 41, 4), (257, 129, 1030, 0), (0, 5, 5, 0),
                                        (5, 5, 0, 0)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_tn(hip, M, N, K, tile, dtype):
    print(f"gemm_tn M={M} N={N} K={K} tile={tile} {dtype}", flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((K, M))
    B = rng.standard_normal((K, N))
    ref = A.T @ B
    got = kernels.gemm_tn(_t(A, hip, dtype), _t(B, hip, dtype), tile=tile)
    tol = 1e-13 if dtype == torch.float64 else 2e-6 * np.sqrt(max(K, 1))
    if M * N:
        assert _rel(got, ref) < tol
    acc = kernels.gemm_tn(_t(A, hip, dtype), _t(B, hip, dtype), out=got.clone(), accumulate=True,
                          tile=tile)
    if M * N:
        assert _rel(acc, 2 * ref) < tol


@pytest.mark.parametrize("M,N,K", [(128, 128, 16), (128, 256, 48), (384, 128, 80), (256, 384, 96),
                                   (256, 256, 64), (256, 256, 40)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_gemm_tn_direct_to_lds(hip, M, N, K, dtype):
    """Shapes that take the direct-to-LDS kernel (M, N multiples of 128, K of its step: one
    stage, a full ring, a ring that wraps), batched with strides and accumulating, through the C
    ABI; K = 40 in float32 falls back to the staged kernel."""
    print(f"gemm_tn dma M={M} N={N} K={K} {dtype}", flush=True)
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    nb = 3
    A = rng.standard_normal((nb, K, M))
    B = rng.standard_normal((nb, K, N))
    ref = np.einsum("bkm,bkn->bmn", A, B)
    Ad, Bd = _t(A, hip, dtype), _t(B, hip, dtype)
    lib = _lib.load()
    for tile in (0, 7):                       # 7: the 8-wave form (the float64 default)
        C = torch.full((nb, M, N), 0.5, dtype=dtype, device=hip)
        for acc in (0, 1):
            rc = lib.tnh_gemm_tn(_lib.dcode(dtype), M, N, K, Ad.data_ptr(), M, Bd.data_ptr(), N,
                                 C.data_ptr(), N, nb, K * M, K * N, M * N, acc, tile,
                                 kernels.stream(hip))
            assert rc == 0, lib.tnh_last_error()
            want = ref if acc == 0 else 2 * ref
            tol = 1e-13 if dtype == torch.float64 else 2e-6 * np.sqrt(K)
            assert _rel(C, want) < tol


@pytest.mark.parametrize("dt", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("M,N,K,ldx", [(128, 128, 16, 0), (128, 256, 48, 0), (384, 128, 80, 0),
                                       (256, 384, 96, 0), (256, 256, 64, 16), (128, 128, 12, 0)])
def test_gemm_tn_complex_direct_to_lds(hip, M, N, K, ldx, dt):
    """Complex shapes that take the direct-to-LDS complex kernel (M, N multiples of 128, K of its
    step, 8 complex128 / 16 complex64 rows: one stage, a full ring, a ring that wraps; padded
    leading dimensions), batched with strides and accumulating, through the C ABI, against
    numpy; K = 12 falls back to the staged kernel. Tile 4 asks for it explicitly."""
    print(f"gemm_tn complex dma M={M} N={N} K={K} ldx={ldx} {dt}", flush=True)
    rng = np.random.default_rng(M + 5 * N + 11 * K)
    nb = 3
    lda, ldb, ldc = M + ldx, N + ldx, N + ldx
    A = rng.standard_normal((nb, K, lda)) + 1j * rng.standard_normal((nb, K, lda))
    B = rng.standard_normal((nb, K, ldb)) + 1j * rng.standard_normal((nb, K, ldb))
    if dt == torch.complex64:
        A, B = A.astype(np.complex64).astype(complex), B.astype(np.complex64).astype(complex)
    ref = np.einsum("bkm,bkn->bmn", A[:, :, :M], B[:, :, :N])
    Ad, Bd = torch.as_tensor(A).to(dt).to(hip), torch.as_tensor(B).to(dt).to(hip)
    tol = 1e-13 if dt == torch.complex128 else 4e-6 * np.sqrt(K)
    lib = _lib.load()
    for tile in (0, 4):
        C = torch.full((nb, M, ldc), 0.5 + 0.25j, dtype=dt, device=hip)
        for acc in (0, 1):
            rc = lib.tnh_gemm_tn(_lib.dcode(dt), M, N, K, Ad.data_ptr(), lda, Bd.data_ptr(), ldb,
                                 C.data_ptr(), ldc, nb, K * lda, K * ldb, M * ldc, acc, tile,
                                 kernels.stream(hip))
            assert rc == 0, lib.tnh_last_error()
            want = ref if acc == 0 else 2 * ref
            got = C[:, :, :N].cpu().numpy()
            assert np.abs(got - want).max() <= tol * np.abs(want).max()
        if ldx:
            assert bool((C[:, :, N:] == 0.5 + 0.25j).all())


@pytest.mark.parametrize("M,N,K,tile", [(1, 1, 1, 0), (17, 33, 9, 0), (128, 128, 128, 1),
                                        (130, 200, 77, 1), (64, 96, 40, 2), (257, 129, 1030, 0),
                                        (300, 520, 64, 0), (0, 5, 5, 0), (5, 5, 0, 0)])
def test_gemm_tn_bf16(hip, M, N, K, tile):
    """bf16 operands on the bf16 matrix cores: with bf16-representable inputs every product is
    exact in float32, so the result must equal the float64 product of the rounded inputs to
    float32 accumulation error; a wrong lane/k layout gives O(1) errors."""
    print(f"gemm_tn bf16 M={M} N={N} K={K} tile={tile}", flush=True)
    rng = np.random.default_rng(5 * M + N + K)
    A = torch.as_tensor(rng.standard_normal((K, M))).to(torch.bfloat16)
    B = torch.as_tensor(rng.standard_normal((K, N))).to(torch.bfloat16)
    ref = A.double().numpy().T @ B.double().numpy()
    got = kernels.gemm_tn(A.to(hip), B.to(hip), tile=tile)
    assert got.dtype == torch.float32 and tuple(got.shape) == (M, N)
    tol = 2e-6 * np.sqrt(max(K, 1))
    if M * N:
        assert _rel(got, ref) < tol
    acc = kernels.gemm_tn(A.to(hip), B.to(hip), out=got.clone(), accumulate=True, tile=tile)
    if M * N:
        assert _rel(acc, 2 * ref) < tol


SHAPES = [  # (a, wl, wm, wr, d1, d2, b) -- tnkit/tests/test_two_site_fused.py plus bulk sizes
    (1, 1, 5, 5, 2, 2, 2), (4, 5, 5, 1, 2, 2, 1), (7, 20, 20, 20, 2, 3, 9), (32, 3, 3, 3, 2, 2, 32),
    (64, 5, 5, 5, 2, 2, 48), (256, 5, 5, 5, 2, 2, 256),
    # a, b multiples of 32 with wl in (3, 5): GEMM1 with the MPO mix in its epilogue, wr != wl
    (64, 5, 4, 3, 2, 2, 96), (96, 3, 7, 8, 2, 2, 64)]


@pytest.mark.parametrize("a,wl,wm,wr,d1,d2,b", SHAPES)
def test_two_site_matvec_f64(hip, a, wl, wm, wr, d1, d2, b):
    print(f"two_site f64 {(a, wl, wm, wr, d1, d2, b)}", flush=True)
    rng = np.random.default_rng(a + 3 * b)
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d1, d1)), rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    ref = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    op = kernels.fused_two_site(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip))
    got = kernels.apply_two_site_fused(_t(th, hip), *op)
    assert got.shape == ref.shape
    assert _rel(got, ref) < 1e-12
    # the operator layouts are the documented permutations
    LT, S, RT = (x.cpu().numpy() for x in op)
    np.testing.assert_array_equal(LT, L.transpose(2, 0, 1))
    np.testing.assert_array_equal(RT, R.transpose(1, 2, 0))
    Sref = D.fused_two_site(L, W1, W2, R)[1]
    assert np.abs(S - Sref).max() <= 1e-13 * max(1, np.abs(Sref).max())


@pytest.mark.parametrize("a,wl,wm,wr,d1,d2,b", SHAPES[3:])
def test_two_site_matvec_f32_operator(hip, a, wl, wm, wr, d1, d2, b):
    """float32 operator with float64 input/output: error at float32 rounding of the sums.
    Shapes with a, b multiples of 32 and wl in (3, 5) take GEMM1 with the MPO mix in its
    epilogue, also with wr != wl."""
    print(f"two_site f32 {(a, wl, wm, wr, d1, d2, b)}", flush=True)
    rng = np.random.default_rng(11 + a)
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d1, d1)), rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    ref = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    op32 = kernels.fused_two_site(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip),
                                  dtype=torch.float32)
    got = kernels.apply_two_site_fused(_t(th, hip), *op32)
    assert got.dtype == torch.float64
    assert _rel(got, ref) < 2e-5
    got32 = kernels.apply_two_site_fused(_t(th, hip, torch.float32), *op32)
    assert got32.dtype == torch.float32
    assert _rel(got32, ref) < 2e-5


@pytest.mark.parametrize("a,wl,wm,wr,d1,d2,b", SHAPES[3:] + [(7, 20, 20, 20, 2, 3, 9)])
def test_two_site_matvec_bf16_operator(hip, a, wl, wm, wr, d1, d2, b):
    """BF16 operator (LT, RT bf16; S float32) with float64 theta/out: theta and the mixed
    intermediate are rounded to bf16 (8-bit mantissa), so the result is within a few parts in
    1e3 of the exact one; the exact layouts of LT/RT are the bf16-rounded permutations."""
    print(f"two_site bf16 {(a, wl, wm, wr, d1, d2, b)}", flush=True)
    rng = np.random.default_rng(23 + a)
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wm, d1, d1)), rng.standard_normal((wm, wr, d2, d2))
    th = rng.standard_normal((a, d1, d2, b))
    ref = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    op = kernels.fused_two_site(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip),
                                dtype=torch.bfloat16)
    LT, S, RT = op
    assert (LT.dtype, S.dtype, RT.dtype) == (torch.bfloat16, torch.float32, torch.bfloat16)
    Lb = torch.as_tensor(L).to(torch.bfloat16).permute(2, 0, 1)
    Rb = torch.as_tensor(R).to(torch.bfloat16).permute(1, 2, 0)
    assert torch.equal(LT.cpu(), Lb) and torch.equal(RT.cpu(), Rb)
    got = kernels.apply_two_site_fused(_t(th, hip), *op)
    assert got.dtype == torch.float64
    assert _rel(got, ref) < 1.5e-2
    with pytest.raises(TypeError):
        kernels.apply_two_site_fused(_t(th, hip, torch.float32), *op)


@pytest.mark.parametrize("a,d,c,wl,wr", [(1, 2, 2, 1, 5), (2, 2, 4, 5, 5), (7, 2, 9, 3, 4),
                                         (64, 2, 48, 5, 5), (200, 2, 256, 5, 5), (5, 3, 11, 4, 1)])
def test_environment_updates_f64(hip, a, d, c, wl, wr):
    print(f"env {(a, d, c, wl, wr)}", flush=True)
    rng = np.random.default_rng(a * 13 + c)
    L = rng.standard_normal((a, wl, a))
    A = rng.standard_normal((a, d, c))
    W = rng.standard_normal((wl, wr, d, d))
    ref = D.update_left_env(L, A, W)
    got = kernels.update_left_env(_t(L, hip), _t(A, hip), _t(W, hip))
    assert got.shape == ref.shape and _rel(got, ref) < 1e-12
    # right: R (a, wr, a) past B (c, d, a)
    R = rng.standard_normal((a, wr, a))
    B = rng.standard_normal((c, d, a))
    ref = D.update_right_env(R, B, W)
    got = kernels.update_right_env(_t(R, hip), _t(B, hip), _t(W, hip))
    assert got.shape == ref.shape and _rel(got, ref) < 1e-12


@pytest.mark.parametrize("n,nlive,j,nsub", [(1, 1, 0, 0), (1000, 5, 4, 1), (4096, 41, 40, 1),
                                            (70001, 11, 10, 10), (300000, 64, 63, 1),
                                            (12345, 3, 2, 0)])
def test_krylov_orth(hip, n, nlive, j, nsub):
    print(f"krylov_orth n={n} nlive={nlive}", flush=True)
    rng = np.random.default_rng(n + nlive)
    V = np.linalg.qr(rng.standard_normal((n, max(nlive, 1))))[0].T.copy() if n >= nlive else \
        rng.standard_normal((nlive, n))
    w = rng.standard_normal(n)
    coef = rng.standard_normal(nsub)
    base = 0 if nsub != 1 else j - 1 if j > 0 else 0
    # numpy restatement of the step (tnkit/eigensolver.py:560-604)
    x = w - (coef @ V[base:base + nsub] if nsub else 0)
    alpha = V[j] @ x
    x = x - alpha * V[j]
    for _ in range(2):
        x = x - (V[:nlive] @ x) @ V[:nlive]
    beta = np.linalg.norm(x)
    Vd, wd = _t(V, hip), _t(w, hip)
    kw = kernels.KrylovWorkspace(n, hip)
    kernels.krylov_orth(Vd, nlive, wd, j, list(coef), base, kw)
    a_, b_ = kw.scal.cpu().tolist()
    assert abs(a_ - alpha) <= 1e-12 * max(1, abs(alpha)) * np.sqrt(n)
    assert abs(b_ - beta) <= 1e-12 * max(1, beta) * np.sqrt(n)
    assert np.abs(wd.cpu().numpy() - x).max() <= 1e-12 * max(1, np.abs(w).max()) * np.sqrt(n)


def test_lincomb_norm_dot_scale(hip):
    rng = np.random.default_rng(0)
    V = rng.standard_normal((7, 100003))
    c = rng.standard_normal(7)
    out = torch.empty(100003, dtype=torch.float64, device=hip)
    kernels.lincomb(_t(V, hip), list(c), out)
    assert _rel(out, c @ V) < 1e-13
    kernels.lincomb(_t(V, hip), list(c), out, beta=2.0)
    assert _rel(out, 3 * (c @ V)) < 1e-13
    x = rng.standard_normal(100003)
    y = rng.standard_normal(100003)
    assert kernels.norm(_t(x, hip)).item() == pytest.approx(np.linalg.norm(x), rel=1e-13)
    assert kernels.dot(_t(x, hip), _t(y, hip)).item() == pytest.approx(x @ y, rel=1e-11)
    z = torch.empty_like(_t(x, hip))
    kernels.scale_by(_t(x, hip), torch.tensor([4.0], dtype=torch.float64, device=hip), 2.0, z)
    np.testing.assert_allclose(z.cpu().numpy(), x * 0.5, rtol=1e-15)


def test_real_mpo_network_matvec(hip):
    """The matvec of a real XXZ local problem (canonical MPS environments) against the oracle."""
    rng = np.random.default_rng(9)
    N, chi = 10, 24
    mpo = xxz_mpo(N)
    M = D.right_canonicalize(D.random_mps(N, chi, 2, rng))
    Rs = [None] * (N + 1)
    Rs[N] = np.ones((1, 1, 1))
    for i in range(N - 1, 4, -1):
        Rs[i] = D.update_right_env(Rs[i + 1], M[i], mpo[i])
    Ls = np.ones((1, 1, 1))
    Md = [_t(m, hip) for m in M]
    Ld = _t(Ls, hip)
    for i in range(3):
        Ls = D.update_left_env(Ls, M[i], mpo[i])
        Ld = kernels.update_left_env(Ld, Md[i], _t(mpo[i], hip))
    assert _rel(Ld, Ls) < 1e-12
    th = np.einsum("asc,ctb->astb", M[3], M[4])
    ref = D.apply_two_site_fused(th, *D.fused_two_site(Ls, mpo[3], mpo[4], Rs[5]))
    op = kernels.fused_two_site(Ld, _t(mpo[3], hip), _t(mpo[4], hip), _t(Rs[5], hip))
    got = kernels.apply_two_site_fused(_t(th, hip), *op)
    assert _rel(got, ref) < 1e-12


@pytest.mark.parametrize("wl", [3, 5])
@pytest.mark.parametrize("low", [False, True])
def test_batched_real_fused_matvec(hip, wl, low):
    """k legs of the real matvec at shapes the fused GEMM1 + mix kernel takes (d = 2, a and b
    multiples of 32), float64 and float32 operators, against the oracle per leg."""
    print(f"batched real fused wl={wl} low={low}", flush=True)
    rng = np.random.default_rng(40 + wl + 2 * low)
    k, a, wm, wr, d, b = 3, 64, 4, 5, 2, 96
    L, R = rng.standard_normal((k, a, wl, a)), rng.standard_normal((k, b, wr, b))
    W1, W2 = rng.standard_normal((k, wl, wm, d, d)), rng.standard_normal((k, wm, wr, d, d))
    th = rng.standard_normal((k, a, d, d, b))
    op = kernels.fused_two_site_b(_t(L, hip), _t(W1, hip), _t(W2, hip), _t(R, hip),
                                  dtype=torch.float32 if low else None)
    got = kernels.apply_two_site_b(_t(th, hip), *op).cpu().numpy()
    for j in range(k):
        ref = D.apply_two_site_fused(th[j], *D.fused_two_site(L[j], W1[j], W2[j], R[j]))
        assert _rel(got[j], ref) < (2e-5 if low else 1e-12)
