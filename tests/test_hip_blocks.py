"""The U(1) block path on the device: run gathers, the grouped GEMM in every tile config, the
block two-site operator against tnkit's recorded matvecs, the float32 operator, and hipGraph
replay."""
import os

import numpy as np
import pytest
import torch

from oracle import blocks as OB
from oracle import dense as D
from tnkit_amd import block_ops as BO
from tnkit_amd import blocks as B

pytestmark = pytest.mark.gpu

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "blocks.npz"))
CASES = ("small", "medium", "odd")


def leg(case, name):
    k = f"{case}/leg_{name}/"
    return B.Leg(tuple(GOLD[k + "charges"].tolist()), tuple(GOLD[k + "dims"].tolist()),
                 int(GOLD[k + "sign"]))


def operands(case, dev):
    a, b, wl, wr, p = (leg(case, n) for n in ("a", "b", "wl", "wr", "p"))
    legs = {"theta": (a, p, p, b.flip()), "L": (a, wl.flip(), a.flip()),
            "R": (b.flip(), wr.flip(), b), "W": (wl, wr, p, p.flip())}
    return {n: BO.BlockTensor(B.canonical(lg, 0),
                              torch.from_numpy(GOLD[f"{case}/{n}/data"]).to(dev))
            for n, lg in legs.items()}


def rel(got, ref):
    got = got.detach().cpu().numpy() if isinstance(got, torch.Tensor) else got
    return float(np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-300))


def test_gather_runs_match_host_expansion(hip):
    rng = np.random.default_rng(1)
    shp = rng.integers(1, 9, size=(40, 3))
    ss = np.stack([B._row_major_rows(s[None])[0][rng.permutation(3)] for s in shp])
    sizes = shp.prod(1)
    src0 = np.cumsum(sizes) - sizes                      # a permuted copy of a packed buffer
    pad = rng.integers(0, 4, size=len(shp))
    dsz = sizes + pad
    dst0 = np.cumsum(dsz) - dsz
    runs = B.box_runs(src0, shp, ss, dst0, B._row_major_rows(shp))
    n = int(dsz.sum())
    g = B.Gather(runs, n, int(sizes.sum()))
    src = torch.randn(int(sizes.sum()), dtype=torch.float64, device=hip)
    idx = B.expand_runs(runs, n)
    ref = np.where(idx >= 0, src.cpu().numpy()[np.maximum(idx, 0)], 0.0)
    for dt in (torch.float64, torch.float32):
        got = BO.gather(g, src, dt).cpu().numpy()
        assert np.array_equal(got, ref.astype(got.dtype))
    got = BO.gather(g, src.float(), torch.float64).cpu().numpy()
    assert np.array_equal(got, ref.astype(np.float32).astype(np.float64))


def long_operands(dev):
    """L and theta with sectors of 40-90 states (K > 32 in the L-theta contraction, the plans
    whose stacks carry K padding for the direct-to-LDS tile configs), random entries."""
    a = B.Leg((-2, -1, 0, 1, 2), (7, 40, 90, 40, 7))
    p = B.Leg((-1, 1), (1, 1))
    wl = B.Leg((-1, 0, 1), (1, 3, 1))
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    out = {}
    for n, lg in (("theta", (a, p, p, a.flip())), ("L", (a, wl.flip(), a.flip()))):
        s = B.canonical(lg, 0)
        out[n] = BO.BlockTensor(s, torch.randn(s.size, dtype=torch.float64, device=dev,
                                               generator=g))
    return out


@pytest.mark.parametrize("case", CASES + ("long",))
@pytest.mark.parametrize("cfg", BO.TILE_CONFIGS)
def test_grouped_gemm_every_tile_config(hip, case, cfg):
    """One contraction (L with theta) in each tile config against per-group torch matmuls."""
    print(f"grouped gemm {case} cfg {cfg}", flush=True)
    ops = long_operands(hip) if case == "long" else operands(case, hip)
    L, th = ops["L"], ops["theta"]
    p = B.plan(L.struct, th.struct, (2,), (0,))
    if cfg in BO.DMA_CONFIGS and p.kpad != B.K_PAD:
        pytest.skip("the direct-to-LDS configs serve plans with K > 32 (padded stacks) only")
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        A = BO.gather(p.dev_a, L.data, dt)
        Bm = BO.gather(p.dev_b, th.data, dt)
        out = torch.full((p.out.size,), float("nan"), dtype=dt, device=hip)
        BO.grouped_gemm(p, A, Bm, out, cfg=cfg)
        ia = torch.from_numpy(p.ref_a.expand()).to(hip)
        ib = torch.from_numpy(p.ref_b.expand()).to(hip)
        Ar, Br = L.data.double()[ia], th.data.double()[ib]
        for pa, pb, po, R, K, C in p.gemms.tolist():
            want = (Ar[pa:pa + R * K].reshape(R, K) @ Br[pb:pb + K * C].reshape(K, C)
                    if K else torch.zeros(R, C, dtype=torch.float64, device=hip))
            got = out[po:po + R * C].reshape(R, C).double()
            assert torch.isfinite(got).all()
            scale = max(float(want.abs().max()) if want.numel() else 0.0, 1.0)
            assert float((got - want).abs().max()) <= tol * scale * max(K, 1)


@pytest.mark.parametrize("case", CASES)
def test_short_k_kernels_on_the_mpo_operator(hip, case):
    """The memory-bound grouped kernels on the two-site operator's middle contraction (W1 and W2
    contracted into one operator, applied to L x theta): the R, K <= 8 kernel where it applies
    and the R, K <= 32 kernel always, against per-group torch matmuls of tnkit's stacking."""
    print(f"short-K kernels {case}", flush=True)
    ops = operands(case, hip)
    X1 = BO.einsum(BO.APPLY_SUBS[0], ops["L"], ops["theta"])
    W12 = BO.einsum(BO.MPO_PAIR_SUB, ops["W"], ops["W"])
    ax_a, ax_b, _ = B.parse_einsum(BO.APPLY_SUBS[1])
    p = B.plan(W12.struct, X1.struct, ax_a, ax_b)
    gs = BO._grouped(p, hip)
    assert gs.mid_ok() and p.kpad == 1
    cfgs = [BO.MID] + ([BO.SMALL] if gs.small_ok() else [])
    ia = torch.from_numpy(p.ref_a.expand()).to(hip)
    ib = torch.from_numpy(p.ref_b.expand()).to(hip)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 2e-6)):
        A = BO.gather(p.dev_a, W12.data, dt)
        Bm = BO.gather(p.dev_b, X1.data, dt)
        Ar, Br = W12.data.double()[ia], X1.data.double()[ib]
        for cfg in cfgs:
            out = torch.full((p.out.size,), float("nan"), dtype=dt, device=hip)
            BO.grouped_gemm(p, A, Bm, out, cfg=cfg)
            for pa, pb, po, R, K, C in p.gemms.tolist():
                want = (Ar[pa:pa + R * K].reshape(R, K) @ Br[pb:pb + K * C].reshape(K, C)
                        if K else torch.zeros(R, C, dtype=torch.float64, device=hip))
                got = out[po:po + R * C].reshape(R, C).double()
                assert torch.isfinite(got).all()
                scale = max(float(want.abs().max()) if want.numel() else 0.0, 1.0)
                assert float((got - want).abs().max()) <= tol * scale * max(K, 1), (cfg, dt)


@pytest.mark.parametrize("case", CASES)
def test_block_two_site_matches_tnkit(hip, case):
    """H_eff theta on blocks equals tnkit's block matvec (recorded) to 1e-12, eagerly in
    float64; the float32 operator with float64 vectors to 5e-6."""
    print(f"block two-site {case}", flush=True)
    ops = operands(case, hip)
    gold = GOLD[f"{case}/final/data"]
    op = BO.BlockTwoSite(ops["L"], ops["W"], ops["W"], ops["R"])
    assert op.in_struct is ops["theta"].struct
    assert op.flops == 2 * sum(int(GOLD[f"{case}/c{c}/macs"]) for c in range(1, 5))
    y = op.apply(ops["theta"].data)
    assert rel(y, gold) < 1e-12
    op32 = BO.BlockTwoSite(ops["L"], ops["W"], ops["W"], ops["R"], dtype=torch.float32)
    y32 = op32.apply(ops["theta"].data)
    assert y32.dtype == torch.float64 and rel(y32, gold) < 5e-6


def _gauss_leg(chi, sigma, sign=1):
    qs = np.arange(-30, 31, 2)
    w = np.exp(-(qs / 2.0) ** 2 / (2 * sigma ** 2))
    dims = np.floor(chi * w / w.sum()).astype(int)
    keep = dims >= 1
    return B.Leg(tuple(int(q) for q in qs[keep]), tuple(int(d) for d in dims[keep]), sign)


def test_block_two_site_random_vs_oracle_and_graph(hip):
    """A Gaussian-sector problem (chi = 160 per bond) against the dense oracle contraction of the
    densified operands; then hipGraph replays, interleaved with eager work, equal eager applies."""
    from tnkit_amd.models import xxz_mpo
    a, b = _gauss_leg(160, 2.0), _gauss_leg(150, 2.2)
    p = B.Leg((-1, 1), (1, 1))
    Wd = np.asarray(xxz_mpo(6)[2])                       # dense bulk W (5, 5, 2, 2)
    # channel charges of the bulk XXZ tensor: I, S+, S-, Sz, I -> 0, 2, -2, 0, 0
    q = np.array([0, 2, -2, 0, 0])
    wleg, order = B.Leg.from_charges(q)
    # physical sector order: charge -1 (down, dense index 1) first
    porder = np.array([1, 0])
    Ws = Wd[np.ix_(order, order, porder, porder)]
    Wlegs = (wleg, wleg.flip(), p, p.flip())
    wdata = OB.dense_to_canonical(Ws, tuple((l.charges, l.dims, l.sign) for l in Wlegs))
    rng = np.random.default_rng(3)
    L = BO.random((a, wleg.flip(), a.flip()), 0, rng, hip)
    R = BO.random((b.flip(), wleg, b), 0, rng, hip)
    W = BO.BlockTensor(B.canonical(Wlegs, 0), torch.from_numpy(wdata).to(hip))
    th = BO.random((a, p, p, b.flip()), 0, rng, hip)
    op = BO.BlockTwoSite(L, W, W, R)
    y = op.apply(th.data)
    T = lambda t: tuple((l.charges, l.dims, l.sign) for l in t.legs)
    dn = lambda t: OB.canonical_to_dense(T(t), t.data.cpu().numpy())
    ref = D.apply_two_site(dn(th), dn(L), Ws, Ws, dn(R))
    refb = OB.dense_to_canonical(ref, T(th))
    assert rel(y, refb) < 1e-12
    run = op.graph()
    xs = [torch.randn(op.size, dtype=torch.float64, device=hip) for _ in range(3)]
    for x in xs:
        eager = op.apply(x)
        torch.cuda.synchronize()
        for _ in range(2):
            g = run(x)
            torch.cuda.synchronize()
            assert torch.equal(g, eager)
        _ = (x * 2.0).sum().item()                       # unrelated eager work between replays


@pytest.mark.parametrize("case", CASES)
def test_small_groups_kernel(hip, case):
    """The MPO contraction (W1 with L theta) through the memory-bound small-group kernel equals
    the grouped MFMA GEMM and per-group torch matmuls."""
    print(f"small groups {case}", flush=True)
    ops = operands(case, hip)
    x1 = BO.tensordot(ops["L"], ops["theta"], ((2,), (0,)))
    W = ops["W"]
    p = B.plan(W.struct, x1.struct, (0, 3), (1, 2))
    assert BO._grouped(p, hip).small_ok()
    for dt, tol in ((torch.float64, 1e-14), (torch.float32, 1e-6)):
        A = BO.gather(p.dev_a, W.data, dt)
        Bm = BO.gather(p.dev_b, x1.data, dt)
        o1 = torch.full((p.out.size,), float("nan"), dtype=dt, device=hip)
        o2 = torch.full((p.out.size,), float("nan"), dtype=dt, device=hip)
        BO.grouped_gemm(p, A, Bm, o1, cfg=BO.SMALL)
        BO.grouped_gemm(p, A, Bm, o2, cfg=4)
        assert torch.isfinite(o1).all()
        ia = torch.from_numpy(p.ref_a.expand()).to(hip)
        ib = torch.from_numpy(p.ref_b.expand()).to(hip)
        Ar, Br = W.data.double()[ia], x1.data.double()[ib]
        ref = torch.zeros(p.out.size, dtype=torch.float64, device=hip)
        for pa, pb, po, R, K, C in p.gemms.tolist():
            if K:
                ref[po:po + R * C] = (Ar[pa:pa + R * K].reshape(R, K)
                                      @ Br[pb:pb + K * C].reshape(K, C)).reshape(-1)
        scale = float(ref.abs().max())
        assert float((o1.double() - ref).abs().max()) <= tol * scale
        assert float((o2.double() - ref).abs().max()) <= tol * scale
