"""The U(1) block maps of tnkit_amd.blocks against tnkit's own, recorded in tests/golden/blocks.npz
(tests/golden/make_block_golden.py): canonical layouts, the four contraction plans of the
two-site matvec, their gather indices and output layouts, bit for bit; plus brute-force checks of
the run decomposition and of the allowed-block enumeration."""
import itertools
import os

import numpy as np
import pytest

from tnkit_amd import blocks as B

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "blocks.npz"))
CASES = ("small", "medium", "odd")
SUBS = ("pwq,qijb->pwijb", "wmsi,pwijb->pmsjb", "mrtj,pmsjb->pstrb", "drb,pstrb->pstd")


def leg(case, name, flip=False):
    k = f"{case}/leg_{name}/"
    lg = B.Leg(tuple(GOLD[k + "charges"].tolist()), tuple(GOLD[k + "dims"].tolist()),
               int(GOLD[k + "sign"]))
    return lg.flip() if flip else lg


def operand_legs(case):
    a, b, wl, wr, p = (leg(case, n) for n in ("a", "b", "wl", "wr", "p"))
    return {"theta": (a, p, p, b.flip()), "L": (a, wl.flip(), a.flip()),
            "R": (b.flip(), wr.flip(), b), "W": (wl, wr, p, p.flip())}


def same_struct(s, case, key):
    g = lambda n: GOLD[f"{case}/{key}/{n}"]
    assert s.size == int(g("size")) and s.charge == int(g("charge"))
    assert np.array_equal(np.array(s.blocks, np.int64).reshape(g("blocks").shape), g("blocks"))
    assert np.array_equal(s.off, g("off"))
    assert np.array_equal(s.st, g("st"))
    assert np.array_equal(s.shp, g("shp"))


@pytest.mark.parametrize("case", CASES)
def test_canonical_layouts(case):
    for name, legs in operand_legs(case).items():
        same_struct(B.canonical(legs, 0), case, name)


def test_allowed_blocks_brute_force():
    """The pruned enumeration equals filtering the full product, in lexicographic order."""
    rng = np.random.default_rng(0)
    for _ in range(40):
        m = int(rng.integers(1, 5))
        legs = []
        for _ in range(m):
            cs = sorted(set(rng.integers(-4, 5, size=int(rng.integers(1, 4))).tolist()))
            legs.append(B.Leg(tuple(cs), tuple([1] * len(cs)), int(rng.choice([-1, 1]))))
        q = int(rng.integers(-3, 4))
        want = tuple(t for t in itertools.product(*(range(len(lg.charges)) for lg in legs))
                     if sum(lg.sign * lg.charges[s] for lg, s in zip(legs, t)) == q)
        assert B.canonical(tuple(legs), q).blocks == want


@pytest.mark.parametrize("case", CASES)
def test_two_site_plans_match_tnkit(case):
    """The four contractions of the two-site matvec: GEMM lists, macs, gather indices and
    output layouts (before and after the einsum's permutation), then the canonical gather."""
    legs = operand_legs(case)
    ops = [B.canonical(legs[n], 0) for n in ("L", "W", "W", "R")]
    x = B.canonical(legs["theta"], 0)
    for c, (sub, A) in enumerate(zip(SUBS, ops)):
        pre = f"c{c + 1}"
        ax_a, ax_b, perm = B.parse_einsum(sub)
        spec = GOLD[f"{case}/{pre}/spec"].tolist()
        assert spec == list(ax_a) + [-1] + list(ax_b) + [-1] + list(perm)
        p = B.plan(A, x, ax_a, ax_b)
        assert np.array_equal(p.gemms, GOLD[f"{case}/{pre}/gemms"])
        assert p.macs == int(GOLD[f"{case}/{pre}/macs"])
        assert p.nmm == int(GOLD[f"{case}/{pre}/nmm"])
        for ref, key, src in ((p.ref_a, "ia", A), (p.ref_b, "ib", x)):
            gold = GOLD[f"{case}/{pre}/{key}"]
            if gold.size == 0:
                assert ref.identity
            else:
                assert not ref.identity
                assert np.array_equal(ref.expand(), gold)
        same_struct(p.out, case, f"{pre}/out")
        x = B.permuted(p.out, perm)
        same_struct(x, case, f"{pre}/perm")
    g = B.to_canonical(x)
    assert np.array_equal(g.expand(), GOLD[f"{case}/final/ix"])
    same_struct(B.canonical(x.legs, x.charge), case, "final")


@pytest.mark.parametrize("case", CASES)
def test_device_stacking_is_the_padded_transpose(case):
    """The device gathers hold the same elements as tnkit's: operand a as the per-group
    transpose (K, R) of tnkit's (R, K) with padded rows, operand b as (K, C) with padded rows;
    every padding position is uncovered (zero-filled by the gather kernel)."""
    legs = operand_legs(case)
    A = B.canonical(legs["L"], 0)
    x = B.canonical(legs["theta"], 0)
    p = B.plan(A, x, (2,), (0,))
    ia, ib = p.ref_a.expand(), p.ref_b.expand()
    da, db = p.dev_a.expand(), p.dev_b.expand()
    covered_a = np.zeros(p.size_a, bool)
    covered_b = np.zeros(p.size_b, bool)
    for (pa, pb, po, R, K, C), (qa, lda, qb, ldb, po2, R2, K2, C2) in zip(p.gemms, p.desc):
        assert (po, R, K, C) == (po2, R2, K2, C2) and lda % 4 == 0 and ldb % 4 == 0
        assert qa % 4 == 0 and qb % 4 == 0 and lda >= R and ldb >= C
        ref = ia[pa:pa + R * K].reshape(R, K).T
        got = da[qa:qa + K * lda].reshape(K, lda)
        assert np.array_equal(got[:, :R], ref)
        covered_a[qa:qa + K * lda].reshape(K, lda)[:, :R] = True
        ref = ib[pb:pb + K * C].reshape(K, C)
        got = db[qb:qb + K * ldb].reshape(K, ldb)
        assert np.array_equal(got[:, :C], ref)
        covered_b[qb:qb + K * ldb].reshape(K, ldb)[:, :C] = True
    assert (da[~covered_a] == -1).all() and (db[~covered_b] == -1).all()


def _explicit(src0, shp, ss, dst0, ds, n):
    idx = np.full(n, -1)
    for b in range(len(shp)):
        for i in np.ndindex(*shp[b]):
            idx[dst0[b] + np.dot(i, ds[b])] = src0[b] + np.dot(i, ss[b])
    return idx


@pytest.mark.parametrize("seed", range(12))
def test_runs_match_explicit_positions(seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, 5))
    shp = rng.integers(1, 5, size=(int(rng.integers(1, 6)), m))
    ss = np.stack([np.array(B._row_major_rows(s[None])[0])[rng.permutation(m)] for s in shp])
    src0 = rng.integers(0, 50, size=len(shp)) * 100
    pad = rng.integers(0, 3, size=len(shp))
    sizes = shp.prod(1) + pad
    dst0 = np.cumsum(sizes) - sizes
    ds = B._row_major_rows(shp)
    runs = B.box_runs(src0, shp, ss, dst0, ds)
    n = int(sizes.sum())
    assert np.array_equal(B.expand_runs(runs, n), _explicit(src0, shp, ss, dst0, ds, n))
    assert (np.diff(runs[2]) > 0).all()                  # destination-sorted, no overlap


def _leg_tuple(lg):
    return (lg.charges, lg.dims, lg.sign)


@pytest.mark.parametrize("case", CASES)
def test_oracle_layouts_and_block_matvec(case):
    """The brute-force oracle reproduces tnkit's canonical layouts, and its dense contraction of
    the densified operands, cut back to blocks, equals tnkit's block matvec (and tnkit's dense
    matvec of the same operands)."""
    from oracle import blocks as OB
    from oracle import dense as D
    legs = operand_legs(case)
    dense = {}
    for name, lg in legs.items():
        t = tuple(_leg_tuple(x) for x in lg)
        blocks, off, _, size = OB.canonical_layout(t)
        assert np.array_equal(np.array(blocks, np.int64).reshape(-1, len(t)),
                              GOLD[f"{case}/{name}/blocks"])
        assert np.array_equal(np.array(off, np.int64), GOLD[f"{case}/{name}/off"])
        assert size == int(GOLD[f"{case}/{name}/size"])
        dense[name] = OB.canonical_to_dense(t, GOLD[f"{case}/{name}/data"])
    y = D.apply_two_site(dense["theta"], dense["L"], dense["W"], dense["W"], dense["R"])
    ref = GOLD[f"{case}/final/dense"]
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
    tl = tuple(_leg_tuple(x) for x in legs["theta"])
    yb = OB.dense_to_canonical(y, tl)
    gold = GOLD[f"{case}/final/data"]
    assert np.abs(yb - gold).max() <= 1e-12 * np.abs(gold).max()


@pytest.mark.parametrize("bm,bn", [(16, 128), (32, 32), (64, 64), (128, 128), (64, 128)])
def test_tile_table_covers_every_tile_once(bm, bn):
    from tnkit_amd.block_ops import tile_table
    rng = np.random.default_rng(bm + bn)
    G = 23
    R = rng.integers(0, 300, G)
    C = rng.integers(0, 700, G)
    K = rng.integers(0, 90, G)
    desc = np.zeros((G, 8), np.int64)
    desc[:, 5], desc[:, 6], desc[:, 7] = R, K, C
    tab = tile_table(desc, bm, bn)
    assert len(tab) % 8 == 0
    live = tab[tab[:, 0] >= 0]
    want = {(g, i, j) for g in range(G) if R[g] and C[g]
            for i in range(-(-R[g] // bm)) for j in range(-(-C[g] // bn))}
    got = [tuple(int(v) for v in r) for r in live]
    assert len(got) == len(set(got)) and set(got) == want


@pytest.mark.parametrize("seed", range(6))
def test_moves_match_explicit_positions(seed):
    """Runs plus transposing tiles cover the same positions as the explicit enumeration, with
    boxes large enough (and transposed enough) for the tile path."""
    rng = np.random.default_rng(100 + seed)
    m = 3
    shp = rng.integers(1, 90, size=(5, m))
    shp[:, 1] = rng.integers(1, 3, size=5)
    ss = np.stack([B._row_major_rows(s[None])[0][rng.permutation(m)] for s in shp])
    src0 = np.cumsum(shp.prod(1)) - shp.prod(1)
    sizes = shp.prod(1) + rng.integers(0, 3, size=5)
    dst0 = np.cumsum(sizes) - sizes
    ds = B._row_major_rows(shp)
    runs, tiles = B.box_moves(src0, shp, ss, dst0, ds)
    n = int(sizes.sum())
    g = B.Gather(runs, n, int(shp.prod(1).sum()), tiles)
    assert np.array_equal(g.expand(), _explicit(src0, shp, ss, dst0, ds, n))
    if seed == 0:
        assert len(tiles) > 0


def test_block_mpo_matches_tnkit():
    """to_block_mpo of the XXZ chain: channel charges, legs and the bulk tensor's buffer equal
    tnkit's (bit for bit)."""
    import torch
    from tnkit_amd import block_mps as BM
    from tnkit_amd.models import xxz_mpo
    mpo, orders = BM.to_block_mpo(xxz_mpo(6), [[1, -1]] * 6, torch.device("cpu"))
    W = mpo[2]
    assert W.legs == (leg("small", "wl"), leg("small", "wr"), leg("small", "p"),
                      leg("small", "p").flip())
    assert np.array_equal(W.data.numpy(), GOLD["small/W/data"])
    chans = BM.mpo_charges(xxz_mpo(6), [[1, -1]] * 6)
    assert [c.tolist() for c in chans] == [[0]] + [[0, 2, -2, 0, 0]] * 5 + [[0]]


def test_random_block_mps_matches_tnkit():
    """The bond sectors, dimensions and random entries of tnkit's random block MPS."""
    import json
    import torch
    from tnkit_amd import block_mps as BM
    from tnkit_amd.models import xxz_mpo
    runs = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "block_dmrg.json")))
    for r in runs:
        mpo, _ = BM.to_block_mpo(xxz_mpo(r["N"]), [[1, -1]] * r["N"], torch.device("cpu"))
        M = BM.random_block_mps(mpo, r["chi"], r["charge"], np.random.default_rng(r["seed"]),
                                torch.device("cpu"))
        assert [[list(t.legs[2].charges), list(t.legs[2].dims)] for t in M] == r["start_bonds"]
        for t, h in zip(M, r["start_norms"]):
            assert abs(float(np.linalg.norm(t.data.numpy())) - float.fromhex(h)) \
                <= 1e-14 * float.fromhex(h)


@pytest.mark.parametrize("case", CASES)
def test_fuse_split_maps_match_tnkit(case):
    """theta fused to (a s1, s2 b), and the factors of tnkit's truncated SVD split back: fused
    legs, layouts and gather indices equal tnkit's; so do the leg positions of the S scaling."""
    legs = operand_legs(case)
    th = B.canonical(legs["theta"], 0)
    rec = B.fuse_record(th.legs, 0, ((0, 1), (2, 3)))
    for g in range(2):
        k = f"{case}/fuse/leg{g}/"
        assert rec.fused[g] == B.Leg(tuple(GOLD[k + "charges"].tolist()),
                                     tuple(GOLD[k + "dims"].tolist()), int(GOLD[k + "sign"]))
    out, g = B.fuse_map(th, rec)
    same_struct(out, case, "fuse")
    assert np.array_equal(g.expand(), GOLD[f"{case}/fuse/ix"])
    bond = B.Leg(tuple(GOLD[f"{case}/tsvd/bond/charges"].tolist()),
                 tuple(GOLD[f"{case}/tsvd/bond/dims"].tolist()), 1)
    U = B.canonical((rec.fused[0], bond.flip()), 0)
    Vh = B.canonical((bond, rec.fused[1]), 0)
    same_struct(U, case, "tsvd/U")
    same_struct(Vh, case, "tsvd/Vh")
    su, gu = B.split_map(U, rec, (0,))
    sv, gv = B.split_map(Vh, rec, (1,))
    same_struct(su, case, "split_u")
    same_struct(sv, case, "split_v")
    assert np.array_equal(gu.expand(), GOLD[f"{case}/split_u/ix"])
    assert np.array_equal(gv.expand(), GOLD[f"{case}/split_v/ix"])
    assert np.array_equal(B.leg_positions(U, 1).expand(), GOLD[f"{case}/scale_u/ix"])
    assert np.array_equal(B.leg_positions(Vh, 0).expand(), GOLD[f"{case}/scale_v/ix"])


def test_trim_equals_one_state_at_a_time():
    """The closed-form trim equals taking one state at a time from the largest sector (the
    first on ties) until the bond holds chi states."""
    from tnkit_amd.block_mps import _trim
    rng = np.random.default_rng(5)
    for _ in range(400):
        n = int(rng.integers(1, 9))
        dims = [int(x) for x in rng.integers(1, 12, n)]
        chi = int(rng.integers(n, sum(dims) + 3))
        ref = list(dims)
        while sum(ref) > chi:
            k = max(range(n), key=lambda i: (ref[i], -i))
            ref[k] -= 1
        assert _trim(dims, chi) == ref, (dims, chi)


def test_mpo_charges_refuses_charge_mixing_channels():
    """X X + Y Y written with real Pauli matrices puts two charges on one channel: refused."""
    from tnkit_amd import block_mps as BM
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.diag([1.0, -1.0])
    Id = np.eye(2)
    W = np.zeros((3, 3, 2, 2))
    W[0, 0], W[0, 1], W[1, 2], W[2, 2], W[0, 2] = Id, X, X, Id, Z
    mpo = [W[:1], W, W[:, 2:]]
    with pytest.raises(ValueError, match="right channel 1"):
        BM.mpo_charges(mpo, [[1, -1]] * 3)
    # a channel the left end never reaches is dead: charge 0, its entries ignored
    W2 = np.zeros((3, 3, 2, 2))
    W2[0, 0], W2[1, 1], W2[0, 1], W2[2, 2] = Id, Id, Z, Id
    W2[2, 0] = X            # would give channel 0 two charges if channel 2 were live
    q = BM.mpo_charges([W2[:1], W2, W2[:, 1:2]], [[1, -1]] * 3)
    assert [c.tolist() for c in q] == [[0], [0, 0, 0], [0, 0, 0], [0]]


@pytest.mark.parametrize("groups", [((0, 1), (2, 3)), ((1, 0), (3, 2)), ((2, 0), (1, 3)),
                                    ((0,), (1, 2, 3)), ((3, 1, 0), (2,))])
def test_fuse_then_full_split_is_identity(groups):
    """Fusing any grouping (permuting ones too) and splitting every fused leg back returns each
    element to its place; a partial split of a leg keeps the other fused leg as it is."""
    legs = (B.Leg((-2, 0, 2), (2, 3, 1)), B.Leg((-1, 1), (1, 2)), B.Leg((-1, 1), (2, 1)),
            B.Leg((-3, -1, 1, 3), (1, 2, 2, 1), -1))
    s = B.canonical(legs, 0)
    rec = B.fuse_record(legs, 0, groups)
    f, g1 = B.fuse_map(s, rec)
    back, g2 = B.split_map(f, rec, (0, 1))
    assert back.legs == legs and back.size == s.size
    assert np.array_equal(g1.expand()[g2.expand()], np.arange(s.size))
    assert sorted(g1.expand().tolist()) == list(range(s.size))
    if rec.perm == (0, 1, 2, 3):
        half, g3 = B.split_map(f, rec, (0,))
        assert half.legs[-1] == rec.fused[1]
        assert sorted(g3.expand().tolist()) == list(range(f.size))
