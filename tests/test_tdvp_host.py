"""Host-side rules of the batched TDVP truncation (no GPU): the eigendecomposition fallback for
padded Gram matrices (one batched SVD with the padding lifted above the spectrum) keeps the
padding out of the kept directions."""
import numpy as np
import pytest
import torch

from tnkit_amd import tdvp as T


def _padded_gram(seed, live, rank, r):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3 * r, rank)) + 1j * rng.standard_normal((3 * r, rank))
    Y = rng.standard_normal((rank, live)) + 1j * rng.standard_normal((rank, live))
    M = np.zeros((3 * r, r), dtype=complex)
    cols = rng.choice(r, size=live, replace=False)
    M[:, cols] = X @ Y * np.logspace(0, -7, live)[None, :]
    G = torch.as_tensor(M.conj().T @ M)
    return 0.5 * (G + G.mH), np.sort(cols)


@pytest.mark.parametrize("live,rank", [(32, 26), (40, 40), (17, 3)])
def test_eigh_by_svd_keeps_padding_out(live, rank):
    r = 128
    G, cols = _padded_gram(live + rank, live, rank, r)
    dg = torch.diagonal(G).real
    shift = -(1.0 + torch.arange(r, dtype=torch.float64) / r)
    add = torch.where(dg == 0, shift, torch.zeros_like(dg))
    w, Z = T._eigh_by_svd(G[None], add[None])
    w, Z = w[0], Z[0]
    Gs = G + torch.diag_embed(add.to(G.dtype))
    scale = max(1.0, float(w.abs().max()))
    assert float((Gs @ Z - Z * w[None, :].to(Z.dtype)).abs().max()) < 1e-12 * scale
    assert float((Z.mH @ Z - torch.eye(r, dtype=Z.dtype)).abs().max()) < 1e-12
    assert bool((w[1:] >= w[:-1]).all())
    # the top `live` directions live on the live columns (also the null ones of a rank-deficient
    # live block), to rounding, and never take a padded unit vector
    top = torch.flip(Z, dims=[1])[:, :live]
    dead = np.setdiff1d(np.arange(r), cols)
    assert float(top[dead].abs().max()) <= 1e-12
    ref = np.linalg.eigvalsh(G.numpy()[np.ix_(cols, cols)])
    assert np.abs(np.sort(w.numpy()[-live:]) - ref).max() < 1e-12 * scale


def _torch_eigh(G, add=None, deferred=False, k=None):
    """The device eigh's contract on the CPU: all eigenvalues ascending, the eigenvectors of the
    k largest (all without k), ascending."""
    Gs = G + (torch.diag_embed(add.to(G.dtype)) if add is not None else 0)
    w, Z = torch.linalg.eigh(Gs)
    if k is not None:
        Z = Z[:, :, Z.shape[2] - k:]
    return (w, Z, None) if deferred else (w, Z)


@pytest.mark.parametrize("iso", ["left", "right"])
@pytest.mark.parametrize("a,b,width", [(12, 10, 16), (12, 30, 16), (30, 10, 16), (30, 30, 50)])
def test_gram_cut_on_the_live_extent(monkeypatch, iso, a, b, width):
    """The truncation formed over the indices some leg fills (rows a < max left rank, columns
    b < max right rank) keeps the same ranks, singular values and product U S Vh as over the
    whole padded bond, with an exact isometry and zero rows on the padding; the eigensolver is
    numpy-grade torch eigh here, the batched device solvers are tested on the GPU."""
    monkeypatch.setattr(T, "_eigh_robust", _torch_eigh)
    rng = np.random.default_rng(a * 100 + b)
    k, d = 5, 2
    ra = np.minimum(np.array([12, 6, 3, 8, 12]), a)
    rb = np.minimum(np.array([7, 10, 4, 5, 2]), b)
    mat = np.zeros((k, a * d, d * b), complex)
    for leg in range(k):
        shp = (ra[leg] * d, d * rb[leg])
        X = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        full = np.zeros((a, d, d, b), complex)
        full[:ra[leg], :, :, :rb[leg]] = X.reshape(ra[leg], d, d, rb[leg])
        mat[leg] = full.reshape(a * d, d * b)
    mt = torch.as_tensor(mat)
    caps = np.array([4, 8, 20, 3, 6])
    ceil = np.minimum(ra * d, rb * d)
    rows = torch.arange(int(ra.max()) * d)
    cols = (torch.arange(d)[:, None] * b + torch.arange(int(rb.max()))[None, :]).reshape(-1)
    U1, S1, V1, r1, _ = T._gram_cut(mt, caps, ceil, 0.0, iso, width)
    U2, S2, V2, r2, _ = T._gram_cut(mt, caps, ceil, 0.0, iso, width, (rows, cols))
    assert U1.shape == U2.shape and V1.shape == V2.shape
    assert (r1 == r2).all()
    for leg in range(k):
        r = int(r1[leg])
        assert float((S1[leg, :r] - S2[leg, :r]).abs().max()) < 1e-12
        p1 = (U1[leg, :, :r] * S1[leg, :r]) @ V1[leg, :r]
        p2 = (U2[leg, :, :r] * S2[leg, :r]) @ V2[leg, :r]
        assert float((p1 - p2).abs().max()) < 1e-10
        iso_m = U2[leg, :, :r] if iso == "left" else V2[leg, :r].mH
        assert float((iso_m.mH @ iso_m - torch.eye(r, dtype=iso_m.dtype)).abs().max()) < 1e-12
        pad_rows = np.ones(a * d, bool)
        pad_rows[:ra[leg] * d] = False
        if iso == "left" and pad_rows.any():            # Householder QR: zero to rounding
            assert float(U2[leg, torch.as_tensor(pad_rows), :r].abs().max()) <= 1e-13


def _mps_vector(M):
    v = M[0]
    for t in M[1:]:
        v = torch.tensordot(v, t, dims=([v.dim() - 1], [0]))
    return v.reshape(-1)


def test_right_canonical_all_batches_legs_of_one_shape():
    """The batched start canonicalisation (one QR per site over the legs whose site shapes agree)
    on the host: every leg ends right-canonical (sum over s, r of M M^H = 1 at every site but the
    first), normalised, and the same state as its input up to normalisation; legs of another
    shape take their own group and the same result as alone."""
    rng = np.random.default_rng(3)
    dims = [1, 2, 4, 6, 4, 2, 1]
    other = [1, 2, 3, 3, 3, 2, 1]
    def rand_mps(ds):
        return [torch.as_tensor(rng.standard_normal((ds[i], 2, ds[i + 1]))
                                + 1j * rng.standard_normal((ds[i], 2, ds[i + 1])))
                for i in range(len(ds) - 1)]
    states = [rand_mps(dims), rand_mps(other), rand_mps(dims), rand_mps(dims)]
    out = T._right_canonical_all(states)
    for M, C in zip(states, out):
        assert [tuple(t.shape) for t in C] == [tuple(t.shape) for t in M]
        for t in C[1:]:
            a, d, b = t.shape
            m = t.reshape(a, d * b)
            assert float((m @ m.mH - torch.eye(a, dtype=m.dtype)).abs().max()) < 1e-13
        v, w = _mps_vector(M), _mps_vector(C)
        assert abs(float(torch.linalg.vector_norm(w)) - 1.0) < 1e-13
        ov = torch.vdot(v, w) / torch.linalg.vector_norm(v)
        assert abs(abs(complex(ov)) - 1.0) < 1e-12
    alone = T._right_canonical(states[1])
    assert all(torch.equal(x, y) for x, y in zip(alone, out[1]))
