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
