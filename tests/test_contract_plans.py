"""The equation planner of HipBackend.einsum (tnkit_amd.contract) on the CPU: every equation
tnkit's drivers send through be.einsum takes the library route it should (GEMM, row projection
or permutation) and, with the device kernels replaced by their torch definitions, gives
torch.einsum's result; the device run of the same equations is in tests/test_hip_backend.py."""
import numpy as np
import pytest
import torch

from tnkit_amd import contract

# (equation, operand shapes, route) for tnkit's equations (file:line in tnkit/)
TNKIT = [
    ("pwq,qijb->pwijb", [(6, 3, 5), (5, 2, 2, 7)], "gemm"),          # _dmrg.py:165
    ("wmsi,pwijb->pmsjb", [(3, 4, 2, 2), (6, 3, 2, 2, 7)], "gemm"),  # :166
    ("mrtj,pmsjb->pstrb", [(4, 3, 2, 2), (6, 4, 2, 2, 7)], "gemm"),  # :167
    ("drb,pstrb->pstd", [(8, 3, 7), (6, 2, 2, 3, 7)], "gemm"),        # :168
    ("wmsi,mrtj->strwij", [(3, 4, 2, 2), (4, 3, 2, 2)], "gemm"),      # :179
    ("aky,ysb->aksb", [(5, 3, 6), (6, 2, 4)], "gemm"),                # :196
    ("klst,aktb->alsb", [(3, 3, 2, 2), (5, 3, 2, 4)], "gemm"),        # :197
    ("bly,alsy->asb", [(4, 3, 4), (5, 3, 2, 4)], "gemm"),             # :198
    ("aky,yb->akb", [(5, 3, 6), (6, 4)], "gemm"),                     # :212
    ("bkz,akz->ab", [(4, 3, 4), (5, 3, 4)], "gemm"),                  # :213
    ("pwq,psd->wqsd", [(5, 3, 5), (5, 2, 6)], "gemm"),                # :219
    ("wqsd,wrsy->qdry", [(3, 5, 2, 6), (3, 4, 2, 2)], "gemm"),        # :220
    ("qdry,qye->dre", [(5, 6, 4, 2), (5, 2, 6)], "gemm"),             # :221
    ("xry,usx->ryus", [(6, 4, 6), (5, 2, 6)], "gemm"),                # :228
    ("ryus,wrsp->yuwp", [(4, 6, 5, 2), (3, 4, 2, 2)], "gemm"),        # :229
    ("yuwp,vpy->uwv", [(6, 5, 3, 2), (5, 2, 6)], "gemm"),             # :230
    ("asc,ctb->astb", [(5, 2, 6), (6, 2, 4)], "gemm"),                # :677, tdvp.py:590
    ("npwq,nqijb->npwijb", [(3, 6, 3, 5), (3, 5, 2, 2, 7)], "gemm"),  # batch_tdvp.py:226
    ("nwmsi,npwijb->npmsjb", [(3, 3, 4, 2, 2), (3, 6, 3, 2, 2, 7)], "gemm"),
    ("nmrtj,npmsjb->npstrb", [(3, 4, 3, 2, 2), (3, 6, 4, 2, 2, 7)], "gemm"),
    ("ndrb,npstrb->npstd", [(3, 8, 3, 7), (3, 6, 2, 2, 3, 7)], "gemm"),
    ("naky,nysb->naksb", [(3, 5, 3, 6), (3, 6, 2, 4)], "gemm"),       # :213
    ("nklst,naktb->nalsb", [(3, 3, 3, 2, 2), (3, 5, 3, 2, 4)], "gemm"),
    ("nbly,nalsy->nasb", [(3, 4, 3, 4), (3, 5, 3, 2, 4)], "gemm"),
    ("naky,nyb->nakb", [(3, 5, 3, 6), (3, 6, 4)], "gemm"),            # :220
    ("nbkz,nakz->nab", [(3, 4, 3, 4), (3, 5, 3, 4)], "gemm"),
    ("npwq,npsd->nwqsd", [(3, 5, 3, 5), (3, 5, 2, 6)], "gemm"),       # :234
    ("nwqsd,nwrsy->nqdry", [(3, 3, 5, 2, 6), (3, 3, 4, 2, 2)], "gemm"),
    ("nqdry,nqye->ndre", [(3, 5, 6, 4, 2), (3, 5, 2, 6)], "gemm"),
    ("nxry,nusx->nryus", [(3, 6, 4, 6), (3, 5, 2, 6)], "gemm"),       # :241
    ("nryus,nwrsp->nyuwp", [(3, 4, 6, 5, 2), (3, 3, 4, 2, 2)], "gemm"),
    ("nyuwp,nvpy->nuwv", [(3, 6, 5, 3, 2), (3, 5, 2, 6)], "gemm"),
    ("nrb,nbsc->nrsc", [(3, 4, 5), (3, 5, 2, 6)], "gemm"),            # :535
    ("ncsa,nar->ncsr", [(3, 4, 2, 5), (3, 5, 6)], "gemm"),            # :568
    ("nasc,nctb->nastb", [(3, 5, 2, 6), (3, 6, 2, 4)], "gemm"),       # :623
    ("rb,bsc->rsc", [(4, 5), (5, 2, 6)], "gemm"),                     # tdvp.py:825
    ("csa,ar->csr", [(4, 2, 5), (5, 6)], "gemm"),                     # tdvp.py:850
    ("nasb->nsba", [(3, 5, 2, 4)], "perm"),                           # batch_tdvp.py:552
    ("sbr->rsb", [(2, 4, 5)], "perm"),                                # tdvp.py:843
    ("ra->ar", [(5, 4)], "perm"),                                     # tdvp.py:845
    ("jn,n->j", [(7, 50), (50,)], "rows_dot"),                        # eigensolver.py:160
    ("j,jn->n", [(7,), (7, 50)], "rows_combine"),                     # :161
    ("r,rn->n", [(7,), (7, 50)], "rows_combine"),                     # :567, :672
    ("kn,n->k", [(7, 50), (50,)], "rows_dot"),                        # :588, :594
    ("k,kn->n", [(7,), (7, 50)], "rows_combine"),                     # :589, :595
    ("rc,rn->cn", [(7, 3), (7, 50)], "gemm"),                         # :634
    ("ij,kj->ik", [(4, 50), (3, 50)], "gemm"),                        # :885
]


@pytest.fixture
def torch_kernels(monkeypatch):
    """The device kernels replaced by their definitions, on CPU tensors."""
    monkeypatch.setattr(contract, "_on_device", lambda ops: True)
    monkeypatch.setattr(contract.kernels, "gemm_tn_batched",
                        lambda A, B: torch.bmm(A.transpose(1, 2), B))
    monkeypatch.setattr(contract.kernels, "rows_dot", lambda V, x: V @ x)
    monkeypatch.setattr(contract.kernels, "rows_combine", lambda c, V: c @ V)


@pytest.mark.parametrize("eq,shapes,route", TNKIT, ids=[t[0] for t in TNKIT])
@pytest.mark.parametrize("dt", [torch.float64, torch.complex128])
def test_tnkit_equations_route_and_value(torch_kernels, eq, shapes, route, dt):
    assert contract.plan_of(eq, tuple(len(s) for s in shapes)) == route
    g = torch.Generator().manual_seed(len(eq))
    ops = [torch.randn(*s, dtype=dt, generator=g) for s in shapes]
    got = contract.einsum(eq, *ops)
    ref = torch.einsum(eq, *ops)
    assert got.shape == ref.shape
    assert torch.allclose(got, ref, rtol=1e-13, atol=1e-13)


def test_conjugated_and_strided_operands(torch_kernels):
    """A lazily conjugated operand (be.conj(A) in the environment update) and a non-contiguous
    one are laid out as stored values."""
    g = torch.Generator().manual_seed(1)
    L = torch.randn(5, 3, 5, dtype=torch.complex128, generator=g)
    A = torch.randn(5, 2, 6, dtype=torch.complex128, generator=g)
    got = contract.einsum("pwq,psd->wqsd", L, A.conj())
    assert torch.allclose(got, torch.einsum("pwq,psd->wqsd", L, A.conj()), atol=1e-13)
    X = torch.randn(6, 2, 5, dtype=torch.float64, generator=g).permute(2, 1, 0)
    got = contract.einsum("asc,ctb->astb", X, torch.randn(6, 2, 4, dtype=torch.float64,
                                                         generator=g))
    assert got.shape == (5, 2, 2, 4)


@pytest.mark.parametrize("eq,nd", [("ii->i", (2,)), ("ij,jk,kl->il", (2, 2, 2)),
                                   ("ij,jk->i", (2, 2)), ("...ij,jk->...ik", (3, 2))])
def test_other_equations_go_to_torch(eq, nd):
    assert contract.plan_of(eq, nd) == "torch"


def test_mixed_real_complex_promotes(torch_kernels):
    a = torch.randn(4, 5, dtype=torch.float64)
    b = torch.randn(5, 3, dtype=torch.complex128)
    got = contract.einsum("ij,jk->ik", a, b)
    assert got.dtype == torch.complex128
    assert torch.allclose(got, a.to(torch.complex128) @ b, atol=1e-13)
    assert np.isfinite(got.abs().max().item())
