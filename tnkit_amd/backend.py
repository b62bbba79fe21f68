"""`backend="hip"`: the MI355X backend object behind tnkit's backend interface.

`get_backend(spec)` mirrors tnkit/backend.py:192-240 for the specs this package serves:
'hip' (complex128), 'hip-real' (float64), 'auto' (real when the problem is real) and a backend
object, which is returned unchanged. Arrays are torch tensors on the current HIP device; the
hot contractions (`fused_two_site`, `apply_two_site_fused`, the environment updates) and the
Krylov-vector work run in libtnkit_hip. There is no CPU path: constructing the backend without
a HIP device or without the built library raises.
"""
from __future__ import annotations

from typing import Any, Optional

import numpy as np

from . import _lib


def _negligible_imag(a, tol=1e-12):
    """(ok, max|imag|): whether the imaginary part is negligible relative to max(1, |real|),
    the rule of tnkit/backend.py:243-276."""
    a = np.asarray(a)
    if not np.iscomplexobj(a):
        return True, 0.0
    im = np.abs(a.imag)
    m = float(im.max(initial=0.0))
    scale = np.maximum(1.0, np.abs(a.real))
    return bool(np.all(im <= tol * scale)), m


def is_real_problem(*groups, tol=1e-12) -> bool:
    """True when every tensor of every group has a negligible imaginary part
    (tnkit/backend.py:163-189)."""
    for g in groups:
        if g is None:
            continue
        for t in g:
            arr = t.detach().cpu().numpy() if hasattr(t, "detach") else t
            if not _negligible_imag(arr, tol)[0]:
                return False
    return True


def get_backend(spec: Any = "hip", *, mpo=None, init_mps=None, real_tol: float = 1e-12):
    """A backend from its name, or `spec` itself when it is already a backend object.

    'hip-real' -> float64, 'hip' -> complex128, 'auto' -> 'hip-real' when `mpo` (and
    `init_mps`) are real to `real_tol`, else 'hip'. Any other string is a ValueError
    (tnkit/backend.py:240)."""
    if not isinstance(spec, str):
        return spec
    if spec == "auto":
        if mpo is None:
            raise ValueError("get_backend(spec='auto') needs mpo=... to inspect")
        spec = "hip-real" if is_real_problem(mpo, init_mps, tol=real_tol) else "hip"
    if spec == "hip-real":
        return HipBackend(real=True)
    if spec == "hip":
        return HipBackend(real=False)
    raise ValueError(f"unknown backend {spec!r}; this package provides 'hip', 'hip-real' and 'auto'")


class HipBackend:
    """tnkit's backend duck-type (tnkit/backend.py:491-816) on the HIP device."""

    is_torch = True

    def __init__(self, real: bool = True, device: Optional[str] = None):
        import torch
        self.torch = torch
        _lib.load()                                  # fails loudly if the library is missing
        if not torch.cuda.is_available():
            raise RuntimeError("backend 'hip' needs a ROCm GPU (torch.cuda.is_available() is "
                               "False); there is no CPU fallback")
        self.device = torch.device(device) if device else torch.device(
            "cuda", torch.cuda.current_device())
        self.is_complex = not real
        self.dtype = torch.complex128 if self.is_complex else torch.float64
        self.rdtype = torch.float64
        self.name = "hip" if self.is_complex else "hip-real"

    # --- creation / movement ------------------------------------------------------------------
    def array(self, x):
        torch = self.torch
        if isinstance(x, torch.Tensor):
            t = x
            if not self.is_complex and t.is_complex():
                ok, m = _negligible_imag(t.detach().cpu().numpy())
                if not ok:
                    raise TypeError(f"real backend received an array with a significant "
                                    f"imaginary part (max|imag|={m:.3e})")
                t = t.real
            return t.to(device=self.device, dtype=self.dtype).contiguous()
        a = np.asarray(x)
        if not self.is_complex and np.iscomplexobj(a):
            ok, m = _negligible_imag(a)
            if not ok:
                raise TypeError(f"real backend received an array with a significant imaginary "
                                f"part (max|imag|={m:.3e}). Use a complex backend.")
            a = a.real
        return torch.as_tensor(np.ascontiguousarray(a)).to(device=self.device, dtype=self.dtype)

    def to_numpy(self, x):
        if isinstance(x, self.torch.Tensor):
            return x.detach().cpu().numpy()
        return np.asarray(x)

    def randn(self, rng, shape):
        """Host numpy draw, then moved: the same state as the numpy backend for the same rng
        (tnkit/backend.py:258-263)."""
        r = rng.standard_normal(shape)
        if self.is_complex:
            r = r + 1j * rng.standard_normal(shape)
        return self.array(r)

    def zeros(self, shape):
        return self.torch.zeros(shape, dtype=self.dtype, device=self.device)

    def ones(self, shape):
        return self.torch.ones(shape, dtype=self.dtype, device=self.device)

    def zeros_like(self, a):
        return self.torch.zeros_like(a)

    def astype(self, x, which):
        """'low' -> float32 (complex64), 'high' -> float64 (complex128). Unlike tnkit's torch
        backend, a real problem stays real in low precision."""
        torch = self.torch
        if which == "low":
            return x.to(torch.complex64 if x.is_complex() else torch.float32)
        return x.to(torch.complex128 if x.is_complex() else torch.float64)

    # --- elementwise / shape ------------------------------------------------------------------
    def conj(self, a):
        return a.conj() if a.is_complex() else a

    def transpose(self, x, perm):
        return x.permute(*perm)

    def diag(self, v):
        return self.torch.diag(v)

    # --- reductions ---------------------------------------------------------------------------
    def vdot(self, a, b):
        torch = self.torch
        if a.dtype == torch.float64 and b.dtype == torch.float64 and not a.is_complex():
            from . import kernels
            return kernels.dot(a.reshape(-1).contiguous(), b.reshape(-1).contiguous())
        return torch.vdot(a.reshape(-1), b.reshape(-1))

    def norm(self, a):
        if a.dtype == self.torch.float64:
            from . import kernels
            return kernels.norm(a.reshape(-1).contiguous())
        return self.torch.linalg.norm(a.reshape(-1))

    def to_scalar(self, x):
        return complex(x.item()) if hasattr(x, "item") else complex(x)

    def norm_axis(self, x, axis=-1):
        return self.torch.linalg.norm(x, dim=axis)

    # --- contraction ----------------------------------------------------------------------------
    def einsum(self, sub, *ops):
        """tnkit's equations on libtnkit_hip: two-operand contractions as one MFMA GEMM, matrix-
        vector projections on the row kernels (`tnkit_amd.contract`)."""
        from . import contract
        return contract.einsum(sub, *ops)

    def mv(self, A, x):
        """A @ x for a (j, n) matrix and a length-n vector: einsum('jn,n->j', A, x)."""
        from . import contract
        return contract.einsum("jn,n->j", A, x)

    def vm(self, c, A):
        """c @ A for a length-j vector and a (j, n) matrix: einsum('j,jn->n', c, A)."""
        from . import contract
        return contract.einsum("j,jn->n", c, A)

    # --- small dense helpers of the device Krylov exponential (tnkit/backend.py:592-751) --------
    def stack(self, xs):
        torch = self.torch
        return torch.stack([x if isinstance(x, torch.Tensor)
                            else torch.as_tensor(x, device=self.device) for x in xs])

    def stack_axis(self, xs, axis=-1):
        return self.torch.stack(list(xs), dim=axis)

    def tridiag(self, d, e):
        t = self.torch
        T = t.diag(d.to(self.dtype))
        if e is not None and int(e.shape[0]) > 0:
            e = e.to(self.dtype)
            T = T + t.diag(e, 1) + t.diag(e, -1)
        return T

    def guard_denom(self, x, tiny=1e-200):
        return self.torch.clamp(x, min=tiny)

    def expm(self, A):
        return self.torch.matrix_exp(A)

    def expm_host(self, A):
        from scipy.linalg import expm as _expm
        return _expm(A.detach().cpu().numpy())

    def expm_taylor(self, A, squarings, order=18):
        """exp(A) by a truncated Taylor series after `squarings` halvings, then squarings: GEMMs
        only, no host branch (capturable)."""
        t = self.torch
        As = A / (2.0 ** int(squarings))
        E = t.eye(A.shape[-1], dtype=A.dtype, device=A.device)
        term = E
        for k in range(1, int(order) + 1):
            term = (term @ As) / k
            E = E + term
        for _ in range(int(squarings)):
            E = E @ E
        return E

    def _sym_tridiag(self, d, e):
        t = self.torch
        dd = (d.real if t.is_complex(d) else d).to(self.rdtype)
        T = t.diag_embed(dd)
        if e is not None and int(e.shape[-1]) > 0:
            ee = (e.real if t.is_complex(e) else e).to(self.rdtype)
            T = T + t.diag_embed(ee, offset=1) + t.diag_embed(ee, offset=-1)
        return T

    def expm_sym_tridiag_eigh_batched(self, d, e, tau):
        """exp(tau T_n) for a stack of real symmetric tridiagonals (k, m): the eigendecomposition
        on the batched Jacobi kernel (m <= 64), exp(tau lam) applied, V diag V^T back."""
        t = self.torch
        T = self._sym_tridiag(d, e).contiguous()
        lam, V = self.eigh(T)
        tdt = t.complex128 if isinstance(tau, complex) or t.is_tensor(tau) and tau.is_complex() \
            else t.float64
        cdt = t.promote_types(self.dtype, tdt)
        ph = t.exp(t.as_tensor(tau, dtype=cdt, device=T.device) * lam.to(cdt))
        Vc = V.to(cdt)
        return (Vc * ph.unsqueeze(-2)) @ Vc.transpose(-1, -2)

    def expm_sym_tridiag_eigh(self, d, e, tau):
        return self.expm_sym_tridiag_eigh_batched(d.unsqueeze(0),
                                                  None if e is None else e.unsqueeze(0), tau)[0]

    # --- factorisations -----------------------------------------------------------------------
    def svd(self, A):
        from . import kernels
        return kernels.thin_svd(A)

    def svd_batched(self, A):
        from . import kernels
        return kernels.thin_svd(A)

    def qr(self, A):
        """Reduced QR: the library's batched Householder QR for tall float64 (rows <= 2048) and
        complex128 (rows <= 1024) device matrices, the vendor QR otherwise."""
        from . import kernels
        if kernels.qr_supported(A):
            return kernels.qr(A)
        return self.torch.linalg.qr(A, mode="reduced")

    def eigh(self, A):
        """Hermitian eigendecomposition (ascending): the batched Jacobi kernel for orders up to
        64 in float64 / complex128, the library's divide-and-conquer solvers for real float64
        matrices of order 65..2048 and (batched) complex128 ones of order 65..1024, the vendor
        solver otherwise or where a library solver reports a failure."""
        from . import kernels
        t = self.torch
        if (A.is_cuda and A.dim() in (2, 3) and A.dtype in (t.float64, t.complex128)
                and 1 <= A.shape[-1] <= kernels.EIGH_JACOBI_MAX_N):
            G = (A.unsqueeze(0) if A.dim() == 2 else A).resolve_conj().contiguous()
            w, Z, sw = kernels.eigh_jacobi(G)
            if bool((sw > 0).all()):
                return (w[0], Z[0]) if A.dim() == 2 else (w, Z)
        if (A.is_cuda and A.dim() in (2, 3) and A.dtype == t.complex128 and A.shape[0] >= 1
                and kernels.EIGH_JACOBI_MAX_N < A.shape[-1] <= kernels.HEEV_MAX_N):
            G = A.unsqueeze(0) if A.dim() == 2 else A
            G = (0.5 * (G + G.mH)).resolve_conj().contiguous()
            n = int(A.shape[-1])
            w, Z, info = kernels.heev_top_async(G, n)
            if int(info.item()) == 0 and bool(t.isfinite(w).all()):
                Z = t.flip(Z, dims=[2])
                return (w[0], Z[0]) if A.dim() == 2 else (w, Z)
        if (A.is_cuda and A.dim() == 2 and A.dtype == t.float64
                and kernels.EIGH_JACOBI_MAX_N < A.shape[-1] <= kernels.SYEV_MAX_N):
            G = 0.5 * (A + A.t())
            n = int(A.shape[-1])
            w, Z = kernels.syev_top(G.contiguous(), n)
            return w, t.flip(Z, dims=[1])
        return t.linalg.eigh(A)

    def sync(self):
        self.torch.cuda.synchronize(self.device)
