"""Functional wrappers of libtnkit_hip over torch tensors on the current HIP device.

torch supplies device memory and the stream; every contraction here is a libtnkit_hip kernel.
Each function checks shapes, dtypes, devices and contiguity on the host before the launch, so a
mismatch raises ValueError/TypeError instead of reaching a kernel.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import torch

from . import _lib
from ._lib import check, dcode, ptr

_REAL = (torch.float64, torch.float32)
_ANY = (torch.float64, torch.float32, torch.complex128, torch.complex64)
_LOW = {torch.float64: torch.float32, torch.complex128: torch.complex64}


def _need(t: torch.Tensor, name: str, ndim: int, dtypes=_ANY) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: must be on the HIP device, got {t.device}")
    if t.dim() != ndim:
        raise ValueError(f"{name}: expected {ndim} axes, got shape {tuple(t.shape)}")
    if t.dtype not in dtypes:
        raise TypeError(f"{name}: dtype {t.dtype} not supported here ({dtypes})")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")
    if t.is_conj() or t.is_neg():
        raise ValueError(f"{name}: a lazily conjugated/negated view; call .resolve_conj() first "
                         "(the kernels read memory as stored)")


def _ws(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def stream(device=None) -> int:
    return _lib.stream_of(device)


# ---------------------------------------------------------------------------------------------
# GEMM
# ---------------------------------------------------------------------------------------------


def gemm_tn(A: torch.Tensor, B: torch.Tensor, out: Optional[torch.Tensor] = None,
            accumulate: bool = False, tile: int = 0) -> torch.Tensor:
    """out[m, n] (+)= sum_k A[k, m] B[k, n] for 2-D contiguous A (K, M), B (K, N). bfloat16
    A and B run on the bf16 matrix cores with float32 accumulation and a float32 out."""
    _need(A, "A", 2, _ANY + (torch.bfloat16,))
    _need(B, "B", 2, _ANY + (torch.bfloat16,))
    if A.dtype != B.dtype or A.shape[0] != B.shape[0]:
        raise ValueError(f"gemm_tn: A {tuple(A.shape)} {A.dtype} and B {tuple(B.shape)} {B.dtype}")
    K, M = A.shape
    N = B.shape[1]
    odt = torch.float32 if A.dtype == torch.bfloat16 else A.dtype
    if out is None:
        out = torch.empty((M, N), dtype=odt, device=A.device)
        accumulate = False
    _need(out, "out", 2)
    if tuple(out.shape) != (M, N) or out.dtype != odt:
        raise ValueError(f"gemm_tn: out {tuple(out.shape)} {out.dtype}, expected {(M, N)}")
    lib = _lib.load()
    check(lib.tnh_gemm_tn(dcode(A.dtype), M, N, K, ptr(A), M, ptr(B), N, ptr(out), N, 1, 0, 0, 0,
                          int(accumulate), int(tile), stream(A.device)), "gemm_tn")
    return out


# ---------------------------------------------------------------------------------------------
# two-site operator
# ---------------------------------------------------------------------------------------------


def fused_two_site(L: torch.Tensor, W1: torch.Tensor, W2: torch.Tensor, R: torch.Tensor,
                   dtype: Optional[torch.dtype] = None) -> Tuple[torch.Tensor, torch.Tensor,
                                                                 torch.Tensor]:
    """(LT, S, RT) of one two-site problem (tnkit/_dmrg.py:172-180) in the device layouts
    LT (q, p, wl), S (d1 d2 wr, wl d1 d2), RT (wr, b, b'). `dtype` (float32) builds the operator
    of the low-precision matvec from float64 inputs; `dtype=torch.bfloat16` builds the BF16
    operator (LT and RT in bfloat16, S in float32)."""
    for t, n in ((L, "L"), (W1, "W1"), (W2, "W2"), (R, "R")):
        _need(t, n, 3 if n in ("L", "R") else 4)
    a, wl, a2 = L.shape
    b, wr, b2 = R.shape
    wl1, wm, d1, d1b = W1.shape
    wm2, wr2, d2, d2b = W2.shape
    if not (a == a2 and b == b2 and wl1 == wl and wm2 == wm and wr2 == wr and d1 == d1b
            and d2 == d2b):
        raise ValueError(f"fused_two_site: inconsistent shapes L{tuple(L.shape)} "
                         f"W1{tuple(W1.shape)} W2{tuple(W2.shape)} R{tuple(R.shape)}")
    dt = dtype or L.dtype
    if dt == torch.bfloat16 and L.dtype.is_complex:
        raise TypeError("fused_two_site: the bfloat16 operator is real")
    if dt != L.dtype:
        L, R = L.to(dt).contiguous(), R.to(dt).contiguous()
    sdt = torch.float32 if dt == torch.bfloat16 else dt
    W1, W2 = W1.to(sdt).contiguous(), W2.to(sdt).contiguous()
    dev = L.device
    LT = torch.empty((a, a, wl), dtype=dt, device=dev)
    S = torch.empty((d1 * d2 * wr, wl * d1 * d2), dtype=sdt, device=dev)
    RT = torch.empty((wr, b, b), dtype=dt, device=dev)
    check(_lib.load().tnh_fused_two_site(dcode(dt), a, wl, wm, wr, d1, d2, b, ptr(L), ptr(W1),
                                         ptr(W2), ptr(R), ptr(LT), ptr(S), ptr(RT), stream(dev)),
          "fused_two_site")
    return LT, S, RT


def apply_two_site_fused(theta: torch.Tensor, LT: torch.Tensor, S: torch.Tensor,
                         RT: torch.Tensor, out: Optional[torch.Tensor] = None,
                         tiles: Tuple[int, int] = (0, 0), ws: Optional[torch.Tensor] = None,
                         stages: int = 7) -> torch.Tensor:
    """H_eff theta (tnkit/_dmrg.py:183-189) for theta (a, d1, d2, b). The operator's dtype sets
    the arithmetic; theta/out may be float64 around a float32 operator (mixed precision) or
    around the bfloat16 operator of `fused_two_site(..., dtype=torch.bfloat16)`."""
    _need(theta, "theta", 4)
    _need(LT, "LT", 3, _ANY + (torch.bfloat16,))
    _need(S, "S", 2)
    _need(RT, "RT", 3, _ANY + (torch.bfloat16,))
    q, d1, d2, b = theta.shape
    a, p, wl = LT.shape
    wr, bb, bp = RT.shape
    if not (a == q and p == q and bb == b and bp == b and S.shape == (d1 * d2 * wr, wl * d1 * d2)):
        raise ValueError(f"apply_two_site_fused: theta{tuple(theta.shape)} LT{tuple(LT.shape)} "
                         f"S{tuple(S.shape)} RT{tuple(RT.shape)} do not match")
    if LT.dtype == torch.bfloat16:
        if not (RT.dtype == torch.bfloat16 and S.dtype == torch.float32
                and theta.dtype == torch.float64):
            raise TypeError("apply_two_site_fused: the bfloat16 operator takes LT, RT bfloat16, "
                            "S float32 and a float64 theta")
    elif not (LT.dtype == S.dtype == RT.dtype):
        raise TypeError("apply_two_site_fused: LT, S, RT must share a dtype")
    elif theta.dtype != LT.dtype and _LOW.get(theta.dtype) != LT.dtype:
        raise TypeError(f"apply_two_site_fused: theta {theta.dtype} with operator {LT.dtype}")
    if out is None:
        out = torch.empty_like(theta)
    elif out.shape != theta.shape or out.dtype != theta.dtype or not out.is_contiguous():
        raise ValueError("apply_two_site_fused: out must match theta")
    lib = _lib.load()
    dt, io = dcode(LT.dtype), dcode(theta.dtype)
    nbytes = lib.tnh_two_site_ws_bytes(dt, a, d1, d2, b, wl, wr)
    if ws is None or ws.numel() < nbytes:
        ws = _ws(nbytes, theta.device)
    check(lib.tnh_apply_two_site_stages(dt, io, a, d1, d2, b, wl, wr, ptr(LT), ptr(S), ptr(RT),
                                        ptr(theta), ptr(out), ptr(ws), ws.numel(),
                                        int(tiles[0]), int(tiles[1]), int(stages),
                                        stream(theta.device)),
          "apply_two_site_fused")
    return out


def two_site_ws(theta: torch.Tensor, LT: torch.Tensor, S: torch.Tensor, RT: torch.Tensor):
    """A workspace tensor for repeated `apply_two_site_fused` calls on one operator."""
    q, d1, d2, b = theta.shape
    nbytes = _lib.load().tnh_two_site_ws_bytes(dcode(LT.dtype), q, d1, d2, b, LT.shape[2],
                                                RT.shape[0])
    return _ws(nbytes, theta.device)


def update_left_env(L: torch.Tensor, A: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """L' (c, wr, c) past a left-canonical A (a, d, c)  (tnkit/_dmrg.py:216-222)."""
    _need(L, "L", 3)
    _need(A, "A", 3)
    _need(W, "W", 4)
    a, wl, a2 = L.shape
    a3, d, c = A.shape
    wl2, wr, d2, d3 = W.shape
    if not (a == a2 == a3 and wl == wl2 and d == d2 == d3) or not (L.dtype == A.dtype == W.dtype):
        raise ValueError(f"update_left_env: L{tuple(L.shape)} A{tuple(A.shape)} W{tuple(W.shape)}")
    lib = _lib.load()
    dt = dcode(L.dtype)
    out = torch.empty((c, wr, c), dtype=L.dtype, device=L.device)
    ws = _ws(lib.tnh_env_ws_bytes(dt, a, d, c, wl, wr), L.device)
    check(lib.tnh_update_left_env(dt, a, d, c, wl, wr, ptr(L), ptr(A), ptr(W), ptr(out), ptr(ws),
                                  ws.numel(), stream(L.device)), "update_left_env")
    return out


def update_right_env(R: torch.Tensor, B: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """R' (c, wl, c) past a right-canonical B (c, d, b)  (tnkit/_dmrg.py:225-231)."""
    _need(R, "R", 3)
    _need(B, "B", 3)
    _need(W, "W", 4)
    b, wr, b2 = R.shape
    c, d, b3 = B.shape
    wl, wr2, d2, d3 = W.shape
    if not (b == b2 == b3 and wr == wr2 and d == d2 == d3) or not (R.dtype == B.dtype == W.dtype):
        raise ValueError(f"update_right_env: R{tuple(R.shape)} B{tuple(B.shape)} W{tuple(W.shape)}")
    lib = _lib.load()
    dt = dcode(R.dtype)
    out = torch.empty((c, wl, c), dtype=R.dtype, device=R.device)
    ws = _ws(lib.tnh_env_ws_bytes(dt, b, d, c, wl, wr), R.device)
    check(lib.tnh_update_right_env(dt, c, d, b, wl, wr, ptr(R), ptr(B), ptr(W), ptr(out), ptr(ws),
                                   ws.numel(), stream(R.device)), "update_right_env")
    return out


# ---------------------------------------------------------------------------------------------
# Krylov vectors (float64)
# ---------------------------------------------------------------------------------------------


class KrylovWorkspace:
    """Device scratch of the Krylov entries: partial sums and the (alpha, beta) pair."""

    def __init__(self, n: int, device):
        lib = _lib.load()
        self.ws = _ws(lib.tnh_krylov_ws_bytes(n, 64), device)
        self.scal = torch.zeros(2, dtype=torch.float64, device=device)


def _coef(vals: Sequence[float]):
    arr = (ctypes.c_double * max(1, len(vals)))(*[float(v) for v in vals])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_double)), arr


def krylov_orth(V: torch.Tensor, nlive: int, w: torch.Tensor, j: int, sub_coef: Sequence[float],
                sub_base: int, kw: KrylovWorkspace) -> None:
    """In place on w: subtract sum_r sub_coef[r] V[sub_base + r], alpha = <V[j], w>, w -= alpha
    V[j], CGS2 against V[:nlive], beta = |w|; (alpha, beta) land in kw.scal (device)."""
    _need(V, "V", 2, (torch.float64,))
    _need(w, "w", 1, (torch.float64,))
    n = w.numel()
    if V.shape[1] != n or nlive > V.shape[0] or sub_base + len(sub_coef) > V.shape[0]:
        raise ValueError("krylov_orth: basis and vector do not match")
    pc, keep = _coef(sub_coef)
    lib = _lib.load()
    check(lib.tnh_krylov_orth(n, nlive, ptr(V), V.stride(0), ptr(w), j, len(sub_coef), sub_base,
                              pc, ptr(kw.scal), ptr(kw.ws), kw.ws.numel(), stream(w.device)),
          "krylov_orth")
    del keep


def rank_update(V: torch.Tensor, x: torch.Tensor, weight: float, y: torch.Tensor,
                kw: Optional[KrylovWorkspace] = None) -> torch.Tensor:
    """In place: y += weight * sum_r <V[r], x> V[r] for the k <= 64 rows of V (float64; x may be
    y itself). The orthogonality penalty and the projector of excited-state DMRG."""
    _need(V, "V", 2, (torch.float64,))
    _need(x, "x", 1, (torch.float64,))
    _need(y, "y", 1, (torch.float64,))
    n = x.numel()
    if V.shape[1] != n or y.numel() != n:
        raise ValueError(f"rank_update: V {tuple(V.shape)}, x {n}, y {y.numel()} do not match")
    if V.shape[0] > 64:
        raise ValueError(f"rank_update: at most 64 vectors, got {V.shape[0]}")
    kw = kw or KrylovWorkspace(n, x.device)
    check(_lib.load().tnh_rank_update(n, V.shape[0], ptr(V), V.stride(0), ptr(x), float(weight),
                                      ptr(y), ptr(kw.ws), kw.ws.numel(), stream(x.device)),
          "rank_update")
    return y


EIGH_JACOBI_MAX_N = 64


def eigh_jacobi(G: torch.Tensor, add: Optional[torch.Tensor] = None, max_sweeps: int = 40,
                tol: float = 1e-16) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Batched eigh of the Hermitian (k, n, n) complex128 or symmetric float64 G + diag(add),
    n <= 64, on the device: (w ascending (k, n) float64, Z (k, n, n) with eigenvectors as
    columns, sweeps (k,) int32, negative where a matrix did not converge in max_sweeps)."""
    _need(G, "G", 3, (torch.complex128, torch.float64))
    k, n, n2 = G.shape
    if n != n2 or not 1 <= n <= EIGH_JACOBI_MAX_N:
        raise ValueError(f"eigh_jacobi: square matrices of order 1..{EIGH_JACOBI_MAX_N}, got "
                         f"{tuple(G.shape)}")
    if add is not None:
        _need(add, "add", 2, (torch.float64,))
        if tuple(add.shape) != (k, n):
            raise ValueError(f"eigh_jacobi: add {tuple(add.shape)} does not match G {tuple(G.shape)}")
    w = torch.empty((k, n), dtype=torch.float64, device=G.device)
    Z = torch.empty_like(G)
    sw = torch.empty(k, dtype=torch.int32, device=G.device)
    if k:
        check(_lib.load().tnh_eigh_jacobi(dcode(G.dtype), k, n, ptr(G), n, n * n,
                                          ptr(add) if add is not None else None, n, ptr(w), n,
                                          ptr(Z), n, n * n, int(max_sweeps), float(tol), ptr(sw),
                                          stream(G.device)), "eigh_jacobi")
    return w, Z, sw


SYEV_MAX_N = 2048


def syev_top(G: torch.Tensor, k: int, check_status: bool = True) -> Tuple[torch.Tensor, torch.Tensor]:
    """Eigendecomposition of one real symmetric (n, n) float64 matrix, 2 <= n <= 2048, on the
    device by the library's own solver (Householder tridiagonalisation in one cooperative launch,
    divide and conquer, WY back-transformation): (w (n,) ascending, Z (n, k) whose column c is the
    eigenvector of the c-th largest eigenvalue). Both triangles of G are read as given, so pass a
    symmetric matrix. With `check_status` the solver's status word is read back (a host sync) and
    a failed reduction raises."""
    _need(G, "G", 2, (torch.float64,))
    n = int(G.shape[0])
    if G.shape[1] != n or not 2 <= n <= SYEV_MAX_N:
        raise ValueError(f"syev_top: square matrices of order 2..{SYEV_MAX_N}, got {tuple(G.shape)}")
    if not 1 <= int(k) <= n:
        raise ValueError(f"syev_top: k = {k} outside 1..{n}")
    lib = _lib.load()
    w = torch.empty(n, dtype=torch.float64, device=G.device)
    Z = torch.empty((n, int(k)), dtype=torch.float64, device=G.device)
    info = torch.zeros(1, dtype=torch.int32, device=G.device)
    ws = _ws(lib.tnh_syev_top_ws_bytes(n, int(k)), G.device)
    check(lib.tnh_syev_top(n, int(k), ptr(G), n, ptr(w), ptr(Z), int(k), ptr(info), ptr(ws),
                           ws.numel(), stream(G.device)), "syev_top")
    if check_status and int(info.item()) != 0:
        raise _lib.HipError(f"syev_top: the solver reported status {int(info.item())} "
                            "(1: a workgroup wait ran out, 2: a leaf did not converge)")
    return w, Z


HEEV_MAX_N = 1024


def heev_top_async(G: torch.Tensor, k: int) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Eigendecomposition of a batch of complex Hermitian (B, n, n) complex128 matrices,
    2 <= n <= 1024, by the library's own solver (complex Householder reduction to a real
    tridiagonal, divide and conquer, complex WY back-transformation), without a host sync:
    (w (B, n) ascending, Z (B, n, k) whose column c is the eigenvector of the c-th largest
    eigenvalue, info (1,) int32 on the device: 0, or bit 0 when a reduction's bounded wait ran
    out, bit 1 when a leaf eigensolve did not converge, over the whole batch). Both triangles are
    read as given, so pass Hermitian matrices."""
    _need(G, "G", 3, (torch.complex128,))
    B, n = int(G.shape[0]), int(G.shape[1])
    if G.shape[2] != n or not 2 <= n <= HEEV_MAX_N or B < 1:
        raise ValueError(f"heev_top: a batch of square matrices of order 2..{HEEV_MAX_N}, "
                         f"got {tuple(G.shape)}")
    if not 1 <= int(k) <= n:
        raise ValueError(f"heev_top: k = {k} outside 1..{n}")
    lib = _lib.load()
    w = torch.empty((B, n), dtype=torch.float64, device=G.device)
    Z = torch.empty((B, n, int(k)), dtype=torch.complex128, device=G.device)
    info = torch.zeros(1, dtype=torch.int32, device=G.device)
    with torch.cuda.device(G.device):
        nb = lib.tnh_heev_top_ws_bytes(B, n, int(k))
    if nb == 0:
        raise _lib.HipError(f"heev_top: no workspace size for batch {B}, order {n}, k {k}")
    ws = _ws(nb, G.device)
    check(lib.tnh_heev_top(B, n, int(k), ptr(G), n, n * n, ptr(w), n, ptr(Z), int(k),
                           n * int(k), ptr(info), ptr(ws), ws.numel(), stream(G.device)),
          "heev_top")
    return w, Z, info


def heev_top(G: torch.Tensor, k: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """`heev_top_async` followed by a read of its status word: raises if the solve failed."""
    w, Z, info = heev_top_async(G, k)
    if int(info.item()) != 0:
        raise _lib.HipError(f"heev_top: the solver reported status {int(info.item())} "
                            "(1: a workgroup wait ran out, 2: a leaf did not converge)")
    return w, Z


QR_MAX_M = {torch.float64: 2048, torch.complex128: 1024}


def qr_supported(A: torch.Tensor) -> bool:
    """Whether `qr` serves A: a device tensor, float64 rows up to 2048 or complex128 rows up to
    1024, 1 <= n <= m."""
    if A.device.type != "cuda" or A.dtype not in QR_MAX_M or A.dim() not in (2, 3):
        return False
    m, n = int(A.shape[-2]), int(A.shape[-1])
    return 1 <= n <= m <= QR_MAX_M[A.dtype] and (A.dim() == 2 or int(A.shape[0]) >= 1)


def qr_any(A: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """Reduced QR on the library's kernels where `qr` serves A, rocSOLVER's otherwise (wide
    matrices, single precision, rows beyond the limits)."""
    if qr_supported(A):
        return qr(A)
    return torch.linalg.qr(A)


def qr(A: torch.Tensor, check_status: bool = True) -> Tuple[torch.Tensor, torch.Tensor]:
    """Thin QR of a tall (m, n) or batched (B, m, n) float64 / complex128 matrix on the library's
    kernels: LAPACK geqrf's reflectors in one cooperative launch per group of resident matrices,
    Q = H_0 ... H_{n-1} [I; 0] by WY blocks. Returns (Q (…, m, n), R (…, n, n) upper with a real
    diagonal), the factors numpy's and torch's QR return up to rounding. With `check_status` the
    status word is read back (a host sync) and a failed reduction raises."""
    if not qr_supported(A):
        raise ValueError(f"qr: float64 (m <= 2048) or complex128 (m <= 1024) with n <= m, got "
                         f"{A.dtype} {tuple(A.shape)}")
    one = A.dim() == 2
    Ab = (A.unsqueeze(0) if one else A).resolve_conj().resolve_neg().contiguous()
    _need(Ab, "A", 3, tuple(QR_MAX_M))
    B, m, n = (int(x) for x in Ab.shape)
    lib = _lib.load()
    Q = torch.empty((B, m, n), dtype=A.dtype, device=A.device)
    R = torch.empty((B, n, n), dtype=A.dtype, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    with torch.cuda.device(A.device):
        nb = lib.tnh_qr_ws_bytes(dcode(A.dtype), B, m, n)
    if nb == 0:
        raise _lib.HipError(f"qr: no workspace size for batch {B}, {m} x {n}")
    ws = _ws(nb, A.device)
    check(lib.tnh_qr(dcode(A.dtype), B, m, n, ptr(Ab), n, m * n, ptr(Q), n, m * n, ptr(R), n,
                     n * n, ptr(info), ptr(ws), ws.numel(), stream(A.device)), "qr")
    if check_status and int(info.item()) != 0:
        raise _lib.HipError("qr: a workgroup wait of the reduction ran out")
    return (Q[0], R[0]) if one else (Q, R)


def eigh_top(G: torch.Tensor, k: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """(w ascending (n,), Z (n, k) for the k largest eigenvalues, descending) of a symmetric
    float64 (n, n) matrix on the device: the batched Jacobi kernel up to order 64, the library's
    divide-and-conquer solver up to 2048. Larger orders raise."""
    n = int(G.shape[0])
    if n <= EIGH_JACOBI_MAX_N:
        w, V, sw = eigh_jacobi(G.unsqueeze(0).contiguous())
        if int(sw[0].item()) < 0:
            raise _lib.HipError("eigh_top: the Jacobi kernel did not converge")
        return w[0], torch.flip(V[0][:, n - int(k):], dims=[1])
    return syev_top(G.contiguous(), int(k))


SVD_JACOBI_LDS_MAX = 163840 - 1024


def svd_jacobi_fits(m: int, n: int, dtype: torch.dtype) -> bool:
    """Whether `svd_jacobi` takes an m x n matrix of this dtype (min(m, n) <= 64 and the work
    matrix with its rotations in one workgroup's LDS)."""
    if dtype not in (torch.float64, torch.complex128) or min(m, n) < 1 or min(m, n) > 64:
        return False
    return _lib.load().tnh_svd_jacobi_lds_bytes(dcode(dtype), m, n) <= SVD_JACOBI_LDS_MAX


def svd_jacobi(A: torch.Tensor, max_sweeps: int = 60, tol: Optional[float] = None
               ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """Thin SVD of an (m, n) or batched (k, m, n) float64 / complex128 matrix on the device by
    one-sided Jacobi: (U (.., m, r), S (.., r) descending, Vh (.., r, n), sweeps), r = min(m, n),
    sweeps negative where a matrix did not converge in max_sweeps. U and Vh are isometries
    (exactly zero singular values get an orthonormal completion)."""
    two = A.dim() == 2
    Ab = A.unsqueeze(0) if two else A
    _need(Ab, "A", 3, (torch.float64, torch.complex128))
    k, m, n = Ab.shape
    if not svd_jacobi_fits(m, n, A.dtype):
        raise ValueError(f"svd_jacobi: {m} x {n} {A.dtype} does not fit one workgroup")
    r = min(m, n)
    tol = 4e-16 * max(m, n) ** 0.5 if tol is None else float(tol)
    S = torch.empty((k, r), dtype=torch.float64, device=A.device)
    U = torch.empty((k, m, r), dtype=A.dtype, device=A.device)
    Vh = torch.empty((k, r, n), dtype=A.dtype, device=A.device)
    sw = torch.empty(k, dtype=torch.int32, device=A.device)
    if k:
        check(_lib.load().tnh_svd_jacobi(dcode(A.dtype), k, m, n, ptr(Ab), n, m * n, ptr(S), r,
                                         ptr(U), r, m * r, ptr(Vh), n, r * n, int(max_sweeps),
                                         tol, ptr(sw), stream(A.device)), "svd_jacobi")
    if two:
        return U[0], S[0], Vh[0], sw
    return U, S, Vh, sw


def thin_svd(A: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """torch.linalg.svd(A, full_matrices=False) for a device matrix: the Jacobi kernel when the
    matrix fits it (a small SVD there costs tens of microseconds against about a millisecond in
    rocSOLVER), rocSOLVER otherwise or for a matrix the kernel leaves unconverged or non-finite
    (which then raises or returns what torch does for it)."""
    if (A.is_cuda and A.dim() in (2, 3) and _JACOBI_SVD
            and svd_jacobi_fits(int(A.shape[-2]), int(A.shape[-1]), A.dtype)):
        Ac = A.resolve_conj().contiguous()
        U, S, Vh, sw = svd_jacobi(Ac)
        if not bool((sw <= 0).any()) and bool(torch.isfinite(S).all()):
            return U, S, Vh
    return torch.linalg.svd(A, full_matrices=False)


_JACOBI_SVD = os.environ.get("TNKIT_AMD_JACOBI", "1") != "0"


def lincomb(V: torch.Tensor, coef: Sequence[float], out: torch.Tensor, beta: float = 0.0) -> torch.Tensor:
    """out = beta * out + sum_r coef[r] V[r] (coefficients from the host)."""
    _need(V, "V", 2, (torch.float64,))
    _need(out, "out", 1, (torch.float64,))
    if len(coef) > V.shape[0] or V.shape[1] != out.numel():
        raise ValueError("lincomb: sizes do not match")
    pc, keep = _coef(coef)
    check(_lib.load().tnh_lincomb(out.numel(), len(coef), ptr(V), V.stride(0), pc, float(beta),
                                  ptr(out), stream(out.device)), "lincomb")
    del keep
    return out


def norm(x: torch.Tensor, kw: Optional[KrylovWorkspace] = None) -> torch.Tensor:
    """|x| as a 0-d float64 device tensor (deterministic two-stage reduction)."""
    _need(x, "x", 1, (torch.float64,))
    kw = kw or KrylovWorkspace(x.numel(), x.device)
    out = torch.empty((), dtype=torch.float64, device=x.device)
    check(_lib.load().tnh_norm(x.numel(), ptr(x), ptr(out), ptr(kw.ws), kw.ws.numel(),
                               stream(x.device)), "norm")
    return out


def dot(x: torch.Tensor, y: torch.Tensor, kw: Optional[KrylovWorkspace] = None) -> torch.Tensor:
    """<x, y> as a 0-d float64 device tensor."""
    _need(x, "x", 1, (torch.float64,))
    _need(y, "y", 1, (torch.float64,))
    if x.numel() != y.numel():
        raise ValueError("dot: sizes differ")
    kw = kw or KrylovWorkspace(x.numel(), x.device)
    out = torch.empty((), dtype=torch.float64, device=x.device)
    check(_lib.load().tnh_dot(x.numel(), ptr(x), ptr(y), ptr(out), ptr(kw.ws), kw.ws.numel(),
                              stream(x.device)), "dot")
    return out


def scale_by(x: torch.Tensor, div: Optional[torch.Tensor], mul: float, out: torch.Tensor) -> torch.Tensor:
    """out = x * mul / max(div, 1e-200), div a 1-element device tensor (or None)."""
    _need(x, "x", 1, (torch.float64,))
    _need(out, "out", 1, (torch.float64,))
    check(_lib.load().tnh_scale_by(x.numel(), ptr(x), ptr(div) if div is not None else None,
                                   float(mul), ptr(out), stream(x.device)), "scale_by")
    return out


# ---------------------------------------------------------------------------------------------
# batched forms: a leading leg axis (tnkit/batch_tdvp.py:211-243)
# ---------------------------------------------------------------------------------------------


def _bstride(t: torch.Tensor) -> int:
    return int(t.stride(0)) if t.shape[0] > 1 else int(t[0].numel())


def _need_b(t: torch.Tensor, name: str, ndim: int) -> None:
    """A batched operand: every leg contiguous, legs at a fixed stride (a row-strided view of a
    Krylov basis qualifies)."""
    if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dim() != ndim:
        raise ValueError(f"{name}: expected a {ndim}-axis tensor on the HIP device")
    if t.dtype not in _ANY:
        raise TypeError(f"{name}: dtype {t.dtype} not supported")
    if not t[0].is_contiguous() or (t.shape[0] > 1 and t.stride(0) < t[0].numel()):
        raise ValueError(f"{name}: each leg must be contiguous and legs must not overlap")
    if t.is_conj() or t.is_neg():
        raise ValueError(f"{name}: a lazily conjugated/negated view; call .resolve_conj() first")


def fused_two_site_b(L: torch.Tensor, W1: torch.Tensor, W2: torch.Tensor, R: torch.Tensor,
                     dtype: Optional[torch.dtype] = None):
    """`fused_two_site` for k legs stacked on axis 0: L (k, a, wl, a), W1 (k, wl, wm, d1, d1),
    W2 (k, wm, wr, d2, d2), R (k, b, wr, b)."""
    for t, n, nd in ((L, "L", 4), (W1, "W1", 5), (W2, "W2", 5), (R, "R", 4)):
        _need(t, n, nd)
    k, a, wl, _ = L.shape
    _, b, wr, _ = R.shape
    _, _, wm, d1, _ = W1.shape
    d2 = W2.shape[3]
    if not (W1.shape[0] == W2.shape[0] == R.shape[0] == k):
        raise ValueError("fused_two_site_b: the leg axes differ")
    dt = dtype or L.dtype
    if dt != L.dtype:
        L, R = L.to(dt), R.to(dt)
    W1, W2 = W1.to(dt).contiguous(), W2.to(dt).contiguous()
    dev = L.device
    LT = torch.empty((k, a, a, wl), dtype=dt, device=dev)
    S = torch.empty((k, d1 * d2 * wr, wl * d1 * d2), dtype=dt, device=dev)
    RT = torch.empty((k, wr, b, b), dtype=dt, device=dev)
    check(_lib.load().tnh_fused_two_site_batched(
        dcode(dt), k, a, wl, wm, wr, d1, d2, b, ptr(L), _bstride(L), ptr(W1), _bstride(W1),
        ptr(W2), _bstride(W2), ptr(R), _bstride(R), ptr(LT), _bstride(LT), ptr(S), _bstride(S),
        ptr(RT), _bstride(RT), stream(dev)), "fused_two_site_b")
    return LT, S, RT


def apply_two_site_b(theta: torch.Tensor, LT: torch.Tensor, S: torch.Tensor, RT: torch.Tensor,
                     out: Optional[torch.Tensor] = None, ws: Optional[torch.Tensor] = None,
                     live: Optional[torch.Tensor] = None):
    """H_eff theta for k legs: theta (k, a, d1, d2, b) with the operator of `fused_two_site_b`.
    `live` (k, 2) int32 on the device: each leg's live left and right bonds in a padded batch
    (theta, L and R zero outside them); complex vectors then go through the ragged entry, whose
    products skip what lies outside each leg's live block."""
    _need_b(theta, "theta", 5)
    k, q, d1, d2, b = theta.shape
    if LT.shape[:2] != (k, q) or RT.shape[0] != k or RT.shape[2] != b:
        raise ValueError("apply_two_site_b: operator and theta do not match")
    if theta.dtype != LT.dtype and _LOW.get(theta.dtype) != LT.dtype:
        raise TypeError(f"apply_two_site_b: theta {theta.dtype} with operator {LT.dtype}")
    wl, wr = LT.shape[3], RT.shape[1]
    lib = _lib.load()
    dt, io = dcode(LT.dtype), dcode(theta.dtype)
    nbytes = k * lib.tnh_two_site_ws_bytes(dt, q, d1, d2, b, wl, wr)
    if ws is None or ws.numel() < nbytes:
        ws = _ws(nbytes, theta.device)
    if out is None:
        out = torch.empty(theta.shape, dtype=theta.dtype, device=theta.device)
    if live is not None and theta.is_complex():
        if (live.dtype != torch.int32 or tuple(live.shape) != (k, 2) or not live.is_contiguous()
                or live.device != theta.device):
            raise ValueError(f"apply_two_site_b: live must be a contiguous (k, 2) int32 device "
                             f"tensor, got {tuple(live.shape)} {live.dtype}")
        check(lib.tnh_apply_two_site_ragged(
            dt, io, k, q, d1, d2, b, wl, wr, ptr(LT), _bstride(LT), ptr(S), _bstride(S), ptr(RT),
            _bstride(RT), ptr(theta), _bstride(theta), ptr(out), _bstride(out), ptr(live),
            ptr(ws), ws.numel(), stream(theta.device)), "apply_two_site_b")
        return out
    check(lib.tnh_apply_two_site_batched(
        dt, io, k, q, d1, d2, b, wl, wr, ptr(LT), _bstride(LT), ptr(S), _bstride(S), ptr(RT),
        _bstride(RT), ptr(theta), _bstride(theta), ptr(out), _bstride(out), ptr(ws), ws.numel(), 0, 0,
        7, stream(theta.device)), "apply_two_site_b")
    return out


def update_left_env_b(L: torch.Tensor, A: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """update_left_env for k legs: L (k, a, wl, a), A (k, a, d, c), W (k, wl, wr, d, d)."""
    _need(L, "L", 4)
    _need(A, "A", 4)
    _need(W, "W", 5)
    k, a, wl, _ = L.shape
    _, _, d, c = A.shape
    wr = W.shape[2]
    lib = _lib.load()
    dt = dcode(L.dtype)
    out = torch.empty((k, c, wr, c), dtype=L.dtype, device=L.device)
    ws = _ws(k * lib.tnh_env_ws_bytes(dt, a, d, c, wl, wr), L.device)
    check(lib.tnh_update_left_env_batched(dt, k, a, d, c, wl, wr, ptr(L), _bstride(L), ptr(A),
                                          _bstride(A), ptr(W), _bstride(W), ptr(out),
                                          _bstride(out), ptr(ws), ws.numel(), stream(L.device)),
          "update_left_env_b")
    return out


def update_right_env_b(R: torch.Tensor, B: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """update_right_env for k legs: R (k, b, wr, b), B (k, c, d, b), W (k, wl, wr, d, d)."""
    _need(R, "R", 4)
    _need(B, "B", 4)
    _need(W, "W", 5)
    k, b, wr, _ = R.shape
    _, c, d, _ = B.shape
    wl = W.shape[1]
    lib = _lib.load()
    dt = dcode(R.dtype)
    out = torch.empty((k, c, wl, c), dtype=R.dtype, device=R.device)
    ws = _ws(k * lib.tnh_env_ws_bytes(dt, b, d, c, wl, wr), R.device)
    check(lib.tnh_update_right_env_batched(dt, k, c, d, b, wl, wr, ptr(R), _bstride(R), ptr(B),
                                           _bstride(B), ptr(W), _bstride(W), ptr(out),
                                           _bstride(out), ptr(ws), ws.numel(), stream(R.device)),
          "update_right_env_b")
    return out


def expm_step_c(V: torch.Tensor, w: torch.Tensor, ab: torch.Tensor, j: int, ws: torch.Tensor,
                ragged=None):
    """Step j of the batched complex Krylov exponential (see tnh_expm_step_c): V (k, m, n),
    w (k, n) complex128, ab (k, 2m) float64 on the device. `ragged` = (live (k, 2) int32 device
    tensor, rowlen, bpad): the legs' live extents in a padded batch (tnh_expm_step_c_ragged)."""
    k, m, n = V.shape
    if ragged is not None:
        live, rowlen, bpad = ragged
        check(_lib.load().tnh_expm_step_c_ragged(k, n, m, j, ptr(V), V.stride(1), V.stride(0),
                                                 ptr(w), w.stride(0), ptr(ab), ab.stride(0),
                                                 ptr(live), int(rowlen), int(bpad), ptr(ws),
                                                 ws.numel(), stream(V.device)), "expm_step_c")
        return
    check(_lib.load().tnh_expm_step_c(k, n, m, j, ptr(V), V.stride(1), V.stride(0), ptr(w),
                                      w.stride(0), ptr(ab), ab.stride(0), ptr(ws), ws.numel(),
                                      stream(V.device)), "expm_step_c")


def expm_ws(k: int, device) -> torch.Tensor:
    return _ws(_lib.load().tnh_expm_ws_bytes(k), device)


def lincomb_c(V: torch.Tensor, c: torch.Tensor, out: torch.Tensor, nrows: int,
              ragged=None) -> torch.Tensor:
    """out[leg] = sum_{r<nrows} c[leg, r] V[leg, r] for complex128 V (k, m, n), c (k, m).
    `ragged` = (live, rowlen, bpad) as for `expm_step_c`: each leg's padding is written as zeros
    without being read."""
    k, m, n = V.shape
    cr = torch.view_as_real(c.contiguous())
    if ragged is not None:
        live, rowlen, bpad = ragged
        check(_lib.load().tnh_lincomb_c_ragged(k, n, nrows, ptr(V), V.stride(1), V.stride(0),
                                               ptr(cr), cr.stride(0), ptr(out), out.stride(0),
                                               ptr(live), int(rowlen), int(bpad),
                                               stream(V.device)), "lincomb_c")
        return out
    check(_lib.load().tnh_lincomb_c(k, n, nrows, ptr(V), V.stride(1), V.stride(0), ptr(cr),
                                    cr.stride(0), ptr(out), out.stride(0), stream(V.device)),
          "lincomb_c")
    return out


# ---------------------------------------------------------------------------------------------
# batched GEMM, row projections, backward entries, graphs
# ---------------------------------------------------------------------------------------------


def gemm_tn_batched(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """out[z, m, n] = sum_k A[z, k, m] B[z, k, n] for contiguous A (Z, K, M), B (Z, K, N) of one
    dtype (f64, f32, c128, c64): `tnh_gemm_tn` batched over z."""
    _need(A, "A", 3)
    _need(B, "B", 3)
    if A.dtype != B.dtype or A.shape[0] != B.shape[0] or A.shape[1] != B.shape[1]:
        raise ValueError(f"gemm_tn_batched: A {tuple(A.shape)} {A.dtype} and B {tuple(B.shape)} "
                         f"{B.dtype}")
    Z, K, M = A.shape
    N = B.shape[2]
    out = torch.empty((Z, M, N), dtype=A.dtype, device=A.device)
    if Z == 0 or M == 0 or N == 0:
        return out
    if K == 0:
        return out.zero_()
    lib = _lib.load()
    check(lib.tnh_gemm_tn(dcode(A.dtype), M, N, K, ptr(A), M, ptr(B), N, ptr(out), N, Z, K * M,
                          K * N, M * N, 0, 0, stream(A.device)), "gemm_tn_batched")
    return out


def rows_dot(V: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """out[r] = sum_i V[r, i] x[i] (no conjugation) for contiguous V (k, n), x (n,)."""
    _need(V, "V", 2)
    _need(x, "x", 1)
    if V.dtype != x.dtype or V.shape[1] != x.shape[0]:
        raise ValueError(f"rows_dot: V {tuple(V.shape)} {V.dtype}, x {tuple(x.shape)} {x.dtype}")
    k, n = V.shape
    out = torch.empty(k, dtype=V.dtype, device=V.device)
    if k == 0:
        return out
    lib = _lib.load()
    ws = _ws(lib.tnh_rows_ws_bytes(dcode(V.dtype)), V.device)
    check(lib.tnh_rows_dot(dcode(V.dtype), k, n, ptr(V), n, ptr(x), ptr(out), ptr(ws), ws.numel(),
                           stream(V.device)), "rows_dot")
    return out


def rows_combine(c: torch.Tensor, V: torch.Tensor) -> torch.Tensor:
    """out[i] = sum_r c[r] V[r, i] for contiguous c (k,), V (k, n), c on the device."""
    _need(c, "c", 1)
    _need(V, "V", 2)
    if V.dtype != c.dtype or V.shape[0] != c.shape[0]:
        raise ValueError(f"rows_combine: c {tuple(c.shape)} {c.dtype}, V {tuple(V.shape)} "
                         f"{V.dtype}")
    k, n = V.shape
    out = torch.empty(n, dtype=V.dtype, device=V.device)
    if k == 0:
        return out.zero_()
    check(_lib.load().tnh_rows_combine(dcode(V.dtype), k, n, ptr(c), ptr(V), n, ptr(out),
                                       stream(V.device)), "rows_combine")
    return out


def matvec2_bwd(L: torch.Tensor, W1: torch.Tensor, W2: torch.Tensor, R: torch.Tensor,
                ybar: torch.Tensor) -> torch.Tensor:
    """theta_bar = H_eff^H ybar for the two-site operator of (L, W1, W2, R), by the library's
    adjoint-operator build and the forward kernels (`tnh_matvec2_bwd`)."""
    for t, n, nd in ((L, "L", 3), (W1, "W1", 4), (W2, "W2", 4), (R, "R", 3), (ybar, "ybar", 4)):
        _need(t, n, nd)
    a, wl, _ = L.shape
    _, wm, d1, _ = W1.shape
    wr, d2 = W2.shape[1], W2.shape[2]
    b = R.shape[0]
    if (tuple(ybar.shape) != (a, d1, d2, b) or len({t.dtype for t in (L, W1, W2, R, ybar)}) != 1
            or W2.shape[0] != wm or R.shape[1] != wr):
        raise ValueError("matvec2_bwd: shapes or dtypes do not match")
    lib = _lib.load()
    dt = dcode(L.dtype)
    ws = _ws(lib.tnh_matvec2_bwd_ws_bytes(dt, a, d1, d2, b, wl, wr), L.device)
    out = torch.empty_like(ybar)
    check(lib.tnh_matvec2_bwd(dt, a, d1, d2, b, wl, wm, wr, ptr(L), ptr(W1), ptr(W2), ptr(R),
                              ptr(ybar), ptr(out), ptr(ws), ws.numel(), stream(L.device)),
          "matvec2_bwd")
    return out


def env_left_bwd(Lbar_out: torch.Tensor, A: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """The gradient of L for L' = update_left_env(L, A, W) given the cotangent of L' (torch's
    convention for complex data), by `tnh_env_left_bwd`."""
    for t, n, nd in ((Lbar_out, "Lbar_out", 3), (A, "A", 3), (W, "W", 4)):
        _need(t, n, nd)
    a, d, c = A.shape
    wl, wr = W.shape[0], W.shape[1]
    if tuple(Lbar_out.shape) != (c, wr, c) or len({Lbar_out.dtype, A.dtype, W.dtype}) != 1:
        raise ValueError("env_left_bwd: shapes or dtypes do not match")
    lib = _lib.load()
    dt = dcode(A.dtype)
    ws = _ws(lib.tnh_env_bwd_ws_bytes(dt, a, d, c, wl, wr), A.device)
    out = torch.empty((a, wl, a), dtype=A.dtype, device=A.device)
    check(lib.tnh_env_left_bwd(dt, a, d, c, wl, wr, ptr(Lbar_out), ptr(A), ptr(W), ptr(out),
                               ptr(ws), ws.numel(), stream(A.device)), "env_left_bwd")
    return out


def env_right_bwd(Rbar_out: torch.Tensor, B: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """The gradient of R for R' = update_right_env(R, B, W) given the cotangent of R'."""
    for t, n, nd in ((Rbar_out, "Rbar_out", 3), (B, "B", 3), (W, "W", 4)):
        _need(t, n, nd)
    c, d, b = B.shape
    wl, wr = W.shape[0], W.shape[1]
    if tuple(Rbar_out.shape) != (c, wl, c) or len({Rbar_out.dtype, B.dtype, W.dtype}) != 1:
        raise ValueError("env_right_bwd: shapes or dtypes do not match")
    lib = _lib.load()
    dt = dcode(B.dtype)
    ws = _ws(lib.tnh_env_bwd_ws_bytes(dt, c, d, b, wl, wr), B.device)
    out = torch.empty((b, wr, b), dtype=B.dtype, device=B.device)
    check(lib.tnh_env_right_bwd(dt, c, d, b, wl, wr, ptr(Rbar_out), ptr(B), ptr(W), ptr(out),
                                ptr(ws), ws.numel(), stream(B.device)), "env_right_bwd")
    return out


class Graph:
    """A hipGraph of library calls recorded through the C ABI (`tnh_graph_begin/end`):

        g = Graph(torch_stream)
        with g.capture():
            ... library calls on torch_stream (no allocation inside) ...
        g.launch()

    The recording keeps the device pointers it saw; replays read and write those buffers."""

    def __init__(self, torch_stream: "torch.cuda.Stream"):
        self.stream = torch_stream
        self.handle = ctypes.c_void_p(None)

    def capture(self):
        import contextlib
        lib = _lib.load()

        @contextlib.contextmanager
        def ctx():
            with torch.cuda.stream(self.stream):
                check(lib.tnh_graph_begin(self.stream.cuda_stream), "graph_begin")
                try:
                    yield self
                finally:
                    h = ctypes.c_void_p(None)
                    check(lib.tnh_graph_end(self.stream.cuda_stream, ctypes.byref(h)),
                          "graph_end")
                    self.handle = h
        return ctx()

    def launch(self, torch_stream: Optional["torch.cuda.Stream"] = None) -> None:
        s = torch_stream if torch_stream is not None else self.stream
        check(_lib.load().tnh_graph_launch(self.handle, s.cuda_stream), "graph_launch")

    def close(self) -> None:
        if self.handle:
            check(_lib.load().tnh_graph_destroy(self.handle), "graph_destroy")
            self.handle = ctypes.c_void_p(None)
