"""ctypes binding of libtnkit_hip.so (declared in include/tnkit_hip.h).

The library is loaded from this package's directory. There is no fallback: if the library is
missing, the import of the HIP backend fails with instructions to build it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_int, c_long, c_size_t, c_void_p
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtnkit_hip.so")

F64, F32, BF16, C128, C64 = 0, 1, 2, 3, 4
OK, ERR_ARG, ERR_WS, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3, 4

_P = c_void_p
_PD = ctypes.POINTER(c_double)

# name -> (restype, argtypes); kept in the order of include/tnkit_hip.h
SIGNATURES = {
    "tnh_version": (c_int, []),
    "tnh_last_error": (ctypes.c_char_p, []),
    "tnh_gemm_tn": (c_int, [c_int, c_int, c_int, c_int, _P, c_long, _P, c_long, _P, c_long, c_int,
                            c_long, c_long, c_long, c_int, c_int, _P]),
    "tnh_fused_two_site": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, _P, _P,
                                   _P, _P, _P, _P, _P, _P]),
    "tnh_two_site_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    "tnh_apply_two_site_fused": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                         _P, _P, _P, _P, _P, _P, c_size_t, c_int, c_int, _P]),
    "tnh_apply_two_site_fused_noalloc": (c_int, [c_int] * 8 + [_P] * 5 + [c_int, c_int, _P]),
    "tnh_apply_two_site_stages": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                          _P, _P, _P, _P, _P, _P, c_size_t, c_int, c_int, c_int,
                                          _P]),
    "tnh_env_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "tnh_update_left_env": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, _P, _P, _P, _P, _P,
                                    c_size_t, _P]),
    "tnh_update_right_env": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, _P, _P, _P, _P, _P,
                                     c_size_t, _P]),
    "tnh_matvec2_bwd_ws_bytes": (c_size_t, [c_int] * 7),
    "tnh_matvec2_bwd": (c_int, [c_int] * 8 + [_P] * 7 + [c_size_t, _P]),
    "tnh_env_bwd_ws_bytes": (c_size_t, [c_int] * 6),
    "tnh_env_left_bwd": (c_int, [c_int] * 6 + [_P] * 5 + [c_size_t, _P]),
    "tnh_env_right_bwd": (c_int, [c_int] * 6 + [_P] * 5 + [c_size_t, _P]),
    "tnh_graph_begin": (c_int, [_P]),
    "tnh_graph_end": (c_int, [_P, ctypes.POINTER(c_void_p)]),
    "tnh_graph_launch": (c_int, [_P, _P]),
    "tnh_graph_destroy": (c_int, [_P]),
    "tnh_fused_two_site_batched": (c_int, [c_int] * 9 + [_P, c_long] * 7 + [_P]),
    "tnh_apply_two_site_batched": (c_int, [c_int] * 9 + [_P, c_long] * 5 + [_P, c_size_t, c_int,
                                                                          c_int, c_int, _P]),
    "tnh_apply_two_site_ragged": (c_int, [c_int] * 9 + [_P, c_long] * 5 + [_P, _P, c_size_t, _P]),
    "tnh_update_left_env_batched": (c_int, [c_int] * 7 + [_P, c_long] * 4 + [_P, c_size_t, _P]),
    "tnh_update_right_env_batched": (c_int, [c_int] * 7 + [_P, c_long] * 4 + [_P, c_size_t, _P]),
    "tnh_expm_ws_bytes": (c_size_t, [c_int]),
    "tnh_expm_step_c": (c_int, [c_int, c_long, c_int, c_int, _P, c_long, c_long, _P, c_long, _P,
                                c_long, _P, c_size_t, _P]),
    "tnh_expm_step_c_ragged": (c_int, [c_int, c_long, c_int, c_int, _P, c_long, c_long, _P, c_long,
                                       _P, c_long, _P, c_int, c_int, _P, c_size_t, _P]),
    "tnh_lincomb_c": (c_int, [c_int, c_long, c_int, _P, c_long, c_long, _P, c_long, _P, c_long,
                              _P]),
    "tnh_lincomb_c_ragged": (c_int, [c_int, c_long, c_int, _P, c_long, c_long, _P, c_long, _P,
                                     c_long, _P, c_int, c_int, _P]),
    "tnh_krylov_ws_bytes": (c_size_t, [c_long, c_int]),
    "tnh_krylov_orth": (c_int, [c_long, c_int, _P, c_long, _P, c_int, c_int, c_int, _PD, _P, _P,
                                c_size_t, _P]),
    "tnh_rank_update": (c_int, [c_long, c_int, _P, c_long, _P, c_double, _P, _P, c_size_t, _P]),
    "tnh_lincomb": (c_int, [c_long, c_int, _P, c_long, _PD, c_double, _P, _P]),
    "tnh_norm": (c_int, [c_long, _P, _P, _P, c_size_t, _P]),
    "tnh_dot": (c_int, [c_long, _P, _P, _P, _P, c_size_t, _P]),
    "tnh_rows_ws_bytes": (c_size_t, [c_int]),
    "tnh_rows_dot": (c_int, [c_int, c_int, c_long, _P, c_long, _P, _P, _P, c_size_t, _P]),
    "tnh_rows_combine": (c_int, [c_int, c_int, c_long, _P, _P, c_long, _P, _P]),
    "tnh_scale_by": (c_int, [c_long, _P, _P, c_double, _P, _P]),
    "tnh_convert": (c_int, [c_int, c_int, c_long, _P, _P, _P]),
    "tnh_gather_chunk": (c_int, []),
    "tnh_gather_runs": (c_int, [c_int, c_int, c_long, _P, _P, c_long, _P, _P, c_int, _P]),
    "tnh_gather_tiles": (c_int, [c_int, c_int, c_long, _P, _P, _P, _P]),
    "tnh_grouped_tile": (c_int, [c_int, c_int, _P, _P, _P]),
    "tnh_grouped_gemm": (c_int, [c_int, _P, c_long, _P, _P, _P, _P, c_int, c_int, _P]),
    "tnh_grouped_small": (c_int, [c_int, _P, c_long, _P, _P, _P, _P, _P]),
    "tnh_grouped_mid": (c_int, [c_int, _P, c_long, _P, _P, _P, _P, _P]),
    "tnh_svd_jacobi_lds_bytes": (c_size_t, [c_int, c_int, c_int]),
    "tnh_svd_jacobi": (c_int, [c_int, c_long, c_int, c_int, _P, c_long, c_long, _P, c_long, _P,
                               c_long, c_long, _P, c_long, c_long, c_int, c_double, _P, _P]),
    "tnh_eigh_jacobi": (c_int, [c_int, c_long, c_int, _P, c_long, c_long, _P, c_long, _P, c_long,
                                _P, c_long, c_long, c_int, c_double, _P, _P]),
    "tnh_sytrd_ws_bytes": (c_size_t, [c_int]),
    "tnh_sytrd": (c_int, [c_int, _P, c_long, _P, _P, _P, _P, _P, _P, c_size_t, _P]),
    "tnh_syev_top_ws_bytes": (c_size_t, [c_int, c_int]),
    "tnh_syev_top": (c_int, [c_int, c_int, _P, c_long, _P, _P, c_long, _P, _P, c_size_t, _P]),
    "tnh_heev_top_ws_bytes": (c_size_t, [c_int, c_int, c_int]),
    "tnh_qr_ws_bytes": (c_size_t, [c_int, c_int, c_int, c_int]),
    "tnh_qr": (c_int, [c_int, c_int, c_int, c_int, _P, c_long, c_long, _P, c_long, c_long, _P,
                       c_long, c_long, _P, _P, c_size_t, _P]),
    "tnh_heev_top": (c_int, [c_int, c_int, c_int, _P, c_long, c_long, _P, c_long, _P, c_long,
                             c_long, _P, _P, c_size_t, _P]),
}


class HipError(RuntimeError):
    """A HIP launch or runtime failure inside libtnkit_hip."""


_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    """Load the library once and bind its signatures. Raises ImportError when it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: the HIP backend has no CPU fallback. Build it with "
            "`python -m tnkit_amd.build` (hipcc, gfx950).")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Map a TNH_* status to the exception types tnkit uses (ValueError for bad arguments)."""
    if rc == OK:
        return
    msg = load().tnh_last_error().decode(errors="replace")
    if rc in (ERR_ARG, ERR_WS):
        raise ValueError(f"{what}: {msg}")
    if rc == ERR_UNSUPPORTED:
        raise TypeError(f"{what}: {msg}")
    raise HipError(f"{what}: {msg}")


def dcode(dtype) -> int:
    """TNH dtype code of a torch dtype."""
    import torch
    table = {torch.float64: F64, torch.float32: F32, torch.bfloat16: BF16,
             torch.complex128: C128, torch.complex64: C64}
    try:
        return table[dtype]
    except KeyError:
        raise TypeError(f"no libtnkit_hip dtype for {dtype}") from None


def ptr(t) -> int:
    return t.data_ptr()


def stream_of(device=None) -> int:
    """The current torch (HIP) stream handle on `device` (the capture stream inside a graph
    capture). torch's raw-stream query skips building a Stream object, which costs several
    microseconds per launch on the latency-bound small-bond paths."""
    import torch
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is None:
        return torch.cuda.current_stream(device).cuda_stream
    if device is None:
        idx = torch.cuda.current_device()
    elif isinstance(device, int):
        idx = device
    else:
        dev = device if isinstance(device, torch.device) else torch.device(device)
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
    return raw(idx)
