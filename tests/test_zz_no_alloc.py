"""tnh_apply_two_site_fused_noalloc through the C ABI: its output equals the workspace entry
tnh_apply_two_site_fused bit for bit (the same kernels on the same inputs), on two calls in a
row whose second is smaller, and the oracle to the matvec tolerances; a shape beyond the fixed
scratch is refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import dense as D
from tnkit_amd import _lib, kernels

pytestmark = pytest.mark.gpu


def _problem(a, b, wl, wr, seed, dev, op_dtype):
    rng = np.random.default_rng(seed)
    L, R = rng.standard_normal((a, wl, a)), rng.standard_normal((b, wr, b))
    W1, W2 = rng.standard_normal((wl, wl, 2, 2)), rng.standard_normal((wl, wr, 2, 2))
    th = rng.standard_normal((a, 2, 2, b))
    ref = D.apply_two_site_fused(th, *D.fused_two_site(L, W1, W2, R))
    t = lambda x: torch.as_tensor(x).to(dev)  # noqa: E731
    LT, S, RT = kernels.fused_two_site(t(L), t(W1), t(W2), t(R), dtype=op_dtype)
    return t(th), LT, S, RT, ref


def _call(lib, name, th, LT, S, RT, out, ws=None):
    a, d1, d2, b = th.shape
    wl, wr = LT.shape[2], RT.shape[0]
    dt, io = _lib.dcode(LT.dtype), _lib.dcode(th.dtype)
    st = _lib.stream_of(th.device)
    args = [dt, io, a, d1, d2, b, wl, wr, LT.data_ptr(), S.data_ptr(), RT.data_ptr(),
            th.data_ptr(), out.data_ptr()]
    if ws is not None:
        args += [ws.data_ptr(), ws.numel()]
    rc = getattr(lib, name)(*args, 0, 0, st)
    _lib.check(rc, name)


@pytest.mark.parametrize("op", [torch.float64, torch.float32])
def test_noalloc_equals_workspace_entry_on_two_calls(hip, op):
    lib = _lib.load()
    shapes = [(256, 192, 5, 5), (64, 96, 3, 3)]   # the second call is smaller
    for k, (a, b, wl, wr) in enumerate(shapes):
        print("noalloc", op, (a, b, wl, wr), flush=True)
        th, LT, S, RT, ref = _problem(a, b, wl, wr, 10 + k, hip, op)
        nbytes = lib.tnh_two_site_ws_bytes(_lib.dcode(op), a, 2, 2, b, wl, wr)
        ws = torch.empty(nbytes, dtype=torch.uint8, device=hip)
        want = torch.full_like(th, float("nan"))
        got = torch.full_like(th, float("nan"))
        _call(lib, "tnh_apply_two_site_fused", th, LT, S, RT, want, ws)
        _call(lib, "tnh_apply_two_site_fused_noalloc", th, LT, S, RT, got)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        err = float(np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max())
        assert err < (1e-12 if op == torch.float64 else 2e-5), err


def test_noalloc_refuses_shapes_beyond_its_scratch(hip):
    lib = _lib.load()
    a = b = 1024
    need = lib.tnh_two_site_ws_bytes(_lib.F64, a, 2, 2, b, 5, 5)
    assert need > 192 * 1024 * 1024
    z = ctypes.c_void_p(0)
    rc = lib.tnh_apply_two_site_fused_noalloc(_lib.F64, _lib.F64, a, 2, 2, b, 5, 5, z, z, z, z,
                                              z, 0, 0, z)
    assert rc == _lib.ERR_WS
    assert b"above the fixed" in lib.tnh_last_error()
