"""The C-ABI library loads on the CPU host and exports every entry that include/tnkit_hip.h
declares, with the Python binding covering each of them. No kernel is launched here."""
import os
import re

from tnkit_amd import _lib

from conftest import ROOT


def _declared():
    with open(os.path.join(ROOT, "include", "tnkit_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(tnh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entries():
    names = _declared()
    assert "tnh_apply_two_site_fused" in names and "tnh_krylov_orth" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_host_side_queries_need_no_device():
    lib = _lib.load()
    assert lib.tnh_version() >= 100
    assert lib.tnh_two_site_ws_bytes(0, 1024, 2, 2, 1024, 5, 5) >= 2 * 4 * 1024 * 1024 * 5 * 8
    assert lib.tnh_krylov_ws_bytes(4096, 41) > 0


def test_bad_arguments_are_refused_before_any_launch():
    lib = _lib.load()
    rc = lib.tnh_apply_two_site_fused(0, 0, 0, 2, 2, 4, 3, 3, None, None, None, None, None, None,
                                      0, 0, 0, None)
    assert rc == _lib.ERR_ARG
    assert b"positive" in lib.tnh_last_error()
    rc = lib.tnh_gemm_tn(0, 8, 8, 8, None, 4, None, 8, None, 8, 1, 0, 0, 0, 0, 0, None)
    assert rc == _lib.ERR_ARG


def test_svd_jacobi_refuses_overlapping_output_strides():
    """A batch whose U (or Vh, S) strides are smaller than one matrix's footprint would have
    workgroups writing over each other: refused before any launch."""
    lib = _lib.load()
    args = dict(m=16, n=8, lda=8, sA=128, S=None, sS=8, U=None, ldu=8, sU=128, Vh=None, ldv=8,
                sV=64)

    def call(**kw):
        a = {**args, **kw}
        return lib.tnh_svd_jacobi(_lib.F64, 4, a["m"], a["n"], None, a["lda"], a["sA"], a["S"],
                                  a["sS"], a["U"], a["ldu"], a["sU"], a["Vh"], a["ldv"], a["sV"],
                                  60, 1e-15, None, None)
    for bad in (dict(sU=100), dict(sV=50), dict(sS=4)):
        assert call(**bad) == _lib.ERR_ARG
        assert b"overlap" in lib.tnh_last_error()


def test_syev_refusals_and_sizes_without_a_device():
    """The symmetric eigensolver's entries refuse bad orders, k and workspaces before any launch,
    and size their workspace on the host."""
    lib = _lib.load()
    assert lib.tnh_syev_top_ws_bytes(2048, 1024) > 3 * 2048 * 2048 * 8
    assert lib.tnh_syev_top_ws_bytes(1, 1) == 0 and lib.tnh_syev_top_ws_bytes(4096, 8) == 0
    assert lib.tnh_sytrd_ws_bytes(2048) > 0 and lib.tnh_sytrd_ws_bytes(4096) == 0
    rc = lib.tnh_syev_top(4096, 8, None, 4096, None, None, 8, None, None, 0, None)
    assert rc == _lib.ERR_UNSUPPORTED
    rc = lib.tnh_syev_top(100, 0, None, 100, None, None, 8, None, None, 0, None)
    assert rc == _lib.ERR_ARG
    rc = lib.tnh_syev_top(100, 10, 1, 100, 1, 1, 10, None, 1, 16, None)
    assert rc == _lib.ERR_WS and b"workspace" in lib.tnh_last_error()
    rc = lib.tnh_sytrd(0, None, 1, None, None, None, None, None, None, 0, None)
    assert rc == _lib.ERR_UNSUPPORTED
