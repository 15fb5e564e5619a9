"""The C ABI: the library loads (no GPU needed) and exports every symbol that
include/*.h declares; host-side validation rejects bad descriptors without
touching the GPU."""

import ctypes
import glob
import os
import re

from tests.conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(sinv_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = _declared()
    for n in ("sinv_rgf_batched", "sinv_zinv_batched", "sinv_zgemm_batched", "sinv_version"):
        assert n in names


def test_library_exports_every_declared_symbol(built_lib):
    missing = [n for n in sorted(_declared()) if not hasattr(built_lib, n)]
    assert not missing, missing


def test_version_and_invalid_descriptor(built_lib):
    from sinv_amd import _native as nat
    assert built_lib.sinv_version() >= 100
    d = nat.RgfDesc(nblocks=0, blocksize=4, batch=1)
    assert built_lib.sinv_rgf_workspace_bytes(ctypes.byref(d)) == 0
    rc = built_lib.sinv_rgf_batched(ctypes.byref(d), None, 0, None)
    assert rc == nat.SINV_E_INVAL
    assert b"invalid" in built_lib.sinv_last_error()


def test_workspace_query_is_host_only(built_lib):
    from sinv_amd import _native as nat
    d = nat.RgfDesc(nblocks=64, blocksize=384, batch=256, precision=0,
                    diag=ctypes.c_void_p(16), upper=ctypes.c_void_p(16), lower=ctypes.c_void_p(16),
                    g_diag=ctypes.c_void_p(16), g_upper=ctypes.c_void_p(16),
                    g_lower=ctypes.c_void_p(16), save_off_diag=1)
    ws = built_lib.sinv_rgf_workspace_bytes(ctypes.byref(d))
    # three b^2 temporaries per instance plus the inverse recursion's scratch
    assert ws >= 3 * 384 * 384 * 16 * 256
    assert ws < 6 * 384 * 384 * 16 * 256
    assert built_lib.sinv_zinv_workspace_bytes(64, 10) == 0
    assert built_lib.sinv_zinv_workspace_bytes(128, 1) == 2 * 64 * 64 * 16


def test_gemm_argument_validation(built_lib):
    from sinv_amd import _native as nat
    z = nat.C128(0, 0)
    rc = built_lib.sinv_zgemm_batched(4, 4, 4, 7, 0, z, ctypes.c_void_p(16), 4, 0,
                                      ctypes.c_void_p(16), 4, 0, z, None, 4, 0,
                                      ctypes.c_void_p(16), 4, 0, None, 1, 0, None)
    assert rc == nat.SINV_E_INVAL
