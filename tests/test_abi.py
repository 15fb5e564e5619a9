"""The C ABI: the library loads (no GPU needed) and exports every symbol that
include/*.h declares; host-side validation rejects bad descriptors without
touching the GPU."""

import ctypes
import glob
import os
import re

import numpy as np
import pytest

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
    # three b^2 temporaries per instance plus the blocked inverse's b^2 copy
    # and its row positions and pivots (2 b ints)
    assert ws >= 4 * 384 * 384 * 16 * 256
    assert ws < 5 * 384 * 384 * 16 * 256
    assert built_lib.sinv_zinv_workspace_bytes(64, 10) == 0
    assert built_lib.sinv_zinv_workspace_bytes(128, 1) == (128 * 128 + 64 + 1) * 16
    assert built_lib.sinv_zinv_workspace_bytes(128, 3) == (3 * 128 * 128 + 192 + 1) * 16


def test_gemm_argument_validation(built_lib):
    from sinv_amd import _native as nat
    z = nat.C128(0, 0)
    rc = built_lib.sinv_zgemm_batched(4, 4, 4, 7, 0, z, ctypes.c_void_p(16), 4, 0,
                                      ctypes.c_void_p(16), 4, 0, z, None, 4, 0,
                                      ctypes.c_void_p(16), 4, 0, None, 1, 0, None)
    assert rc == nat.SINV_E_INVAL


def test_gemm_tune_table_host_side(built_lib):
    """The tuned-shape table and the config queries are host-only state."""
    import ctypes as ct
    n = built_lib.sinv_gemm_config_count()
    assert n >= 2
    bm, bn = ct.c_int(), ct.c_int()
    tiles = []
    for c in range(n):
        assert built_lib.sinv_gemm_config_tile(c, ct.byref(bm), ct.byref(bn)) == 0
        tiles.append((bm.value, bn.value))
    assert tiles[0] == (64, 64)
    assert built_lib.sinv_gemm_config_tile(n, ct.byref(bm), ct.byref(bn)) != 0
    key = (200, 120, 64, 7, 0, 2, 1, 0)
    built_lib.sinv_gemm_tune_clear()
    default = built_lib.sinv_gemm_select(*key)
    other = (default + 1) % n
    assert built_lib.sinv_gemm_tune_set(*key, other) == 0
    assert built_lib.sinv_gemm_select(*key) == other
    assert built_lib.sinv_gemm_select(200, 120, 64, 8, 0, 2, 1, 0) == default  # other batch
    assert built_lib.sinv_gemm_force_config(n - 1) == 0
    assert built_lib.sinv_gemm_select(*key) == n - 1
    assert built_lib.sinv_gemm_force_config(-1) == 0
    assert built_lib.sinv_gemm_tune_set(*key, n) != 0
    built_lib.sinv_gemm_tune_clear()
    assert built_lib.sinv_gemm_select(*key) == default
    # untuned: large grids get the 64 x 64 tile, small ones the 32 x 32 tile,
    # and products too small to give every CU a 32 x 32 tile split k four ways
    assert tiles[built_lib.sinv_gemm_select(512, 512, 512, 1, 0, 0, 0, 0)] == (32, 32)
    assert tiles[built_lib.sinv_gemm_select(384, 384, 384, 256, 0, 0, 0, 0)] == (64, 64)
    small = built_lib.sinv_gemm_select(128, 128, 128, 1, 0, 0, 0, 0)
    assert tiles[small] == (16, 16) and built_lib.sinv_gemm_config_ksplit(small) == 4
    assert built_lib.sinv_gemm_config_ksplit(0) == 1
    assert built_lib.sinv_gemm_config_ksplit(n) < 0


def test_autotune_table_save_load(tmp_path, built_lib):
    import json
    from sinv_amd import autotune
    table = {(64, 64, 64, 256, 0, 0, 1, 0): {"best": 1, "ms": {0: 0.03, 1: 0.02}}}
    p = tmp_path / "t.json"
    rows = [dict(key=list(k), best=v["best"], ms={str(c): t for c, t in v["ms"].items()})
            for k, v in table.items()]
    p.write_text(json.dumps(dict(device="x", entries=rows)))
    assert autotune.load(str(p), install_now=False) == table


def test_converters_reject_bad_arguments_on_the_host(built_lib):
    from sinv_amd import _native as nat
    buf = ctypes.c_void_p(16)  # never dereferenced: validation fails first
    assert built_lib.sinv_dense_to_blocks(buf, 100, 100, 10000, 30, 1, buf, buf, buf, None) \
        == nat.SINV_E_INVAL  # 100 is not a multiple of 30
    assert b"multiple" in built_lib.sinv_last_error()
    assert built_lib.sinv_dense_to_blocks(None, 90, 90, 8100, 30, 1, buf, buf, buf, None) \
        == nat.SINV_E_INVAL
    assert built_lib.sinv_blocks_to_dense(buf, None, None, 3, 8, 1, buf, 24, 576, None) \
        == nat.SINV_E_INVAL  # upper/lower missing for 3 blocks
    assert built_lib.sinv_scatter_blocks(buf, 0, 1, buf, buf, 2, 8, 1, buf, 8, 256, None) \
        == nat.SINV_E_INVAL  # lda smaller than the matrix


def test_int8x_host_validation(built_lib):
    """int8x entry points reject bad arguments on the host (no GPU call), and the
    RGF and adjoint workspaces grow by the digit planes in int8x mode."""
    from sinv_amd import _native as nat
    L = built_lib
    assert L.sinv_zgemm_int8x_workspace_bytes(64, 64, 64, 1, 0, 2, 4) == 0
    assert L.sinv_zgemm_int8x_workspace_bytes(64, 64, 64, 1, 0, 2, 8) == 0
    ws7 = L.sinv_zgemm_int8x_workspace_bytes(100, 70, 130, 1, 0, 3, 7)
    # digit planes: 3 copies of op(A) (128 x 192 padded) and 1 of op(B) (128 x 192), 3 parts x 7
    assert ws7 >= (3 + 1) * 3 * 7 * 128 * 192
    assert L.sinv_zgemm_int8x_workspace_bytes(100, 70, 130, 1, 0, 3, 5) < ws7
    one = nat.C128.of(1)
    rc = L.sinv_zgemm_int8x_batched(4, 4, 4, 0, 0, one, None, 4, 0, None, 4, 0, nat.C128.of(0),
                                    None, 4, 0, None, 4, 0, None, 1, 4, None, 0, None)
    assert rc == nat.SINV_E_INVAL
    rc = L.sinv_zgemm_int8x_batched(4, 4, 20000, 0, 0, one, ctypes.c_void_p(16), 20000, 0,
                                    ctypes.c_void_p(16), 4, 0, nat.C128.of(0), None, 4, 0,
                                    ctypes.c_void_p(16), 4, 0, None, 1, 7, None, 0, None)
    assert rc == nat.SINV_E_INVAL and b"exceeds" in L.sinv_last_error()
    assert L.sinv_int8x_force_variant(3) == nat.SINV_E_INVAL
    assert L.sinv_int8x_force_variant(-1) == nat.SINV_OK
    assert L.sinv_gj_force_variant(3) == nat.SINV_E_INVAL
    assert L.sinv_gj_force_variant(2) == nat.SINV_OK
    d = nat.RgfDesc(nblocks=8, blocksize=96, batch=4, precision=0, diag=1, upper=1, lower=1,
                    g_diag=1, g_upper=1, g_lower=1, save_off_diag=1)
    w64 = L.sinv_rgf_workspace_bytes(ctypes.byref(d))
    d.precision = 2
    w8 = L.sinv_rgf_workspace_bytes(ctypes.byref(d))
    assert w64 > 0 and w8 >= w64 + 2 * 4 * 3 * 6 * 128 * 128  # 6 digit planes per operand
    d.precision = 3   # mixed with FP64 refinement: the FP64 workspace
    assert L.sinv_rgf_workspace_bytes(ctypes.byref(d)) == w64
    d.precision = 4   # no such precision
    assert L.sinv_rgf_workspace_bytes(ctypes.byref(d)) == 0


def test_inconsistent_shapes_raise_before_any_launch():
    """Every batched entry point checks block counts, block sizes and batch
    sizes on the host before anything reaches the device (a mismatch would
    otherwise read past the end of a buffer).  Runs without a GPU: the checks
    come before the device transfer."""
    from sinv_amd import core
    d = np.zeros((4, 3, 3), complex)
    u = np.zeros((3, 3, 3), complex)
    bad_u = np.zeros((2, 3, 3), complex)
    bad_b = np.zeros((3, 3, 2), complex)
    for fn, extra in ((core.rgf_batched, ()), (core.bcr_batched, ()),
                      (core.pdiv_batched, (2,))):
        for uu, ll in ((bad_u, u), (u, bad_u), (bad_b, u)):
            with pytest.raises(ValueError):
                fn(d, uu, ll, *extra)
    d12 = np.zeros((12, 3, 3), complex)
    u11 = np.zeros((11, 3, 3), complex)
    with pytest.raises(ValueError):
        core.psr_batched(d12, u11, np.zeros((10, 3, 3), complex), 2)
    with pytest.raises(ValueError):
        core.psr_batched(np.zeros((12, 3, 2), complex), u11, u11, 2)
    # the adjoint: gradient blocks of the wrong count or size
    for gd, gu, gl in ((np.zeros((3, 3, 3), complex), u, u), (d, bad_u, u), (d, u, bad_b)):
        with pytest.raises(ValueError):
            core.rgf_adjoint_batched(d, u, u, gd, gu, gl)
    with pytest.raises(ValueError):
        core.rgf_adjoint_ragged([np.zeros((3, 3))] * 2, [np.zeros((3, 3))], [np.zeros((3, 3))],
                                [np.zeros((2, 2))] * 2, [np.zeros((2, 2))], [np.zeros((2, 2))])
    # GEMM: batches must be equal or 1; C0 and shift must match
    A = np.zeros((2, 4, 5), complex)
    B = np.zeros((3, 5, 6), complex)
    with pytest.raises(ValueError, match="batch sizes"):
        core.zgemm_batched(A, B)
    B2 = np.zeros((2, 5, 6), complex)
    with pytest.raises(ValueError, match="C0"):
        core.zgemm_batched(A, B2, C0=np.zeros((2, 4, 7), complex))
    with pytest.raises(ValueError, match="C0"):
        core.zgemm_batched(A, B2, C0=np.zeros((3, 4, 6), complex))
    with pytest.raises(ValueError, match="shift"):
        core.zgemm_batched(A, B2, shift=np.zeros(3, complex))
    with pytest.raises(ValueError, match="inner"):
        core.zgemm_batched(A, np.zeros((2, 4, 6), complex))
    with pytest.raises(ValueError):
        core.zgemm_batched(A, B2, opA="X")


def test_inverse_method_host_side(built_lib):
    """The per-thread inverse method: argument checks, the previous method
    returned, and workspace queries sized for the Newton-Schulz buffers only
    while it is selected (host-only state, no GPU)."""
    from sinv_amd import _native as nat
    L = built_lib
    assert L.sinv_inverse_method(5, 0) == -1
    assert L.sinv_inverse_method(1, 17) == -1
    assert L.sinv_inverse_method(1, -1) == -1
    piv = L.sinv_zinv_workspace_bytes(512, 1)
    with nat.inverse_method("newton", 3):
        assert L.sinv_inverse_method(1, 0) == 1
        n, m = 512, 4  # Y and the in-place copy of M for up to 4 blocks, 2 x 4 words
        assert L.sinv_zinv_workspace_bytes(512, 1) == piv + (2 * n * n * m + 2 * m) * 16
        assert L.sinv_zinv_workspace_bytes(64, 1) == 0
        d = nat.RgfDesc(nblocks=8, blocksize=512, batch=1, precision=0,
                        diag=ctypes.c_void_p(16), upper=ctypes.c_void_p(16),
                        lower=ctypes.c_void_p(16), g_diag=ctypes.c_void_p(16),
                        g_upper=ctypes.c_void_p(16), g_lower=ctypes.c_void_p(16), save_off_diag=1)
        ws_newton = L.sinv_rgf_workspace_bytes(ctypes.byref(d))
    assert L.sinv_inverse_method(0, 0) == 0
    ws_piv = L.sinv_rgf_workspace_bytes(ctypes.byref(d))
    assert ws_newton == ws_piv + (2 * 512 * 512 * 4 + 8) * 16
    # a workspace sized under the elimination is refused under the iteration
    with nat.inverse_method("newton"):
        rc = L.sinv_rgf_batched(ctypes.byref(d), ctypes.c_void_p(256), ws_piv, None)
    assert rc == nat.SINV_E_INVAL and b"workspace too small" in L.sinv_last_error()
    with pytest.raises(ValueError):
        nat.inverse_method("lu")
    rc = L.sinv_column_dominance(None, None, None)
    assert rc == nat.SINV_E_INVAL
    # per-block sizes are refused before anything is launched
    sizes = (ctypes.c_int * 8)(*([512] * 8))
    d.bsizes = sizes
    assert L.sinv_column_dominance(ctypes.byref(d), ctypes.c_void_p(256), None) == nat.SINV_E_INVAL
