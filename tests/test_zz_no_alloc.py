"""sinv_rgf_batched_noalloc (workspace in a library-held device array)
against sinv_rgf_batched (caller's workspace), through the C ABI: the same
bits on two calls in a row, the second smaller than the first, and a clear
rejection of a call that needs more workspace than the array holds."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(entry, H, E, eta, precision="fp64"):
    import torch
    from sinv_amd import _native as nat, core
    d, u, l = (core.to_dev(x) for x in H)
    sh = core.to_dev(np.asarray(E) + 1j * eta)
    B, N, b = sh.numel(), d.shape[0], d.shape[-1]
    dev = d.device
    Gd = torch.full((B, N, b, b), float("nan"), dtype=torch.complex128, device=dev)
    Gu = torch.full((B, N - 1, b, b), float("nan"), dtype=torch.complex128, device=dev)
    Gl = torch.full((B, N - 1, b, b), float("nan"), dtype=torch.complex128, device=dev)
    info = torch.empty(1, dtype=torch.int64, device=dev)
    desc = nat.RgfDesc(
        nblocks=N, blocksize=b, bsizes=None, batch=B, sym_mat=0, save_off_diag=1,
        precision=nat.PREC[precision], diag=core._p(d), diag_stride=0, upper=core._p(u),
        upper_stride=0, lower=core._p(l), lower_stride=0, dscale=nat.C128.of(-1.0),
        oscale=nat.C128.of(-1.0), shift=core._p(sh), g_diag=core._p(Gd), g_upper=core._p(Gu),
        g_lower=core._p(Gl), info=core._info_ptr(info))
    L = nat.lib()
    if entry == "noalloc":
        rc = L.sinv_rgf_batched_noalloc(ctypes.byref(desc), core._stream())
    else:
        ws = core._workspace(L.sinv_rgf_workspace_bytes(ctypes.byref(desc)))
        rc = L.sinv_rgf_batched(ctypes.byref(desc), core._p(ws), ws.numel(), core._stream())
    nat.check(rc, entry)
    nat.check_info(info)
    return Gd, Gu, Gl


def test_noalloc_matches_workspace_entry_point(built_lib):
    import torch
    from sinv_amd.utils import configs
    # first call: blocks above 64 (the blocked inverse's workspace in use)
    print("noalloc: 12 blocks of 160, 8 energies", flush=True)
    H1 = configs.hermitian_H(12, 160, seed=41)
    E1 = configs.energies(8, seed=42)
    a = _solve("noalloc", H1, E1, 1e-3)
    r = _solve("workspace", H1, E1, 1e-3)
    torch.cuda.synchronize()
    for x, y in zip(a, r):
        assert torch.equal(x, y)
    # second call, smaller, after the first left its data in the array
    print("noalloc: 7 blocks of 48, 3 energies", flush=True)
    H2 = configs.hermitian_H(7, 48, seed=43)
    E2 = configs.energies(3, seed=44)
    a = _solve("noalloc", H2, E2, 1e-2)
    r = _solve("workspace", H2, E2, 1e-2)
    torch.cuda.synchronize()
    for x, y in zip(a, r):
        assert torch.equal(x, y)
    assert not any(torch.isnan(x).any().item() for x in a)


def test_noalloc_rejects_calls_beyond_its_array(built_lib):
    """A call needing more workspace than SINV_NOALLOC_SCRATCH_BYTES is refused
    on the host, with the sizes in the error text; nothing is launched."""
    from sinv_amd import _native as nat
    L = nat.lib()
    desc = nat.RgfDesc(nblocks=4, blocksize=512, batch=64, precision=0, save_off_diag=1,
                       diag=ctypes.c_void_p(16), upper=ctypes.c_void_p(16),
                       lower=ctypes.c_void_p(16), g_diag=ctypes.c_void_p(16),
                       g_upper=ctypes.c_void_p(16), g_lower=ctypes.c_void_p(16))
    assert L.sinv_rgf_workspace_bytes(ctypes.byref(desc)) > 64 << 20
    assert L.sinv_rgf_batched_noalloc(ctypes.byref(desc), None) == nat.SINV_E_INVAL
    assert b"more than the 67108864" in L.sinv_last_error()
