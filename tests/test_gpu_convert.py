"""GPU dense <-> block-tridiagonal converters (matrices_utils.py:115-189)
against the host converters, which follow the reference line for line
(tests/test_host_layer.py pins those), and the device-resident path of the
dense drop-in bcr_serial against its host path."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *s):
    return rng.normal(size=s) + 1j * rng.normal(size=s)


@pytest.mark.parametrize("B,n,b", [(1, 40, 40), (2, 120, 40), (3, 96, 8), (1, 1000, 125)])
def test_dense_to_blocks_matches_host(built_lib, B, n, b):
    from sinv_amd import core
    from sinv_amd.utils import matu
    print(f"dense_to_blocks B={B} n={n} b={b}", flush=True)
    rng = np.random.default_rng(n + b)
    A = _rand(rng, B, n, n)
    d, u, l = (x.cpu().numpy() for x in core.dense_to_blocks(A, b))
    for z in range(B):
        hd, hu, hl = matu.convertDenseToBlkTridiag(A[z], b)
        assert np.array_equal(d[z], hd) and np.array_equal(u[z], hu) and np.array_equal(l[z], hl)


@pytest.mark.parametrize("B,N,b", [(1, 1, 16), (2, 5, 24), (1, 9, 64)])
def test_blocks_to_dense_matches_host(built_lib, B, N, b):
    from sinv_amd import core
    from sinv_amd.utils import matu
    rng = np.random.default_rng(N * b)
    d, u, l = _rand(rng, B, N, b, b), _rand(rng, B, max(N - 1, 0), b, b), _rand(rng, B, max(N - 1, 0), b, b)
    A = core.blocks_to_dense(d, u, l).cpu().numpy()
    for z in range(B):
        assert np.array_equal(A[z], matu.convertBlkTridiagToDense(d[z], u[z], l[z]))


def test_scatter_blocks_and_skip_out_of_range(built_lib):
    import torch
    from sinv_amd import core
    rng = np.random.default_rng(2)
    N, b = 6, 8
    A = torch.zeros((1, N * b, N * b), dtype=torch.complex128, device="cuda")
    blocks = _rand(rng, 1, 3, b, b)
    core.scatter_blocks(A, blocks, [(0, 5), (4, 1), (6, 0)])  # last one is outside the grid
    got = A[0].cpu().numpy()
    ref = np.zeros((N * b, N * b), complex)
    ref[0:b, 5 * b:6 * b] = blocks[0, 0]
    ref[4 * b:5 * b, b:2 * b] = blocks[0, 1]
    assert np.array_equal(got, ref)


def test_matu_dispatches_device_tensors(built_lib):
    import torch
    from sinv_amd.utils import matu
    A = matu.generateBandedDiagonalMatrix(96, 20, is_complex=True, seed=63)
    At = torch.as_tensor(A).cuda()
    d, u, l = matu.convertDenseToBlkTridiag(At, 16)
    hd, hu, hl = matu.convertDenseToBlkTridiag(A, 16)
    assert d.is_cuda and np.array_equal(d.cpu().numpy(), hd)
    assert np.array_equal(u.cpu().numpy(), hu) and np.array_equal(l.cpu().numpy(), hl)
    back = matu.convertBlkTridiagToDense(d, u, l)
    assert back.is_cuda
    assert np.array_equal(back.cpu().numpy(), matu.convertBlkTridiagToDense(hd, hu, hl))


@pytest.mark.parametrize("n,b,cplx", [(128, 16, True), (160, 32, False), (252, 36, True)])
def test_bcr_serial_device_path_matches_host_path(built_lib, n, b, cplx):
    import torch
    from sinv_amd.algorithms import bcr_s
    from sinv_amd.utils import matu
    A = matu.generateBandedDiagonalMatrix(n, 2 * b - 1, is_complex=cplx, seed=63)
    A = A + (4 * b) * np.eye(n)
    host = bcr_s.bcr_serial(A, b)
    dev = bcr_s.bcr_serial(torch.as_tensor(A).cuda(), b)
    assert dev.is_cuda and dev.dtype == torch.as_tensor(A).dtype
    assert np.array_equal(dev.cpu().numpy(), host)
