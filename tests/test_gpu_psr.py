"""GPU parity of PSR on one GPU (psr_seqsolve.py:20-72, partitions in place of
ranks) against the oracle's single-process restatement (oracle/psr_ref.py,
pinned to the dense inverse in tests/test_oracle_psr.py).  Shapes and
generator of psr_test.py:40-80 (SEED = 63) for 3..4 partitions, plus larger
and batched cases."""

import numpy as np
import pytest

from oracle import dense, psr_ref
from oracle import matu_ref as mu

pytestmark = pytest.mark.gpu

SEED = 63
PSR_SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12)]


@pytest.mark.parametrize("n,b", PSR_SHAPES)
@pytest.mark.parametrize("P", [3, 4])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_psr_reference_shapes(built_lib, n, b, P, cplx, sym):
    from sinv_amd import core
    if n // b < 3 * P:
        pytest.skip("fewer than 3 block rows per partition (psr_test.py:71 skips these too)")
    print(f"psr n={n} b={b} P={P} cplx={cplx} sym={sym}", flush=True)
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    d, u, l = mu.convertDenseToBlkTridiag(A, b)
    got = tuple(x[0].cpu().numpy() for x in core.psr_batched(d, u, l, P))
    ref = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(got, ref):
        assert np.allclose(g, r)
    O = psr_ref.psr_blocks(d.astype(complex), u.astype(complex), l.astype(complex), P)
    assert dense.rel_err(got, O) < 1e-15 * np.linalg.cond(A) * 10


@pytest.mark.parametrize("N,b,P", [(12, 16, 2), (30, 24, 5), (64, 64, 8), (40, 96, 3)])
def test_psr_dominant(built_lib, N, b, P):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(N, b, seed=N + P)
    got = tuple(x[0].cpu().numpy() for x in core.psr_batched(d, u, l, P))
    assert dense.rel_err(got, psr_ref.psr_blocks(d, u, l, P)) < 1e-12
    assert dense.rel_err(got, dense.selected_inverse_dense(d, u, l)) < 1e-12


def test_psr_energies_agrees_with_rgf(built_lib):
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(24, 48, seed=8)
    E = configs.energies(3, seed=9) + 1e-2j
    Gp = core.psr_batched(*H, 4, shift=E, dscale=-1.0, oscale=-1.0)
    Gr = core.rgf_batched(*H, shift=E, dscale=-1.0, oscale=-1.0)
    for k in range(len(E)):
        p_ = tuple(x[k].cpu().numpy() for x in Gp)
        r_ = tuple(x[k].cpu().numpy() for x in Gr)
        assert dense.rel_err(p_, r_) < 1e-11


def test_psr_rejects_too_many_partitions(built_lib):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(8, 4, seed=0)
    with pytest.raises(ValueError):
        core.psr_batched(d, u, l, 3)


RGF2S_SHAPES = [(4, 1), (8, 2), (12, 3), (128, 8), (128, 16), (128, 32)]


@pytest.mark.parametrize("n,b", RGF2S_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("save", [True, False])
def test_rgf2sided_dropin(built_lib, n, b, cplx, sym, save):
    from oracle import rgf2s_ref
    from sinv_amd.algorithms import rgf2s
    print(f"rgf2sided n={n} b={b} cplx={cplx} sym={sym} save={save}", flush=True)
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    d, u, l = mu.convertDenseToBlkTridiag(A, b)
    got = rgf2s.rgf2sided(d, u, l, sym_mat=sym, save_off_diag=save)
    ref = rgf2s_ref.rgf2sided_blocks(d, u, l, sym_mat=sym, save_off_diag=save)
    assert all(g.dtype == d.dtype for g in got)
    cond = np.linalg.cond(A)
    assert dense.rel_err(got, ref) < 1e-15 * cond * 10
    if save:
        R = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
        for g, r in zip(got, R):
            assert np.allclose(g, r)


@pytest.mark.parametrize("n,b", [(9, 1), (24, 2), (144, 12), (256, 8)])
def test_psr_seqsolve_dropin(built_lib, n, b):
    from sinv_amd.algorithms import psr_s
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), True, False, SEED)
    G = psr_s.psr_seqsolve(A, b)
    got = mu.convertDenseToBlkTridiag(G, b)
    ref = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(got, ref):
        assert np.allclose(g, r)


@pytest.mark.parametrize("P,first_rows", [(4, 0), (2, 13)])
@pytest.mark.parametrize("inverse", ["pivoted", "newton"])
def test_psr_graph_capture_replays(built_lib, P, first_rows, inverse):
    """PSR forks its top and bottom partitions onto side streams and joins them
    back to the caller's stream.  Captured into a graph (after one eager call,
    as graph capture requires), every replay -- with eager work and
    synchronisations in between -- equals the eager result bit for bit, with
    either block-inverse method (the iteration's state is reset by a kernel)."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print(f"psr graph P={P} first_rows={first_rows} inverse={inverse}", flush=True)
    d, u, l = (core.to_dev(x) for x in configs.dominant_blocks(28, 96, seed=5))
    ref = core.psr_batched(d, u, l, P, first_rows=first_rows, inverse=inverse)
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        core.psr_batched(d, u, l, P, first_rows=first_rows, check=False, inverse=inverse)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = core.psr_batched(d, u, l, P, first_rows=first_rows, check=False, inverse=inverse)
    x = torch.randn(512, 512, dtype=torch.float64, device=d.device)
    for rep in range(3):
        y = (x @ x).sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.isfinite(y)
        for o, r in zip(out, ref):
            assert torch.equal(o, r), rep
