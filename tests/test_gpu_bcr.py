"""GPU parity of BCR-serial (bcr_serial.py:19-466) against the oracle.

The oracle's dense restatement (oracle/bcr_ref.py) is pinned to the dense
inverse in tests/test_oracle_bcr.py; its compressed schedule is bitwise the
dense one.  Inputs are the reference test's generator and shapes (SEED = 63,
bcr_serial_test.py:40-65).
"""

import numpy as np
import pytest

from oracle import bcr_ref, dense
from oracle import matu_ref as mu

pytestmark = pytest.mark.gpu

SEED = 63
BCR_SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12)]


@pytest.mark.parametrize("n,b", BCR_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_bcr_serial_dense_dropin(built_lib, n, b, cplx, sym):
    from sinv_amd.algorithms import bcr_s
    print(f"bcr ref-shape n={n} b={b} cplx={cplx} sym={sym}", flush=True)
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    G = bcr_s.bcr_serial(A, b)
    assert G.dtype == A.dtype and G.shape == A.shape
    # the reference's acceptance test (bcr_serial_test.py:66-73)
    got = mu.convertDenseToBlkTridiag(G, b)
    ref = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(got, ref):
        assert np.allclose(g, r)
    # every block of the reference's dense output, fill-in included
    Gref = bcr_ref.bcr_serial_dense(A, b)
    cond = np.linalg.cond(A)
    err = np.linalg.norm(G - Gref) / np.linalg.norm(Gref)
    assert err < 1e-15 * cond * 10, err
    # nothing outside the blocks the reduction produces
    from sinv_amd import core
    N = n // b
    mask = np.zeros(G.shape, bool)
    pos = [(i, i) for i in range(N)] + [(i, i + 1) for i in range(N - 1)] + \
        [(i + 1, i) for i in range(N - 1)] + core.bcr_fill_positions(N)
    for r, c in pos:
        mask[r * b:(r + 1) * b, c * b:(c + 1) * b] = True
    assert not np.any(G[~mask])


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5, 7, 8, 15, 16, 17, 33])
def test_bcr_block_counts(built_lib, N):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(N, 24, seed=N)
    Gd, Gu, Gl = core.bcr_batched(d, u, l)
    got = tuple(x[0].cpu().numpy() for x in (Gd, Gu, Gl))
    ref = bcr_ref.bcr_blocks(d, u, l)
    assert dense.rel_err(got, ref) < 1e-13


@pytest.mark.parametrize("N,b", [(7, 8), (16, 12), (21, 64), (33, 96)])
def test_bcr_fill_blocks_are_blocks_of_the_inverse(built_lib, N, b):
    """The stride-2^l blocks bcr_serial returns besides the tridiagonal ones
    (fill-in of the reduction, bcr_serial.py:137-169, 308-383) against the
    same blocks of np.linalg.inv, and the tridiagonal bands too, at
    1e-15 cond(A) n."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(N, b, seed=N * b)
    Gd, Gu, Gl, Gf = core.bcr_batched(d, u, l, return_fill=True)
    A = dense.assemble(d, u, l)
    inv = np.linalg.inv(A)
    tol = 1e-15 * np.linalg.cond(A) * A.shape[0] * np.abs(inv).max()
    pos = core.bcr_fill_positions(N)
    Gf = Gf[0].cpu().numpy()
    assert Gf.shape[0] == len(pos)
    for k, (r, c) in enumerate(pos):
        assert np.abs(Gf[k] - inv[r * b:(r + 1) * b, c * b:(c + 1) * b]).max() <= tol, (r, c)
    sel = dense.select(inv, b)
    for g, ref in zip((Gd, Gu, Gl), sel):
        assert np.abs(g[0].cpu().numpy() - np.stack(ref)).max() <= tol


def test_bcr_dominant_16x64(built_lib):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(16, 64, seed=1)
    got = tuple(x[0].cpu().numpy() for x in core.bcr_batched(d, u, l))
    assert dense.rel_err(got, bcr_ref.bcr_blocks(d, u, l)) < 1e-12
    assert dense.rel_err(got, dense.selected_inverse_dense(d, u, l)) < 1e-12


def test_bcr_energies_agrees_with_rgf(built_lib):
    """Batched over energies with A(E) formed on the fly; BCR and RGF agree."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(13, 72, seed=2)
    E = configs.energies(4, seed=3) + 1e-2j
    Gb = core.bcr_batched(*H, shift=E, dscale=-1.0, oscale=-1.0)
    Gr = core.rgf_batched(*H, shift=E, dscale=-1.0, oscale=-1.0)
    for k in range(len(E)):
        b_ = tuple(x[k].cpu().numpy() for x in Gb)
        r_ = tuple(x[k].cpu().numpy() for x in Gr)
        assert dense.rel_err(b_, r_) < 1e-11
        d, u, l = configs.shifted(H, E[k])
        assert dense.residuals(d, u, l, *b_) < 1e-11


BCRP_SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12)]


@pytest.mark.parametrize("n,b", BCRP_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
@pytest.mark.parametrize("nparts", [1, 2])
def test_bcr_parallel_dropin_reference_shapes(built_lib, n, b, cplx, sym, nparts):
    """bcr_parallel_test.py:57-79 (run there on 2 ranks): the tridiagonal
    blocks of the result against the dense inverse (the reference's
    allclose), and bit-identical to bcr_serial, whose arithmetic it shares."""
    from sinv_amd.algorithms import bcr_p, bcr_s
    from sinv_amd.utils import matu
    print(f"bcr_parallel n={n} b={b} cplx={cplx} sym={sym} nparts={nparts}", flush=True)
    A = matu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, 63)
    G = bcr_p.bcr_parallel(A, b, nparts)
    R = np.linalg.inv(A)
    for x, r in zip(matu.convertDenseToBlkTridiag(G, b), matu.convertDenseToBlkTridiag(R, b)):
        assert np.allclose(x, r)
    assert np.array_equal(G, bcr_s.bcr_serial(A, b))


@pytest.mark.parametrize("N,P", [(31, 2), (31, 4), (20, 3), (45, 8)])
def test_bcr_parallel_launches_follow_the_rank_split(built_lib, N, P):
    """With P partitions every level's eliminated-row inverses run as one
    launch per partition that owns eliminated rows of that level (the ranks'
    rows of bcr_utils.divide_matrix on the padded rows), and the produced
    blocks equal BCR-serial's bit for bit."""
    import torch
    from sinv_amd import _native as nat, core
    from sinv_amd.algorithms import bcr_p
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(N, 16, seed=N + P)
    ref = core.bcr_batched(d, u, l, return_fill=True)
    torch.cuda.synchronize()
    with nat.LaunchProfiler() as prof:
        got = core.bcr_batched(d, u, l, return_fill=True, npart=P)
        torch.cuda.synchronize()
    for x, y in zip(got, ref):
        assert torch.equal(x, y)
    inv_launches = sum(1 for kind, _, _ in prof.records() if kind == "inv_leaf")
    npad = 2 ** int(np.ceil(np.log2(N + 1))) - 1
    starts, sizes = bcr_p.divide_matrix(npad, P)
    stride = sizes[0]
    expect = 1  # the root
    s = 1
    while 2 * s - 1 < npad:
        elim = [m * s + s - 1 for m in range(0, N // s, 2)]   # eliminated real rows of the level
        if N // s > 1:
            expect += len({min(r // stride, P - 1) for r in elim})
        s *= 2
    assert inv_launches == expect, (inv_launches, expect)
