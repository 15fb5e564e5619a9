"""Pins of the PSR oracle (oracle/psr_ref.py, restating psr_seqsolve.py on one
process): the dense inverse on psr_test.py's shapes and generator, several
partition counts, and agreement with the RGF oracle."""

import numpy as np
import pytest

from oracle import dense, psr_ref, rgf_ref
from oracle import matu_ref as mu

PSR_SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12)]


@pytest.mark.parametrize("n,b", PSR_SHAPES)
@pytest.mark.parametrize("P", [2, 3, 4])
@pytest.mark.parametrize("cplx", [False, True])
def test_psr_oracle_vs_dense(n, b, P, cplx):
    if n // b < 3 * P:
        pytest.skip("fewer than 3 block rows per partition")
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, False, 63)
    d, u, l = mu.convertDenseToBlkTridiag(A, b)
    G = psr_ref.psr_blocks(d, u, l, P)
    R = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(G, R):
        assert np.allclose(g, r)


def test_partitions_follow_divide_matrix():
    # psr_utils.py:86-99: the first partition takes the remainder
    assert psr_ref.partitions(32, 3) == ([0, 12, 22], [12, 10, 10])
    assert psr_ref.partitions(12, 4) == ([0, 3, 6, 9], [3, 3, 3, 3])


def test_psr_oracle_agrees_with_rgf_oracle():
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(40, 6, seed=3)
    for P in (2, 3, 5, 13):
        assert dense.rel_err(psr_ref.psr_blocks(d, u, l, P), rgf_ref.rgf_blocks(d, u, l)) < 1e-14


def test_psr_oracle_rejects_small_partitions():
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(8, 2, seed=0)
    with pytest.raises(ValueError):
        psr_ref.psr_blocks(d, u, l, 3)


# --- RGF-2sided restatement (oracle/rgf2s_ref.py)
RGF2S_SHAPES = [(4, 1), (8, 2), (12, 3), (128, 8), (128, 16), (128, 32)]


@pytest.mark.parametrize("n,b", RGF2S_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_rgf2s_oracle_vs_dense(n, b, cplx, sym):
    """rgf2sided_test.py:36-80: all three bands against the dense inverse."""
    from oracle import rgf2s_ref
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, 63)
    d, u, l = mu.convertDenseToBlkTridiag(A, b)
    G = rgf2s_ref.rgf2sided_blocks(d, u, l, sym_mat=sym)
    R = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(G, R):
        assert np.allclose(g, r)


def test_rgf2s_oracle_unsaved_off_diagonals_are_zero_but_the_middle():
    from oracle import rgf2s_ref
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(10, 3, seed=2)
    Gd, Gu, Gl = rgf2s_ref.rgf2sided_blocks(d, u, l, save_off_diag=False)
    R = dense.selected_inverse_dense(d, u, l)
    assert np.allclose(Gd, R[0])
    for i in range(9):
        if i == 4:
            assert np.allclose(Gu[i], R[1][i]) and np.allclose(Gl[i], R[2][i])
        else:
            assert not Gu[i].any() and not Gl[i].any()
