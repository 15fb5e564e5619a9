"""Pins of the BCR oracle (oracle/bcr_ref.py, restating bcr_serial.py:19-466
and bcr_utils.py:17-71)."""

import math

import numpy as np
import pytest

from oracle import bcr_ref, dense
from oracle import matu_ref as mu

SEED = 63
BCR_SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12)]


@pytest.mark.parametrize("n,b", BCR_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_bcr_dense_restatement_vs_inverse(n, b, cplx, sym):
    """bcr_serial_test.py:54-73: all three bands against the dense inverse."""
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    G = bcr_ref.bcr_serial_dense(A, b)
    assert G.shape == A.shape
    got = mu.convertDenseToBlkTridiag(G, b)
    ref = mu.convertDenseToBlkTridiag(np.linalg.inv(A), b)
    for g, r in zip(got, ref):
        assert np.allclose(g, r)


@pytest.mark.parametrize("n,b", BCR_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
def test_compressed_schedule_is_bitwise_the_dense_one(n, b, cplx):
    """Skipping the identity padding and deduplicating inv(A_tt) changes no bit
    of any produced block, including the stride-2^l fill-in blocks, and the
    dense output has no other nonzero blocks."""
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, False, SEED)
    G = bcr_ref.bcr_serial_dense(A, b)
    d, u, l = mu.convertDenseToBlkTridiag(A, b)
    _, _, _, Gfill = bcr_ref.bcr_blocks(d, u, l, return_fill=True)
    mask = np.zeros(G.shape, bool)
    for (r, c), blk in Gfill.items():
        assert np.array_equal(blk, G[r * b:(r + 1) * b, c * b:(c + 1) * b]), (r, c)
        mask[r * b:(r + 1) * b, c * b:(c + 1) * b] = True
    assert not np.any(G[~mask])


@pytest.mark.parametrize("n,b", BCR_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_fill_in_blocks_are_blocks_of_the_inverse(n, b, cplx, sym):
    """bcr_serial returns the stride-2^l coupling blocks the reduction creates
    (bcr_serial.py:137-169, 308-383); each is the same block of A^-1.  Closed
    form, independent of any BCR restatement: every fill-in block of the
    oracle's dense output against np.linalg.inv, at 1e-15 cond(A) n on the
    diagonally dominant version of the reference's test matrix and at the
    reference's own allclose on the raw one."""
    from sinv_amd.core import bcr_fill_positions
    N = n // b
    raw = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    for A, strict in ((raw + 3 * b * np.eye(n), True), (raw, False)):
        G = bcr_ref.bcr_serial_dense(A, b)
        inv = np.linalg.inv(A)
        tol = 1e-15 * np.linalg.cond(A) * n
        pos = bcr_fill_positions(N)
        assert len(pos) == len(set(pos))
        for r, c in pos:
            got = G[r * b:(r + 1) * b, c * b:(c + 1) * b]
            ref = inv[r * b:(r + 1) * b, c * b:(c + 1) * b]
            if strict:
                assert np.abs(got - ref).max() <= tol * np.abs(inv).max(), (r, c)
            else:
                assert np.allclose(got, ref), (r, c)


def test_fill_in_pin_catches_a_wrong_block():
    """The pin above is not vacuous: a fill-in block computed with the sign of
    its update flipped (a plausible slip in center_produce) fails it."""
    from sinv_amd.core import bcr_fill_positions
    n, b = 144, 12
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), True, False, SEED) + 3 * b * np.eye(n)
    G = bcr_ref.bcr_serial_dense(A, b)
    inv = np.linalg.inv(A)
    r, c = bcr_fill_positions(n // b)[0]
    blk = G[r * b:(r + 1) * b, c * b:(c + 1) * b]
    assert np.abs(-blk - inv[r * b:(r + 1) * b, c * b:(c + 1) * b]).max() > 1e-6
    assert np.abs(blk - inv[r * b:(r + 1) * b, c * b:(c + 1) * b]).max() < 1e-14


def test_padding_helpers():
    # bcr_utils.py:35-40: 2^ceil(log2(N+1)) - 1 - N
    for N, pad in [(1, 0), (2, 1), (3, 0), (9, 6), (12, 3), (16, 15), (31, 0), (32, 31),
                   (64, 63), (96, 31), (128, 127)]:
        assert bcr_ref.distance_to_power_of_two(N) == pad
    A = np.arange(4.0).reshape(2, 2) + 1j
    P = bcr_ref.identity_padding(A, 3)
    assert P.shape == (5, 5) and np.array_equal(P[:2, :2], A)
    assert np.array_equal(P[2:, 2:], np.eye(3)) and not P[:2, 2:].any()


def test_schedule_shape():
    height, root, levels = bcr_ref.bcr_schedule(16)
    assert height == 4 and root == 15
    s, elim, keep = levels[0]
    assert s == 1 and elim == list(range(0, 16, 2)) and keep == list(range(1, 16, 2))
    # every real row is eliminated exactly once, except the root
    seen = sorted(t for _, e, _ in levels for t in e)
    assert seen == sorted(set(range(16)) - {root})
    for N in (1, 2, 5, 31, 32, 100):
        h, r, lv = bcr_ref.bcr_schedule(N)
        assert h == int(math.log2(N + bcr_ref.distance_to_power_of_two(N)))
        assert sorted(t for _, e, _ in lv for t in e) == sorted(set(range(N)) - {r})


def test_bcr_agrees_with_rgf_and_residuals():
    from oracle import rgf_ref
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(21, 8, seed=5)
    Gb = bcr_ref.bcr_blocks(d, u, l)
    Gr = rgf_ref.rgf_blocks(d, u, l)
    assert dense.rel_err(Gb, Gr) < 1e-14
    assert dense.residuals(d, u, l, *Gb) < 1e-14


def test_bcr_flop_count_reference_schedule():
    """The compressed schedule does no more work than the padded reference."""
    n_inv, n_gemm = bcr_ref.bcr_flops(16, 64)
    assert n_inv == 16 and n_gemm < 444   # reference padded schedule: 83 inv / 444 gemm
