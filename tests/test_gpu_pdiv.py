"""GPU PDIV (sinv_pdiv_batched) against the PDIV oracle and the dense inverse,
on the reference's test shapes (pdiv_test.py:28-56) for 1, 2, 4 and 8
partitions, batched instances, and agreement with RGF at larger sizes."""

import numpy as np
import pytest

from oracle import dense, pdiv_ref

pytestmark = pytest.mark.gpu

SHAPES = [(2, 1), (4, 2), (6, 3), (3, 1), (6, 2), (9, 3), (128, 8), (128, 16), (128, 32)]


def _cases():
    for n, b in SHAPES:
        for P in (1, 2, 4, 8):
            if P <= n // b:
                yield n, b, P


@pytest.mark.parametrize("n,b,P", list(_cases()))
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_pdiv_reference_shapes(built_lib, n, b, P, cplx, sym):
    from sinv_amd.algorithms import pdiv_lm
    from sinv_amd.utils import matu
    print(f"pdiv n={n} b={b} P={P} cplx={cplx} sym={sym}", flush=True)
    A = matu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, 63) + 3 * b * np.eye(n)
    got = pdiv_lm.pdiv(A, b, P)
    ref = pdiv_ref.pdiv_blocks(A, b, P)
    for (gD, gU, gL), (rD, rU, rL) in zip(got, ref):
        for g, r in ((gD, rD), (gU, rU), (gL, rL)):
            g, r = np.stack(g), np.stack(r)
            assert np.abs(g - r).max() <= 1e-12 * max(np.abs(r).max(), 1e-300)
    G = pdiv_ref.stacked(got)
    assert dense.rel_err(G, dense.select(np.linalg.inv(A), b)) < 1e-12
    # the raw reference matrix, with the reference's own acceptance (allclose vs the dense inverse)
    Araw = matu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, 63)
    Gd, Gu, Gl = pdiv_ref.stacked(pdiv_lm.pdiv(Araw, b, P))
    Rd, Ru, Rl = dense.select(np.linalg.inv(Araw), b)
    assert np.allclose(Gd, Rd) and np.allclose(Gu, Ru) and np.allclose(Gl, Rl)


@pytest.mark.parametrize("N,b,P", [(64, 48, 8), (40, 64, 4), (33, 80, 2)])
def test_pdiv_batched_agrees_with_rgf(built_lib, N, b, P):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(N, b, seed=N + b)
    rng = np.random.default_rng(P)
    D = np.stack([d + 0.1 * k * np.eye(b) * (1 + 1j) for k in range(3)])  # 3 instances
    Gp = core.pdiv_batched(D, u, l, P)
    Gr = core.rgf_batched(D, u, l)
    num = sum(float(((a - r).abs() ** 2).sum()) for a, r in zip(Gp, Gr))
    den = sum(float((r.abs() ** 2).sum()) for r in Gr)
    assert np.sqrt(num / den) < 1e-12


def test_pdiv_rejects_bad_partition_counts(built_lib):
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(6, 4, seed=1)
    with pytest.raises(ValueError, match="power of 2"):
        core.pdiv_batched(d, u, l, 3)
    with pytest.raises(ValueError, match="greater than the number of processes"):
        core.pdiv_batched(d, u, l, 8)
