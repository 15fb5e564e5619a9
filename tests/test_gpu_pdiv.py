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


@pytest.mark.parametrize("n,b,P", [(9, 3, 2), (128, 8, 4), (128, 16, 8), (128, 32, 2), (6, 3, 2)])
@pytest.mark.parametrize("cplx", [False, True])
def test_pdiv_localmap_per_rank_workflow(built_lib, n, b, P, cplx):
    """The reference's per-rank workflow (pdiv_test.py:81-108): check, divide,
    partition, then every rank calls pdiv_lm.pdiv_localmap(K_i[rank], Bu_i,
    Bl_i, blocksize) -- here P ranks as threads of one process
    (pdiv_u.LocalComm).  Each rank's lists match the oracle's per-rank output
    and the reference's allclose against the dense inverse, with zeros in the
    last rank's bridge slots."""
    from sinv_amd.algorithms import pdiv_lm, pdiv_u
    from sinv_amd.utils import matu
    print(f"pdiv_localmap n={n} b={b} P={P} cplx={cplx}", flush=True)
    A = matu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, False, 63) + 3 * b * np.eye(n)
    pdiv_u.check_multiprocessing(P)
    pdiv_u.check_input(A, b, P)
    starts, sizes = pdiv_u.divide_matrix(A, P, b)
    K_i, Bu_i, Bl_i = pdiv_u.partition_subdomain(A, starts, sizes, b)
    out = pdiv_u.LocalComm(P).run(
        lambda comm: pdiv_lm.pdiv_localmap(K_i[comm.Get_rank()], Bu_i, Bl_i, b, comm=comm))
    ref = pdiv_ref.pdiv_blocks(A, b, P)
    R = np.linalg.inv(A)
    nblocks = n // b
    for rank, ((D, U, L), (rD, rU, rL)) in enumerate(zip(out, ref)):
        assert len(D) == len(U) == len(L) == sizes[rank]
        assert np.asarray(D).dtype == A.dtype
        for g, r in ((D, rD), (U, rU), (L, rL)):
            assert np.abs(np.stack(g) - np.stack(r)).max() <= 1e-12 * max(np.abs(r).max(), 1e-300)
        for i in range(sizes[rank]):
            ip = starts[rank] + i
            assert np.allclose(D[i], R[ip * b:(ip + 1) * b, ip * b:(ip + 1) * b])
            if ip < nblocks - 1:
                assert np.allclose(U[i], R[ip * b:(ip + 1) * b, (ip + 1) * b:(ip + 2) * b])
                assert np.allclose(L[i], R[(ip + 1) * b:(ip + 2) * b, ip * b:(ip + 1) * b])
            else:
                assert not np.any(U[i]) and not np.any(L[i])


def test_pdiv_localmap_single_rank(built_lib):
    """With no communicator (one rank, as a serial caller of the reference
    has) the partition is the whole matrix and there are no bridges."""
    from sinv_amd.algorithms import pdiv_lm
    from sinv_amd.utils import matu
    A = matu.generateBandedDiagonalMatrix(48, 4, True, False, 63) + 24 * np.eye(48)
    D, U, L = pdiv_lm.pdiv_localmap(A, [], [], 8)
    sel = dense.select(np.linalg.inv(A), 8)
    assert np.abs(np.stack(D) - sel[0]).max() < 1e-14
    assert np.abs(np.stack(U[:-1]) - sel[1]).max() < 1e-14 and not np.any(U[-1])
