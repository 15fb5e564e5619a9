"""Host-side logic of the drop-in layer (no GPU): the bsparse-like container,
the converters and generators of matrices_utils.py, and the workload
generators."""

import numpy as np
import pytest

from oracle import matu_ref as mu
from sinv_amd.bsparse import BlockTridiag
from sinv_amd.utils import configs, matu


def test_generators_match_reference_streams():
    for cplx in (False, True):
        for sym in (False, True):
            a = matu.generateRandomNumpyMat(17, cplx, sym, 63)
            b = mu.generateRandomNumpyMat(17, cplx, sym, 63)
            assert np.array_equal(a, b)
            a = matu.generateBandedDiagonalMatrix(40, 3, cplx, sym, 63)
            b = mu.generateBandedDiagonalMatrix(40, 3, cplx, sym, 63)
            assert np.array_equal(a, b)


def _golden_generators():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "generator_vectors.json")
    with open(path) as f:
        return json.load(f)


def _unpack(x):
    a = np.asarray(x["re"])
    return a + 1j * np.asarray(x["im"]) if "im" in x else a


def test_generators_match_the_references_frozen_streams():
    """Frozen outputs of the reference's own matrices_utils.py on its tests'
    shapes (tests/golden/make_generator_vectors.py): every generated matrix of
    sinv_amd.utils.matu and of the oracle has the same bytes (SHA-256), dtype
    and shape, for the banded generator (with the tests' float bandwidth
    ceil(b / 2) as well) and the dense one."""
    import hashlib
    g = _golden_generators()
    assert len(g["banded"]) == 72 and len(g["random"]) == 16
    for c in g["banded"]:
        for impl in (matu, mu):
            for bw in (c["bandwidth"], float(c["bandwidth"])):
                A = impl.generateBandedDiagonalMatrix(c["size"], bw, c["is_complex"],
                                                      c["is_symmetric"], c["seed"])
                assert str(A.dtype) == c["dtype"] and list(A.shape) == c["shape"]
                assert hashlib.sha256(np.ascontiguousarray(A).tobytes()).hexdigest() == c["sha256"], c
    for c in g["random"]:
        for impl in (matu, mu):
            A = impl.generateRandomNumpyMat(c["size"], c["is_complex"], c["is_symmetric"], c["seed"])
            assert hashlib.sha256(np.ascontiguousarray(A).tobytes()).hexdigest() == c["sha256"], c
            assert np.array_equal(A[0], _unpack(c["first_row"]))


def test_converters_match_the_references_frozen_outputs():
    g = _golden_generators()
    for c in g["convert"]:
        A = _unpack(c["A"])
        d, u, l = matu.convertDenseToBlkTridiag(A, c["blocksize"])
        for got, key in ((d, "diag"), (u, "upper"), (l, "lower")):
            assert np.array_equal(np.asarray(got), _unpack(c[key])), key
        assert np.array_equal(matu.convertBlkTridiagToDense(d, u, l), _unpack(c["dense"]))


def test_golden_pin_catches_a_shifted_stream():
    """The pin is not vacuous: drawing the imaginary part before the real one
    (a plausible slip in the complex generator) changes the bytes."""
    import hashlib
    c = next(x for x in _golden_generators()["banded"] if x["is_complex"] and x["size"] == 36)
    np.random.seed(c["seed"])
    n = c["size"]
    im = np.random.rand(n, n)
    re = np.random.rand(n, n)
    A = re + 1j * im
    assert hashlib.sha256(A.tobytes()).hexdigest() != c["sha256"]


def test_banded_mask_against_loop():
    """matrices_utils.py:86-89 loops over all entries; the vectorised mask agrees."""
    A = matu.generateBandedDiagonalMatrix(23, 4.0, True, False, 5)
    B = mu.generateRandomNumpyMat(23, True, False, 5)
    for i in range(23):
        for j in range(23):
            if i - j > 4 or j - i > 4:
                B[i, j] = 0
    assert np.array_equal(A, B)


def test_converters_round_trip():
    A = matu.generateBandedDiagonalMatrix(48, 6, True, False, 1)
    d, u, l = matu.convertDenseToBlkTridiag(A, 8)
    assert d.shape == (6, 8, 8) and u.shape == (5, 8, 8)
    assert np.array_equal(matu.convertBlkTridiagToDense(d, u, l), A)
    d2, u2, l2 = mu.convertDenseToBlkTridiag(A, 8)
    assert np.array_equal(d, d2) and np.array_equal(u, u2) and np.array_equal(l, l2)


def test_container_surface():
    d, u, l = configs.dominant_blocks(5, 3, seed=0)
    A = BlockTridiag(d, u, l)
    assert A.blockorder == 5 and A.blocksize == 3 and A.shape == (15, 15)
    assert np.array_equal(A.blocks[2, 3], u[2]) and np.array_equal(A.blocks[3, 2], l[2])
    assert not A.blocks[0, 4].any()
    G = A.copy() * np.nan
    assert np.isnan(G.diag).all() and not np.isnan(A.diag).any()
    G.blocks[1, 1] = d[1]
    assert np.array_equal(G.diag[1], d[1])
    with pytest.raises(IndexError):
        G.blocks[0, 2] = d[0]
    assert np.array_equal(A.toarray(), matu.convertBlkTridiagToDense(d, u, l))
    B = BlockTridiag.from_dense(A.toarray(), 3)
    assert np.array_equal(B.toarray(), A.toarray())


def test_workload_generators():
    d, u, l = configs.dominant_blocks(4, 8, seed=1)
    A = matu.convertBlkTridiagToDense(d, u, l)
    assert np.array_equal(A, A.T)                      # complex symmetric
    off = np.abs(A).sum(1) - np.abs(np.diag(A))
    assert (np.abs(np.diag(A)) > off).all()             # strictly diagonally dominant
    H = configs.hermitian_H(4, 8, seed=1)
    Hd = matu.convertBlkTridiagToDense(*H)
    assert np.allclose(Hd, Hd.conj().T)
    assert configs.rgf_flops(16, 64) == 8 * 64 ** 3 * (9 * 16 - 8)
    assert configs.rgf_flops_ragged([64] * 16) == configs.rgf_flops(16, 64)


def test_product_path_has_no_oracle_import():
    """The product package never imports the oracle (it is test infrastructure)."""
    import pathlib
    root = pathlib.Path(__file__).resolve().parents[1] / "sinv_amd"
    for f in root.rglob("*.py"):
        txt = f.read_text()
        assert "import oracle" not in txt and "from oracle" not in txt, f


def test_bcr_parallel_partitioning_helpers():
    """bcr_utils.py:133-184: equal strides, the last partition takes the rest."""
    from sinv_amd.algorithms import bcr_p
    assert bcr_p.divide_matrix(16, 4) == ([0, 4, 8, 12], [4, 4, 4, 4])
    assert bcr_p.divide_matrix(15, 4) == ([0, 3, 6, 9], [3, 3, 3, 6])
    assert bcr_p.get_process_rowblock_index(6, 3) == (6, 9)


def test_bcr_parallel_schedule_known_answer():
    """15 padded block rows over 2 ranks (rows [0, 7) and [7, 15)), worked by
    hand from bcr_parallel.py:227-256 (a rank reduces the rows of i_elim in
    its range) and :733-758 (it produces the rows of i_prod in its range)."""
    from sinv_amd.algorithms import bcr_p
    red = bcr_p.reduction_schedule(15, 2)
    assert [r[0] for r in red] == [[1, 3, 5, 7, 9, 11, 13], [3, 7, 11], [7]]
    assert [r[1] for r in red] == [[(0, 2), (3, 6)], [(0, 0), (1, 2)], [None, (0, 0)]]
    prod = bcr_p.production_schedule(15, 2)
    assert [p[0] for p in prod] == [[3, 11], [1, 5, 9, 13], [0, 2, 4, 6, 8, 10, 12, 14]]
    assert [p[1] for p in prod] == [[[0], [1]], [[0, 1], [2, 3]], [[0, 1, 2, 3], [4, 5, 6, 7]]]


@pytest.mark.parametrize("nbp", [1, 3, 7, 15, 31, 63, 127])
def test_bcr_parallel_schedule_covers_every_row_once(nbp):
    """For every rank count, each level's reduced and produced rows are split
    among the ranks without gap or overlap, so one level-batched launch over
    all partitions does the ranks' combined work."""
    from sinv_amd.algorithms import bcr_p
    for P in range(1, min(nbp, 9) + 1):
        bcr_p.check_schedule_covers(nbp, P)
        for i_elim, per in bcr_p.reduction_schedule(nbp, P):
            bounds = [bcr_p.get_process_rowblock_index(s, z)
                      for s, z in zip(*bcr_p.divide_matrix(nbp, P))]
            for (top, bottom), r in zip(bounds, per):
                if r is not None:
                    assert all(top <= i_elim[q] < bottom for q in range(r[0], r[1] + 1))


def test_bcr_parallel_schedule_check_catches_a_gap():
    """The coverage check fails when a level's row is left to no partition
    (a planted mistake: one rank's range dropped)."""
    from sinv_amd.algorithms import bcr_p
    orig = bcr_p.reduction_schedule

    def dropped(nbp, P):
        out = orig(nbp, P)
        i_elim, per = out[0]
        return [(i_elim, [None] + per[1:])] + out[1:]

    bcr_p.reduction_schedule = dropped
    try:
        with pytest.raises(AssertionError):
            bcr_p.check_schedule_covers(15, 2)
    finally:
        bcr_p.reduction_schedule = orig


def test_newton_iteration_count_bound():
    """newton_iterations(rho) is the first step k with rho^(2^k) <= 1e-9 (the
    acceptance bound; the 1-norm residual after k steps is at most rho^(2^k))."""
    from sinv_amd import core
    assert core.newton_iterations(1.0) is None and core.newton_iterations(1.5) is None
    assert core.newton_iterations(float("nan")) is None
    assert core.newton_iterations(0.0) == 1
    for rho in np.linspace(0.001, 0.995, 200):
        k = core.newton_iterations(rho)
        assert rho ** (2 ** k) <= 1e-9 * (1 + 1e-12)
        if k > 1:
            assert rho ** (2 ** (k - 1)) > 1e-9


def test_newton_schulz_on_dominant_schur_complements():
    """The facts "auto" rests on, on the CPU: for a strictly column diagonally
    dominant block-tridiagonal A, every Schur complement M_i that RGF inverts
    (rgf.py:44-48, formed by the oracle's recurrence) has a column ratio no
    larger than A's, and the iteration from X_1 = 2 D^-1 - D^-1 M D^-1 reaches
    the acceptance bound within newton_iterations(rho_A) steps and then matches
    numpy.linalg.inv.  A planted non-dominant block fails the first fact."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(5, 48, seed=3)

    def ratio(M):
        a = np.abs(M)
        dg = np.diag(a)
        return ((a.sum(0) - dg) / dg).max()

    b = 48
    A = np.zeros((5 * b, 5 * b), complex)
    for i in range(5):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = d[i]
        if i < 4:
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = u[i]
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = l[i]
    rho_a = ratio(A)
    assert rho_a < 1
    kmax = core.newton_iterations(rho_a)
    g = None
    for i in range(5):
        M = d[i] if i == 0 else d[i] - l[i - 1] @ g @ u[i - 1]
        assert ratio(M) <= rho_a
        dg = np.diag(M)
        X = 2 * np.diag(1 / dg) - (M / dg[:, None]) / dg[None, :]
        # X_1 is one Newton-Schulz step from the Jacobi start D^-1
        X0 = np.diag(1 / dg)
        assert np.allclose(X, X0 @ (2 * np.eye(b) - M @ X0), rtol=1e-14, atol=0)
        for k in range(1, kmax + 1):
            T = 2 * np.eye(b) - M @ X
            delta = np.abs(T - np.eye(b)).max()
            X = X @ T
            if delta <= 1e-9:
                break
        assert delta <= 1e-9
        ref = np.linalg.inv(M)
        assert np.linalg.norm(X - ref) / np.linalg.norm(ref) < 1e-13
        g = X
    # planted: a block with a weak diagonal breaks dominance (ratio above 1)
    bad = d[0] - 2.9 * b * np.eye(b)
    assert ratio(bad) > 1
