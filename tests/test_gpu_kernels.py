"""GPU parity of the two kernels under the hot path, through the C ABI.

K2 (batched complex GEMM) against numpy's matmul, K1 (batched block inverse)
against numpy.linalg.inv (the two library calls the reference makes,
rgf.py:44-62, bcr_serial.py:121-383).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _op(X, op):
    return X if op == "N" else (np.swapaxes(X, -1, -2) if op == "T" else np.conj(np.swapaxes(X, -1, -2)))


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (16, 16, 4), (64, 64, 64), (65, 33, 17),
                                   (128, 96, 200), (3, 130, 70)])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T")])
def test_zgemm_fp64(built_lib, M, N, K, opA, opB):
    from sinv_amd import core
    print(f"zgemm M={M} N={N} K={K} {opA}{opB}", flush=True)
    rng = np.random.default_rng(M * 1000 + N * 10 + K)
    batch = 3
    A = _rand(rng, batch, *((M, K) if opA == "N" else (K, M)))
    B = _rand(rng, batch, *((K, N) if opB == "N" else (N, K)))
    C0 = _rand(rng, batch, M, N)
    sh = _rand(rng, batch)
    alpha, beta = 0.7 - 0.2j, -1.3 + 0.5j
    C = core.zgemm_batched(A, B, opA=opA, opB=opB, alpha=alpha, beta=beta, C0=C0, shift=sh)
    ref = alpha * (_op(A, opA) @ _op(B, opB)) + beta * C0
    for z in range(batch):
        ref[z][np.diag_indices(min(M, N))] += sh[z]
    got = C.cpu().numpy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-14 * max(K, 1) ** 0.5 * 10, err


def test_zgemm_shared_operand_and_mixed(built_lib):
    from sinv_amd import core
    rng = np.random.default_rng(5)
    A = _rand(rng, 100, 80)          # shared by the batch (stride 0)
    B = _rand(rng, 4, 80, 120)
    ref = A @ B
    C = core.zgemm_batched(A, B).cpu().numpy()
    assert np.abs(C - ref).max() / np.abs(ref).max() < 1e-14
    Cm = core.zgemm_batched(A, B, precision="mixed").cpu().numpy()
    em = np.abs(Cm - ref).max() / np.abs(ref).max()
    assert 1e-9 < em < 1e-5, em   # complex64 operands, FP32 accumulation


@pytest.mark.parametrize("n", [1, 2, 3, 8, 17, 63, 64, 65, 100, 128, 200, 384, 416])
def test_zinv_dominant(built_lib, n):
    from sinv_amd import core
    print(f"zinv dominant n={n}", flush=True)
    rng = np.random.default_rng(n)
    batch = 4
    A = _rand(rng, batch, n, n) + 2.0 * n * np.eye(n)
    X = core.zinv_batched(A).cpu().numpy()
    ref = np.linalg.inv(A)
    err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
    assert err < 1e-14, err


@pytest.mark.parametrize("n", [2, 5, 16, 33, 64, 65, 96, 128, 200, 384, 416])
def test_zinv_needs_pivoting(built_lib, n):
    """Every block size pivots like LAPACK, over all rows: a random
    non-dominant block, one with a zero leading entry and a permutation invert
    to numpy's result."""
    from sinv_amd import core
    rng = np.random.default_rng(100 + n)
    A = _rand(rng, 3, n, n)
    A[1, 0, 0] = 0.0
    A[2] = np.roll(np.eye(n), 1, axis=0) * (1 + 1j)  # permutation: every pivot off-diagonal
    X = core.zinv_batched(A).cpu().numpy()
    ref = np.linalg.inv(A)
    for z in range(3):
        err = np.linalg.norm(X[z] - ref[z]) / np.linalg.norm(ref[z])
        assert err < 1e-15 * np.linalg.cond(A[z]) * n, (z, err)


@pytest.mark.parametrize("n", [37, 64])
def test_zinv_leaf_both_panel_widths(built_lib, n):
    """The leaf kernel uses 16 waves of 4 columns for launches under 512
    matrices and 8 waves of 8 columns above: both give numpy's inverse, and
    agree with each other to rounding."""
    from sinv_amd import core
    rng = np.random.default_rng(7 + n)
    A = _rand(rng, 600, n, n)
    X_big = core.zinv_batched(A).cpu().numpy()      # 600 matrices: 8-wave panels
    X_small = core.zinv_batched(A[:8]).cpu().numpy()  # 8 matrices: 16-wave panels
    ref = np.linalg.inv(A[:8])
    for z in range(8):
        tol = 1e-15 * np.linalg.cond(A[z]) * n
        assert np.linalg.norm(X_small[z] - ref[z]) / np.linalg.norm(ref[z]) < tol, z
        assert np.linalg.norm(X_big[z] - ref[z]) / np.linalg.norm(ref[z]) < tol, z
    zs = rng.choice(600, 16, replace=False)
    err = max(np.linalg.norm(X_big[z] @ A[z] - np.eye(n)) for z in zs)
    assert err < 1e-11, err


def test_zinv_near_tie_pivot(built_lib):
    """Two pivot candidates whose moduli differ by 1e-7 relative (inside the
    kernel's 2^-15 comparison window, where it takes the earlier row and
    LAPACK the larger one): the inverse is as accurate as numpy's."""
    from sinv_amd import core
    rng = np.random.default_rng(11)
    n = 48
    A = _rand(rng, 2, n, n)
    A[:, 5, 0] = 3.0
    A[:, 9, 0] = 3.0 * (1 + 1e-7)
    A[1, :, 1] *= 1e-3
    A[1, 20, 1] = 2.0 - 1e-9j
    A[1, 30, 1] = -(2.0 + 1e-9)
    X = core.zinv_batched(A).cpu().numpy()
    ref = np.linalg.inv(A)
    for z in range(2):
        err = np.linalg.norm(X[z] - ref[z]) / np.linalg.norm(ref[z])
        assert err < 1e-15 * np.linalg.cond(A[z]) * n, (z, err)


@pytest.mark.parametrize("variant", ["panel16", "panel8", "panel16f", "panel8f", "rows8", "lds"])
def test_zinv_every_leaf_kernel(built_lib, variant):
    """Every selectable leaf kernel inverts random, zero-leading-entry and
    permutation blocks to numpy's result and flags a singular one; rows8 and
    lds take LAPACK's exact first maximum as pivot, the panel kernels the
    earliest row within 2^-15 of it."""
    from sinv_amd import _native as nat, core
    print(f"leaf variant {variant}", flush=True)
    nat.force_leaf_variant(variant)
    try:
        for n in (2, 5, 33, 64):  # n = 1 with a zero entry would be singular
            rng = np.random.default_rng(200 + n)
            A = _rand(rng, 3, n, n)
            A[1, 0, 0] = 0.0
            A[2] = np.roll(np.eye(n), 1, axis=0) * (1 + 1j)
            X = core.zinv_batched(A).cpu().numpy()
            ref = np.linalg.inv(A)
            for z in range(3):
                err = np.linalg.norm(X[z] - ref[z]) / np.linalg.norm(ref[z])
                assert err < 1e-15 * np.linalg.cond(A[z]) * max(n, 2), (n, z, err)
        S = _rand(np.random.default_rng(9), 2, 20, 20)
        S[1, :, 7] = 0.0
        with pytest.raises(np.linalg.LinAlgError) as ei:
            core.zinv_batched(S)
        assert ei.value.instance == 1
    finally:
        nat.force_leaf_variant("auto")


def test_panel_leaf_variants_agree_bitwise(built_lib):
    """The panel kernels differ only in how panels are handed from wave to
    wave (barrier or flags, apply order row- or column-wise, 8 or 16 waves):
    every element receives the same update sequence, so all four give the
    same bits, for full and ragged panels."""
    from sinv_amd import _native as nat, core
    rng = np.random.default_rng(23)
    outs = {}
    try:
        for v in ("panel16", "panel16f", "panel8", "panel8f"):
            nat.force_leaf_variant(v)
            outs[v] = [core.zinv_batched(_rand(np.random.default_rng(n), 6, n, n)).cpu().numpy()
                       for n in (64, 61, 37, 8)]
    finally:
        nat.force_leaf_variant("auto")
    for v in ("panel16f", "panel8", "panel8f"):
        for x, y in zip(outs[v], outs["panel16"]):
            assert np.array_equal(x, y), v


def test_leaf_every_size_both_widths(built_lib):
    """Every leaf size n = 1 .. 64 (full and partial last panels for both
    widths), random non-dominant blocks that need pivoting: numpy's inverse
    to the conditioning, with 4 blocks (16-wave panels) and 600 blocks
    (8-wave panels) per launch."""
    from sinv_amd import core
    for n in range(1, 65):
        rng = np.random.default_rng(1000 + n)
        A = _rand(rng, 600, n, n)
        X_big = core.zinv_batched(A).cpu().numpy()
        X_small = core.zinv_batched(A[:4]).cpu().numpy()
        ref = np.linalg.inv(A[:4])
        for z in range(4):
            tol = 1e-15 * np.linalg.cond(A[z]) * max(n, 2) + 1e-300
            rn = np.linalg.norm(ref[z])
            assert np.linalg.norm(X_small[z] - ref[z]) / rn < tol, (n, z)
            assert np.linalg.norm(X_big[z] - ref[z]) / rn < tol, (n, z)


def test_leaf_handoff_wait_is_bounded(built_lib):
    """Failure handling of the leaf's flag hand-off, forced from the host: with
    the wait bound set to zero polls the waves that would wait give up at
    once, the kernel still exits, and the blocks report as singular; with the
    default bound restored the same blocks invert normally."""
    from sinv_amd import _native as nat, core
    rng = np.random.default_rng(17)
    A = _rand(rng, 4, 64, 64) + 128 * np.eye(64)
    nat.force_leaf_variant("panel16f")
    try:
        assert built_lib.sinv_leaf_spin_limit(0) == 0
        with pytest.raises(np.linalg.LinAlgError):
            core.zinv_batched(A)
    finally:
        assert built_lib.sinv_leaf_spin_limit(-1) == 0
        nat.force_leaf_variant("auto")
    X = core.zinv_batched(A).cpu().numpy()
    assert np.abs(X - np.linalg.inv(A)).max() < 1e-13


def test_leaf_variant_rejects_unknown(built_lib):
    from sinv_amd import _native as nat
    assert built_lib.sinv_leaf_force_variant(7) == nat.SINV_E_INVAL
    assert built_lib.sinv_leaf_force_variant(nat.LEAF_VARIANTS["auto"]) == 0


def test_zinv_singular_raises(built_lib):
    from sinv_amd import core
    A = np.zeros((2, 8, 8), complex)
    A[0] = np.eye(8)
    with pytest.raises(np.linalg.LinAlgError) as ei:
        core.zinv_batched(A)
    assert ei.value.instance == 1


@pytest.mark.parametrize("n,zc,batch", [(64, 40, 4), (37, 20, 4), (64, 63, 600)])
def test_zinv_singular_late_pivot(built_lib, n, zc, batch):
    """A zero column stays exactly zero through the earlier steps, so the leaf
    meets an exactly zero pivot at step zc: the inverse raises, naming the
    instance (both panel widths: 4 and 600 blocks per launch)."""
    from sinv_amd import core
    rng = np.random.default_rng(3)
    A = _rand(rng, batch, n, n)
    A[batch - 1, :, zc] = 0.0
    with pytest.raises(np.linalg.LinAlgError) as ei:
        core.zinv_batched(A)
    assert ei.value.instance == batch - 1


def test_zinv_nan_input_propagates(built_lib):
    """NaN entries propagate to NaN results without a fault (no comparable
    pivot candidate keeps the current row)."""
    from sinv_amd import core
    A = np.tile(np.eye(8, dtype=complex), (3, 1, 1))
    A[1, 2, 3] = np.nan
    A[2] = np.nan
    X = core.zinv_batched(A, check=False).cpu().numpy()
    assert np.allclose(X[0], np.eye(8))
    assert np.isnan(X[1]).any() and np.isnan(X[2]).all()


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (65, 33, 17), (128, 96, 200), (130, 140, 40)])
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "C"), ("C", "T")])
def test_zgemm_every_tile_config(built_lib, M, N, K, opA, opB):
    """Every tile configuration against numpy, and configs with the same k-split
    bit-identical to each other: they differ only in which workgroup computes
    which outputs, the summation over k is the same chunked MFMA sequence in
    each.  Configs with different k-splits sum the same products in another
    order and agree to rounding."""
    from sinv_amd import autotune, core
    import torch
    rng = np.random.default_rng(M + 7 * N + 13 * K)
    batch = 2
    A = _rand(rng, batch, *((M, K) if opA == "N" else (K, M)))
    B = _rand(rng, batch, *((K, N) if opB == "N" else (N, K)))
    C0 = _rand(rng, batch, M, N)
    ref = 0.5 * (_op(A, opA) @ _op(B, opB)) - C0
    outs = {}
    for cfg, tile in autotune.configs():
        print(f"zgemm cfg={cfg} tile={tile} M={M} N={N} K={K} {opA}{opB}", flush=True)
        built_lib.sinv_gemm_force_config(cfg)
        try:
            for prec in ("fp64", "mixed"):
                C = core.zgemm_batched(A, B, opA=opA, opB=opB, alpha=0.5, beta=-1.0, C0=C0,
                                       precision=prec)
                got = C.cpu().numpy()
                err = np.abs(got - ref).max() / np.abs(ref).max()
                assert err < (1e-13 if prec == "fp64" else 1e-5), (cfg, prec, err)
                if prec == "fp64":
                    outs[cfg] = C
        finally:
            built_lib.sinv_gemm_force_config(-1)
    first = {}
    for cfg, C in outs.items():
        ks = built_lib.sinv_gemm_config_ksplit(cfg)
        assert ks >= 1
        if ks in first:
            assert torch.equal(C, outs[first[ks]]), (cfg, first[ks])
        else:
            first[ks] = cfg
    assert len(first) >= 2   # sequential and split-k configs were both exercised


def test_autotune_installs_and_keeps_results(built_lib):
    """Tuning a small batched RGF installs one config per GEMM shape (the one
    that timed fastest).  The sweep's results change at most by rounding (a
    tuned config may split k differently from the default one) and are
    bit-identical from run to run under the tuned table."""
    import torch
    from sinv_amd import autotune, core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(6, 96, seed=3)
    E = configs.energies(4, seed=4) + 1e-2j
    plan = core.RgfPlan(*H, shift=E, dscale=-1.0, oscale=-1.0)
    autotune.clear()
    plan.run()
    torch.cuda.synchronize()
    before = tuple(x.clone() for x in plan.outputs())
    table = autotune.tune(plan.run, reps=1)
    try:
        assert table
        for key, v in table.items():
            assert autotune.select(key) == v["best"]
            assert set(v["ms"]) == {c for c, _ in autotune.configs()}
        plan.run()
        torch.cuda.synchronize()
        plan.check()
        tuned = tuple(x.clone() for x in plan.outputs())
        for a, b in zip(tuned, before):
            assert float((a - b).abs().max()) <= 1e-12 * float(b.abs().max())
        plan.run()
        torch.cuda.synchronize()
        for a, b in zip(plan.outputs(), tuned):
            assert torch.equal(a, b)
    finally:
        autotune.clear()
