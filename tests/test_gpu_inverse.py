"""GPU parity of the blocked inverse for blocks larger than 64 (Gauss-Jordan
with partial pivoting over all n rows) against numpy.linalg.inv, the call the
reference makes for every block inverse (rgf.py:44,48; bcr_serial.py:121,
148,249).

Tolerance: 1e-15 * cond(A) * n relative (Frobenius) to numpy's inverse, the
forward-error scale of any backward-stable inversion.  The cases are chosen so
that an unpivoted or block-locally pivoted elimination fails them: a
permutation matrix (every leading block singular), a well-conditioned matrix
whose leading 64 x 64 and n/2 x n/2 blocks are nearly singular, and random
non-dominant blocks at sizes covering every panel configuration (64 < n <=
2048 register panels of 8 to 64 columns, n > 2048 the L2-resident panel).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(rng, *shape):
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def _check(X, A, note):
    ref = np.linalg.inv(A)
    n = A.shape[-1]
    err = np.linalg.norm(X - ref) / np.linalg.norm(ref)
    tol = 1e-15 * np.linalg.cond(A) * n
    assert err < tol, (note, err, tol)
    return err


@pytest.mark.parametrize("n", [65, 128, 256, 384, 416, 512])
def test_zinv_permutation(built_lib, n):
    """Cyclic shifts: every leading sub-block is singular, so only pivoting
    across all rows inverts them; the inverse is the transpose, exactly."""
    from sinv_amd import core
    print(f"zinv permutation n={n}", flush=True)
    P = np.roll(np.eye(n), 1, axis=0).astype(np.complex128)
    Q = np.roll(np.eye(n), n // 2 + 1, axis=1) * (1 - 2j)
    A = np.stack([P, Q])
    X = core.zinv_batched(A).cpu().numpy()
    assert np.array_equal(X[0], P.T)
    assert np.abs(X[1] - np.linalg.inv(Q)).max() <= 1e-15


@pytest.mark.parametrize("n", [128, 200, 384])
def test_zinv_near_singular_leading_block(built_lib, n):
    """A = [[A11, I], [I, 0]] with A11 of condition ~1e13: A itself is well
    conditioned (its inverse is [[0, I], [I, -A11]]), but any elimination that
    pivots only inside the leading block divides by A11's tiny singular value."""
    from sinv_amd import core
    print(f"zinv near-singular A11 n={n}", flush=True)
    rng = np.random.default_rng(n)
    h = n // 2
    U, _ = np.linalg.qr(_rand(rng, h, h))
    V, _ = np.linalg.qr(_rand(rng, h, h))
    s = np.logspace(0, -13, h)
    A11 = (U * s) @ V.conj().T
    A = np.zeros((n, n), complex)
    A[:h, :h] = A11
    A[:h, h:2 * h] = np.eye(h)
    A[h:2 * h, :h] = np.eye(h)
    if n > 2 * h:
        A[2 * h:, 2 * h:] = 1.0
    assert np.linalg.cond(A) < 1e3 and np.linalg.cond(A11) > 1e12
    X = core.zinv_batched(A[None]).cpu().numpy()[0]
    _check(X, A, n)
    inv = np.linalg.inv(A)
    assert np.abs(X - inv).max() < 1e-12


@pytest.mark.parametrize("n", [65, 96, 100, 128, 130, 191, 200, 256, 300, 384, 416, 512,
                               700, 1100, 1500])
def test_zinv_random_non_dominant(built_lib, n):
    """Random blocks with no diagonal dominance, a zero leading entry and a
    zero leading min(64, n/2)-square block, against numpy.linalg.inv."""
    from sinv_amd import core
    print(f"zinv random n={n}", flush=True)
    rng = np.random.default_rng(1000 + n)
    A = _rand(rng, 3, n, n)
    A[1, 0, 0] = 0.0
    k = min(64, n // 2)
    A[2, :k, :k] = 0.0
    X = core.zinv_batched(A).cpu().numpy()
    for z in range(3):
        _check(X[z], A[z], (n, z))


@pytest.mark.parametrize("n", [2100, 2600])
def test_zinv_l2_resident_panels(built_lib, n):
    """Above 2048 rows the panels stay in L2 (16 columns, positions in LDS)."""
    from sinv_amd import core
    print(f"zinv l2-panel n={n}", flush=True)
    rng = np.random.default_rng(n)
    A = _rand(rng, 1, n, n)
    A[0, :64, :64] = 0.0
    X = core.zinv_batched(A).cpu().numpy()
    _check(X[0], A[0], n)


def test_zinv_many_blocks(built_lib):
    """300 blocks of 96 in one launch sequence (the batched path of the RGF
    drivers), including exact pivot ties (a block of +-1 entries)."""
    from sinv_amd import core
    rng = np.random.default_rng(3)
    n = 96
    A = _rand(rng, 300, n, n)
    A[7] = rng.choice([-1.0, 1.0], (n, n)) + 0j
    while np.linalg.cond(A[7]) > 1e6:
        A[7] = rng.choice([-1.0, 1.0], (n, n)) + 0j
    X = core.zinv_batched(A).cpu().numpy()
    for z in (0, 7, 150, 299):
        _check(X[z], A[z], z)
    err = max(np.linalg.norm(X[z] @ A[z] - np.eye(n)) / np.linalg.cond(A[z]) for z in range(300))
    assert err < 1e-13, err


@pytest.mark.parametrize("n", [80, 300, 2100])
def test_zinv_singular_reported(built_lib, n):
    """A zero column makes every pivot candidate of that column zero: the
    instance is reported through info (LinAlgError, like numpy)."""
    from sinv_amd import core
    rng = np.random.default_rng(n)
    A = _rand(rng, 2, n, n)
    A[1, :, n // 3] = 0.0
    with pytest.raises(np.linalg.LinAlgError) as ei:
        core.zinv_batched(A)
    assert ei.value.instance == 1


@pytest.mark.parametrize("n", [100, 256, 384, 416, 512, 700, 1000, 1100])
def test_zinv_all_panel_kernels(built_lib, n):
    """The flow, per-column and per-panel tall-panel kernels (different panel
    widths and update orders, so different sweep shapes and rounding) all give
    numpy's inverse, pivoting over all rows, and agree with each other to
    rounding."""
    from sinv_amd import _native as nat, core
    L = nat.lib()
    rng = np.random.default_rng(77 + n)
    A = _rand(rng, 2, n, n)
    A[1, 0, :] *= 1e-3
    out = {}
    try:
        for v in (1, 0, 2):   # SINV_GJ_PANEL, SINV_GJ_STEP, SINV_GJ_FLOW
            assert L.sinv_gj_force_variant(v) == 0
            out[v] = core.zinv_batched(A).cpu().numpy()
    finally:
        L.sinv_gj_force_variant(2)
    for z in range(2):
        for v in (0, 1, 2):
            _check(out[v][z], A[z], (n, z, v))
        for v in (0, 2):
            assert np.linalg.norm(out[v][z] - out[1][z]) / np.linalg.norm(out[1][z]) < \
                1e-15 * np.linalg.cond(A[z]) * n


def test_zinv_flow_wait_bound_reports(built_lib):
    """A flow-kernel wave whose hand-off wait runs out stops and reports its
    matrix through info (forced from the host with a zero poll budget), so the
    kernel exits; the default budget restores normal results."""
    from sinv_amd import _native as nat, core
    L = nat.lib()
    rng = np.random.default_rng(5)
    A = _rand(rng, 2, 200, 200)
    try:
        assert L.sinv_gj_force_variant(2) == 0
        assert L.sinv_leaf_spin_limit(0) == 0
        with pytest.raises(np.linalg.LinAlgError):
            core.zinv_batched(A)
    finally:
        L.sinv_leaf_spin_limit(-1)
    _check(core.zinv_batched(A).cpu().numpy()[1], A[1], "after")
