"""Pins of the RGF oracle (oracle/rgf_ref.py, restating rgf.py:40-62).

Each check is chosen so that a plausible slip in the restatement (a dropped
Schur term, a transposed coupling, a wrong sign in the backward sweep, the
symmetric shortcut on a non-symmetric matrix) fails at least one of them --
see test_planted_mistakes_are_caught.
"""

import numpy as np
import pytest

from oracle import dense, rgf_ref
from oracle import matu_ref as mu

SEED = 63
RGF_SHAPES = [(1, 1), (2, 2), (3, 3), (2, 1), (4, 2), (6, 3), (3, 1), (6, 2), (9, 3),
              (128, 8), (128, 16), (128, 32)]


def _banded_blocks(n, b, cplx, sym):
    A = mu.generateBandedDiagonalMatrix(n, np.ceil(b / 2), cplx, sym, SEED)
    return mu.convertDenseToBlkTridiag(A, b)


@pytest.mark.parametrize("n,b", RGF_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_rgf_oracle_equals_dense_inverse(n, b, cplx, sym):
    """The reference's acceptance test (rgf_test.py:81-94), on all three bands
    (the reference's own test never asserts the diagonal, rgf_test.py:87)."""
    d, u, l = _banded_blocks(n, b, cplx, sym)
    G = rgf_ref.rgf_blocks(d, u, l, sym_mat=sym)
    R = dense.selected_inverse_dense(d, u, l)
    for g, r in zip(G, R):
        assert np.allclose(g, r)
    cond = np.linalg.cond(dense.assemble(d, u, l))
    assert dense.rel_err(G, R) < 1e-15 * cond * 10


def test_closed_form_scalar_laplacian():
    """b = 1, A = tridiag(-1, 2, -1): G_ij = min(i,j)(n+1-max(i,j))/(n+1) (1-based)."""
    n = 40
    d = np.full((n, 1, 1), 2.0)
    u = np.full((n - 1, 1, 1), -1.0)
    l = np.full((n - 1, 1, 1), -1.0)
    Gd, Gu, Gl = rgf_ref.rgf_blocks(d, u, l)
    i = np.arange(1, n + 1)
    exact_d = i * (n + 1 - i) / (n + 1)
    exact_off = i[:-1] * (n + 1 - (i[:-1] + 1)) / (n + 1)
    assert np.allclose(Gd[:, 0, 0], exact_d, rtol=1e-13)
    assert np.allclose(Gu[:, 0, 0], exact_off, rtol=1e-13)
    assert np.allclose(Gl[:, 0, 0], exact_off, rtol=1e-13)


def test_decoupled_blocks_and_single_block():
    rng = np.random.default_rng(0)
    d = rng.normal(size=(5, 4, 4)) + 4 * np.eye(4)
    z = np.zeros((4, 4, 4))
    Gd, Gu, Gl = rgf_ref.rgf_blocks(d, z, z)
    assert np.allclose(Gd, np.linalg.inv(d), rtol=1e-14)
    assert np.allclose(Gu, 0) and np.allclose(Gl, 0)
    Gd1, _, _ = rgf_ref.rgf_blocks(d[:1], z[:0], z[:0])
    assert np.allclose(Gd1[0], np.linalg.inv(d[0]), rtol=1e-14)


def test_residual_identities_and_symmetry():
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(12, 16, seed=3)
    G = rgf_ref.rgf_blocks(d, u, l)
    assert dense.residuals(d, u, l, *G) < 1e-14
    # A = A^T  =>  G = G^T: G_ii symmetric, G_{i,i+1} = G_{i+1,i}^T
    assert np.allclose(G[0], np.swapaxes(G[0], 1, 2), rtol=1e-12)
    assert np.allclose(G[1], np.swapaxes(G[2], 1, 2), rtol=1e-12)


def test_save_off_diag_false_leaves_nan():
    d, u, l = _banded_blocks(24, 3, True, False)
    Gd, Gu, Gl = rgf_ref.rgf_blocks(d, u, l, save_off_diag=False)
    assert np.isnan(Gu).all() and np.isnan(Gl).all()
    assert np.allclose(Gd, dense.selected_inverse_dense(d, u, l)[0])


def test_dos_positivity():
    """For eta > 0 and Hermitian H: -Im tr G_ii >= 0 on every block."""
    from sinv_amd.utils import configs
    H = configs.hermitian_H(6, 12, seed=1)
    d, u, l = configs.shifted(H, 0.3 + 0.05j)
    Gd, _, _ = rgf_ref.rgf_blocks(d, u, l)
    assert (-np.imag(np.trace(Gd, axis1=1, axis2=2)) > 0).all()


def _mistaken_rgf(kind, d, u, l):
    inv = np.linalg.inv
    N = d.shape[0]
    Gd = np.empty_like(d)
    Gu = np.empty_like(u)
    Gl = np.empty_like(l)
    Gd[0] = inv(d[0])
    for i in range(1, N):
        if kind == "dropped_schur":
            Gd[i] = inv(d[i])
        elif kind == "transposed_coupling":
            Gd[i] = inv(d[i] - u[i - 1] @ Gd[i - 1] @ l[i - 1])
        else:
            Gd[i] = inv(d[i] - l[i - 1] @ Gd[i - 1] @ u[i - 1])
    for i in range(N - 2, -1, -1):
        g = Gd[i].copy()
        X = Gd[i + 1] @ l[i] @ g
        Gl[i] = X if kind == "sign" else -X
        Gu[i] = Gl[i].T if kind == "sym_shortcut" else -g @ u[i] @ Gd[i + 1]
        Gd[i] = g + g @ u[i] @ X
    return Gd, Gu, Gl


@pytest.mark.parametrize("kind", ["dropped_schur", "transposed_coupling", "sign", "sym_shortcut"])
def test_planted_mistakes_are_caught(kind):
    """A restatement with one planted slip fails the dense-inverse pin."""
    d, u, l = _banded_blocks(128, 8, True, False)
    G = _mistaken_rgf(kind, d, u, l)
    R = dense.selected_inverse_dense(d, u, l)
    assert dense.rel_err(G, R) > 1e-6


def test_reassociated_order_is_an_equally_valid_evaluation():
    """rgf_blocks(reassociate=True), used to measure an instance's rounding
    floor, evaluates the same recurrences: it matches the dense inverse as
    closely as the reference order does, and differs from it only by
    rounding."""
    from sinv_amd.utils import configs
    H = configs.hermitian_H(12, 24, seed=4)
    d, u, l = configs.shifted(H, 0.3 + 1e-3j)
    ref = rgf_ref.rgf_blocks(d, u, l)
    alt = rgf_ref.rgf_blocks(d, u, l, reassociate=True)
    exact = dense.selected_inverse_dense(d, u, l)
    e_ref, e_alt = dense.rel_err(ref, exact), dense.rel_err(alt, exact)
    assert e_alt < 10 * max(e_ref, 1e-15) and e_ref < 1e-10
    assert 0 < dense.rel_err(alt, ref) < 1e-10
