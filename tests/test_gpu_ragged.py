"""GPU parity with per-block sizes b_i (ragged block-tridiagonal matrices):
RGF and its adjoint against the dense inverse and the list-based oracle."""

import numpy as np
import pytest

from oracle import dense, rgf_ref

pytestmark = pytest.mark.gpu


def _rand_lists(rng, bs, shift):
    c = lambda r, k: rng.normal(size=(r, k)) + 1j * rng.normal(size=(r, k))
    d = [c(b, b) + shift * b * np.eye(b) for b in bs]
    u = [c(bs[i], bs[i + 1]) for i in range(len(bs) - 1)]
    l = [c(bs[i + 1], bs[i]) for i in range(len(bs) - 1)]
    return d, u, l


@pytest.mark.parametrize("bs", [[3, 5, 2, 8, 4], [70, 33, 96, 64, 130, 17], [256, 416, 256]])
def test_rgf_ragged_vs_dense(built_lib, bs):
    from sinv_amd import core
    print(f"rgf ragged bs={bs}", flush=True)
    rng = np.random.default_rng(len(bs))
    d, u, l = _rand_lists(rng, bs, 3.0)
    Gd, Gu, Gl = core.rgf_ragged(d, u, l)
    got = ([x[0].cpu().numpy() for x in Gd], [x[0].cpu().numpy() for x in Gu],
           [x[0].cpu().numpy() for x in Gl])
    R = dense.select_lists(np.linalg.inv(dense.assemble_lists(d, u, l)), bs)
    assert dense.rel_err_lists(got, R) < 1e-12
    O = rgf_ref.rgf_lists(d, u, l)
    assert dense.rel_err_lists(got, O) < 1e-12


def test_rgf_ragged_energies(built_lib):
    """A(E) = (E + i eta) I - H with b in {256, 416} mixed inside one instance."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    bs = [256, 416, 416, 256, 416, 256]
    H = configs.hermitian_H_ragged(bs, seed=3)
    E = configs.energies(3, seed=4) + 1e-2j
    Gd, Gu, Gl = core.rgf_ragged(*H, shift=E, dscale=-1.0, oscale=-1.0)
    for k in range(len(E)):
        d, u, l = configs.shifted_lists(H, E[k])
        O = rgf_ref.rgf_lists(d, u, l)
        got = ([x[k].cpu().numpy() for x in Gd], [x[k].cpu().numpy() for x in Gu],
               [x[k].cpu().numpy() for x in Gl])
        assert dense.rel_err_lists(got, O) < 1e-11


@pytest.mark.parametrize("every", [0, 2])
def test_adjoint_ragged_vs_dense(built_lib, every):
    from sinv_amd import core
    bs = [40, 72, 24, 65, 33]
    rng = np.random.default_rng(9)
    d, u, l = _rand_lists(rng, bs, 3.0)
    g = _rand_lists(rng, bs, 0.0)
    Ad, Au, Al = core.rgf_adjoint_ragged(d, u, l, *g, checkpoint_every=every)
    got = ([x[0].cpu().numpy() for x in Ad], [x[0].cpu().numpy() for x in Au],
           [x[0].cpu().numpy() for x in Al])
    G = np.linalg.inv(dense.assemble_lists(d, u, l))
    gG = dense.assemble_lists(*g)
    R = dense.select_lists(-G.conj().T @ gG @ G.conj().T, bs)
    assert dense.rel_err_lists(got, R) < 1e-12


def test_adjoint_ragged_sum_batch(built_lib):
    """sum_batch=True returns the gradient blocks summed over the batch (the
    energies' dL/dA added up), equal to summing the per-instance blocks, with
    G_ii still per instance."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    bs = [40, 72, 24, 65]
    H = configs.hermitian_H_ragged(bs, seed=5)
    E = configs.energies(5, seed=6) + 1e-2j
    gd = [(-1j / np.pi) * np.eye(b) for b in bs]
    zu = [np.zeros((bs[i], bs[i + 1]), complex) for i in range(len(bs) - 1)]
    zl = [np.zeros((bs[i + 1], bs[i]), complex) for i in range(len(bs) - 1)]
    kw = dict(shift=E, dscale=-1.0, oscale=-1.0, checkpoint_every=2, return_g_diag=True)
    per = core.rgf_adjoint_ragged(*H, gd, zu, zl, **kw)
    tot = core.rgf_adjoint_ragged(*H, gd, zu, zl, sum_batch=True, **kw)
    for part_per, part_tot in zip(per[:3], tot[:3]):
        for a, s in zip(part_per, part_tot):
            assert s.shape == a.shape[1:]
            ref = a.sum(0)
            assert float((s - ref).abs().max()) <= 1e-13 * max(float(ref.abs().max()), 1.0)
    for a, s in zip(per[3], tot[3]):
        assert torch.equal(a, s)
