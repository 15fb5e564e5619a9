"""GPU parity of the RGF hot path (rgf.py:17-64) against the oracle.

The oracle (oracle/rgf_ref.py) is the op-for-op numpy restatement, itself
pinned to the dense inverse in tests/test_oracle_rgf.py.  Inputs are the
reference tests' own generators and shapes (SEED = 63, rgf_test.py:17,
bcr_serial_test.py:40-53) and the diagonally dominant 16 x 64 workload.
"""

import numpy as np
import pytest

from oracle import dense, rgf_ref
from oracle import matu_ref as mu

pytestmark = pytest.mark.gpu

SEED = 63
REF_SHAPES = [(1, 1), (2, 2), (3, 3), (2, 1), (4, 2), (6, 3), (3, 1), (6, 2), (9, 3),
              (128, 8), (128, 16), (128, 32)]


def _blocks(n, b, cplx, sym):
    A = mu.generateBandedDiagonalMatrix(n, int(np.ceil(b / 2)) if b > 1 else 1, cplx, sym, SEED)
    return mu.convertDenseToBlkTridiag(A, b)


@pytest.mark.parametrize("n,b", REF_SHAPES)
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("sym", [False, True])
def test_rgf_reference_shapes(built_lib, n, b, cplx, sym):
    from sinv_amd.algorithms import rgf
    from sinv_amd.bsparse import BlockTridiag
    print(f"rgf ref-shape n={n} b={b} cplx={cplx} sym={sym}", flush=True)
    d, u, l = _blocks(n, b, cplx, sym)
    A = BlockTridiag(d, u, l)
    G = rgf.rgf(A, sym)
    ref = rgf_ref.rgf_blocks(d, u, l, sym_mat=sym)
    dref = dense.selected_inverse_dense(d, u, l)
    assert G.dtype == d.dtype
    got = (G.diag, G.upper, G.lower)
    # the reference's own acceptance criterion, on all three block bands
    for g, r in zip(got, dref):
        assert np.allclose(g, r)
    cond = np.linalg.cond(A.toarray())
    assert dense.rel_err(got, ref) < 1e-15 * cond * 10


def test_rgf_dominant_16x64(built_lib):
    """16 blocks of 64, complex128, diagonally dominant: 1e-12 vs oracle."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    d, u, l = configs.dominant_blocks(16, 64, seed=1)
    Gd, Gu, Gl = core.rgf_batched(d, u, l)
    got = tuple(x[0].cpu().numpy() for x in (Gd, Gu, Gl))
    ref = rgf_ref.rgf_blocks(d, u, l)
    assert dense.rel_err(got, ref) < 1e-12
    assert dense.rel_err(got, dense.selected_inverse_dense(d, u, l)) < 1e-12
    # sym_mat path: A = A^T for this generator
    Gs = core.rgf_batched(d, u, l, sym_mat=True)
    got_s = tuple(x[0].cpu().numpy() for x in Gs)
    assert dense.rel_err(got_s, rgf_ref.rgf_blocks(d, u, l, sym_mat=True)) < 1e-12


def test_rgf_save_off_diag_false_is_nan(built_lib):
    from sinv_amd.algorithms import rgf
    from sinv_amd.bsparse import BlockTridiag
    d, u, l = _blocks(24, 3, True, False)
    G = rgf.rgf(BlockTridiag(d, u, l), save_off_diag=False)
    ref = rgf_ref.rgf_blocks(d, u, l, save_off_diag=False)
    assert np.isnan(G.upper).all() and np.isnan(G.lower).all()
    assert np.abs(G.diag - ref[0]).max() / np.abs(ref[0]).max() < 1e-12


def test_rgf_energies_batched(built_lib):
    """A(E) = (E + i eta) I - H over a batch of energies: each instance equals
    the oracle on the explicitly formed A(E)."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(6, 96, seed=3)
    E = configs.energies(5, seed=4)
    eta = 1e-2
    Gd, Gu, Gl = core.rgf_energies(*H, E, eta)
    for k, e in enumerate(E):
        d, u, l = configs.shifted(H, e + 1j * eta)
        ref = rgf_ref.rgf_blocks(d, u, l)
        got = (Gd[k].cpu().numpy(), Gu[k].cpu().numpy(), Gl[k].cpu().numpy())
        assert dense.rel_err(got, ref) < 1e-11, k


def test_rgf_singular_block_raises(built_lib):
    from sinv_amd import core
    d, u, l = _blocks(12, 3, True, False)
    d = d.copy()
    d[0] = 0.0
    with pytest.raises(np.linalg.LinAlgError) as ei:
        core.rgf_batched(d, u, l)
    assert ei.value.block == 0


def test_rgf_graph_replays_match_eager(built_lib):
    """The batched sweep captured into a HIP graph: every replay, with eager
    work and synchronisations in between, reproduces the eager result."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(9, 80, seed=12)
    E = configs.energies(6, seed=13) + 1e-2j
    plan = core.RgfPlan(*H, shift=E, dscale=-1.0, oscale=-1.0)
    plan.run()
    torch.cuda.synchronize()
    ref = tuple(x.clone() for x in plan.outputs())
    g = plan.capture()
    for rep in range(4):
        for x in plan.outputs():
            x.zero_()
        other = core.rgf_batched(*H, shift=E[:2], dscale=-1.0, oscale=-1.0)  # eager work between
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, r in zip(plan.outputs(), ref):
            assert torch.equal(a, r), rep
        plan.check()
        del other


def test_planset_concurrent_matches_sequential_and_replays(built_lib):
    """Plans of different block sizes run concurrently on their own streams give
    the same bits as the plans run one after another, eagerly and from a
    captured graph (several replays with eager work in between)."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    plans, refs = [], []
    for gi, b in enumerate((24, 64, 72)):
        H = configs.hermitian_H(7, b, seed=40 + gi)
        E = configs.energies(5, seed=50 + gi) + 1e-2j
        p = core.RgfPlan(*H, shift=E, dscale=-1.0, oscale=-1.0)
        p.run()
        plans.append(p)
    torch.cuda.synchronize()
    refs = [tuple(x.clone() for x in p.outputs()) for p in plans]
    ps = core.PlanSet(plans)

    def check_all(tag):
        for p, ref in zip(plans, refs):
            for a, r in zip(p.outputs(), ref):
                assert torch.equal(a, r), tag
            p.check()

    for p in plans:
        for x in p.outputs():
            x.zero_()
    ps.run()
    torch.cuda.synchronize()
    check_all("eager")
    g = ps.capture()
    for rep in range(3):
        for p in plans:
            for x in p.outputs():
                x.zero_()
        core.rgf_batched(*configs.hermitian_H(3, 16, seed=1))  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        check_all(f"replay {rep}")


@pytest.mark.parametrize("b", [128, 384])
def test_rgf_headline_configuration_fp64(built_lib, b):
    """The batched-energies workload's configuration: 64 blocks of b, A(E) =
    (E + 1e-3 i) I - H with H from the workload's Hermitian generator, three
    energies inside the band [-2, 2], FP64.  Each instance's selected blocks
    against the oracle (rgf.py:40-62 restated, numpy inv / @), relative
    Frobenius over all diag/upper/lower blocks.  Bound: max(1e-12, 4 x the
    instance's rounding floor), the floor being the distance between the
    oracle and the same recurrences with every triple product re-associated
    (an equally valid FP64 evaluation order); the test prints both."""
    from oracle import rgf_ref
    from sinv_amd import core
    from sinv_amd.utils import configs
    eta = 1e-3
    H = configs.hermitian_H(64, b, seed=10)
    E = np.array([-1.3, 0.05, 0.9])
    Gd, Gu, Gl = core.rgf_energies(*H, E, eta)
    for k, e in enumerate(E):
        d, u, l = configs.shifted(H, e + 1j * eta)
        ref = rgf_ref.rgf_blocks(d, u, l)
        alt = rgf_ref.rgf_blocks(d, u, l, reassociate=True)
        got = (Gd[k].cpu().numpy(), Gu[k].cpu().numpy(), Gl[k].cpu().numpy())
        err = dense.rel_err(got, ref)
        floor = dense.rel_err(alt, ref)
        print(f"b={b} E={e}: GPU vs oracle {err:.3e}, rounding floor {floor:.3e}", flush=True)
        assert err <= max(1e-12, 4 * floor), (e, err, floor)
