"""GPU parity of the adjoint sweep (gradient of the selected inversion).

Pins: the dense formula grad_A = -G^H grad_G G^H (oracle/adjoint_ref.py,
pinned to finite differences on the CPU in tests/test_oracle_adjoint.py),
central finite differences of the GPU forward itself, and torch autograd
through a dense torch.linalg.inv (float64, on the CPU).
"""

import numpy as np
import pytest

from oracle import adjoint_ref

pytestmark = pytest.mark.gpu


def _rand_blocks(rng, N, b):
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    return c(N, b, b), c(max(N - 1, 0), b, b), c(max(N - 1, 0), b, b)


@pytest.mark.parametrize("N,b", [(1, 8), (2, 5), (5, 16), (9, 33), (6, 80)])
def test_adjoint_vs_dense_formula(built_lib, N, b):
    from sinv_amd import core
    from sinv_amd.utils import configs
    print(f"adjoint dense N={N} b={b}", flush=True)
    rng = np.random.default_rng(N * 100 + b)
    d, u, l = configs.dominant_blocks(N, b, seed=N + b)
    gd, gu, gl = _rand_blocks(rng, N, b)
    Ad, Au, Al = (x[0].cpu().numpy() for x in core.rgf_adjoint_batched(d, u, l, gd, gu, gl))
    Rd, Ru, Rl = adjoint_ref.grad_dense(d, u, l, gd, gu, gl)
    num = sum(np.linalg.norm(a - r) ** 2 for a, r in zip((Ad, Au, Al), (Rd, Ru, Rl)))
    den = sum(np.linalg.norm(r) ** 2 for r in (Rd, Ru, Rl))
    assert np.sqrt(num / den) < 1e-12


@pytest.mark.parametrize("every", [0, 2, 3, 5])
def test_adjoint_checkpointing_is_exact(built_lib, every):
    """Recomputed segments reproduce the stored ones bit for bit."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(11, 24, seed=4)
    E = configs.energies(3, seed=5) + 1e-2j
    rng = np.random.default_rng(1)
    g = _rand_blocks(rng, 11, 24)
    ref = core.rgf_adjoint_batched(*H, *g, shift=E, dscale=-1, oscale=-1, checkpoint_every=0)
    got = core.rgf_adjoint_batched(*H, *g, shift=E, dscale=-1, oscale=-1, checkpoint_every=every)
    for a, r in zip(got, ref):
        assert (a == r).all()


def test_adjoint_energies_vs_dense(built_lib):
    """A(E) = (E + i eta) I - H for several energies, grad_G shared."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(7, 40, seed=6)
    E = configs.energies(4, seed=7) + 1e-2j
    gd = adjoint_ref.dos_grad_G((7, 40))
    gu = np.zeros((6, 40, 40), complex)
    gl = np.zeros((6, 40, 40), complex)
    Ad, Au, Al = core.rgf_adjoint_batched(*H, gd, gu, gl, shift=E, dscale=-1, oscale=-1,
                                          checkpoint_every=3)
    for k, e in enumerate(E):
        d, u, l = configs.shifted(H, e)
        R = adjoint_ref.grad_dense(d, u, l, gd, gu, gl)
        for a, r in zip((Ad[k], Au[k], Al[k]), R):
            a = a.cpu().numpy()
            assert np.linalg.norm(a - r) <= 1e-11 * np.linalg.norm(r) + 1e-300


def test_dos_gradient_32x128(built_lib):
    """32 blocks of 128, E = 0.3 + 0.01i, L = -(1/pi) sum Im tr G_ii:
    dL/dH through the autograd Function vs torch autograd through a dense
    torch.linalg.inv (CPU, float64) and vs central finite differences."""
    import torch
    from sinv_amd import autograd as sag
    from sinv_amd.utils import configs, matu
    N, b = 32, 128
    H = configs.hermitian_H(N, b, seed=11)
    E = 0.3
    eta = 0.01
    dev = torch.device("cuda")
    Hd, Hu, Hl = (torch.tensor(x, device=dev, requires_grad=True) for x in H)
    Gd, Gu, Gl = sag.greens_function(Hd, Hu, Hl, [E], eta, checkpoint_every=8)
    L = sag.dos_loss(Gd)
    L.backward()
    gd, gu, gl = (x.grad.cpu().numpy() for x in (Hd, Hu, Hl))

    # torch autograd through the dense inverse (CPU float64)
    Hdense = torch.tensor(matu.convertBlkTridiagToDense(*H), requires_grad=True)
    n = N * b
    A = (E + 1j * eta) * torch.eye(n, dtype=torch.complex128) - Hdense
    Ginv = torch.linalg.inv(A)
    Ld = -torch.diagonal(Ginv).imag.sum() / torch.pi
    Ld.backward()
    print(f"DOS loss: rgf {L.item():.15e} dense {Ld.item():.15e}", flush=True)
    assert abs(Ld.item() - L.item()) < 1e-12 * abs(Ld.item())
    gdense = Hdense.grad.numpy()
    rd, ru, rl = matu.convertDenseToBlkTridiag(gdense, b)
    num = np.linalg.norm(gd - rd) ** 2 + np.linalg.norm(gu - ru) ** 2 + np.linalg.norm(gl - rl) ** 2
    den = np.linalg.norm(rd) ** 2 + np.linalg.norm(ru) ** 2 + np.linalg.norm(rl) ** 2
    print(f"dL/dH rel err vs dense torch autograd: {np.sqrt(num / den):.3e}", flush=True)
    assert np.sqrt(num / den) < 1e-10

    # central finite differences (h = 1e-6, Re and Im separately) on sampled entries
    from sinv_amd import core

    def loss(d_, u_, l_):
        G = core.rgf_batched(d_, u_, l_, shift=np.array([E + 1j * eta]), dscale=-1.0,
                             oscale=-1.0, save_off_diag=False)
        return adjoint_ref.dos_loss(G[0][0].cpu().numpy())

    rng = np.random.default_rng(0)
    for which, grad in (("d", gd), ("u", gu), ("l", gl)):
        for _ in range(2):
            blk = rng.integers(0, N - 1)
            i, j = rng.integers(0, b, size=2)
            fd = adjoint_ref.fd_grad(loss, *H, which, (blk, i, j), h=1e-6)
            assert abs(fd - grad[blk, i, j]) < 1e-5 * max(1.0, abs(grad[blk, i, j])), (which, fd, grad[blk, i, j])


def test_adjoint_returns_forward_diagonal(built_lib):
    """The adjoint's sweep over A^H yields G_ii as well: equal to the forward RGF."""
    from sinv_amd import core
    from sinv_amd.utils import configs
    H = configs.hermitian_H(10, 48, seed=21)
    E = configs.energies(3, seed=22) + 1e-2j
    gd = adjoint_ref.dos_grad_G((10, 48))
    z = np.zeros((9, 48, 48), complex)
    Ad, Au, Al, Gd = core.rgf_adjoint_batched(*H, gd, z, z, shift=E, dscale=-1, oscale=-1,
                                              checkpoint_every=3, return_g_diag=True)
    Gf = core.rgf_batched(*H, shift=E, dscale=-1, oscale=-1)[0]
    err = float(((Gd - Gf).abs() ** 2).sum().sqrt() / (Gf.abs() ** 2).sum().sqrt())
    assert err < 1e-12
