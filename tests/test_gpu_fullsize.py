"""Parity at the benchmark configurations' full sizes, through properties that
do not need a dense or CPU reference there:

* the selected-block residual identities of G = A^-1 on block row i,
  A_{i,i-1} G_{i-1,i} + A_ii G_ii + A_{i,i+1} G_{i+1,i} = I, and on block
  column i, G_{i,i-1} A_{i-1,i} + G_ii A_ii + G_{i,i+1} A_{i+1,i} = I;
* RGF and BCR, two different eliminations, give the same blocks;
* for A(E) = (E + i eta) I - H with H Hermitian and eta > 0 the density of
  states -Im tr G_ii / pi is non-negative for every block;
* sampled columns of A^-1 from LAPACK's banded LU (an independent pivoted
  solver that fits the host at n = 65536) match the selected blocks.

Sizes: the single large instance (128 blocks of 512, diagonally dominant
generator) and the batched-energies workload's largest block size (64
blocks of 384, Hermitian H, eta = 1e-3) at 8 energies.  The residuals are
formed on the GPU with torch (complex128 products), so nothing of the 1.6 GB
result leaves the device."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _residuals(d, u, l, Gd, Gu, Gl):
    """Largest ||(A G)_ii - I||_F / sqrt(b) and ||(G A)_ii - I||_F / sqrt(b)
    over instances and block rows; inputs (B, N, b, b) / (B, N-1, b, b)."""
    import torch
    b = d.shape[-1]
    eye = torch.eye(b, dtype=d.dtype, device=d.device)
    R = torch.matmul(d, Gd)
    R[:, 1:] += torch.matmul(l, Gu)    # A_{i,i-1} G_{i-1,i}
    R[:, :-1] += torch.matmul(u, Gl)   # A_{i,i+1} G_{i+1,i}
    C = torch.matmul(Gd, d)
    C[:, 1:] += torch.matmul(Gl, u)    # G_{i,i-1} A_{i-1,i}
    C[:, :-1] += torch.matmul(Gu, l)   # G_{i,i+1} A_{i+1,i}
    rn = torch.linalg.matrix_norm(R - eye).max().item() / np.sqrt(b)
    cn = torch.linalg.matrix_norm(C - eye).max().item() / np.sqrt(b)
    return rn, cn


def _residuals_normed(d, u, l, Gd, Gu, Gl):
    """Largest ||(A G)_ii - I||_F / (||A_{i,:}||_F ||G_{:,i}||_F) and
    ||(G A)_ii - I||_F / (||G_{i,:}||_F ||A_{:,i}||_F) over instances and block
    rows (a backward-error scale, comparable across conditioning); inputs are
    torch tensors (B, N, b, b) / (B, N-1, b, b) on any device."""
    import torch
    b = d.shape[-1]
    eye = torch.eye(b, dtype=d.dtype, device=d.device)
    sq = lambda x: torch.linalg.matrix_norm(x) ** 2   # noqa: E731
    R = torch.matmul(d, Gd)
    R[:, 1:] += torch.matmul(l, Gu)
    R[:, :-1] += torch.matmul(u, Gl)
    C = torch.matmul(Gd, d)
    C[:, 1:] += torch.matmul(Gl, u)
    C[:, :-1] += torch.matmul(Gu, l)
    a_row, a_col, g_row, g_col = sq(d), sq(d), sq(Gd), sq(Gd)
    a_row[:, 1:] += sq(l)
    a_row[:, :-1] += sq(u)
    a_col[:, 1:] += sq(u)
    a_col[:, :-1] += sq(l)
    g_col[:, 1:] += sq(Gu)
    g_col[:, :-1] += sq(Gl)
    g_row[:, 1:] += sq(Gl)
    g_row[:, :-1] += sq(Gu)
    rn = (torch.linalg.matrix_norm(R - eye) / torch.sqrt(a_row * g_col)).max().item()
    cn = (torch.linalg.matrix_norm(C - eye) / torch.sqrt(g_row * a_col)).max().item()
    return rn, cn


def _rel(x, y):
    import torch
    num = sum(float(torch.linalg.vector_norm(a - c) ** 2) for a, c in zip(x, y))
    den = sum(float(torch.linalg.vector_norm(c) ** 2) for c in y)
    return np.sqrt(num / den)


def test_single_large_instance_rgf_and_bcr(built_lib):
    """128 blocks of 512 (n = 65536), batch 1: both selected inversions satisfy
    the residual identities and agree with each other."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("full size: 128 blocks of 512, rgf + bcr", flush=True)
    d, u, l = (core.to_dev(x)[None] for x in configs.dominant_blocks(128, 512, seed=2))
    G_rgf = core.rgf_batched(d, u, l)
    G_bcr = core.bcr_batched(d, u, l)
    torch.cuda.synchronize()
    for G in (G_rgf, G_bcr):
        rn, cn = _residuals(d, u, l, *G)
        assert rn < 1e-12 and cn < 1e-12, (rn, cn)
    assert _rel(G_bcr, G_rgf) < 1e-12


def test_single_large_instance_against_banded_columns(built_lib):
    """128 blocks of 512 (n = 65536): columns of A^-1 from an independent
    pivoted solver -- LAPACK's banded LU (zgbsv through
    scipy.linalg.solve_banded, bandwidths 2b - 1) on A e_j -- in a first, a
    middle and a last block column; the GPU's selected blocks in those columns
    (block rows c - 1, c, c + 1) must match them to 1e-12 relative."""
    import scipy.linalg as sl
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("full size: 128 blocks of 512, rgf vs banded-LU columns", flush=True)
    N, b = 128, 512
    dh, uh, lh = configs.dominant_blocks(N, b, seed=2)
    d, u, l = (core.to_dev(x)[None] for x in (dh, uh, lh))
    Gd, Gu, Gl = (x[0] for x in core.rgf_batched(d, u, l))
    torch.cuda.synchronize()
    n, kl = N * b, 2 * b - 1
    ab = np.zeros((2 * kl + 1, n), complex)   # ab[kl + r - c, c] = A[r, c]
    cols = np.arange(b)[None, :]
    rows = np.arange(b)[:, None]
    for i in range(N):
        for j, blk in ((i, dh[i]), (i + 1, uh[i] if i + 1 < N else None),
                       (i - 1, lh[i - 1] if i >= 1 else None)):
            if blk is None:
                continue
            r, c = i * b + rows, j * b + cols
            ab[kl + r - c, np.broadcast_to(c, (b, b))] = blk
    picks = [(0, 3), (N // 2, 257), (N - 1, 510)]     # (block column, column inside it)
    rhs = np.zeros((n, len(picks)), complex)
    for q, (c, k) in enumerate(picks):
        rhs[c * b + k, q] = 1.0
    X = sl.solve_banded((kl, kl), ab, rhs, check_finite=False)
    num = den = 0.0
    for q, (c, k) in enumerate(picks):
        for r, blk in ((c, Gd[c]), (c - 1, Gu[c - 1] if c >= 1 else None),
                       (c + 1, Gl[c] if c + 1 < N else None)):
            if blk is None:
                continue
            got = blk[:, k].cpu().numpy()
            ref = X[r * b:(r + 1) * b, q]
            num += np.linalg.norm(got - ref) ** 2
            den += np.linalg.norm(ref) ** 2
    err = np.sqrt(num / den)
    print(f"selected blocks vs banded-LU columns: {err:.3e}", flush=True)
    assert err < 1e-12


def test_energies_full_blocks_residuals_and_dos(built_lib):
    """64 blocks of 384, A(E) = (E + 1e-3 i) I - H at 8 energies in [-2, 2]
    (inside the band, where A is least well conditioned): the residual
    identities, normalised by the norms of A's block row and G's block column,
    stay within 10x the oracle's own on the same instances, and the DOS of
    every block is non-negative."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("full size: 64 blocks of 384, 8 energies", flush=True)
    eta = 1e-3
    H = configs.hermitian_H(64, 384, seed=12)
    E = configs.energies(8, seed=22)
    Gd, Gu, Gl = core.rgf_energies(*H, E, eta)
    Hd, Hu, Hl = (core.to_dev(x) for x in H)
    sh = torch.as_tensor(E + 1j * eta, dtype=torch.complex128, device=Hd.device)
    eye = torch.eye(384, dtype=torch.complex128, device=Hd.device)
    d = sh[:, None, None, None] * eye - Hd[None]
    u = (-Hu)[None].expand(len(E), -1, -1, -1).contiguous()
    l = (-Hl)[None].expand(len(E), -1, -1, -1).contiguous()
    rn, cn = _residuals_normed(d, u, l, Gd, Gu, Gl)
    print(f"normalised residuals over 8 energies: row {rn:.3e}, column {cn:.3e}", flush=True)
    assert rn < 1e-12 and cn < 1e-12, (rn, cn)
    # against the oracle's own normalised residual on the same instances (CPU)
    from oracle import rgf_ref
    for k in (0, 5):
        hd, hu, hl = configs.shifted(H, E[k] + 1j * eta)
        ref = rgf_ref.rgf_blocks(hd, hu, hl)
        orn, ocn = _residuals_normed(*[torch.as_tensor(x)[None] for x in (hd, hu, hl, *ref)])
        grn, gcn = _residuals_normed(d[k:k + 1], u[k:k + 1], l[k:k + 1], Gd[k:k + 1],
                                     Gu[k:k + 1], Gl[k:k + 1])
        print(f"energy {k}: GPU {grn:.3e} / {gcn:.3e}, oracle {orn:.3e} / {ocn:.3e}", flush=True)
        assert grn < 10 * orn and gcn < 10 * ocn, (k, grn, gcn, orn, ocn)
    dos = -torch.diagonal(Gd, dim1=-2, dim2=-1).imag.sum(-1) / np.pi   # (energies, blocks)
    assert float(dos.min()) >= 0.0, float(dos.min())


def test_adjoint_full_blocks_directional_derivative(built_lib):
    """The forward+adjoint workload's instance (96 blocks, sizes 256 / 416 mixed
    inside the instance, g checkpointed every 8 blocks) at 4 energies, eta =
    1e-3, L = -(1/pi) sum_E sum_i Im tr G_ii: the adjoint gradient's
    directional derivative Re <dL/dH, D> along a random block-tridiagonal D
    equals the central difference of L along D (h = 1e-6)."""
    import torch
    from sinv_amd import core
    from sinv_amd.utils import configs
    print("full size: 96 ragged blocks (256/416), 4 energies, adjoint", flush=True)
    rng = np.random.default_rng(31)
    bs = [int(x) for x in rng.choice([256, 416], size=96)]
    H = configs.hermitian_H_ragged(bs, seed=32)
    E = configs.energies(4, seed=33) + 1e-3j
    dev = torch.device("cuda")
    Hd = [core.to_dev(x) for x in H[0]]
    Hu = [core.to_dev(x) for x in H[1]]
    Hl = [core.to_dev(x) for x in H[2]]
    gd = [(-1j / np.pi) * torch.eye(b, dtype=torch.complex128, device=dev) for b in bs]
    zu = [torch.zeros_like(x) for x in Hu]
    zl = [torch.zeros_like(x) for x in Hl]
    Ad, Au, Al, Gd = core.rgf_adjoint_ragged(Hd, Hu, Hl, gd, zu, zl, shift=E, dscale=-1.0,
                                             oscale=-1.0, checkpoint_every=8,
                                             return_g_diag=True)
    grad = [-a.sum(0) for a in Ad + Au + Al]   # A = (E + i eta) I - H
    drng = np.random.default_rng(5)
    D = [torch.as_tensor(drng.standard_normal(x.shape) + 1j * drng.standard_normal(x.shape),
                         device=dev) for x in Hd + Hu + Hl]
    ddir = sum(float((torch.conj(g) * d).real.sum()) for g, d in zip(grad, D))

    def loss(t):
        nd, nu = len(Hd), len(Hu)
        X = [h + t * d for h, d in zip(Hd + Hu + Hl, D)]
        G = core.rgf_ragged(X[:nd], X[nd:nd + nu], X[nd + nu:], shift=E, dscale=-1.0,
                            oscale=-1.0, save_off_diag=False)[0]
        return sum(-float(torch.diagonal(g, dim1=-2, dim2=-1).imag.sum()) / np.pi for g in G)

    L0 = sum(-float(torch.diagonal(g, dim1=-2, dim2=-1).imag.sum()) / np.pi for g in Gd)
    assert abs(loss(0.0) - L0) < 1e-11 * abs(L0)
    h = 1e-6
    fd = (loss(h) - loss(-h)) / (2 * h)
    print(f"directional derivative: adjoint {ddir:.12e}, central difference {fd:.12e}", flush=True)
    assert abs(fd - ddir) < 1e-5 * abs(ddir), (fd, ddir)
