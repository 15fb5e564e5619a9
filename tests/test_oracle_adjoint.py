"""Pins of the gradient oracle (oracle/adjoint_ref.py) on the CPU.

grad_dense (the closed form -G^H grad_G G^H on the pattern) is checked against
central finite differences of the oracle's own RGF, perturbing real and
imaginary parts separately; a planted sign / conjugation slip fails.
"""

import numpy as np
import pytest

from oracle import adjoint_ref, rgf_ref


def _setup(seed=0, N=4, b=3):
    from sinv_amd.utils import configs
    H = configs.hermitian_H(N, b, seed=seed)
    d, u, l = configs.shifted(H, 0.2 + 0.3j)
    rng = np.random.default_rng(seed)
    c = lambda *s: rng.normal(size=s) + 1j * rng.normal(size=s)
    W = (c(N, b, b), c(N - 1, b, b), c(N - 1, b, b))
    return d, u, l, W


def _loss(W):
    def f(d, u, l):
        G = rgf_ref.rgf_blocks(d, u, l)
        return float(sum(np.real(np.vdot(w, g)) for w, g in zip(W, G)))
    return f


def test_grad_dense_matches_fd():
    d, u, l, W = _setup()
    # L = Re <W, G>  =>  grad_G = W
    R = adjoint_ref.grad_dense(d, u, l, *W)
    f = _loss(W)
    rng = np.random.default_rng(3)
    for which, arr in (("d", R[0]), ("u", R[1]), ("l", R[2])):
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in arr.shape)
            fd = adjoint_ref.fd_grad(f, d, u, l, which, idx, h=1e-6)
            assert abs(fd - arr[idx]) < 1e-6 * max(1, abs(arr[idx]))


@pytest.mark.parametrize("slip", ["no_conj", "sign"])
def test_planted_slips_fail_fd(slip):
    from oracle.dense import assemble, select
    d, u, l, W = _setup(seed=1)
    G = np.linalg.inv(assemble(d, u, l))
    gG = assemble(*W)
    bad = (-G @ gG @ G) if slip == "no_conj" else (G.conj().T @ gG @ G.conj().T)
    R = select(bad, d.shape[1])
    f = _loss(W)
    idx = (1, 0, 1)
    fd = adjoint_ref.fd_grad(f, d, u, l, "d", idx, h=1e-6)
    assert abs(fd - R[0][idx]) > 1e-3 * abs(fd)


def test_dos_grad_G_is_the_loss_gradient():
    d, u, l, _ = _setup(seed=2)
    gd = adjoint_ref.dos_grad_G(d.shape)
    f = lambda d_, u_, l_: adjoint_ref.dos_loss(rgf_ref.rgf_blocks(d_, u_, l_)[0])
    R = adjoint_ref.grad_dense(d, u, l, gd, np.zeros_like(u), np.zeros_like(l))
    fd = adjoint_ref.fd_grad(f, d, u, l, "u", (1, 2, 0), h=1e-6)
    assert abs(fd - R[1][1, 2, 0]) < 1e-6 * max(1, abs(fd))
