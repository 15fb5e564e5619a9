"""Gradient oracles for the adjoint of the selected inversion.

Test infrastructure only (see ``oracle/__init__.py``).

The reference has no adjoint.  For a real loss L of G = A^-1 and torch's complex
gradient convention (grad z = dL/dRe z + i dL/dIm z), the gradient of A is

    grad_A = -G^H grad_G G^H                                     (dense formula)

restricted here to the block-tridiagonal pattern.  For A(E) = (E + i eta) I - H,
grad_H = -grad_A.  ``grad_dense`` evaluates it with a dense inverse;
``fd_grad`` is the central finite difference of a loss, perturbing real and
imaginary parts separately.
"""

import numpy as np

from .dense import assemble, select


def grad_dense(diag, upper, lower, gGd, gGu, gGl):
    b = diag.shape[1]
    G = np.linalg.inv(assemble(diag, upper, lower))
    gG = assemble(gGd, gGu, gGl)
    GH = G.conj().T
    return select(-GH @ gG @ GH, b)


def dos_loss(Gd):
    """L = -(1/pi) sum_i Im tr G_ii (density of states)."""
    return float(-np.imag(np.trace(Gd, axis1=1, axis2=2).sum()) / np.pi)


def dos_grad_G(Gd_shape, dtype=np.complex128):
    """dL/dG for the DOS loss: -(i/pi) I on every diagonal block, 0 elsewhere."""
    N, b = Gd_shape[0], Gd_shape[1]
    gd = np.zeros((N, b, b), dtype) + (-1j / np.pi) * np.eye(b)
    return gd


def fd_grad(loss_of_blocks, diag, upper, lower, which, idx, h=1e-6):
    """Central FD of loss_of_blocks(diag, upper, lower) w.r.t. one entry.

    which in {'d','u','l'}, idx = (block, row, col).  Returns the complex number
    dL/dRe + i dL/dIm (torch's convention)."""
    arrs = {'d': diag, 'u': upper, 'l': lower}
    out = 0j
    for unit in (1.0, 1j):
        vals = []
        for sgn in (+1, -1):
            mod = {k: v.copy() for k, v in arrs.items()}
            mod[which][idx] += sgn * h * unit
            vals.append(loss_of_blocks(mod['d'], mod['u'], mod['l']))
        out += unit * (vals[0] - vals[1]) / (2 * h)
    return out
