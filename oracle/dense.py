"""Closed-form pins for the selected inverse.

Test infrastructure only (see ``oracle/__init__.py``).

The reference's own acceptance test is the dense inverse restricted to the
block-tridiagonal pattern (rgf_test.py:81-94, bcr_serial_test.py:66-73).  The
functions here assemble that check, plus the exact selected-block residual
identities that hold for G = A^-1 and involve only selected blocks:

  A_{i,i-1} G_{i-1,i} + A_ii G_ii + A_{i,i+1} G_{i+1,i} = I      (row form)
  G_{i,i-1} A_{i-1,i} + G_ii A_ii + G_{i,i+1} A_{i+1,i} = I      (column form)

They are usable where the dense inverse is too large to form.
"""

import numpy as np


def assemble(diag, upper, lower):
    N, b = diag.shape[0], diag.shape[1]
    A = np.zeros((N * b, N * b), dtype=diag.dtype)
    for i in range(N):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = diag[i]
        if i < N - 1:
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = upper[i]
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = lower[i]
    return A


def select(M, b):
    N = M.shape[0] // b
    d = np.stack([M[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(N)])
    if N > 1:
        u = np.stack([M[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] for i in range(N - 1)])
        l = np.stack([M[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(N - 1)])
    else:
        u = np.zeros((0, b, b), M.dtype)
        l = np.zeros((0, b, b), M.dtype)
    return d, u, l


def selected_inverse_dense(diag, upper, lower):
    """sel(inv(A)) via numpy.linalg.inv of the assembled matrix."""
    return select(np.linalg.inv(assemble(diag, upper, lower)), diag.shape[1])


def rel_err(G, Gref):
    """Frobenius relative error over all selected blocks (diag, upper, lower)."""
    num = sum(np.linalg.norm((np.asarray(a) - np.asarray(r)).ravel()) ** 2 for a, r in zip(G, Gref))
    den = sum(np.linalg.norm(np.asarray(r).ravel()) ** 2 for r in Gref)
    return float(np.sqrt(num / den))


def residuals(diag, upper, lower, Gd, Gu, Gl):
    """Max over block rows of the two exact selected-block residual identities,
    relative to ||A_i|| * ||G_i|| of the row.  Zero for the exact inverse."""
    N, b = diag.shape[0], diag.shape[1]
    eye = np.eye(b)
    worst = 0.0
    for i in range(N):
        R = diag[i] @ Gd[i] - eye
        C = Gd[i] @ diag[i] - eye
        scale = np.linalg.norm(diag[i]) * np.linalg.norm(Gd[i])
        if i > 0:
            R = R + lower[i - 1] @ Gu[i - 1]
            C = C + Gl[i - 1] @ upper[i - 1]
        if i < N - 1:
            R = R + upper[i] @ Gl[i]
            C = C + Gu[i] @ lower[i]
        worst = max(worst, np.linalg.norm(R) / scale, np.linalg.norm(C) / scale)
    return worst


def assemble_lists(diag, upper, lower):
    """Dense matrix of a block-tridiagonal matrix with per-block sizes."""
    bs = [d.shape[0] for d in diag]
    off = np.concatenate([[0], np.cumsum(bs)])
    A = np.zeros((off[-1], off[-1]), dtype=np.result_type(*diag))
    for i in range(len(bs)):
        A[off[i]:off[i + 1], off[i]:off[i + 1]] = diag[i]
        if i < len(bs) - 1:
            A[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = upper[i]
            A[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = lower[i]
    return A


def select_lists(M, bs):
    off = np.concatenate([[0], np.cumsum(bs)])
    n = len(bs)
    d = [M[off[i]:off[i + 1], off[i]:off[i + 1]] for i in range(n)]
    u = [M[off[i]:off[i + 1], off[i + 1]:off[i + 2]] for i in range(n - 1)]
    l = [M[off[i + 1]:off[i + 2], off[i]:off[i + 1]] for i in range(n - 1)]
    return d, u, l


def rel_err_lists(G, R):
    num = sum(np.linalg.norm(a - r) ** 2 for ga, gr in zip(G, R) for a, r in zip(ga, gr))
    den = sum(np.linalg.norm(r) ** 2 for gr in R for r in gr)
    return float(np.sqrt(num / den))
