"""Restatement of the reference's 2-sided RGF on one process.

Test infrastructure only (see ``oracle/__init__.py``).

Follows src/sinv/algorithms/rgf/rgf2sided.py: the upper half (blocks
[0, N//2), rgf2sided_upperprocess :88-179) and the lower half
(rgf2sided_lowerprocess :183-274) run one after the other, the MPI swap
(:141-142, :236-237) becomes an array read, and the result is assembled as
rgf2sided does on rank 0 (:57-81): zeros where nothing is written, the middle
coupling always from the upper side.  Needs N >= 4 (both halves at least two
blocks; the reference's own tests use nblocks >= 4).
"""

import numpy as np

inv = np.linalg.inv


def _upper(D, U, L, sym_mat, save_off_diag):
    n = D.shape[0]
    G = np.zeros((n + 1,) + D.shape[1:], D.dtype)
    Gu = np.zeros_like(U)
    Gl = np.zeros_like(L)
    G[0] = inv(D[0])
    for i in range(1, n):
        G[i] = inv(D[i] - L[i - 1] @ G[i - 1] @ U[i - 1])
    return G, Gu, Gl


def _upper_finish(D, U, L, G, Gu, Gl, sym_mat, save_off_diag):
    n = D.shape[0]
    G[n - 1] = inv(D[n - 1] - L[n - 2] @ G[n - 2] @ U[n - 2] - U[n - 1] @ G[n] @ L[n - 1])
    Gl[n - 1] = -G[n] @ L[n - 1] @ G[n - 1]
    Gu[n - 1] = Gl[n - 1].T if sym_mat else -G[n - 1] @ U[n - 1] @ G[n]
    for i in range(n - 2, -1, -1):
        g = G[i].copy()
        X = G[i + 1] @ L[i] @ g
        if save_off_diag:
            Gl[i] = -X
            Gu[i] = Gl[i].T if sym_mat else -g @ U[i] @ G[i + 1]
        G[i] = g + g @ U[i] @ X
    return G[:n], Gu, Gl


def _lower(D, U, L):
    n = D.shape[0]
    G = np.zeros((n + 1,) + D.shape[1:], D.dtype)
    G[-1] = inv(D[-1])
    for i in range(n - 1, 0, -1):
        G[i] = inv(D[i - 1] - U[i] @ G[i + 1] @ L[i])
    return G


def _lower_finish(D, U, L, G, sym_mat, save_off_diag):
    n = D.shape[0]
    Gu = np.zeros_like(U)
    Gl = np.zeros_like(L)
    G[1] = inv(D[0] - L[0] @ G[0] @ U[0] - U[1] @ G[2] @ L[1])
    Gl[0] = G[1] @ L[0] @ G[0]
    Gu[0] = Gl[0].T if sym_mat else G[0] @ U[0] @ G[1]
    for i in range(1, n):
        g = G[i + 1].copy()
        X = g @ L[i] @ G[i]
        if save_off_diag:
            Gl[i] = -X
            Gu[i] = Gl[i].T if sym_mat else -G[i] @ U[i] @ g
        G[i + 1] = g + X @ U[i] @ g
    return G[1:], Gu, Gl


def rgf2sided_blocks(diag, upper, lower, sym_mat=False, save_off_diag=True):
    N = diag.shape[0]
    if N < 4:
        raise ValueError("the 2-sided restatement needs at least 4 blocks")
    h = N // 2                                           # rgf2sided.py:49
    Gup, Guu, Glu = _upper(diag[:h], upper[:h], lower[:h], sym_mat, save_off_diag)
    Glo = _lower(diag[h:], upper[h - 1:], lower[h - 1:])
    gl_left, gr_right = Gup[h - 1].copy(), Glo[1].copy()  # the swap
    Gup[h] = gr_right
    Glo[0] = gl_left
    Gd_u, Gu_u, Gl_u = _upper_finish(diag[:h], upper[:h], lower[:h], Gup, Guu, Glu,
                                     sym_mat, save_off_diag)
    Gd_l, Gu_l, Gl_l = _lower_finish(diag[h:], upper[h - 1:], lower[h - 1:], Glo,
                                     sym_mat, save_off_diag)
    Gd = np.concatenate([Gd_u, Gd_l])
    Gu = np.concatenate([Gu_u, Gu_l[1:]])
    Gl = np.concatenate([Gl_u, Gl_l[1:]])
    return Gd, Gu, Gl
