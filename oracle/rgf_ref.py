"""Op-for-op numpy restatement of the reference's serial RGF.

Test infrastructure only (see ``oracle/__init__.py``).

Follows src/sinv/algorithms/rgf/rgf.py:40-62 on the stacked-block layout of
matrices_utils.py:141-151 (``diag (N,b,b)``, ``upper[i] = A_{i,i+1}``,
``lower[i] = A_{i+1,i}``):

  * G starts as NaN everywhere (rgf.py:41, ``A.copy() * np.nan``)
  * g_0 = inv(A_00)                                        (rgf.py:44)
  * g_i = inv(A_ii - (A_{i,i-1} @ g_{i-1}) @ A_{i-1,i})   (rgf.py:47-48)
  * backward, i = N-2 .. 0                                 (rgf.py:50-62):
      X = (G_{i+1,i+1} @ A_{i+1,i}) @ g_i
      G_{i+1,i} = -X
      G_{i,i+1} = G_{i+1,i}.T (sym)  or  ((-g_i) @ A_{i,i+1}) @ G_{i+1,i+1}
      G_ii = g_i + (g_i @ A_{i,i+1}) @ X
    with the two off-diagonal writes skipped when ``save_off_diag`` is False.

``rgf_blocks`` is the restatement; ``rgf_lists`` runs the same recurrences
on lists of blocks of different sizes (b_i per block row).
"""

import numpy as np

inv = np.linalg.inv


def rgf_blocks(diag, upper, lower, sym_mat=False, save_off_diag=True, reassociate=False):
    """rgf.py:40-62.  ``reassociate=True`` evaluates every triple product as
    X @ (Y @ Z) instead of the reference's (X @ Y) @ Z: an equally valid FP64
    evaluation order, whose distance from the reference order measures the
    instance's rounding floor (what two correct FP64 implementations can
    differ by); it is not the reference's arithmetic."""
    N = diag.shape[0]
    Gd = np.full_like(diag, np.nan)
    Gu = np.full_like(upper, np.nan)
    Gl = np.full_like(lower, np.nan)

    def mm3(X, Y, Z):
        return X @ (Y @ Z) if reassociate else X @ Y @ Z

    Gd[0] = inv(diag[0])
    for i in range(1, N):
        Gd[i] = inv(diag[i] - mm3(lower[i - 1], Gd[i - 1], upper[i - 1]))

    for i in range(N - 2, -1, -1):
        g_ii = Gd[i].copy()
        X = mm3(Gd[i + 1], lower[i], g_ii)
        if save_off_diag:
            Gl[i] = -X
            if sym_mat:
                Gu[i] = Gl[i].T
            else:
                Gu[i] = mm3(-g_ii, upper[i], Gd[i + 1])
        Gd[i] = g_ii + mm3(g_ii, upper[i], X)
    return Gd, Gu, Gl


def rgf_lists(diag, upper, lower, sym_mat=False, save_off_diag=True):
    """rgf.py:40-62 on lists: diag[i] b_i x b_i, upper[i] b_i x b_{i+1},
    lower[i] b_{i+1} x b_i."""
    N = len(diag)
    Gd = [None] * N
    Gu = [np.full_like(u, np.nan) for u in upper]
    Gl = [np.full_like(l, np.nan) for l in lower]
    Gd[0] = inv(diag[0])
    for i in range(1, N):
        Gd[i] = inv(diag[i] - lower[i - 1] @ Gd[i - 1] @ upper[i - 1])
    for i in range(N - 2, -1, -1):
        g_ii = Gd[i].copy()
        X = Gd[i + 1] @ lower[i] @ g_ii
        if save_off_diag:
            Gl[i] = -X
            Gu[i] = Gl[i].T if sym_mat else -g_ii @ upper[i] @ Gd[i + 1]
        Gd[i] = g_ii + g_ii @ upper[i] @ X
    return Gd, Gu, Gl
