"""Drop-in for ``sinv.algorithms.rgf.rgf`` (src/sinv/algorithms/rgf/rgf.py:17-64).

Same signature, keyword meaning and result container: ``G = A.copy() * nan``
with the diagonal and (if ``save_off_diag``) the two off-diagonal block bands
overwritten by the selected inverse (rgf.py:41-62).  The arithmetic runs on
the GPU (libsinv_hip); the container stays on the host, as in the reference.

Real input is computed in complex128 with zero imaginary parts (every product
and reciprocal of such values keeps the imaginary part exactly zero) and
returned real, as the reference returns A's dtype.
"""

import numpy as np

from .. import core
from ..bsparse import BlockTridiag


def _stacked(A):
    if isinstance(A, BlockTridiag):
        return A.stacked()
    N = A.blockorder
    d = np.stack([np.asarray(A.blocks[i, i]) for i in range(N)])
    b = d.shape[1]
    if N > 1:
        u = np.stack([np.asarray(A.blocks[i, i + 1]) for i in range(N - 1)])
        l = np.stack([np.asarray(A.blocks[i + 1, i]) for i in range(N - 1)])
    else:
        u = l = np.zeros((0, b, b), d.dtype)
    return d, u, l


def rgf(A, sym_mat=False, save_off_diag=True, precision="fp64"):
    d, u, l = _stacked(A)
    real = not np.iscomplexobj(d) and not np.iscomplexobj(u) and not np.iscomplexobj(l)
    Gd, Gu, Gl = core.rgf_batched(d, u, l, sym_mat=sym_mat, save_off_diag=save_off_diag,
                                  precision=precision)
    Gd, Gu, Gl = (x[0].cpu().numpy() for x in (Gd, Gu, Gl))
    if real:
        Gd, Gu, Gl = Gd.real, Gu.real, Gl.real
    G = A.copy() * np.nan
    N = A.blockorder
    if isinstance(G, BlockTridiag):
        G.diag[...] = Gd
        if save_off_diag and N > 1:
            G.upper[...] = Gu
            G.lower[...] = Gl
        return G
    for i in range(N):
        G.blocks[i, i] = Gd[i]
    if save_off_diag:
        for i in range(N - 1):
            G.blocks[i + 1, i] = Gl[i]
            G.blocks[i, i + 1] = Gu[i]
    return G


def rgf_blocks(diag, upper, lower, sym_mat=False, save_off_diag=True, precision="fp64"):
    """Stacked-block form (the layout of rgf2sided.py:18-24): numpy in, numpy out."""
    real = not any(np.iscomplexobj(np.asarray(x)) for x in (diag, upper, lower))
    Gd, Gu, Gl = core.rgf_batched(diag, upper, lower, sym_mat=sym_mat,
                                  save_off_diag=save_off_diag, precision=precision)
    out = tuple(x[0].cpu().numpy() for x in (Gd, Gu, Gl))
    return tuple(o.real for o in out) if real else out
