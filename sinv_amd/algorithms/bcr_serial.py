"""Drop-in for ``sinv.algorithms.bcr_s.bcr_serial`` (bcr_serial.py:19-62).

``bcr_serial(A, blocksize)`` takes the dense matrix and returns a dense
matrix of the same kind as the reference: the selected inverse's diagonal,
upper and lower blocks plus the other blocks the cyclic reduction produces
(G between the active rows of each level, at distance 2^l), zeros elsewhere,
cropped to ``(n // blocksize) * blocksize`` (bcr_serial.py:39,60).  Only the
block-tridiagonal part of A is read, as in the reference.  The reduction and
production run on the GPU (libsinv_hip ``sinv_bcr_batched``).

``bcr_blocks`` is the stacked-block form (numpy in, numpy out).  A matrix
already in HBM (a torch device tensor) stays there: its blocks are gathered,
and the dense result assembled, by the device converters.
"""

import numpy as np

from .. import core


def _blocks_of(A, b, N):
    d = np.empty((N, b, b), dtype=A.dtype)
    u = np.empty((max(N - 1, 0), b, b), dtype=A.dtype)
    l = np.empty((max(N - 1, 0), b, b), dtype=A.dtype)
    for i in range(N):
        d[i] = A[i * b:(i + 1) * b, i * b:(i + 1) * b]
        if i < N - 1:
            u[i] = A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b]
            l[i] = A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b]
    return d, u, l


def _bcr_serial_device(A, b, npart=1):
    # A resident in HBM (torch device tensor): gather, solve and assemble the
    # dense result on the device; returns a device tensor
    import torch
    N = A.shape[0] // b
    Ad = A[:N * b, :N * b].to(torch.complex128)
    d, u, l = core.dense_to_blocks(Ad, b)
    Gd, Gu, Gl, Gf = core.bcr_batched(d, u, l, return_fill=True, npart=npart)
    G = core.blocks_to_dense(Gd, Gu, Gl)
    core.scatter_blocks(G, Gf, core.bcr_fill_positions(N))
    G = G[0]
    return G.real.to(A.dtype) if not A.is_complex() else G.to(A.dtype)


def bcr_serial(A, blocksize, _npart=1):
    from ..utils.matu import _on_device
    if _on_device(A):
        return _bcr_serial_device(A, int(blocksize), _npart)
    A = np.asarray(A)
    b = int(blocksize)
    N = A.shape[0] // b
    real = not np.iscomplexobj(A)
    d, u, l = _blocks_of(A, b, N)
    Gd, Gu, Gl, Gf = core.bcr_batched(d, u, l, return_fill=True, npart=_npart)
    Gd, Gu, Gl, Gf = (x[0].cpu().numpy() for x in (Gd, Gu, Gl, Gf))
    G = np.zeros((N * b, N * b), dtype=np.complex128)
    for i in range(N):
        G[i * b:(i + 1) * b, i * b:(i + 1) * b] = Gd[i]
        if i < N - 1:
            G[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = Gu[i]
            G[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = Gl[i]
    for k, (r, c) in enumerate(core.bcr_fill_positions(N)):
        G[r * b:(r + 1) * b, c * b:(c + 1) * b] = Gf[k]
    return G.real.astype(A.dtype) if real else G.astype(A.dtype)


def bcr_blocks(diag, upper, lower):
    real = not any(np.iscomplexobj(np.asarray(x)) for x in (diag, upper, lower))
    out = tuple(x[0].cpu().numpy() for x in core.bcr_batched(diag, upper, lower))
    return tuple(o.real for o in out) if real else out
