"""Drop-in for ``sinv.algorithms.rgf2s.rgf2sided`` (rgf2sided.py:18-84) on one GPU.

The reference runs the upper half (rows [0, N//2)) and the lower half on two
MPI ranks: each sweeps towards the middle (:132-137, :223-232), they swap one
connected block (:141-142), connect the middle (:146-162, :241-255) and sweep
back (:165-176, :260-271).  Here the two halves are the two partitions of the
GPU PSR driver (``sinv_psr_batched`` with npart = 2, first_rows = N//2): its
two reductions are the half-sweeps and its 2-block reduced system is the
connection.  Result semantics follow the reference: arrays start at zero
(:52-54, not NaN); with ``save_off_diag=False`` only the middle coupling
(computed unconditionally, :157-162) is kept among the off-diagonals; with
``sym_mat`` the upper blocks are the plain transposes of the lower ones.
"""

import numpy as np

from .. import core


def rgf2sided(A_bloc_diag, A_bloc_upper, A_bloc_lower, sym_mat=False, save_off_diag=True):
    d = np.asarray(A_bloc_diag)
    N = d.shape[0]
    if N < 2:
        raise ValueError("rgf2sided needs at least 2 blocks (one per side)")
    real = not any(np.iscomplexobj(np.asarray(x)) for x in (A_bloc_diag, A_bloc_upper, A_bloc_lower))
    h = N // 2
    Gd, Gu, Gl = (x[0].cpu().numpy() for x in
                  core.psr_batched(A_bloc_diag, A_bloc_upper, A_bloc_lower, 2, first_rows=h))
    if sym_mat:
        Gu = np.ascontiguousarray(np.swapaxes(Gl, 1, 2))
    if not save_off_diag:
        keep = h - 1
        mask = np.zeros(N - 1, bool)
        mask[keep] = True
        Gu = np.where(mask[:, None, None], Gu, 0)
        Gl = np.where(mask[:, None, None], Gl, 0)
    out = (Gd, Gu, Gl)
    return tuple(o.real for o in out) if real else out
