"""Drop-in for ``sinv.algorithms.psr_s.psr_seqsolve`` (psr_seqsolve.py:20-72)
on one GPU, with ``npart`` partitions in place of the MPI ranks.

``psr_seqsolve(A, blocksize, npart=3)`` takes the dense matrix and returns a
dense matrix holding the block-tridiagonal part of A^-1 (the blocks the
reference's production writes and gathers, :693-942), zeros elsewhere.  Input
checks follow psr_utils.check_input (:44-55).  The partition-local Schur
reduction and production run as batched kernels over partitions
(``sinv_psr_batched``).
"""

import numpy as np

from .. import core


def psr_seqsolve(A, blocksize, npart=3):
    A = np.asarray(A)
    b = int(blocksize)
    # psr_utils.py:44-55
    if A.shape[0] % b != 0:
        raise ValueError("The matrix size must be a multiple of the blocksize.")
    if b > A.shape[0]:
        raise ValueError("The blocksize must be smaller than the matrix size.")
    if b < 1:
        raise ValueError("The blocksize must be greater than 0.")
    N = A.shape[0] // b
    if N < 3 * npart:
        raise ValueError("The number of blocks is to low. There should be at least 3 blockrows per process")
    d = np.stack([A[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(N)])
    u = np.stack([A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] for i in range(N - 1)])
    l = np.stack([A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(N - 1)])
    Gd, Gu, Gl = (x[0].cpu().numpy() for x in core.psr_batched(d, u, l, npart))
    G = np.zeros((N * b, N * b), dtype=np.complex128)
    for i in range(N):
        G[i * b:(i + 1) * b, i * b:(i + 1) * b] = Gd[i]
        if i < N - 1:
            G[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = Gu[i]
            G[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = Gl[i]
    return G.real.astype(A.dtype) if not np.iscomplexobj(A) else G.astype(A.dtype)
