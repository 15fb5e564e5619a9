"""Drop-in for ``sinv.algorithms.pdiv_lm`` / ``pdiv_u`` (PDIV, P-Division) on one
GPU, with ``npart`` partitions in place of the MPI ranks.

The reference runs ``pdiv_localmap(K_local, Bu_i, Bl_i, blocksize)`` on each
of P ranks (pdiv_localmap.py:27-90) after ``divide_matrix`` and
``partition_subdomain`` (pdiv_utils.py:84-170); rank r returns three lists for
its partition: diagonal, upper and lower blocks, whose last upper/lower
entries are the bridges to the next partition (zero on the last rank).
``pdiv(A, blocksize, npart)`` returns that per-rank output for every rank at
once (a list of (diag, upper, lower) tuples); the work runs on the GPU
(``sinv_pdiv_batched``: the partitions' dense inverses and their selected-block
production batched over partitions).  Input checks follow pdiv_utils.py:16-80.
"""

import math

import numpy as np

from .. import core


def check_multiprocessing(comm_size):
    """pdiv_utils.py:16-35."""
    if not math.log2(comm_size).is_integer():
        raise ValueError("The number of processes must be a power of 2.")


def check_input(A, blocksize, comm_size):
    """pdiv_utils.py:39-80."""
    if A.shape[0] % blocksize != 0:
        raise ValueError("The matrix size must be a multiple of the blocksize.")
    if blocksize > A.shape[0]:
        raise ValueError("The blocksize must be smaller than the matrix size.")
    if blocksize < 1:
        raise ValueError("The blocksize must be greater than 0.")
    if A.shape[0] // blocksize < comm_size:
        raise ValueError("The number of blocks must be greater than the number of processes.")


def divide_matrix(A, n_partitions, blocksize):
    """pdiv_utils.py:84-122: (start block row, block count) of each partition."""
    nblocks = A.shape[0] // blocksize
    m = nblocks // n_partitions
    sizes = [nblocks - m * (n_partitions - 1)] + [m] * (n_partitions - 1)
    starts = [0]
    for i in range(1, n_partitions):
        starts.append(starts[-1] + sizes[i - 1])
    return starts, sizes


def partition_subdomain(A, l_start_blockrow, l_partitions_blocksizes, blocksize):
    """pdiv_utils.py:126-170: dense partitions and the upper/lower bridges."""
    K, Bu, Bl = [], [], []
    b = blocksize
    for i, (s, n) in enumerate(zip(l_start_blockrow, l_partitions_blocksizes)):
        s0, s1 = s * b, (s + n) * b
        K.append(A[s0:s1, s0:s1])
        if i < len(l_start_blockrow) - 1:
            Bu.append(A[s1 - b:s1, s1:s1 + b])
            Bl.append(A[s1:s1 + b, s1 - b:s1])
    return K, Bu, Bl


def pdiv(A, blocksize, npart):
    A = np.asarray(A)
    b = int(blocksize)
    check_multiprocessing(npart)
    check_input(A, b, npart)
    N = A.shape[0] // b
    real = not np.iscomplexobj(A)
    d = np.stack([A[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(N)])
    if N > 1:
        u = np.stack([A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] for i in range(N - 1)])
        l = np.stack([A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(N - 1)])
    else:
        u = l = np.zeros((0, b, b), dtype=A.dtype)
    Gd, Gu, Gl = (x[0].cpu().numpy() for x in core.pdiv_batched(d, u, l, npart))
    if real:
        Gd, Gu, Gl = Gd.real, Gu.real, Gl.real
    starts, sizes = divide_matrix(A, npart, b)
    zero = np.zeros((b, b), dtype=Gd.dtype)
    out = []
    for s, n in zip(starts, sizes):
        D = [Gd[s + i] for i in range(n)]
        U = [Gu[s + i] if s + i < N - 1 else zero.copy() for i in range(n)]
        L = [Gl[s + i] if s + i < N - 1 else zero.copy() for i in range(n)]
        out.append((D, U, L))
    return out


def pdiv_blocks(diag, upper, lower, npart):
    """Stacked-block form: (diag, upper, lower) of A^-1's tridiagonal part."""
    real = not any(np.iscomplexobj(np.asarray(x)) for x in (diag, upper, lower))
    out = tuple(x[0].cpu().numpy() for x in core.pdiv_batched(diag, upper, lower, npart))
    return tuple(o.real for o in out) if real else out
