"""Drop-in for ``sinv.algorithms.pdiv_lm`` (PDIV, P-Division,
src/sinv/algorithms/pdiv/pdiv_localmap.py) on one GPU.

``pdiv_localmap(K_local, l_upperbridges, l_lowerbridges, blocksize, comm)``
is the reference's per-rank entry (pdiv_localmap.py:27-90): rank r passes its
dense partition K_local and every bridge block, and gets its partition's
diagonal, upper and lower blocks of A^-1 back, the last upper / lower entries
being the bridges to the next partition (zeros on the last rank).  The ranks'
partitions are gathered through ``comm`` (mpi4py-like; the reference reads
MPI.COMM_WORLD, here ``pdiv_u.LocalComm`` runs the ranks as threads, and the
default is a single rank), and the whole selected inversion runs on the GPU
(``sinv_pdiv_batched``: the partitions' dense inverses and their
selected-block production batched over partitions, coupled by a Woodbury
correction).  ``pdiv(A, blocksize, npart)`` returns every rank's output at
once.  The matrix is block tridiagonal, as the reference's algorithm assumes
(the bridges are the only coupling between partitions).
"""

import numpy as np

from .. import core
from . import pdiv_utils as pdiv_u

check_multiprocessing = pdiv_u.check_multiprocessing
check_input = pdiv_u.check_input
divide_matrix = pdiv_u.divide_matrix
partition_subdomain = pdiv_u.partition_subdomain


def _stacks(parts, l_upperbridges, l_lowerbridges, b):
    """Stacked diag / upper / lower blocks of the whole matrix from the
    partitions (dense, block tridiagonal) and the bridges between them."""
    d, u, l = [], [], []
    P = len(parts)
    for p, K in enumerate(parts):
        K = np.asarray(K)
        if K.ndim != 2 or K.shape[0] != K.shape[1] or K.shape[0] % b or K.shape[0] == 0:
            raise ValueError("each partition must be a square matrix of whole blocks")
        n = K.shape[0] // b
        for i in range(n):
            d.append(K[i * b:(i + 1) * b, i * b:(i + 1) * b])
            if i < n - 1:
                u.append(K[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b])
                l.append(K[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b])
        if p < P - 1:
            u.append(np.asarray(l_upperbridges[p]))
            l.append(np.asarray(l_lowerbridges[p]))
    dt = np.result_type(*[x.dtype for x in d + u + l])
    zero = np.zeros((0, b, b), dtype=dt)
    return (np.stack(d).astype(dt), np.stack(u).astype(dt) if u else zero,
            np.stack(l).astype(dt) if l else zero)


def _rank_output(Gd, Gu, Gl, start, n, N):
    zero = np.zeros(Gd.shape[-2:], dtype=Gd.dtype)
    D = [Gd[start + i] for i in range(n)]
    U = [Gu[start + i] if start + i < N - 1 else zero.copy() for i in range(n)]
    L = [Gl[start + i] if start + i < N - 1 else zero.copy() for i in range(n)]
    return D, U, L


def pdiv_localmap(K_local, l_upperbridges, l_lowerbridges, blocksize, comm=None):
    """pdiv_localmap.py:27-90: this rank's diagonal, upper and lower blocks of
    A^-1 (lists of b x b arrays, as many as the partition has block rows)."""
    comm = pdiv_u.SELF if comm is None else comm
    P, rank = comm.Get_size(), comm.Get_rank()
    check_multiprocessing(P)
    b = int(blocksize)
    if len(l_upperbridges) != P - 1 or len(l_lowerbridges) != P - 1:
        raise ValueError("need one upper and one lower bridge between consecutive partitions")
    parts = comm.allgather(np.asarray(K_local))
    d, u, l = _stacks(parts, l_upperbridges, l_lowerbridges, b)
    N = d.shape[0]
    real = not np.iscomplexobj(d)
    Gd, Gu, Gl = (x[0].cpu().numpy() for x in core.pdiv_batched(d, u, l, P))
    if real:
        Gd, Gu, Gl = Gd.real, Gu.real, Gl.real
    sizes = [np.asarray(K).shape[0] // b for K in parts]
    start = sum(sizes[:rank])
    return _rank_output(Gd, Gu, Gl, start, sizes[rank], N)


def pdiv(A, blocksize, npart):
    """Every rank's ``pdiv_localmap`` output for the dense block-tridiagonal A
    split into ``npart`` partitions by ``divide_matrix`` (pdiv_test.py:81-87
    for all ranks): a list of (diag, upper, lower) lists, one per rank."""
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
    return [_rank_output(Gd, Gu, Gl, s, n, N) for s, n in zip(starts, sizes)]


def pdiv_blocks(diag, upper, lower, npart):
    """Stacked-block form: (diag, upper, lower) of A^-1's tridiagonal part."""
    real = not any(np.iscomplexobj(np.asarray(x)) for x in (diag, upper, lower))
    out = tuple(x[0].cpu().numpy() for x in core.pdiv_batched(diag, upper, lower, npart))
    return tuple(o.real for o in out) if real else out
