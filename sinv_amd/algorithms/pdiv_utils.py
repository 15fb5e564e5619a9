"""Drop-in for ``sinv.algorithms.pdiv_u`` (src/sinv/algorithms/pdiv/pdiv_utils.py):
the input checks and the matrix partitioning the PDIV workflow runs before
``pdiv_lm.pdiv_localmap`` (pdiv_test.py:81-87), with the reference's names,
arguments and errors.

``LocalComm`` stands in for the MPI communicator: P ranks run as P threads of
one process (each calling ``pdiv_localmap`` with its own rank), so the
per-rank workflow runs on one GPU without MPI.  Any communicator with
mpi4py's ``Get_rank`` / ``Get_size`` / ``allgather`` works in its place.
"""

import math
import threading

import numpy as np


def check_multiprocessing(comm_size):
    """pdiv_utils.py:16-35: the number of processes must be a power of 2."""
    if comm_size < 1 or not math.log2(comm_size).is_integer():
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
    """pdiv_utils.py:84-122: (start block row, block count) of each partition;
    the first partition takes the remainder."""
    nblocks = A.shape[0] // blocksize
    m = nblocks // n_partitions
    sizes = [nblocks - m * (n_partitions - 1)] + [m] * (n_partitions - 1)
    starts = [0]
    for i in range(1, n_partitions):
        starts.append(starts[-1] + sizes[i - 1])
    return starts, sizes


def partition_subdomain(A, l_start_blockrow, l_partitions_blocksizes, blocksize):
    """pdiv_utils.py:126-170: the dense partitions K_i and the upper / lower
    bridge blocks between consecutive partitions."""
    K, Bu, Bl = [], [], []
    b = blocksize
    for i, (s, n) in enumerate(zip(l_start_blockrow, l_partitions_blocksizes)):
        s0, s1 = s * b, (s + n) * b
        K.append(A[s0:s1, s0:s1])
        if i < len(l_start_blockrow) - 1:
            Bu.append(A[s1 - b:s1, s1:s1 + b])
            Bl.append(A[s1:s1 + b, s1 - b:s1])
    return K, Bu, Bl


class _Rank:
    def __init__(self, world, rank):
        self._world, self._rank = world, rank

    def Get_rank(self):
        return self._rank

    def Get_size(self):
        return self._world.size

    def allgather(self, obj):
        return self._world._allgather(self._rank, obj)


class _Single:
    """The one-rank communicator (what a serial caller has)."""

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    def allgather(self, obj):
        return [obj]


SELF = _Single()


class LocalComm:
    """P ranks of one process: ``LocalComm(P).run(fn)`` calls ``fn(comm)`` on P
    threads, comm being rank r's communicator (Get_rank, Get_size, allgather),
    and returns the P results in rank order.  allgather is collective: every
    rank must call it, as with MPI."""

    def __init__(self, size, timeout=600.0):
        check_multiprocessing(size)
        self.size = size
        self._timeout = timeout
        self._barrier = threading.Barrier(size, timeout=timeout)
        self._slots = [None] * size

    def _allgather(self, rank, obj):
        self._slots[rank] = obj
        self._barrier.wait()
        out = list(self._slots)
        self._barrier.wait()  # nobody overwrites a slot before every rank has read it
        return out

    def run(self, fn):
        results = [None] * self.size
        errors = []

        def body(r):
            try:
                results[r] = fn(_Rank(self, r))
            except BaseException as e:  # noqa: BLE001 -- re-raised in the caller
                errors.append(e)
                self._barrier.abort()

        threads = [threading.Thread(target=body, args=(r,)) for r in range(self.size)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(self._timeout)
        if errors:
            raise errors[0]
        return results
