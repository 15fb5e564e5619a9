"""Drop-in for ``sinv.algorithms.bcr_p.bcr_parallel`` (bcr_parallel.py:21-78) on
one GPU.

The reference pads to 2^k - 1 blocks, splits the block rows of every level
among the ranks (bcr_utils.divide_matrix, :133-172), reduces and produces with
point-to-point exchanges between neighbouring ranks, and gathers each rank's
rows of G on rank 0 (agregate_result_on_root :1019-1060).  Its arithmetic is
block cyclic reduction's, the same as bcr_serial's.  On one device every
rank's share of a level is processed by one batched launch (the level-batched
``sinv_bcr_batched``), so the distribution over ranks has no counterpart and
the call returns rank 0's gathered result: the dense matrix with the
selected inverse's tridiagonal blocks and the fill-in blocks of the cyclic
reduction, cropped to (n // blocksize) * blocksize.  Which extra off-band
blocks a rank holds when the reference gathers is a detail of its message
pattern, not checked by its test (bcr_parallel_test.py:70-79 compares the
tridiagonal blocks); here they are the fill-in blocks of bcr_serial.
"""

from . import bcr_serial as _serial


def divide_matrix(n_blocks, n_partitions):
    """bcr_utils.py:133-172: equal strides, the last partition takes the rest."""
    stride = n_blocks // n_partitions
    starts = [p * stride for p in range(n_partitions)]
    sizes = [stride] * (n_partitions - 1) + [n_blocks - (n_partitions - 1) * stride]
    return starts, sizes


def get_process_rowblock_index(start_blockrow, partitions_blocksizes):
    """bcr_utils.py:176-184."""
    return start_blockrow, start_blockrow + partitions_blocksizes


def bcr_parallel(A, blocksize):
    return _serial.bcr_serial(A, blocksize)
