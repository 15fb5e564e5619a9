"""Drop-in for ``sinv.algorithms.bcr_p.bcr_parallel`` (bcr_parallel.py:21-78) on
one GPU.

The reference pads to 2^k - 1 blocks, splits the block rows among the ranks
(bcr_utils.divide_matrix, :133-172), and at every level each rank reduces
(bcr_parallel.py:227-256) and later produces (:1004-1015, produce :684-758)
the rows of that level that fall in its range [top, bottom), exchanging
halo rows with its neighbours between levels and gathering each rank's rows
of G on rank 0 (agregate_result_on_root :1019-1060).  Its arithmetic is
block cyclic reduction's, the same as bcr_serial's.

On one device the partitions play the part of the ranks: the device BCR
runs with ``n_partitions`` (``sinv_bcr_parallel_batched``), issuing every
level operation as one launch per partition over the rows the partition owns
-- the rank split of bcr_utils.divide_matrix on the padded rows, reduction of
its kept rows and production of its rows as the ranks do -- and each
operation family for all partitions before the next, which is where the
ranks exchange halo rows (here the rows stay resident in HBM).
``reduction_schedule`` and ``production_schedule`` restate the reference's
per-level, per-rank row ranges; they cover every row of a level exactly once
(checked on every call), and the device's per-partition launches follow the
same ownership (tests compare the two).  The call returns rank 0's gathered
result: the dense matrix with the selected inverse's tridiagonal blocks and
the fill-in blocks of the cyclic reduction, cropped to (n // blocksize) *
blocksize; the arithmetic equals bcr_serial's bit for bit.  Which extra
off-band blocks a rank holds when the reference gathers is a detail of its
message pattern, not checked by its test (bcr_parallel_test.py:70-79
compares the tridiagonal blocks); here they are the fill-in blocks of
bcr_serial.
"""

import math

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


def _partition_bounds(nblocks_padded, n_partitions):
    starts, sizes = divide_matrix(nblocks_padded, n_partitions)
    return [get_process_rowblock_index(s, z) for s, z in zip(starts, sizes)]


def reduction_schedule(nblocks_padded, n_partitions):
    """Per level l = 0 .. height-1: for each partition, the inclusive index
    range (first, last) into that level's reduced rows
    i_elim = [2^(l+1) - 1, ..., step 2^(l+1)] that the partition reduces, or
    None when none of them lies in its rows [top, bottom)
    (bcr_parallel.py:227-256)."""
    height = int(math.log2(nblocks_padded))
    out = []
    for level in range(height):
        i_elim = list(range(2 ** (level + 1) - 1, nblocks_padded, 2 ** (level + 1)))
        per = []
        for top, bottom in _partition_bounds(nblocks_padded, n_partitions):
            idx = [i for i, r in enumerate(i_elim) if top <= r < bottom]
            per.append((idx[0], idx[-1]) if idx else None)
        out.append((i_elim, per))
    return out


def production_schedule(nblocks_padded, n_partitions):
    """Per level l = height-1 .. 0: for each partition, the indices into that
    level's produced rows i_prod (compute_i_prod, bcr_utils.py:99-129) whose
    block row k_to lies in the partition's rows (produce, bcr_parallel.py:
    733-758)."""
    height = int(math.log2(nblocks_padded))
    out = []
    for level in range(height - 1, -1, -1):
        s = 2 ** level
        i_from = list(range(2 * s - 1, nblocks_padded, 2 * s))
        i_prod = []
        for n_, f in enumerate(i_from):
            if n_ == 0:
                i_prod += [f - s, f + s]
            else:
                if i_prod[n_] != f - s:
                    i_prod.append(f - s)
                i_prod.append(f + s)
        per = [[q for q, k in enumerate(i_prod) if top <= k < bottom]
               for top, bottom in _partition_bounds(nblocks_padded, n_partitions)]
        out.append((i_prod, per))
    return out


def check_schedule_covers(nblocks_padded, n_partitions):
    """Every row of every level is reduced (produced) by exactly one partition:
    the condition under which one level-batched launch over all partitions is
    the ranks' combined work of that level.  Raises AssertionError otherwise."""
    for i_elim, per in reduction_schedule(nblocks_padded, n_partitions):
        got = [q for r in per if r is not None for q in range(r[0], r[1] + 1)]
        assert sorted(got) == list(range(len(i_elim))) and len(set(got)) == len(got)
    for i_prod, per in production_schedule(nblocks_padded, n_partitions):
        got = [q for p in per for q in p]
        assert sorted(got) == list(range(len(i_prod))) and len(set(got)) == len(got)


def bcr_parallel(A, blocksize, n_partitions=1):
    """bcr_parallel.py:21-78 with the MPI communicator's size given as
    ``n_partitions`` (ranks become partitions of one level-batched launch
    sequence)."""
    if n_partitions < 1:
        raise ValueError("n_partitions must be at least 1")
    nblocks = A.shape[0] // blocksize
    nbp = nblocks + (2 ** math.ceil(math.log2(nblocks + 1)) - 1 - nblocks)
    if n_partitions > nbp:
        raise ValueError("more partitions than padded block rows")
    check_schedule_covers(nbp, n_partitions)
    return _serial.bcr_serial(A, blocksize, _npart=n_partitions)
