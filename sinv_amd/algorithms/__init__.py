"""Namespace mirroring ``sinv.algorithms`` (src/sinv/algorithms/__init__.py:3-9):
``rgf`` (``rgf.rgf``), ``bcr_s`` (``bcr_s.bcr_serial``), and the MPI variants'
partition-local work on one GPU: ``rgf2s`` (``rgf2s.rgf2sided``) and
``psr_s`` (``psr_s.psr_seqsolve``), ``pdiv_lm`` (``pdiv_lm.pdiv_localmap``
per rank, and ``pdiv_lm.pdiv`` for every rank at once) and ``pdiv_u``
(pdiv_utils.py's checks and partitioning), with partitions in place of ranks, and ``bcr_p`` (``bcr_p.bcr_parallel``: every
rank's rows of a level are one batched launch, rank 0's gathered result is
returned)."""

from . import rgf as rgf
from . import bcr_serial as bcr_s
from . import bcr_parallel as bcr_p
from . import rgf2sided as rgf2s
from . import psr_seqsolve as psr_s
from . import pdiv_localmap as pdiv_lm
from . import pdiv_utils as pdiv_u

__all__ = ["rgf", "bcr_s", "bcr_p", "rgf2s", "psr_s", "pdiv_lm", "pdiv_u"]
