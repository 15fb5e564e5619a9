"""Namespace mirroring ``sinv.algorithms`` (src/sinv/algorithms/__init__.py:3-9)
for the serial hot path: ``rgf`` (module with ``rgf.rgf``) and ``bcr_s``
(module with ``bcr_s.bcr_serial``).  The MPI variants are not provided."""

from . import rgf as rgf
from . import bcr_serial as bcr_s

__all__ = ["rgf", "bcr_s"]
