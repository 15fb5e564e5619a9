"""sinv_amd -- MI355X-native selected inversion of block-tridiagonal matrices.

The serial hot path of Nano-TCAD/SINV (RGF sweep and BCR-serial) as
hand-written gfx950 HIP kernels behind a C ABI (include/sinv.h,
libsinv_hip.so), with the reference's Python call surface:

    from sinv_amd.algorithms import rgf, bcr_s
    G = rgf.rgf(A, sym_mat=False, save_off_diag=True)     # rgf.py:17-64
    G = bcr_s.bcr_serial(A_dense, blocksize)               # bcr_serial.py:19-62
    from sinv_amd.utils import matu                        # matrices_utils.py

plus batched, energy-resolved, mixed-precision and differentiable entry
points in ``sinv_amd.core`` and ``sinv_amd.autograd``.
"""

from .bsparse import BlockTridiag  # noqa: F401

__version__ = "0.1.0"
