"""Test-matrix generators and dense <-> stacked-block converters.

Same names, arguments and results as src/sinv/utils/matrices_utils.py
(generateRandomNumpyMat :12-51, generateBandedDiagonalMatrix :55-91,
transformToSymmetric :95-111, convertDenseToBlkTridiag :115-151,
convertBlkTridiagToDense :155-189).  The random streams are numpy's legacy
``np.random.seed`` / ``rand`` ones, so a given seed reproduces the
reference's matrices; the banded mask is vectorised (the reference loops over
all n^2 entries in Python).
"""

import numpy as np


def generateRandomNumpyMat(matrice_size, is_complex=False, is_symmetric=False, seed=None):
    if seed is not None:
        np.random.seed(seed)
    A = np.random.rand(matrice_size, matrice_size)
    if is_complex:
        A = A + 1j * np.random.rand(matrice_size, matrice_size)
    if is_symmetric:
        A = A + A.T
    return A


def generateBandedDiagonalMatrix(matrice_size, matrice_bandwidth, is_complex=False,
                                 is_symmetric=False, seed=None):
    A = generateRandomNumpyMat(matrice_size, is_complex, is_symmetric, seed)
    off = np.abs(np.subtract.outer(np.arange(matrice_size), np.arange(matrice_size)))
    A[off > matrice_bandwidth] = 0
    return A


def transformToSymmetric(A):
    return A + A.T


def _on_device(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def convertDenseToBlkTridiag(A, blocksize):
    # a matrix already in HBM (torch device tensor) is gathered on the device;
    # returns device tensors (N,b,b), (N-1,b,b), (N-1,b,b)
    if _on_device(A):
        from .. import core
        d, u, l = core.dense_to_blocks(A, blocksize)
        return d[0], u[0], l[0]
    # N = ceil(n / b) as in the reference; a ragged last block raises numpy's
    # broadcast ValueError there and here alike
    N = int(np.ceil(A.shape[0] / blocksize))
    b = blocksize
    d = np.zeros((N, b, b), dtype=A.dtype)
    u = np.zeros((max(N - 1, 0), b, b), dtype=A.dtype)
    l = np.zeros((max(N - 1, 0), b, b), dtype=A.dtype)
    for i in range(N):
        d[i] = A[i * b:(i + 1) * b, i * b:(i + 1) * b]
        if i < N - 1:
            u[i] = A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b]
            l[i] = A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b]
    return d, u, l


def convertBlkTridiagToDense(A_diagblk, A_upperblk, A_lowerblk):
    if _on_device(A_diagblk):
        from .. import core
        return core.blocks_to_dense(A_diagblk, A_upperblk, A_lowerblk)[0]
    N, b = A_diagblk.shape[0], A_diagblk.shape[1]
    A = np.zeros((N * b, N * b), dtype=A_diagblk.dtype)
    for i in range(N):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = A_diagblk[i]
        if i < N - 1:
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = A_upperblk[i]
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = A_lowerblk[i]
    return A
