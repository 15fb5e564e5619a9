"""Restatement of the reference's test-matrix generators and block converters.

Test infrastructure only (see ``oracle/__init__.py``).

Follows src/sinv/utils/matrices_utils.py:
  * generateRandomNumpyMat        :12-51  (legacy ``np.random.seed`` + ``rand``;
                                           complex = two consecutive draws; sym = A + A.T)
  * generateBandedDiagonalMatrix  :55-91  (mask |i-j| > bw -> 0; the reference's
                                           O(n^2) Python loop is vectorised here,
                                           same result)
  * transformToSymmetric          :95-111
  * convertDenseToBlkTridiag      :115-151 (N = ceil(n/b), stacked layout)
  * convertBlkTridiagToDense      :155-189
"""

import numpy as np


def generateRandomNumpyMat(matrice_size, is_complex=False, is_symmetric=False, seed=None):
    # matrices_utils.py:37-49: reseed, one or two consecutive rand() draws, A + A.T
    if seed is not None:
        np.random.seed(seed)
    if is_complex:
        re = np.random.rand(matrice_size, matrice_size)
        im = np.random.rand(matrice_size, matrice_size)
        A = re + 1j * im
    else:
        A = np.random.rand(matrice_size, matrice_size)
    if is_symmetric:
        A = A + A.T
    return A


def generateBandedDiagonalMatrix(matrice_size, matrice_bandwidth, is_complex=False,
                                 is_symmetric=False, seed=None):
    # matrices_utils.py:84-91: zero every entry with i - j > bw or j - i > bw
    A = generateRandomNumpyMat(matrice_size, is_complex, is_symmetric, seed)
    i = np.arange(matrice_size)[:, None]
    j = np.arange(matrice_size)[None, :]
    A[(i - j > matrice_bandwidth) | (j - i > matrice_bandwidth)] = 0
    return A


def transformToSymmetric(A):
    # matrices_utils.py:111
    return A + A.T


def convertDenseToBlkTridiag(A, blocksize):
    # matrices_utils.py:139-151
    nblocks = int(np.ceil(A.shape[0] / blocksize))
    d = np.zeros((nblocks, blocksize, blocksize), dtype=A.dtype)
    u = np.zeros((nblocks - 1, blocksize, blocksize), dtype=A.dtype)
    l = np.zeros((nblocks - 1, blocksize, blocksize), dtype=A.dtype)
    for i in range(nblocks):
        d[i] = A[i * blocksize:(i + 1) * blocksize, i * blocksize:(i + 1) * blocksize]
        if i < nblocks - 1:
            u[i] = A[i * blocksize:(i + 1) * blocksize, (i + 1) * blocksize:(i + 2) * blocksize]
            l[i] = A[(i + 1) * blocksize:(i + 2) * blocksize, i * blocksize:(i + 1) * blocksize]
    return d, u, l


def convertBlkTridiagToDense(d, u, l):
    # matrices_utils.py:177-189
    b = d.shape[1]
    N = d.shape[0]
    A = np.zeros((N * b, N * b), dtype=d.dtype)
    for i in range(N):
        A[i * b:(i + 1) * b, i * b:(i + 1) * b] = d[i]
        if i < N - 1:
            A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = u[i]
            A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = l[i]
    return A
