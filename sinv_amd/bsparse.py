"""Block-tridiagonal container duck-typed to the part of ``quasi.bsparse`` that
the reference's RGF uses.

rgf.py (src/sinv/algorithms/rgf/rgf.py:13,41,44-62) and rgf_test.py:75-94 use
``A.blocks[i, j]`` (get/set), ``A.blockorder``, ``A.copy()``, ``A * scalar``
(``A.copy() * np.nan``) and ``A.toarray()``.  quasi is not vendored with the
reference and is not installed, so this container provides exactly that
surface over the stacked layout of matrices_utils.py:141-151.  Blocks outside
the tridiagonal pattern read as zeros and cannot be written.
"""

import numpy as np


class _BlockView:
    def __init__(self, owner):
        self._o = owner

    def _where(self, key):
        i, j = key
        n = self._o.blockorder
        if not (0 <= i < n and 0 <= j < n):
            raise IndexError(f"block ({i}, {j}) outside a {n} x {n} block matrix")
        if i == j:
            return self._o.diag, i
        if j == i + 1:
            return self._o.upper, i
        if i == j + 1:
            return self._o.lower, j
        return None, None

    def __getitem__(self, key):
        arr, k = self._where(key)
        if arr is None:
            b = self._o.blocksize
            return np.zeros((b, b), dtype=self._o.dtype)
        return arr[k]

    def __setitem__(self, key, value):
        arr, k = self._where(key)
        if arr is None:
            raise IndexError(f"block {key} is outside the block-tridiagonal pattern")
        arr[k] = value


class BlockTridiag:
    """A block-tridiagonal matrix stored as stacked diag/upper/lower blocks."""

    def __init__(self, diag, upper=None, lower=None):
        diag = np.asarray(diag)
        if diag.ndim != 3 or diag.shape[1] != diag.shape[2]:
            raise ValueError("diag must have shape (N, b, b)")
        N, b = diag.shape[0], diag.shape[1]
        empty = np.zeros((max(N - 1, 0), b, b), dtype=diag.dtype)
        upper = empty.copy() if upper is None else np.asarray(upper)
        lower = empty.copy() if lower is None else np.asarray(lower)
        dt = np.result_type(diag, upper, lower)
        self.diag = np.array(diag, dtype=dt)
        self.upper = np.array(upper, dtype=dt).reshape(max(N - 1, 0), b, b)
        self.lower = np.array(lower, dtype=dt).reshape(max(N - 1, 0), b, b)
        self.blocks = _BlockView(self)

    # -- the quasi.bsparse surface used by rgf.py
    @property
    def blockorder(self):
        return self.diag.shape[0]

    @property
    def blocksize(self):
        return self.diag.shape[1]

    @property
    def dtype(self):
        return self.diag.dtype

    @property
    def shape(self):
        n = self.blockorder * self.blocksize
        return (n, n)

    def copy(self):
        return BlockTridiag(self.diag.copy(), self.upper.copy(), self.lower.copy())

    def __mul__(self, s):
        return BlockTridiag(self.diag * s, self.upper * s, self.lower * s)

    __rmul__ = __mul__

    def toarray(self):
        N, b = self.blockorder, self.blocksize
        A = np.zeros((N * b, N * b), dtype=self.dtype)
        for i in range(N):
            A[i * b:(i + 1) * b, i * b:(i + 1) * b] = self.diag[i]
            if i < N - 1:
                A[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] = self.upper[i]
                A[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] = self.lower[i]
        return A

    # -- constructors
    @classmethod
    def from_dense(cls, A, blocksize):
        from .utils.matu import convertDenseToBlkTridiag
        return cls(*convertDenseToBlkTridiag(np.asarray(A), blocksize))

    def stacked(self):
        return self.diag, self.upper, self.lower
