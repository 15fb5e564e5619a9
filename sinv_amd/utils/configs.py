"""Synthetic block-tridiagonal inputs of the benchmark workloads.

* ``dominant_blocks``   -- complex128 blocks with re/im ~ U[-1,1) from a seed,
  A_{i+1,i} = A_{i,i+1}^T, symmetric diagonal blocks, diagonal shifted by
  +3*b*I (diagonally dominant; A = A^T so ``sym_mat=True`` is valid).
  Workloads: 16 blocks of 64 (n = 1024) and 128 blocks of 512 (n = 65536).
* ``hermitian_H``       -- Hermitian H with re/im ~ N(0, 1/sqrt(b)):
  H_ii = (X + X^H)/2, couplings scaled by 0.5 and H_{i+1,i} = H_{i,i+1}^H.
  The solved matrices are A(E) = (E + i*eta) I - H.
  Workloads: 256 energies in [-2, 2], eta = 1e-3, 64 blocks of b in
  {128, 256, 384}; the density-of-states gradient at E = 0.3 + 0.01i on 32
  blocks of 128; 512 energies on 96 blocks with b in {256, 416}.
* ``energies``          -- energies uniform in [-2, 2] from a seed.

All generators use ``numpy.random.default_rng(seed)`` and return numpy arrays
(host), stacked as diag (N,b,b), upper (N-1,b,b), lower (N-1,b,b).
"""

import numpy as np


def dominant_blocks(N, b, seed=0, dtype=np.complex128):
    rng = np.random.default_rng(seed)

    def u(*shape):
        x = rng.uniform(-1.0, 1.0, size=shape)
        if np.issubdtype(dtype, np.complexfloating):
            x = x + 1j * rng.uniform(-1.0, 1.0, size=shape)
        return x.astype(dtype)

    X = u(N, b, b)
    diag = 0.5 * (X + np.swapaxes(X, 1, 2)) + 3.0 * b * np.eye(b, dtype=dtype)
    upper = u(max(N - 1, 0), b, b)
    lower = np.ascontiguousarray(np.swapaxes(upper, 1, 2))
    return diag, upper, lower


def hermitian_H(N, b, seed=0, coupling=0.5):
    rng = np.random.default_rng(seed)
    s = 1.0 / np.sqrt(b)

    def g(*shape):
        return rng.normal(0.0, s, size=shape) + 1j * rng.normal(0.0, s, size=shape)

    X = g(N, b, b)
    diag = 0.5 * (X + np.conj(np.swapaxes(X, 1, 2)))
    upper = coupling * g(max(N - 1, 0), b, b)
    lower = np.ascontiguousarray(np.conj(np.swapaxes(upper, 1, 2)))
    return diag, upper, lower


def energies(n, seed=0, lo=-2.0, hi=2.0):
    return np.random.default_rng(seed).uniform(lo, hi, size=n)


def shifted(H, E):
    """Dense-block form of A(E) = E I - H for one complex energy E (host)."""
    d, u, l = H
    b = d.shape[1]
    return E * np.eye(b) - d, -u, -l


def rgf_flops(N, b, sym=False):
    """Algorithmic real flops of one RGF solve (inverse and complex GEMM = 8 b^3)."""
    return 8.0 * b ** 3 * ((7 * N - 6) if sym else (9 * N - 8))


def rgf_flops_ragged(bs):
    """Per-block sum for ragged block sizes (non-symmetric, reference op order)."""
    bs = [int(x) for x in bs]
    N = len(bs)
    f = sum(8.0 * b ** 3 for b in bs)
    f += sum(8.0 * (bs[i] * bs[i - 1] ** 2 + bs[i] ** 2 * bs[i - 1]) for i in range(1, N))
    f += sum(8.0 * (2 * bs[i] * bs[i + 1] ** 2 + 4 * bs[i] ** 2 * bs[i + 1]) for i in range(N - 1))
    return f


def hermitian_H_ragged(bs, seed=0, coupling=0.5):
    """Hermitian block-tridiagonal H with per-block sizes bs (lists of blocks):
    re/im ~ N(0, 1/sqrt(b_i)) on the diagonal blocks, couplings ~ N(0,
    1/sqrt(sqrt(b_i b_{i+1}))) scaled by `coupling`, H_{i+1,i} = H_{i,i+1}^H."""
    rng = np.random.default_rng(seed)
    diag, upper, lower = [], [], []
    for i, b in enumerate(bs):
        s = 1.0 / np.sqrt(b)
        X = rng.normal(0, s, (b, b)) + 1j * rng.normal(0, s, (b, b))
        diag.append(0.5 * (X + X.conj().T))
    for i in range(len(bs) - 1):
        s = 1.0 / np.sqrt(np.sqrt(bs[i] * bs[i + 1]))
        U = coupling * (rng.normal(0, s, (bs[i], bs[i + 1])) + 1j * rng.normal(0, s, (bs[i], bs[i + 1])))
        upper.append(U)
        lower.append(np.ascontiguousarray(U.conj().T))
    return diag, upper, lower


def shifted_lists(H, E):
    d, u, l = H
    return [E * np.eye(x.shape[0]) - x for x in d], [-x for x in u], [-x for x in l]
