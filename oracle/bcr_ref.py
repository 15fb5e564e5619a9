"""Restatement of the reference's serial block cyclic reduction (BCR).

Test infrastructure only (see ``oracle/__init__.py``).

Two restatements of src/sinv/algorithms/bcr/bcr_serial.py:

``bcr_serial_dense(A, blocksize)`` -- op for op on dense storage, as the
reference runs it:
  * nblocks = n // b (silently truncating, bcr_serial.py:39)
  * pad to 2^p - 1 blocks with a real identity (bcr_utils.py:35-40, :63-71)
  * dense L, U, G of the padded size (bcr_serial.py:47-49)
  * reduce_bcr / reduce   (bcr_serial.py:173-216, :66-169)
  * invert_block(root)    (bcr_serial.py:220-249)
  * produce_bcr with corner_produce / center_produce
                          (bcr_serial.py:387-466, :253-304, :308-383)
  * crop to nblocks*b     (bcr_serial.py:60)
  Out-of-range fill-in indices are the reference's empty-slice no-ops
  (bcr_serial.py:137-143, :163-169).

``bcr_blocks(diag, upper, lower)`` -- the same schedule on block-compressed
storage (a dict of b x b blocks), with the identity padding skipped.  Padding
rows have zero coupling, so every contribution they make to a real row is an
exact zero; skipping them changes no real block.  Each eliminated row's
``inv(A_tt)`` is computed once per level and reused in the production (the
reference computes it up to three times on identical input).  This is the
schedule the GPU implementation follows, level by level.
"""

import math

import numpy as np

inv = np.linalg.inv


# --------------------------------------------------------------------------
# bcr_utils.py:17-71
def distance_to_power_of_two(matrice_size):
    p = math.ceil(math.log2(matrice_size + 1))
    return 2 ** p - 1 - matrice_size


def identity_padding(A, padding_size):
    n = A.shape[0]
    Ap = np.zeros((n + padding_size, n + padding_size), dtype=A.dtype)
    Ap[:n, :n] = A
    Ap[n:, n:] = np.eye(padding_size)
    return Ap


# --------------------------------------------------------------------------
# dense, op-for-op
def bcr_serial_dense(A, blocksize):
    b = blocksize
    nb0 = A.shape[0] // b
    A = identity_padding(A, distance_to_power_of_two(nb0) * b)
    nbp = A.shape[0] // b
    L = np.zeros((nbp * b, nbp * b), dtype=A.dtype)
    U = np.zeros_like(L)
    G = np.zeros_like(L)

    def blk(M, r, c):
        # numpy slice semantics of the reference, including negative/oob starts
        return M[r * b:(r + 1) * b, c * b:(c + 1) * b]

    def setblk(M, r, c, v):
        M[r * b:(r + 1) * b, c * b:(c + 1) * b] = v

    height = int(math.log2(nbp))
    root = 0
    # reduce_bcr (bcr_serial.py:204-216)
    for lvl in range(height):
        s = 2 ** lvl
        keep = list(range(2 * s - 1, nbp, 2 * s))
        for j in keep:
            i, k = j - s, j + s
            if i >= 0:                                   # :119-143
                Aii_inv = inv(blk(A, i, i))
                setblk(U, i, j, Aii_inv @ blk(A, i, j))
                setblk(L, j, i, blk(A, j, i) @ Aii_inv)
                setblk(A, j, j, blk(A, j, j) - blk(L, j, i) @ blk(A, i, j))
                if i != keep[0]:
                    h = i - s
                    tgt = A[j * b:(j + 1) * b, h * b:(h + 1) * b]
                    if tgt.size:
                        tgt[...] = -blk(L, j, i) @ A[i * b:(i + 1) * b, h * b:(h + 1) * b]
            if k <= nbp - 1:                             # :146-169
                Akk_inv = inv(blk(A, k, k))
                setblk(U, k, j, Akk_inv @ blk(A, k, j))
                setblk(L, j, k, blk(A, j, k) @ Akk_inv)
                setblk(A, j, j, blk(A, j, j) - blk(L, j, k) @ blk(A, k, j))
                if k != keep[-1]:
                    l_ = k + s
                    tgt = A[j * b:(j + 1) * b, l_ * b:(l_ + 1) * b]
                    if tgt.size:
                        tgt[...] = -blk(L, j, k) @ A[k * b:(k + 1) * b, l_ * b:(l_ + 1) * b]
        root = keep[-1]
    # invert_block (bcr_serial.py:248-249)
    setblk(G, root, root, inv(blk(A, root, root)))

    def corner(kf, kt):                                  # :293-304
        setblk(G, kf, kt, -1 * blk(G, kf, kf) @ blk(L, kf, kt))
        setblk(G, kt, kf, -blk(U, kt, kf) @ blk(G, kf, kf))
        setblk(G, kt, kt, inv(blk(A, kt, kt)) - blk(G, kt, kf) @ blk(L, kf, kt))

    def center(ka, kt, kb):                              # :354-383
        setblk(G, ka, kt, -blk(G, ka, ka) @ blk(L, ka, kt) - blk(G, ka, kb) @ blk(L, kb, kt))
        setblk(G, kb, kt, -blk(G, kb, ka) @ blk(L, ka, kt) - blk(G, kb, kb) @ blk(L, kb, kt))
        setblk(G, kt, ka, -blk(U, kt, ka) @ blk(G, ka, ka) - blk(U, kt, kb) @ blk(G, kb, ka))
        setblk(G, kt, kb, -blk(U, kt, ka) @ blk(G, ka, kb) - blk(U, kt, kb) @ blk(G, kb, kb))
        setblk(G, kt, kt, inv(blk(A, kt, kt)) - blk(G, kt, ka) @ blk(L, ka, kt)
               - blk(G, kt, kb) @ blk(L, kb, kt))

    # produce_bcr (bcr_serial.py:417-466)
    for lvl in range(height - 1, -1, -1):
        s = 2 ** lvl
        i_from = list(range(2 * s - 1, nbp, 2 * s))
        i_prod = []
        for n_, f in enumerate(i_from):
            if n_ == 0:
                i_prod += [f - s, f + s]
            else:
                if i_prod[n_] != f - s:
                    i_prod.append(f - s)
                i_prod.append(f + s)
        last = len(i_prod) - 1
        for q, kt in enumerate(i_prod):
            if q == 0:
                corner(kt + s, kt)
            if q != 0 and q == last:
                if i_prod[-1] <= nbp - s - 1:
                    center(kt - s, kt, kt + s)
                else:
                    corner(kt - s, kt)
            if q != 0 and q != last:
                center(kt - s, kt, kt + s)
    return G[:nb0 * b, :nb0 * b]


# --------------------------------------------------------------------------
# block-compressed, padding skipped
def bcr_schedule(N):
    """Level schedule shared by the compressed restatement and the GPU driver.

    Returns (height, root, levels) with levels[l] = (s, elim_rows, keep_rows)
    restricted to real rows (< N).  The padded size is 2^p - 1 blocks
    (bcr_utils.py:35-38); height = log2 of it, floored (bcr_serial.py:204).
    """
    nbp = N + distance_to_power_of_two(N)
    height = int(math.log2(nbp))
    levels = []
    for lvl in range(height):
        s = 2 ** lvl
        elim = list(range(s - 1, N, 2 * s))
        keep = list(range(2 * s - 1, N, 2 * s))
        levels.append((s, elim, keep))
    root = 2 ** height - 1 if height > 0 else 0
    return height, root, levels


def bcr_blocks(diag, upper, lower, return_fill=False):
    N = diag.shape[0]
    A = {}
    for i in range(N):
        A[i, i] = diag[i].copy()
        if i < N - 1:
            A[i, i + 1] = upper[i].copy()
            A[i + 1, i] = lower[i].copy()
    L, U, G, Ainv = {}, {}, {}, {}
    height, root, levels = bcr_schedule(N)

    for s, elim, keep in levels:
        for t in elim:
            Ainv[t] = inv(A[t, t])
        for j in keep:
            i, k = j - s, j + s
            U[i, j] = Ainv[i] @ A[i, j]
            L[j, i] = A[j, i] @ Ainv[i]
            A[j, j] = A[j, j] - L[j, i] @ A[i, j]
            if i - s >= 0:
                A[j, i - s] = -L[j, i] @ A[i, i - s]
            if k < N:
                U[k, j] = Ainv[k] @ A[k, j]
                L[j, k] = A[j, k] @ Ainv[k]
                A[j, j] = A[j, j] - L[j, k] @ A[k, j]
                if k + s < N:
                    A[j, k + s] = -L[j, k] @ A[k, k + s]
    G[root, root] = inv(A[root, root])

    for s, elim, keep in reversed(levels):
        for t in elim:
            a, c = t - s, t + s
            if a < 0 or c >= N:
                f = c if a < 0 else a
                G[f, t] = -1 * G[f, f] @ L[f, t]
                G[t, f] = -U[t, f] @ G[f, f]
                G[t, t] = Ainv[t] - G[t, f] @ L[f, t]
            else:
                G[a, t] = -G[a, a] @ L[a, t] - G[a, c] @ L[c, t]
                G[c, t] = -G[c, a] @ L[a, t] - G[c, c] @ L[c, t]
                G[t, a] = -U[t, a] @ G[a, a] - U[t, c] @ G[c, a]
                G[t, c] = -U[t, a] @ G[a, c] - U[t, c] @ G[c, c]
                G[t, t] = Ainv[t] - G[t, a] @ L[a, t] - G[t, c] @ L[c, t]

    Gd = np.stack([G[i, i] for i in range(N)])
    Gu = np.stack([G[i, i + 1] for i in range(N - 1)]) if N > 1 else np.zeros((0,) + diag.shape[1:], diag.dtype)
    Gl = np.stack([G[i + 1, i] for i in range(N - 1)]) if N > 1 else np.zeros((0,) + diag.shape[1:], diag.dtype)
    if return_fill:
        return Gd, Gu, Gl, G
    return Gd, Gu, Gl


def bcr_flops(N, b):
    """Real flops of the compressed schedule (inverse and complex GEMM = 8b^3 each)."""
    height, root, levels = bcr_schedule(N)
    n_inv, n_gemm = 1, 0
    for s, elim, keep in levels:
        n_inv += len(elim)
        for j in keep:
            i, k = j - s, j + s
            n_gemm += 3 + (1 if i - s >= 0 else 0)
            if k < N:
                n_gemm += 3 + (1 if k + s < N else 0)
    for s, elim, keep in levels:
        for t in elim:
            a, c = t - s, t + s
            n_gemm += 3 if (a < 0 or c >= N) else 10
    return n_inv, n_gemm
