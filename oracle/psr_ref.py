"""Restatement of the reference's PSR (parallel Schur reduction) on one process.

Test infrastructure only (see ``oracle/__init__.py``).

Follows src/sinv/algorithms/psr/psr_seqsolve.py with the P ranks' work done
one after the other on dense block storage:
  * divide_matrix (psr_utils.py:61-99): partition m = N // P blocks, the
    first partition takes the remainder N - m (P - 1)
  * reduce_schur_topleftcorner      :127-169   (partition 0, downwards)
  * reduce_schur_central            :222-309   (partitions 1..P-2)
  * reduce_schur_bottomrightcorner  :173-218   (partition P-1, upwards)
  * aggregate_reduced_system        :313-383   (2(P-1) boundary rows)
  * inverse_reduced_system          :470-491   (numpy.linalg.inv of it)
  * produce_schur_topleftcorner / _bottomrightcorner / _central :693-898
The MPI exchanges become array reads.  Returns the block-tridiagonal part
(diag, upper, lower) of the produced G.
"""

import numpy as np

inv = np.linalg.inv


def partitions(N, P):
    m = N // P
    n0 = N - m * (P - 1)
    starts = [0] + [n0 + p * m for p in range(P - 1)]
    sizes = [n0] + [m] * (P - 1)
    return starts, sizes


def psr_blocks(diag, upper, lower, P):
    N, b = diag.shape[0], diag.shape[1]
    if N < 3 * P:
        raise ValueError("The number of blocks is to low. There should be at least 3 blockrows per process")
    starts, sizes = partitions(N, P)
    A = {}
    for i in range(N):
        A[i, i] = diag[i].copy()
        if i < N - 1:
            A[i, i + 1] = upper[i].copy()
            A[i + 1, i] = lower[i].copy()
    Z = np.zeros((b, b), dtype=diag.dtype)
    get = lambda r, c: A.get((r, c), Z)
    L, U, G = {}, {}, {}

    for p in range(P):
        s, n = starts[p], sizes[p]
        e = s + n - 1
        if p == 0:                                   # psr_seqsolve.py:157-167
            for i in range(s + 1, s + n):
                Ainv = inv(get(i - 1, i - 1))
                L[i, i - 1] = get(i, i - 1) @ Ainv
                U[i - 1, i] = Ainv @ get(i - 1, i)
                A[i, i] = get(i, i) - L[i, i - 1] @ get(i - 1, i)
        elif p == P - 1:                             # psr_seqsolve.py:203-216
            for i in range(e - 1, s - 1, -1):
                Ainv = inv(get(i + 1, i + 1))
                L[i, i + 1] = get(i, i + 1) @ Ainv
                U[i + 1, i] = Ainv @ get(i + 1, i)
                A[i, i] = get(i, i) - L[i, i + 1] @ get(i + 1, i)
        else:                                        # psr_seqsolve.py:260-307
            for i in range(s + 2, s + n):
                Ainv = inv(get(i - 1, i - 1))
                L[i, i - 1] = get(i, i - 1) @ Ainv
                L[s, i - 1] = get(s, i - 1) @ Ainv
                U[i - 1, i] = Ainv @ get(i - 1, i)
                U[i - 1, s] = Ainv @ get(i - 1, s)
                A[i, i] = get(i, i) - L[i, i - 1] @ get(i - 1, i)
                A[s, s] = get(s, s) - L[s, i - 1] @ get(i - 1, s)
                A[i, s] = get(i, s) - L[i, i - 1] @ get(i - 1, s)
                A[s, i] = get(s, i) - L[s, i - 1] @ get(i - 1, i)

    # reduced system over the boundary rows (psr_seqsolve.py:313-383)
    red = [starts[0] + sizes[0] - 1]
    for p in range(1, P - 1):
        red += [starts[p], starts[p] + sizes[p] - 1]
    red.append(starts[P - 1])
    R = len(red)
    As = np.zeros((R * b, R * b), dtype=diag.dtype)
    for x, rx in enumerate(red):
        for y, ry in enumerate(red):
            As[x * b:(x + 1) * b, y * b:(y + 1) * b] = get(rx, ry)
    Gs = inv(As)                                     # psr_seqsolve.py:489
    for x, rx in enumerate(red):
        for y, ry in enumerate(red):
            G[rx, ry] = Gs[x * b:(x + 1) * b, y * b:(y + 1) * b]

    for p in range(P):
        s, n = starts[p], sizes[p]
        e = s + n - 1
        if p == 0:                                   # psr_seqsolve.py:731-740
            for i in range(e, s, -1):
                G[i, i - 1] = -G[i, i] @ L[i, i - 1]
                G[i - 1, i] = -U[i - 1, i] @ G[i, i]
                G[i - 1, i - 1] = inv(get(i - 1, i - 1)) - U[i - 1, i] @ G[i, i - 1]
        elif p == P - 1:                             # psr_seqsolve.py:777-786
            for i in range(s, e):
                G[i, i + 1] = -G[i, i] @ L[i, i + 1]
                G[i + 1, i] = -U[i + 1, i] @ G[i, i]
                G[i + 1, i + 1] = inv(get(i + 1, i + 1)) - U[i + 1, i] @ G[i, i + 1]
        else:                                        # psr_seqsolve.py:823-898
            G[e, e - 1] = -1 * G[e, s] @ L[s, e - 1] - G[e, e] @ L[e, e - 1]
            G[e - 1, e] = -1 * U[e - 1, e] @ G[e, e] - U[e - 1, s] @ G[s, e]
            for i in range(e - 1, s, -1):
                G[s, i] = -1 * G[s, s] @ L[s, i] - G[s, i + 1] @ L[i + 1, i]
                G[i, s] = -1 * U[i, i + 1] @ G[i + 1, s] - U[i, s] @ G[s, s]
            for i in range(e - 1, s + 1, -1):
                G[i, i] = inv(get(i, i)) - U[i, s] @ G[s, i] - U[i, i + 1] @ G[i + 1, i]
                G[i - 1, i] = -1 * U[i - 1, s] @ G[s, i] - U[i - 1, i] @ G[i, i]
                G[i, i - 1] = -1 * G[i, s] @ L[s, i - 1] - G[i, i] @ L[i, i - 1]
            G[s + 1, s + 1] = (inv(get(s + 1, s + 1)) - U[s + 1, s] @ G[s, s + 1]
                               - U[s + 1, s + 2] @ G[s + 2, s + 1])

    Gd = np.stack([G[i, i] for i in range(N)])
    Gu = np.stack([G[i, i + 1] for i in range(N - 1)])
    Gl = np.stack([G[i + 1, i] for i in range(N - 1)])
    return Gd, Gu, Gl
