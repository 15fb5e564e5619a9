"""Restatement of the reference's PDIV (P-Division, pairwise) selected inversion
on one process.

Test infrastructure only (see ``oracle/__init__.py``).

Follows src/sinv/algorithms/pdiv/pdiv_localmap.py with the P ranks simulated
in a loop; every MPI send becomes a read of the sender's state as it was when
the send happened:
  * divide_matrix / partition_subdomain (pdiv_utils.py:84-170): the first
    partition takes the remainder N - m (P - 1), m = N // P; bridges
    Bu_i = A[e_i, e_i + 1], Bl_i = A[e_i + 1, e_i]
  * invert_partition                 pdiv_localmap.py:94-115 (dense inverse)
  * initialize_matrixmaps / crossmaps                 :119-184
  * update_maps, reduction steps 1..log2(P)           :188-246
      get_middle_process :250-271, get_U :275-356 (corner blocks from the
      middle rank and the one after it), get_J :360-400,
      get_nextprocess_matrixmap :404-466 (maps of rank + 1 before this
      step's updates), update_crossmap :556-816, update_matrixmap :820-985
  * produce_partition :989-1041, produce_*_element :1045-1241,
    produce_bridges :1291-1345, produce_upper/lower_bridge :1349-1436
Returns, per partition, the reference's three lists (diag, upper, lower),
whose last upper/lower entries are the bridges to the next partition (zero
on the last partition).
"""

import math

import numpy as np

inv = np.linalg.inv


def check_input(A, blocksize, npart):
    """pdiv_utils.py:16-80."""
    if not math.log2(npart).is_integer():
        raise ValueError("The number of processes must be a power of 2.")
    if A.shape[0] % blocksize != 0:
        raise ValueError("The matrix size must be a multiple of the blocksize.")
    if blocksize > A.shape[0]:
        raise ValueError("The blocksize must be smaller than the matrix size.")
    if blocksize < 1:
        raise ValueError("The blocksize must be greater than 0.")
    if A.shape[0] // blocksize < npart:
        raise ValueError("The number of blocks must be greater than the number of processes.")


def divide_matrix(A, npart, b):
    N = A.shape[0] // b
    m = N // npart
    sizes = [N - m * (npart - 1)] + [m] * (npart - 1)
    starts = [0]
    for p in range(1, npart):
        starts.append(starts[-1] + sizes[p - 1])
    return starts, sizes


def partition_subdomain(A, starts, sizes, b):
    K, Bu, Bl = [], [], []
    for i in range(len(starts)):
        s0 = starts[i] * b
        s1 = s0 + sizes[i] * b
        K.append(A[s0:s1, s0:s1].copy())
        if i < len(starts) - 1:
            Bu.append(A[s1 - b:s1, s1:s1 + b].copy())
            Bl.append(A[s1:s1 + b, s1 - b:s1].copy())
    return K, Bu, Bl


def _blk(K, r, c, b):
    return K[r * b:(r + 1) * b, c * b:(c + 1) * b]


def _last(K, b):
    return K.shape[0] // b - 1


def _elem(r, c, K, M, b):
    """produce_matrix_elements (pdiv_localmap.py:1194-1241)."""
    L = _last(K, b)
    R0 = K[r * b:(r + 1) * b, 0:b]
    RL = K[r * b:(r + 1) * b, L * b:(L + 1) * b]
    C0 = K[0:b, c * b:(c + 1) * b]
    CL = K[L * b:(L + 1) * b, c * b:(c + 1) * b]
    return (K[r * b:(r + 1) * b, c * b:(c + 1) * b] + R0 @ M[8] @ C0 + R0 @ M[9] @ CL
            + RL @ M[10] @ C0 + RL @ M[11] @ CL)


def pdiv_blocks(A, b, npart):
    check_input(A, b, npart)
    starts, sizes = divide_matrix(A, npart, b)
    K, Bu, Bl = partition_subdomain(A, starts, sizes, b)
    P = npart
    dt = A.dtype if np.iscomplexobj(A) else np.float64
    eye = np.identity(b, dtype=dt)
    Z = lambda: np.zeros((b, b), dtype=dt)

    K = [inv(k) for k in K]                                           # :94-115
    M = [[eye.copy() if i in (0, 3, 4, 7) else Z() for i in range(12)] for _ in range(P)]
    Mn = [[Z() for _ in range(12)] for _ in range(P)]                 # l_M_ip1
    C = [[Z() for _ in range(12)] for _ in range(P)]

    for step in range(1, int(math.log2(P)) + 1):                     # :188-246
        stride = 2 ** step
        for start in range(0, P, stride):
            end = start + stride - 1
            mid = start - 1 + math.ceil((end - start) / 2)
            Km, Mm, Kd, Md = K[mid], M[mid], K[mid + 1], M[mid + 1]
            Lm, Ld = _last(Km, b), _last(Kd, b)
            UUR = Mm[0] @ _blk(Km, 0, Lm, b) + Mm[1] @ _blk(Km, Lm, Lm, b)      # :1045-1081
            ULL = _blk(Km, Lm, 0, b) @ Mm[4] + _blk(Km, Lm, Lm, b) @ Mm[5]      # :1123-1157
            ULR = _elem(Lm, Lm, Km, Mm, b)
            DUL = _elem(0, 0, Kd, Md, b)
            DUR = _blk(Kd, 0, 0, b) @ Md[2] + _blk(Kd, 0, Ld, b) @ Md[3]        # :1085-1119
            DLL = Md[6] @ _blk(Kd, 0, 0, b) + Md[7] @ _blk(Kd, Ld, 0, b)        # :1161-1190
            Bu_m, Bl_m = Bu[mid], Bl[mid]
            J = np.zeros((2 * b, 2 * b), dtype=dt)                              # :360-400
            J[:b, :b] = eye
            J[:b, b:] = -DUL @ Bl_m
            J[b:, :b] = -ULR @ Bu_m
            J[b:, b:] = eye
            J = inv(J)
            J11, J12, J21, J22 = J[:b, :b], J[:b, b:], J[b:, :b], J[b:, b:]
            # maps of rank r + 1 as sent before this step's updates (:404-552)
            sent = [[x.copy() for x in M[r]] for r in range(P)]
            for r in range(start, end):
                keys = (0, 1, 4, 5) if r >= mid else (2, 3, 6, 7)
                for k in keys:
                    Mn[r][k] = sent[r + 1][k]
            for r in range(start, end + 1):
                Mr, Nr, Cr = M[r], Mn[r], C[r]
                if r < mid:                                                      # :604-658
                    for i, (a, bb) in enumerate([(Mr[2], Nr[6]), (Mr[2], Nr[7]), (Mr[3], Nr[6]),
                                                 (Mr[3], Nr[7]), (Nr[2], Mr[6]), (Nr[2], Mr[7]),
                                                 (Nr[3], Mr[6]), (Nr[3], Mr[7])]):
                        Cr[i] = Cr[i] - a @ Bu_m @ J12 @ bb
                elif r == mid:                                                   # :662-719
                    for i, (a, bb) in enumerate([(Mr[2], Nr[0]), (Mr[2], Nr[1]), (Mr[3], Nr[0]),
                                                 (Mr[3], Nr[1])]):
                        Cr[i] = Cr[i] - a @ Bu_m @ J11 @ bb
                    for i, (a, bb) in enumerate([(Nr[4], Mr[6]), (Nr[4], Mr[7]), (Nr[5], Mr[6]),
                                                 (Nr[5], Mr[7])]):
                        Cr[4 + i] = Cr[4 + i] - a @ Bl_m @ J22 @ bb
                elif r < end:                                                    # :723-776
                    for i, (a, bb) in enumerate([(Mr[4], Nr[0]), (Mr[4], Nr[1]), (Mr[5], Nr[0]),
                                                 (Mr[5], Nr[1]), (Nr[4], Mr[0]), (Nr[4], Mr[1]),
                                                 (Nr[5], Mr[0]), (Nr[5], Mr[1])]):
                        Cr[i] = Cr[i] - a @ Bl_m @ J21 @ bb
                if r <= mid:                                                     # :874-928
                    Mr[8] = Mr[8] + Mr[2] @ Bu_m @ J12 @ Mr[6]
                    Mr[9] = Mr[9] + Mr[2] @ Bu_m @ J12 @ Mr[7]
                    Mr[10] = Mr[10] + Mr[3] @ Bu_m @ J12 @ Mr[6]
                    Mr[11] = Mr[11] + Mr[3] @ Bu_m @ J12 @ Mr[7]
                    Mr[0] = Mr[0] + UUR @ Bu_m @ J12 @ Mr[6]
                    Mr[1] = Mr[1] + UUR @ Bu_m @ J12 @ Mr[7]
                    Mr[4] = Mr[4] + Mr[2] @ Bu_m @ J12 @ ULL
                    Mr[5] = Mr[5] + Mr[3] @ Bu_m @ J12 @ ULL
                    Mr[2] = Mr[2] @ Bu_m @ J11 @ DUR
                    Mr[3] = Mr[3] @ Bu_m @ J11 @ DUR
                    Mr[6] = DLL @ Bl_m @ J22 @ Mr[6]
                    Mr[7] = DLL @ Bl_m @ J22 @ Mr[7]
                else:                                                            # :932-985
                    Mr[8] = Mr[8] + Mr[4] @ Bl_m @ J21 @ Mr[0]
                    Mr[9] = Mr[9] + Mr[4] @ Bl_m @ J21 @ Mr[1]
                    Mr[10] = Mr[10] + Mr[5] @ Bl_m @ J21 @ Mr[0]
                    Mr[11] = Mr[11] + Mr[5] @ Bl_m @ J21 @ Mr[1]
                    Mr[2] = Mr[2] + Mr[4] @ Bl_m @ J21 @ DUR
                    Mr[3] = Mr[3] + Mr[5] @ Bl_m @ J21 @ DUR
                    Mr[4] = Mr[4] @ Bl_m @ J22 @ ULL
                    Mr[5] = Mr[5] @ Bl_m @ J22 @ ULL
                    Mr[6] = Mr[6] + DLL @ Bl_m @ J21 @ Mr[0]
                    Mr[7] = Mr[7] + DLL @ Bl_m @ J21 @ Mr[1]
                    Mr[0] = UUR @ Bu_m @ J11 @ Mr[0]
                    Mr[1] = UUR @ Bu_m @ J11 @ Mr[1]

    out = []
    for r in range(P):                                                   # :989-1041
        n = sizes[r]
        D = [_elem(i, i, K[r], M[r], b) for i in range(n)]
        U = [_elem(i, i + 1, K[r], M[r], b) if i < n - 1 else Z() for i in range(n)]
        L = [_elem(i + 1, i, K[r], M[r], b) if i < n - 1 else Z() for i in range(n)]
        if r < P - 1:                                                    # :1291-1436
            Lr, K2 = _last(K[r], b), K[r + 1]
            L2 = _last(K2, b)
            p1_1N, p1_N1, p1_NN = _blk(K[r], 0, Lr, b), _blk(K[r], Lr, 0, b), _blk(K[r], Lr, Lr, b)
            p2_11, p2_N1, p2_1N = _blk(K2, 0, 0, b), _blk(K2, L2, 0, b), _blk(K2, 0, L2, b)
            c = C[r]
            U[-1] = p1_N1 @ c[0] @ p2_11 + p1_N1 @ c[1] @ p2_N1 + p1_NN @ c[2] @ p2_11 \
                + p1_NN @ c[3] @ p2_N1
            L[-1] = p2_11 @ c[4] @ p1_1N + p2_11 @ c[5] @ p1_NN + p2_1N @ c[6] @ p1_1N \
                + p2_1N @ c[7] @ p1_NN
        out.append((D, U, L))
    return out


def stacked(out):
    """Per-partition lists -> global stacked (diag (N,b,b), upper (N-1,b,b), lower (N-1,b,b))."""
    D = [x for d, _, _ in out for x in d]
    U = [x for _, u, _ in out for x in u][:-1]
    L = [x for _, _, l in out for x in l][:-1]
    return np.stack(D), np.stack(U), np.stack(L)
