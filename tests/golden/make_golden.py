"""Writes tests/golden/selected_inverse_cases.npz: inputs and expected outputs of
the reference's own acceptance tests, frozen.

The reference ships no data files; its tests (rgf_test.py:41-94,
bcr_serial_test.py:40-73, psr_test.py, pdiv_test.py:28-115) build their
matrices with matrices_utils.generateBandedDiagonalMatrix(size, bandwidth,
is_complex, is_symmetric, seed=63) and accept a result if it is allclose to
numpy.linalg.inv of the matrix restricted to the block-tridiagonal pattern.
This script regenerates a few of those inputs with the same legacy-RNG
generator (sinv_amd.utils.matu, pinned to the reference's stream in
tests/test_host_layer.py) and stores each input with the selected blocks of
its dense inverse.

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from sinv_amd.utils import matu  # noqa: E402

# (tag, size, blocksize, bandwidth, is_complex, is_symmetric); every reference
# test uses bandwidth ceil(b / 2) (bcr_serial_test.py:61, psr_test.py:69,
# rgf2sided_test.py:60, pdiv_test.py:70), which keeps A block-tridiagonal
CASES = [
    ("bcr_128_32_cplx", 128, 32, 16, True, False),
    ("bcr_96_16_real_sym", 96, 16, 8, False, True),
    ("pdiv_128_16_real_sym", 128, 16, 8, False, True),
    ("pdiv_128_8_cplx", 128, 8, 4, True, False),
    ("rgf2s_64_8_cplx_sym", 64, 8, 4, True, True),
]


def main():
    out = {}
    for tag, n, b, bw, cplx, sym in CASES:
        A = matu.generateBandedDiagonalMatrix(n, bw, cplx, sym, 63)
        X = np.linalg.inv(A)
        N = n // b
        out[f"{tag}__A"] = A
        out[f"{tag}__b"] = np.array(b)
        out[f"{tag}__Xd"] = np.stack([X[i * b:(i + 1) * b, i * b:(i + 1) * b] for i in range(N)])
        out[f"{tag}__Xu"] = np.stack([X[i * b:(i + 1) * b, (i + 1) * b:(i + 2) * b] for i in range(N - 1)])
        out[f"{tag}__Xl"] = np.stack([X[(i + 1) * b:(i + 2) * b, i * b:(i + 1) * b] for i in range(N - 1)])
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selected_inverse_cases.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
