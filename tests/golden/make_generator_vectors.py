"""Writes tests/golden/generator_vectors.json: the reference's own matrix
generators and converters (src/sinv/utils/matrices_utils.py, numpy only),
run on the shapes its tests use, frozen as fixtures.

The reference's tests build every input with
generateBandedDiagonalMatrix(size, ceil(b / 2), is_complex, is_symmetric,
seed=63) (bcr_serial_test.py:61, pdiv_test.py:70, psr_test.py:69,
rgf2sided_test.py:60).  No other check pins these generators, so this script
imports the reference's module in the build container (by file path, so that
the package's other imports -- mpi4py, quasi -- are not needed) and records,
per case, the SHA-256 of the generated matrix's bytes (an exact pin of the
legacy-RNG stream), its first row, its diagonal and its Frobenius norm; for a
few small cases it also stores the converters' outputs in full.  Only the
numbers are kept; nothing of the reference's source is.

    python tests/golden/make_generator_vectors.py
"""

import hashlib
import importlib.util
import json
import math
import os

import numpy as np

REF = "/root/reference/src/sinv/utils/matrices_utils.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "generator_vectors.json")

# (size, blocksize) of bcr_serial_test.py:40-65 and pdiv_test.py:28-56
SHAPES = [(9, 1), (18, 2), (27, 3), (12, 1), (24, 2), (36, 3), (256, 8), (240, 10), (144, 12),
          (2, 1), (4, 2), (6, 3), (3, 1), (6, 2), (9, 3), (128, 8), (128, 16), (128, 32)]
RANDOM_SIZES = [1, 5, 17, 64]
CONVERT = [(12, 3, True, False), (9, 3, False, True), (8, 4, True, True)]


def digest(A):
    return hashlib.sha256(np.ascontiguousarray(A).tobytes()).hexdigest()


def pack(A):
    A = np.asarray(A)
    if np.iscomplexobj(A):
        return {"re": A.real.tolist(), "im": A.imag.tolist()}
    return {"re": A.tolist()}


def summary(A):
    return {"dtype": str(A.dtype), "shape": list(A.shape), "sha256": digest(A),
            "first_row": pack(A[0]), "diag": pack(np.diag(A)),
            "fro": float(np.linalg.norm(A))}


def main():
    spec = importlib.util.spec_from_file_location("ref_matrices_utils", REF)
    mu = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mu)
    out = {"banded": [], "random": [], "convert": []}
    for n, b in SHAPES:
        bw = math.ceil(b / 2)
        for cplx in (False, True):
            for sym in (False, True):
                A = mu.generateBandedDiagonalMatrix(n, bw, cplx, sym, 63)
                out["banded"].append(dict(size=n, bandwidth=bw, is_complex=cplx,
                                          is_symmetric=sym, seed=63, **summary(A)))
    for n in RANDOM_SIZES:
        for cplx in (False, True):
            for sym in (False, True):
                A = mu.generateRandomNumpyMat(n, cplx, sym, 63)
                out["random"].append(dict(size=n, is_complex=cplx, is_symmetric=sym, seed=63,
                                          **summary(A)))
    for n, b, cplx, sym in CONVERT:
        A = mu.generateBandedDiagonalMatrix(n, math.ceil(b / 2), cplx, sym, 63)
        d, u, l = mu.convertDenseToBlkTridiag(A, b)
        back = mu.convertBlkTridiagToDense(d, u, l)
        out["convert"].append(dict(size=n, blocksize=b, is_complex=cplx, is_symmetric=sym, seed=63,
                                   A=pack(A), diag=pack(np.asarray(d)), upper=pack(np.asarray(u)),
                                   lower=pack(np.asarray(l)), dense=pack(back)))
    with open(OUT, "w") as f:
        json.dump(out, f)
    print(OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
