"""Frozen acceptance cases of the reference's tests (tests/golden/, written by
tests/golden/make_golden.py): seed-63 banded matrices and the selected blocks
of their dense inverse.  The inputs must regenerate bit for bit, and every
oracle restatement must meet the reference's own acceptance (numpy.allclose
against the dense inverse) on them."""

import os

import numpy as np
import pytest

from oracle import bcr_ref, dense, pdiv_ref, psr_ref, rgf2s_ref, rgf_ref
from sinv_amd.utils import matu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "selected_inverse_cases.npz")


def load_cases():
    z = np.load(GOLDEN)
    tags = sorted({k.split("__")[0] for k in z.files})
    return {t: {k: z[f"{t}__{k}"] for k in ("A", "b", "Xd", "Xu", "Xl")} for t in tags}


CASES = load_cases()


def close(G, c):
    return (np.allclose(G[0], c["Xd"]) and np.allclose(G[1], c["Xu"])
            and np.allclose(G[2], c["Xl"]))


def test_inputs_regenerate_bit_for_bit():
    from tests.golden.make_golden import CASES as SPEC
    for tag, n, b, bw, cplx, sym in SPEC:
        assert np.array_equal(matu.generateBandedDiagonalMatrix(n, bw, cplx, sym, 63),
                              CASES[tag]["A"]), tag


@pytest.mark.parametrize("tag", sorted(CASES))
def test_oracles_meet_the_reference_acceptance(tag):
    c = CASES[tag]
    A, b = c["A"], int(c["b"])
    d, u, l = dense.select(A, b)
    assert close(rgf_ref.rgf_blocks(d, u, l), c)
    assert close(bcr_ref.bcr_blocks(d, u, l), c)
    assert close(dense.select(bcr_ref.bcr_serial_dense(A, b), b), c)
    assert close(rgf2s_ref.rgf2sided_blocks(d, u, l), c)
    N = A.shape[0] // b
    if N >= 6:
        assert close(psr_ref.psr_blocks(d, u, l, 2), c)
    for P in (2, 4):
        if P <= N:
            assert close(pdiv_ref.stacked(pdiv_ref.pdiv_blocks(A, b, P)), c)
