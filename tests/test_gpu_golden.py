"""The GPU drop-ins on the frozen acceptance cases of the reference's tests
(tests/golden/): each must meet the reference's own criterion (numpy.allclose
against the dense inverse's selected blocks)."""

import numpy as np
import pytest

from oracle import dense, pdiv_ref
from tests.test_golden import CASES, close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag", sorted(CASES))
def test_dropins_meet_the_reference_acceptance(built_lib, tag):
    from sinv_amd.algorithms import bcr_s, pdiv_lm, psr_s, rgf, rgf2s
    print(f"golden {tag}", flush=True)
    c = CASES[tag]
    A, b = c["A"], int(c["b"])
    d, u, l = dense.select(A, b)
    assert close(rgf.rgf_blocks(d, u, l), c)
    assert close(dense.select(bcr_s.bcr_serial(A, b), b), c)
    assert close(rgf2s.rgf2sided(d, u, l), c)
    N = A.shape[0] // b
    if N >= 6:
        assert close(dense.select(psr_s.psr_seqsolve(A, b, 2), b), c)
    for P in (2, 4):
        if P <= N:
            assert close(pdiv_ref.stacked(pdiv_lm.pdiv(A, b, P)), c)
