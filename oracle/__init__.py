"""CPU oracle for the selected-inversion hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker / the timed CPU baseline.  The product path
(``sinv_amd``) never imports it and has no CPU fallback.

Contents (each function cites the reference file:line it restates):

* ``rgf_ref``   -- op-for-op numpy restatement of ``rgf`` (rgf.py:40-62),
                   including the NaN initialisation and ``(X@Y)@Z`` association.
* ``bcr_ref``   -- op-for-op numpy restatement of ``bcr_serial``
                   (bcr_serial.py:39-466, bcr_utils.py:17-71) on dense storage,
                   plus a block-compressed restatement of the same schedule
                   without the identity padding.
* ``matu_ref``  -- the reference's generators and dense <-> stacked-block
                   converters (matrices_utils.py:12-189).
* ``dense``     -- the closed-form pin: selected blocks of ``numpy.linalg.inv``
                   of the assembled dense matrix, and the exact selected-block
                   residual identities.
* ``adjoint_ref`` -- gradient oracles for the adjoint sweep: the dense formula
                   ``dA = -G^H dG G^H`` restricted to the pattern, and central
                   finite differences.

How each function is pinned (closed forms, brute force, invariants) is listed in
the test files ``tests/test_oracle_*.py``.
"""
