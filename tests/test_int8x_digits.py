"""CPU checks of the digit split behind the int8x GEMM (sinv_amd/csrc/int8x.hip).

The kernel takes the S signed 7-bit digits of y in (-1, 1) as the base-128
digits of floor(|y| 128^S) with y's sign, from a high part of 4 digits and a
low part of S - 4 digits.  These tests restate both that form and the
sequential form (y <- 128 y; q = trunc(y); y <- y - q) in numpy and check that
they agree digit for digit, that every digit fits int8, and that the digits
reconstruct y to within 128^-S; then that the digit-pair product with pairs of
index sum >= S dropped stays inside the bound the GPU tests use.
"""

import numpy as np
import pytest


def digits_sequential(y, S):
    q = np.zeros(y.shape + (S,), dtype=np.int64)
    r = y.copy()
    for t in range(S):
        r = r * 128.0
        d = np.trunc(r)
        q[..., t] = d.astype(np.int64)
        r = r - d
    return q


def digits_integer(y, S):
    lb = 7 * (S - 4)
    a = np.ldexp(np.abs(y), 7 * S)
    hd = np.trunc(np.ldexp(a, -lb))
    hi = hd.astype(np.uint64)
    lo = np.trunc(a - np.ldexp(hd, lb)).astype(np.uint64)
    q = np.zeros(y.shape + (S,), dtype=np.int64)
    for t in range(4):
        q[..., t] = ((hi >> np.uint64(7 * (3 - t))) & np.uint64(127)).astype(np.int64)
    for t in range(4, S):
        q[..., t] = ((lo >> np.uint64(7 * (S - 1 - t))) & np.uint64(127)).astype(np.int64)
    return np.where((y < 0)[..., None], -q, q)


def _samples(rng, n):
    y = rng.uniform(-1, 1, n)
    # magnitudes across the whole exponent range below 1, exact digit boundaries and signed zeros
    y = np.concatenate([y, rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-60, 0, n),
                        np.array([0.0, -0.0, 1 - 2 ** -53, -(1 - 2 ** -53), 2 ** -7, -2 ** -7,
                                  127 / 128, -127 / 128, 2 ** -49, -2 ** -49, 2 ** -50])])
    return y


@pytest.mark.parametrize("S", [5, 6, 7])
def test_integer_digits_equal_sequential_digits(S):
    y = _samples(np.random.default_rng(S), 20000)
    a = digits_sequential(y, S)
    b = digits_integer(y, S)
    assert np.array_equal(a, b)
    assert np.abs(a).max() <= 127
    w = 128.0 ** -np.arange(1, S + 1)
    rec = (a * w).sum(axis=-1)
    assert np.all(np.abs(y - rec) < 128.0 ** -S)
    # the remainder has y's sign (truncation toward zero)
    assert np.all((y - rec) * y >= 0)


def test_planted_rounding_slip_is_caught():
    """Rounding instead of truncating the first digit overflows int8 near 1 and
    breaks the agreement: the check above would fail on such a slip."""
    y = np.array([0.999, -0.9999, 0.5])
    q = digits_sequential(y, 7)
    slipped = q.copy()
    slipped[:, 0] = np.rint(y * 128).astype(np.int64)
    assert not np.array_equal(q, slipped)
    assert np.abs(slipped).max() > 127


@pytest.mark.parametrize("S", [5, 7])
def test_truncated_digit_product_inside_bound(S):
    """sum_k a_k b_k from row/column scales and digit pairs with t + u < S
    (exact integer sums, as on the int8 matrix cores), against the bound
    4 (S + 1) 2^(-7S) K max|a| max|b| of the GPU tests."""
    rng = np.random.default_rng(3)
    K = 200
    for _ in range(20):
        a = rng.normal(size=K) * 2.0 ** rng.integers(-30, 30)
        b = rng.normal(size=K) * 2.0 ** rng.integers(-30, 30)
        ea = np.frexp(np.abs(a).max())[1]
        eb = np.frexp(np.abs(b).max())[1]
        qa = digits_integer(np.ldexp(a, -ea), S)
        qb = digits_integer(np.ldexp(b, -eb), S)
        total = 0.0
        for d in range(S):
            acc = 0
            for t in range(d + 1):
                acc += int(np.dot(qa[:, t], qb[:, d - t]))  # exact integer digit-pair sums
            total += acc * 2.0 ** (-7 * (d + 2))
        got = np.ldexp(total, ea + eb)
        exact = float(np.dot(a.astype(np.longdouble), b.astype(np.longdouble)))
        bound = 4 * (S + 1) * 2.0 ** (-7 * S) * K * np.abs(a).max() * np.abs(b).max()
        assert abs(got - exact) <= bound
