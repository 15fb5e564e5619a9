"""CPU checks of the digit split behind the int8x GEMM (sinv_amd/csrc/int8x.hip).

Each row (column) gets a power-of-two scale 2^e with |x| / 2^e <= 127/128, and
y = x / 2^e becomes the integer N = rint(y 2^(8S-1)) (ties to even),
|N| <= 127 * 2^(8S-8), written with S balanced base-256 digits: S - 1 low
digits in [-128, 127] and a top digit in [-127, 127]; digit t weighs
2^(-7-8t).  The kernel reads the digits off the bytes of m = N + B_S
(B_S = 128 in each of the S - 1 low bytes: a biased low digit is its byte, the
top digit is the top byte), and for S <= 6 forms m with one floating-point
addition, y 2^(8S-1) + (1.5 2^52 + B_S), whose low 48 bits are m mod 2^48.
These tests restate both steps, check them against two independent
constructions of the balanced digits (peeling from the low end, and top-down
with Python integers), that every digit fits int8, that the digits reconstruct
y to within 2^(-8S), and that the digit-pair product with pairs of index sum
>= S dropped stays inside the bound the GPU tests use.
"""

import numpy as np
import pytest


def scale_exp(mx):
    """Exponent e of the scale 2^e of a row whose largest magnitude is mx."""
    f, e = np.frexp(mx)
    return e + 1 if f > 127 / 128 else e


def top_integer(y, S):
    return np.rint(np.ldexp(y, 8 * S - 1)).astype(np.int64)


def bias(S):
    """128 in each of the S - 1 low bytes."""
    return sum(128 << (8 * t) for t in range(S - 1))


def digits_bytes(y, S):
    """The kernel's method: the bytes of m = N + B_S (byte S - 1 - t is digit
    t, minus 128 for the biased low digits)."""
    m = top_integer(y, S) + bias(S)
    q = np.zeros(y.shape + (S,), dtype=np.int64)
    for t in range(S):
        byte = (m >> (8 * (S - 1 - t))) & 255
        q[..., t] = byte - 128 if t > 0 else np.where(byte >= 128, byte - 256, byte)
    return q


def magic_low_bits(y, S):
    """For S <= 6: the low 48 bits of y 2^(8S-1) + (1.5 2^52 + B_S) as a double."""
    magic = 1.5 * 2.0 ** 52 + bias(S)
    t = np.ldexp(y, 8 * S - 1) + magic
    return t.view(np.int64) & ((1 << 48) - 1)


def digits_peel(y, S):
    """The kernel's order: low byte balanced into [-128, 127], then shift."""
    n = top_integer(y, S)
    q = np.zeros(y.shape + (S,), dtype=np.int64)
    for t in range(S - 1, 0, -1):
        r = n & 255
        d = np.where(r >= 128, r - 256, r)
        q[..., t] = d
        n = (n - d) >> 8
    q[..., 0] = n
    return q


def digits_topdown(y, S):
    """Independent construction: the top digit first, as the smallest value
    that leaves a remainder the S - 1 lower balanced digits can express."""
    out = np.zeros(y.shape + (S,), dtype=np.int64)
    for idx, v in np.ndenumerate(y):
        n = int(np.rint(np.ldexp(v, 8 * S - 1)))
        for t in range(S):
            w = 256 ** (S - 1 - t)
            lo_max = 127 * (w - 1) // 255      # largest value of the lower digits
            d = -((lo_max - n) // w)            # ceil((n - lo_max) / w)
            out[idx + (t,)] = d
            n -= d * w
        assert n == 0
    return out


def _samples(rng, n):
    y = rng.uniform(-127 / 128, 127 / 128, n)
    # magnitudes across the whole exponent range below the bound, byte
    # boundaries of the integer, and signed zeros
    y = np.concatenate([y, rng.uniform(-1, 1, n) * 2.0 ** rng.integers(-60, -1, n),
                        np.array([0.0, -0.0, 127 / 128, -127 / 128, 2 ** -7, -2 ** -7,
                                  0.5, -0.5, 2 ** -47, -2 ** -47, 2 ** -48, 1.5 * 2 ** -8,
                                  -(1.5 * 2 ** -8), 127.5 / 256, -127.5 / 256])])
    return y


@pytest.mark.parametrize("S", [5, 6, 7])
def test_byte_digits_equal_peeled_and_topdown_digits(S):
    y = _samples(np.random.default_rng(S), 4000)
    a = digits_bytes(y, S)
    assert np.array_equal(a, digits_peel(y, S))
    assert np.array_equal(a, digits_topdown(y, S))
    assert a[..., 1:].min() >= -128 and a[..., 1:].max() <= 127
    assert np.abs(a[..., 0]).max() <= 127
    w = 2.0 ** (-7 - 8 * np.arange(S))
    rec = (a * w).sum(axis=-1)
    # rint: the remainder is at most half of the last digit's weight
    assert np.all(np.abs(y - rec) <= 2.0 ** -(8 * S))


@pytest.mark.parametrize("S", [5, 6])
def test_magic_addition_gives_the_biased_integer(S):
    """One rounded double addition yields m = N + B_S in its low 48 bits,
    ties to even like np.rint (B_S is even), across the whole range."""
    y = _samples(np.random.default_rng(10 + S), 4000)
    # exact ties of the integer rounding: half-way points between integers
    half = (np.arange(-40, 40) + 0.5) * 2.0 ** -(8 * S - 1)
    y = np.concatenate([y, half])
    m = (top_integer(y, S) + bias(S)) & ((1 << 48) - 1)
    assert np.array_equal(magic_low_bits(y, S), m)
    # planted: truncating instead of rounding differs on the ties
    trunc = (np.trunc(np.ldexp(half, 8 * S - 1)).astype(np.int64) + bias(S)) & ((1 << 48) - 1)
    assert not np.array_equal(magic_low_bits(half, S), trunc)


def test_scale_keeps_the_top_digit_in_int8():
    """Without the 127/128 rule a row maximum just below a power of two would
    give a top digit of 128: the rule moves such rows to the next exponent."""
    for mx in (1.0 - 2 ** -53, 0.999, 127.5 / 128, 0.75, 2.0 ** -30 * 0.9999):
        e = scale_exp(mx)
        y = np.array([mx, -mx]) / 2.0 ** e
        assert np.all(np.abs(y) <= 127 / 128)
        q = digits_peel(y, 6)
        assert np.abs(q).max() <= 127
    # planted: the frexp exponent alone overflows the top digit near 1
    y = np.array([1.0 - 2 ** -53])
    assert digits_peel(y, 6)[0, 0] == 128


def test_planted_unbalanced_digits_are_caught():
    """Plain unsigned bytes (no balancing) reconstruct the same integer but
    leave digits above 127, which the range check rejects."""
    y = np.array([0.49, 0.99 * 127 / 128, 0.3 + 2 ** -30])
    n = top_integer(np.abs(y), 6)
    unbalanced = np.stack([(n >> (8 * (5 - t))) & 255 for t in range(6)], axis=-1)
    assert unbalanced.max() > 127
    assert not np.array_equal(unbalanced, digits_peel(np.abs(y), 6))


@pytest.mark.parametrize("S", [5, 6, 7])
def test_truncated_digit_product_inside_bound(S):
    """sum_k a_k b_k from row/column scales and digit pairs with t + u < S
    (exact integer sums, as on the int8 matrix cores), against the bound
    4.1 (S + 4) 2^(-8S) K max|a| max|b| of the GPU tests."""
    rng = np.random.default_rng(3)
    K = 200
    worst = 0.0
    for _ in range(20):
        a = rng.normal(size=K) * 2.0 ** rng.integers(-30, 30)
        b = rng.normal(size=K) * 2.0 ** rng.integers(-30, 30)
        ea = scale_exp(np.abs(a).max())
        eb = scale_exp(np.abs(b).max())
        qa = digits_peel(np.ldexp(a, -ea), S)
        qb = digits_peel(np.ldexp(b, -eb), S)
        total = 0.0
        for s in range(S):
            acc = 0
            for t in range(s + 1):
                acc += int(np.dot(qa[:, t], qb[:, s - t]))  # exact integer digit-pair sums
            assert abs(acc) < 2 ** 31
            total += acc * 2.0 ** (-14 - 8 * s)
        got = np.ldexp(total, ea + eb)
        exact = float(np.dot(a.astype(np.longdouble), b.astype(np.longdouble)))
        bound = 4.1 * (S + 4) * 2.0 ** (-8 * S) * K * np.abs(a).max() * np.abs(b).max()
        assert abs(got - exact) <= bound
        worst = max(worst, abs(got - exact) / bound)
    assert worst > 0  # the check measures something
