"""CPU: the Ozaki-scheme specification (oracle/ozaki.py).  The scheme has no
reference counterpart; these tests pin it by exact arithmetic: the digits
are exactly the balanced base-256 expansion of rint(x' * 2^(8s)), slice
products are exact integers, and the error falls ~2^-8 per slice."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z


def _rint_exact(v, s):
    x = Fraction(float(v)) * 2 ** (8 * s)
    fl = x.numerator // x.denominator
    r = x - fl
    return fl + (1 if r > Fraction(1, 2) or (r == Fraction(1, 2) and fl % 2 == 1) else 0)


@pytest.mark.parametrize("s", [1, 2, 5, 8, 9, 12, 16])
def test_digits_are_exact_balanced_expansion(s):
    rnd = np.random.default_rng(s)
    X = Z.phi_matrix(rnd, (14, 40), 4.0)
    X[2] = 0.0
    X[4, :6] = [0.24999999999999997, -0.2, 1e-300, 5e-324, -7.0, 3.0]
    # lines whose max sits at the threshold of the scale rule, on both sides,
    # and lines of values whose N lands on the top digit's limits
    X[5, :3] = [253 / 256, -0.5, 0.25]
    X[6, :3] = [-np.nextafter(253 / 256, 1), 0.5, 1e-3]
    X[7, :4] = [-(253 / 256), 253 / 256, -np.nextafter(253 / 256, 0), 0.9999999999999999]
    S, E, zero = Z.split8(X, s, axis=1)
    assert zero[2] and not zero[4]
    a = np.ldexp(X, -E.reshape(-1, 1))
    a[zero] = 0
    assert np.all(np.abs(a) <= 253 / 512)
    for (i, j), v in np.ndenumerate(a):
        digits = [int(S[t][i, j]) for t in range(s)]
        assert all(-128 <= d <= 127 for d in digits)
        assert sum(d * 256 ** (s - 1 - t) for t, d in enumerate(digits)) == _rint_exact(v, s)


def test_column_split_is_row_split_of_transpose():
    rnd = np.random.default_rng(1)
    B = Z.phi_matrix(rnd, (30, 20), 2.0)
    S1, E1, _ = Z.split8(B, 6, axis=0)
    S2, E2, _ = Z.split8(B.T.copy(), 6, axis=1)
    assert np.array_equal(E1, E2)
    assert all(np.array_equal(a, b.T) for a, b in zip(S1, S2))


def test_integer_inputs_are_exact_with_enough_slices():
    """Small integers need few digits: with s covering them the scheme is exact."""
    rnd = np.random.default_rng(2)
    A = rnd.integers(-1000, 1001, (40, 300)).astype(np.float64)
    B = rnd.integers(-1000, 1001, (300, 30)).astype(np.float64)
    assert np.array_equal(Z.ozaki_dgemm8(A, B, 3), A @ B)


def test_error_falls_about_eight_bits_per_slice():
    rnd = np.random.default_rng(3)
    m, n, k = 48, 40, 2048
    A, B = Z.phi_matrix(rnd, (m, k), 0.5), Z.phi_matrix(rnd, (k, n), 0.5)
    Af, Bf = np.asfortranarray(A), np.asfortranarray(B)
    ii, jj = np.repeat(np.arange(m), n), np.tile(np.arange(n), m)
    ref, ab = O.dd_sample(0, 0, m, n, k, Af, m, Bf, k, ii, jj)
    errs = [np.max(np.abs(Z.ozaki_dgemm8(A, B, s)[ii, jj] - ref) / ab) for s in range(2, 8)]
    for e0, e1 in zip(errs, errs[1:]):
        assert e1 < e0 * 2.0 ** -5          # at least 5 of the 8 bits per slice
    assert errs[-1] < 2e-15


def test_zero_lines_give_zero():
    A = np.ones((5, 7)); A[1] = 0.0
    B = np.ones((7, 4)); B[:, 2] = 0.0
    C = Z.ozaki_dgemm8(A, B, 2)
    assert np.all(C[1] == 0.0) and np.all(C[:, 2] == 0.0) and C[0, 0] == 7.0


def test_k_chunking_rule():
    assert Z.kchunk8(8192, 8) == 8192
    assert Z.kchunk8(20000, 16) == 7936          # floor(8191.99 / 256) * 256
    assert Z.kchunk8(1, 3) == 256
    assert Z.kchunk8(100000, 8) == 16128
    for s in range(1, 17):
        kc = Z.kchunk8(10 ** 7, s)
        assert s * kc * 128 * 128 < 2 ** 31 and kc % 256 == 0
        assert s * (kc + 256) * 128 * 128 >= 2 ** 31     # the largest such multiple


def _bias_window_digits(v, E, s):
    """The device's digit rule (ozaki_slice_kernel): digit i of the balanced
    expansion of X is byte i of U = X + 0x8080...80 (mod 2^64) xor 0x80.
    Restated here with Python integers, independently of split8's carries."""
    a = float(np.ldexp(v, -int(E)))
    if a == 0.0:
        return [0] * s
    f, ex = np.frexp(a)
    M = int(np.ldexp(f, 53))
    z = int(ex) - 53 + 8 * s
    if z < 0:
        X, off = int(np.rint(np.ldexp(float(M), z))), 0
    else:
        X, off = M << (z % 8), z // 8
    U = (X + 0x8080808080808080) % 2 ** 64
    out = []
    for t in range(s):
        i = s - 1 - t - off
        b = ((U >> (8 * i)) & 255) ^ 0x80 if 0 <= i < 8 else 0
        out.append(b - 256 if b >= 128 else b)
    return out


@pytest.mark.parametrize("s", [1, 3, 8, 11, 16])
def test_device_bias_window_rule_equals_split8(s):
    rnd = np.random.default_rng(100 + s)
    X = Z.phi_matrix(rnd, (10, 30), 4.0)
    X[1, :8] = [0.24999999999999997, -0.2, 1e-300, 5e-324, -7.0, 3.0, -0.0, 2.0 ** -40]
    S, E, zero = Z.split8(X, s, axis=1)
    for (i, j), v in np.ndenumerate(X):
        want = [int(S[t][i, j]) for t in range(s)]
        got = [0] * s if zero[i] else _bias_window_digits(v, E[i], s)
        assert got == want, (i, j, v)


def test_scale_rule_threshold():
    """E = e + 1 up to a line-max mantissa of 253/256, e + 2 above it; the
    digits of lines at the threshold stay within int8."""
    X = np.array([[253 / 256, 0.1], [np.nextafter(253 / 256, 1), 0.1], [0.5, -0.5], [-1.0, 0.0],
                  [3.0, 1.0], [0.0, 0.0]])
    E, zero = Z.line_scale8(X, 1)
    assert list(E[:5]) == [1, 2, 1, 2, 3] and list(zero) == [False] * 5 + [True]
    # a' = 253/512 exactly: N = 126.5 * 256^(s-1), i.e. 126 (ties to even)
    # for s = 1 and the balanced 127 * 256 - 128 for s >= 2
    for s, want in ((1, [126]), (2, [127, -128]), (3, [127, -128, 0])):
        S, E2, _ = Z.split8(X, s, 1)
        assert np.array_equal(E2, E)
        assert [int(q[0, 0]) for q in S] == want
