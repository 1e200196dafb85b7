"""CPU: the CRT form of the Ozaki scheme (oracle/ozaki.py, ozaki2_dgemm).
No reference counterpart; pinned by exact integer arithmetic: residues
equal Python's integer remainders, Garner's digits rebuild the integer,
the line scale keeps the Cauchy-Schwarz bound that makes C' recoverable,
a brute-force Python-integer product (no residues at all) gives the same
bits, small integer inputs come out exact, and the error obeys the
rounding bound of the scaled integers."""
from math import gcd

import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z


def _bal(x, p):
    r = x % p
    hi = (p - 1) // 2 if p % 2 else p // 2 - 1
    return r - p if r > hi else r


def test_moduli_and_budget():
    ms = Z.CRT_MODULI
    assert all(2 <= p <= 256 for p in ms)
    assert all(gcd(a, b) == 1 for i, a in enumerate(ms) for b in ms[i + 1:])
    for N in range(Z.CRT_MIN, Z.CRT_MAX + 1):
        P, T = Z.crt_product(N), Z.crt_T(N)
        assert 2 ** (2 * T + 1) < P                 # |C'| <= 2^2T < P/2
        assert 2 ** (2 * T + 5) > P                 # within two bits of the largest such T
        assert 4 ** T >= Z.CRT_KMAX                 # rounding slack sqrt(k)/2 <= 2^(T-1)
        assert Z.CRT_SLACK ** 2 >= Z.CRT_KMAX > (Z.CRT_SLACK - 1) ** 2 and Z.CRT_SLACK < 2 ** T


@pytest.mark.parametrize("N", [4, 9, 16, 20, 25])
def test_residues_equal_integer_remainders(N):
    rnd = np.random.default_rng(N)
    T = Z.crt_T(N)
    vals = []
    for _ in range(400):
        bits = int(rnd.integers(0, T))
        mant = int(rnd.integers(0, 2 ** 53))
        v = (mant >> max(0, 53 - bits)) << max(0, bits - 53)          # < 2^T, 53 significant bits at most
        vals.append(-v if rnd.integers(0, 2) else v)
    vals += [0, 1, -1, 2 ** (T - 1), -(2 ** (T - 1)), 255, -256, 128, -128]
    X = np.array([float(v) for v in vals])
    assert all(float(v) == v for v in vals)
    R = Z.crt_residues(X, N)
    for t, p in enumerate(Z.CRT_MODULI[:N]):
        assert R[t].dtype == np.int8
        assert [int(r) for r in R[t]] == [_bal(v, p) for v in vals], p


@pytest.mark.parametrize("N", [4, 7, 14, 20, 25])
def test_garner_digits_rebuild_the_integer(N):
    rnd = np.random.default_rng(100 + N)
    P = Z.crt_product(N)
    xs = [int(rnd.integers(0, 2 ** 62)) * int(rnd.integers(0, 2 ** 62)) * int(rnd.integers(0, 2 ** 40)) % P - P // 2
          for _ in range(300)]
    xs += [-(P // 2), P // 2 - 1, 0, 1, -1]
    Zs = [np.array([_bal(x, p) for x in xs], dtype=np.int64) for p in Z.CRT_MODULI[:N]]
    v = Z.crt_garner(Zs, N)
    Qs = Z.crt_composites(N)
    prodQ = 1
    for Q in Qs:
        prodQ *= Q
    assert len(v) == (N + 1) // 2 and prodQ == P
    for idx, x in enumerate(xs):
        acc, M = 0, 1
        for i, Q in enumerate(Qs):
            d = int(v[i][idx])
            assert -(Q // 2) <= d <= (Q - 1) // 2
            acc += d * M
            M *= Q
        assert acc == x


@pytest.mark.parametrize("N", [4, 12, 16, 20, 25])
@pytest.mark.parametrize("phi", [0.5, 4.0])
def test_line_scale_keeps_the_norm_bound(N, phi):
    rnd = np.random.default_rng(int(phi * 10) + N)
    X = Z.phi_matrix(rnd, (40, 300), phi)
    X[3] = 0.0
    X[5, :] = 0.0
    X[5, 7] = -3.0                               # single-element line
    X[6, :4] = [1e-300, 2.0 ** -1000, 1.0, 0.75]
    X[7] *= 2.0 ** 600
    T = Z.crt_T(N)
    Ai, e, zero = Z.crt_scaled(X, 1, N)
    assert zero[3] and not zero[5]
    for i in range(X.shape[0]):
        if zero[i]:
            assert np.all(Ai[i] == 0)
            continue
        ints = [int(v) for v in Ai[i]]
        assert all(float(v) == w for v, w in zip(ints, Ai[i]))
        assert sum(v * v for v in ints) <= 4 ** T       # ||a'||_2 <= 2^T
        # not wasteful: the scaled line reaches at least 2^(T-2) in norm
        nrm = np.sqrt(np.sum(np.ldexp(X[i], int(e[i])) ** 2))
        assert nrm > 2.0 ** (T - 2)
    # columns of op(B) follow the same rule as rows of op(B)^T
    Bi, eb, zb = Z.crt_scaled(X.T.copy(), 0, N)
    assert np.array_equal(eb[~zb], e[~zero]) and np.array_equal(Bi.T, Ai)


def _bruteforce(opA, opB, N, alpha, beta, Cin):
    """No residues: exact Python-integer dot products, then the balanced
    mixed-radix digits of C' itself over the pair products p_0 p_1,
    p_2 p_3, ... (p_N-1 alone last for odd N), then the same Horner sum."""
    Ai, eA, zA = Z.crt_scaled(opA, 1, N)
    Bi, eB, zB = Z.crt_scaled(opB, 0, N)
    m, k = opA.shape
    n = opB.shape[1]
    ps = Z.CRT_MODULI[:N]
    Qs = [ps[t] * (ps[t + 1] if t + 1 < N else 1) for t in range(0, N, 2)]
    C = np.zeros((m, n))
    for i in range(m):
        for j in range(n):
            x = sum(int(Ai[i, l]) * int(Bi[l, j]) for l in range(k))
            assert abs(x) < Z.crt_product(N) // 2
            digits = []
            for Q in Qs:
                d = _bal(x, Q)
                digits.append(d)
                x = (x - d) // Q
            assert x == 0
            acc = float(digits[-1])
            for t in range(len(Qs) - 2, -1, -1):
                acc = acc * float(Qs[t])
                acc = acc + float(digits[t])
            v = 0.0 if (zA[i] or zB[j]) else np.ldexp(np.ldexp(acc, -int(eA[i])), -int(eB[j]))
            v = alpha * v
            if beta != 0.0:
                v = v + beta * Cin[i, j]
            C[i, j] = v
    return C


@pytest.mark.parametrize("N", [4, 9, 16, 20, 25])
def test_matches_bruteforce_integers(N):
    rnd = np.random.default_rng(7 * N)
    opA = Z.phi_matrix(rnd, (5, 9), 2.0)
    opB = Z.phi_matrix(rnd, (9, 7), 2.0)
    opA[1] = 0.0
    opB[:, 2] = 0.0
    Cin = rnd.standard_normal((5, 7))
    got = Z.ozaki2_dgemm(opA, opB, N, alpha=-0.75, beta=0.5, Cin=Cin)
    assert np.array_equal(got, _bruteforce(opA, opB, N, -0.75, 0.5, Cin))


def test_small_integer_inputs_are_exact():
    rnd = np.random.default_rng(3)
    A = rnd.integers(-1024, 1025, (33, 64)).astype(np.float64)
    B = rnd.integers(-1024, 1025, (64, 21)).astype(np.float64)
    assert np.array_equal(Z.ozaki2_dgemm(A, B, 6), A @ B)


@pytest.mark.parametrize("N", [8, 12, 16])
def test_error_within_rounding_bound(N):
    """|C - C_exact| <= 2^-(eA+fB) (||b'||_1/2 + ||a'||_1/2 + 3k/4) + 2N u |C|:
    a' = 2^e a + r with |r| <= 1/2, so |a'b' - 2^(e+f) ab| <= |r_a||b'| +
    |a'||r_b| + |r_a r_b| summed over k, plus the Horner sum's roundings."""
    rnd = np.random.default_rng(N)
    m, k, n = 24, 500, 20
    opA, opB = Z.phi_matrix(rnd, (m, k), 2.0), Z.phi_matrix(rnd, (k, n), 2.0)
    got = Z.ozaki2_dgemm(opA, opB, N)
    Ai, eA, _ = Z.crt_scaled(opA, 1, N)
    Bi, eB, _ = Z.crt_scaled(opB, 0, N)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ref, _ = O.dd_sample(0, 0, m, n, k, np.asfortranarray(opA), m, np.asfortranarray(opB), k,
                         ii.ravel(), jj.ravel())
    ref = ref.reshape(m, n)
    bound = (np.ldexp(np.ldexp(0.5 * np.sum(np.abs(Bi), axis=0)[None, :] + 0.5 * np.sum(np.abs(Ai), axis=1)[:, None]
                               + 3 * k / 4, -eA[:, None]), -eB[None, :])
             + 2 * N * 2.0 ** -53 * np.abs(ref))
    assert np.all(np.abs(got - ref) <= bound * (1 + 1e-6) + 1e-300)


def test_error_falls_with_moduli():
    rnd = np.random.default_rng(11)
    m, k, n = 32, 2000, 32
    opA, opB = Z.phi_matrix(rnd, (m, k), 0.5), Z.phi_matrix(rnd, (k, n), 0.5)
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    ref, ab = O.dd_sample(0, 0, m, n, k, np.asfortranarray(opA), m, np.asfortranarray(opB), k,
                          ii.ravel(), jj.ravel())
    errs = []
    for N in (6, 8, 10, 12, 14):
        C = Z.ozaki2_dgemm(opA, opB, N).ravel()
        errs.append(np.max(np.abs(C - ref) / ab))
    # each extra modulus adds ~8 bits to P, ~4 bits to each operand's scale
    for a, b in zip(errs, errs[1:]):
        assert b < a / 100
    assert errs[-1] < 1e-14


@pytest.mark.parametrize("T", [15, 31, 54, 58, 80, 95])
def test_scale_rule_is_the_tightest_power_of_two(T):
    """h is minimal, and the bound it certifies, 4^(T-h) S <= (2^T - 2897/2)^2,
    holds exactly: e = T + 20 - E - h is the largest exponent the rule admits."""
    rnd = np.random.default_rng(T)
    for _ in range(3000):
        S = max(1, int(rnd.integers(1, 2 ** 62)) >> int(rnd.integers(0, 62)))
        h = Z.crt_scale_h(S, T)
        assert 4 * 4 ** (T - h) * S <= (2 ** (T + 1) - Z.CRT_SLACK) ** 2
        if h > 0:                                       # one less fails the rule
            sh = 2 * (h - 1) - T
            sub = Z.CRT_SLACK << sh if sh >= 0 else -((-Z.CRT_SLACK) >> (-sh))
            assert S > 4 ** (h - 1) - sub


def _fma_magic_low_word(x, p):
    """Low 32 bits of fl(x * fl(1/p) + 1.5 2^52), FP64 round to nearest even:
    the sum lies in [2^52, 2^53), where doubles are the integers."""
    from fractions import Fraction
    y = (3 << 51) + x * Fraction(1.0 / p)
    f = y.numerator // y.denominator
    rem = y - f
    t = f + (1 if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and f % 2) else 0)
    return (t - (3 << 51)) & 0xFFFFFFFF


def _s24(v):
    v &= 0xFFFFFF
    return v - (1 << 24) if v & 0x800000 else v


def test_product_epilogue_reduction_low_byte():
    """The CRT product's epilogue reduces an INT32 accumulator x modulo an odd
    p as x + mul24(low word of fma(x, 1/p, 1.5 2^52), -p) and keeps the low
    byte.  That byte must be the balanced residue's for every INT32 x: checked
    at the extremes and at the values closest to a half-integer quotient,
    where a wrong rounding would show, and at random values."""
    rng = np.random.default_rng(5)
    for p in [q for q in Z.CRT_MODULI if q % 2]:
        xs = [-(1 << 31), (1 << 31) - 1, 0, 1, -1, p // 2, -(p // 2)]
        for c in rng.integers(-(1 << 31) // p, (1 << 31) // p, 60):
            for dlt in ((p - 1) // 2, -(p - 1) // 2, (p + 1) // 2):
                x = int(c) * p + dlt
                if -(1 << 31) <= x < (1 << 31):
                    xs.append(x)
        xs += [int(v) for v in rng.integers(-(1 << 31), 1 << 31, 60)]
        for x in xs:
            q = _fma_magic_low_word(x, p)
            got = (x + _s24(q) * _s24(-p)) & 0xFF
            assert got == _bal(x, p) & 0xFF, (p, x)
