"""CPU: the oracle's data-movement and arithmetic references, pinned by the
compiled reference (copy/transpose), IEEE identities and exact arithmetic."""
import numpy as np
import pytest

from oracle import oracle as O


def test_copy_block_matches_compiled_reference():
    r = O.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built")
    rnd = np.random.default_rng(0)
    for dt in (np.float64, np.float32):
        for nrow, ncol, lds, ldd in [(7, 5, 9, 6), (1, 1, 1, 1), (33, 17, 40, 17), (0, 4, 4, 4)]:
            src = rnd.standard_normal(max(1, nrow * lds)).astype(dt)
            a = np.zeros(max(1, nrow * ldd), dt)
            b = np.zeros_like(a)
            O.copy_block(src, nrow, ncol, lds, a, ldd)
            r.copy_matrix_block(src.itemsize, nrow, ncol, src.ctypes.data, lds, b.ctypes.data, ldd, 0)
            assert np.array_equal(a, b)


def test_transpose_matches_compiled_reference_and_is_involution():
    rnd = np.random.default_rng(1)
    r = O.ref_lib()
    for nrow, ncol, ldA, ldAT in [(5, 3, 3, 5), (40, 17, 20, 41), (1, 9, 9, 1)]:
        A = rnd.standard_normal(ldA * nrow)
        AT = np.zeros(ldAT * ncol)
        O.transpose_cm(nrow, ncol, A, ldA, AT, ldAT)
        Mat = A.reshape(nrow, ldA)[:, :ncol]            # A is ncol x nrow col-major
        assert np.array_equal(AT.reshape(ncol, ldAT)[:, :nrow], Mat.T)
        back = np.zeros(ldA * nrow)
        O.transpose_cm(ncol, nrow, AT, ldAT, back, ldA)
        assert np.array_equal(back.reshape(nrow, ldA)[:, :ncol], Mat)
        if r is not None:
            AT2 = np.zeros_like(AT)
            r.transpose_cm_mat(nrow, ncol, A.ctypes.data, ldA, AT2.ctypes.data, ldAT)
            assert np.array_equal(AT, AT2)


def test_op_sum_types():
    rnd = np.random.default_rng(2)
    for dt in (np.float64, np.float32, np.int32):
        a = (rnd.standard_normal(1000) * 100).astype(dt)
        b = (rnd.standard_normal(1000) * 100).astype(dt)
        want = b + a
        O.op_sum(a, b)
        assert np.array_equal(b, want)


def test_fold_sum_order_and_exactness():
    rnd = np.random.default_rng(3)
    P = [rnd.standard_normal((37, 23)).T.copy().ravel() for _ in range(3)]  # 23x37 col-major
    dst = np.zeros(23 * 37)
    O.fold_sum(P, 23, 37, [23] * 3, dst, 23)
    assert np.array_equal(dst, (P[0] + P[1]) + P[2])           # left-to-right fold
    # a wrong association order is visible on these inputs
    assert not np.array_equal(dst, P[0] + (P[1] + P[2]))
    # kp = 2: order-free
    d2 = np.zeros_like(dst)
    O.fold_sum([P[1], P[0]], 23, 37, [23, 23], d2, 23)
    assert np.array_equal(d2, P[0] + P[1])


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_integer_inputs_exact(ta, tb):
    """Integer-valued operands make every summation order exact
    (the idea of extra/test_cannon.c:59,68), so C must equal numpy's."""
    rnd = np.random.default_rng(4)
    m, n, k = 29, 31, 37
    A = rnd.integers(-8, 9, (k, m) if ta else (m, k)).astype(np.float64)
    B = rnd.integers(-8, 9, (n, k) if tb else (k, n)).astype(np.float64)
    Cm = rnd.integers(-8, 9, (m, n)).astype(np.float64)
    Af, Bf, Cf = (np.asfortranarray(x) for x in (A, B, Cm))
    O.dgemm(ta, tb, m, n, k, 2.0, Af, Af.shape[0], Bf, Bf.shape[0], -1.0, Cf, m)
    want = 2.0 * ((A.T if ta else A) @ (B.T if tb else B)) - Cm
    assert np.array_equal(Cf, want)


def test_dgemm_linear_fill_closed_form():
    """examples/example_AB.c:107-128 fill; C_ij has a closed form."""
    m, n, k = 40, 30, 50
    i = np.arange(m)[:, None]; l = np.arange(k)[None, :]
    A = np.asfortranarray(0.19 * i + 0.24 * l)
    l2 = np.arange(k)[:, None]; j = np.arange(n)[None, :]
    B = np.asfortranarray(0.11 * l2 + 0.12 * j)
    Cm = np.zeros((m, n), order="F")
    O.dgemm(0, 0, m, n, k, 1.0, A, m, B, k, 0.0, Cm, m)
    S1, S2 = k * (k - 1) / 2, k * (k - 1) * (2 * k - 1) / 6
    I, J = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    closed = k * 0.19 * 0.12 * I * J + 0.19 * 0.11 * I * S1 + 0.24 * 0.12 * J * S1 + 0.24 * 0.11 * S2
    assert np.max(np.abs(Cm - closed) / np.abs(closed).clip(1e-300)) < 1e-12


def test_dd_sample_exact_on_integers_and_beats_double():
    rnd = np.random.default_rng(5)
    m, n, k = 8, 9, 4000
    A = np.asfortranarray(rnd.integers(-2**20, 2**20, (m, k)).astype(np.float64))
    B = np.asfortranarray(rnd.integers(-2**20, 2**20, (k, n)).astype(np.float64))
    ii, jj = np.repeat(np.arange(m), n), np.tile(np.arange(n), m)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    exact = [sum(int(A[i, l]) * int(B[l, j]) for l in range(k)) for i, j in zip(ii, jj)]
    # products reach 2^40 and sums 2^52: a double-double holds them exactly
    assert all(float(e) == r for e, r in zip(exact, ref))
    assert np.allclose(ab, (np.abs(A) @ np.abs(B))[ii, jj], rtol=1e-14)
    # cancellation case where plain double accumulation loses everything
    a = np.asfortranarray(np.array([[1e16, 1.0, -1e16]]))
    b = np.asfortranarray(np.array([[1.0], [1.0], [1.0]]))
    r1, _ = O.dd_sample(0, 0, 1, 1, 3, a, 1, b, 3, [0], [0])
    assert r1[0] == 1.0
