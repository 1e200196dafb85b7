"""GPU: the B^T B Gram kernel (upper triangle, split-K) and the CholeskyQR
pieces (potrf, trtri, trsm) against numpy/scipy and exact arithmetic, and
CholeskyQR2 end to end (examples/example_CholeskyQR.c, two passes)."""
import numpy as np
import pytest
import scipy.linalg as sla

from ca3dmm_amd import cholqr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k", [(1, 1), (5, 3), (128, 128), (130, 1000), (300, 20000), (512, 70000)])
def test_gram_integer_exact_and_symmetric(hip, n, k):
    rnd = np.random.default_rng(n + k)
    A = rnd.integers(-3, 4, (k, n)).astype(np.float64)
    dA = hip.DeviceArray.from_numpy(A)
    dC = hip.DeviceArray.filled((n, n), 0xFF)
    hip.gram(n, k, dA, dC, mirror=True)
    C = dC.numpy()
    assert np.array_equal(C, A.T @ A)
    assert np.array_equal(C, C.T)


class _At:
    """A device address inside a DeviceArray, for hip.gram's A."""
    def __init__(self, ptr):
        self.ptr = ptr


@pytest.mark.parametrize("n,k,pad,off", [(130, 1001, 0, 0), (257, 3000, 1, 0), (96, 2048, 0, 1), (300, 777, 2, 3)])
def test_gram_odd_leading_dimension_and_offset(hip, monkeypatch, n, k, pad, off):
    """A with an odd leading dimension or starting an odd number of doubles
    into its buffer: 16-byte direct-to-LDS loads from 8-byte aligned
    addresses give the register-staged kernel's bits, exact on integers."""
    rnd = np.random.default_rng(n + k + pad + off)
    lda = k + pad
    for kind in ("f", "int"):
        A = rnd.standard_normal((k, n)) if kind == "f" else rnd.integers(-3, 4, (k, n)).astype(np.float64)
        buf = np.zeros(off + lda * n)
        for c in range(n):
            buf[off + c * lda:off + c * lda + k] = A[:, c]
        dA = hip.DeviceArray.from_numpy(buf)
        got, ref = hip.DeviceArray.filled((n, n), 0xFF), hip.DeviceArray.filled((n, n), 0xFF)
        monkeypatch.setenv("CA3DMM_DGEMM_LOADS", "reg")
        hip.gram(n, k, _At(dA.at(off)), ref, lda=lda)
        monkeypatch.delenv("CA3DMM_DGEMM_LOADS", raising=False)
        hip.gram(n, k, _At(dA.at(off)), got, lda=lda)
        assert np.array_equal(got.numpy(), ref.numpy()), (n, k, lda, off, kind)
        if kind == "int":
            assert np.array_equal(got.numpy(), A.T @ A)


def test_gram_upper_only_and_beta(hip):
    rnd = np.random.default_rng(3)
    n, k = 200, 3000
    A = rnd.standard_normal((k, n))
    C0 = rnd.standard_normal((n, n))
    dA, dC = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(C0)
    hip.gram(n, k, dA, dC, alpha=2.0, beta=0.5, mirror=False)
    C = dC.numpy()
    want = 2.0 * (A.T @ A) + 0.5 * C0
    iu = np.triu_indices(n)
    il = np.tril_indices(n, -1)
    assert np.allclose(C[iu], want[iu], rtol=1e-13, atol=1e-10)
    assert np.array_equal(C[il], C0[il])            # lower triangle untouched


def test_potrf_matches_scipy_and_reports_info(hip):
    rnd = np.random.default_rng(4)
    n = 300
    X = rnd.standard_normal((2 * n, n))
    S = X.T @ X
    dS = hip.DeviceArray.from_numpy(S)
    assert hip.potrf_upper(n, dS) == 0
    R = np.triu(dS.numpy())
    assert np.allclose(R, sla.cholesky(S, lower=False), rtol=1e-10, atol=1e-12)
    assert np.linalg.norm(R.T @ R - S) <= 1e-13 * np.linalg.norm(S)
    # leading minor of order 3 not positive definite -> info = 3 (LAPACK)
    B = np.eye(6) * 4.0
    B[2, 2] = -1.0
    dB = hip.DeviceArray.from_numpy(B)
    assert hip.potrf_upper(6, dB) == 3


def test_potrf_blocked_info_and_lower_part(hip):
    """The blocked factorization (64-wide block columns): the strictly lower
    part of A is not touched, R matches scipy across several blocks, and a
    pivot that fails in a later block reports LAPACK's info."""
    rnd = np.random.default_rng(14)
    for n in (64, 65, 200, 513):
        X = rnd.standard_normal((2 * n, n))
        S = X.T @ X
        L0 = rnd.standard_normal((n, n))
        A = np.triu(S) + np.tril(L0, -1)            # junk below the diagonal
        dA = hip.DeviceArray.from_numpy(A)
        assert hip.potrf_upper(n, dA) == 0
        got = dA.numpy()
        assert np.array_equal(np.tril(got, -1), np.tril(L0, -1)), n
        R = np.triu(got)
        assert np.allclose(R, sla.cholesky(S, lower=False), rtol=1e-10, atol=1e-12), n
    n = 300
    X = rnd.standard_normal((2 * n, n))
    S = X.T @ X
    S[199, :] = 0.0
    S[:, 199] = 0.0                                 # leading minor of order 200 singular
    dS = hip.DeviceArray.from_numpy(S)
    assert hip.potrf_upper(n, dS) == 200


def _chol_factor(rnd, n):
    """Upper R of a well-conditioned SPD matrix (kappa(R) of a few units)."""
    X = rnd.standard_normal((4 * n, n))
    return sla.cholesky(X.T @ X / (4 * n) + 0.5 * np.eye(n), lower=False)


def test_trtri_and_trsm(hip):
    rnd = np.random.default_rng(5)
    n, m = 257, 5000
    R = _chol_factor(rnd, n)
    dR = hip.DeviceArray.from_numpy(R)
    dX = hip.DeviceArray.filled((n, n), 0xFF)
    hip.trtri_upper(n, dR, dX)
    X = dX.numpy()
    assert np.array_equal(np.tril(X, -1), np.zeros((n, n)))
    assert np.linalg.norm(R @ X - np.eye(n)) < 1e-12
    B = rnd.standard_normal((m, n))
    dB = hip.DeviceArray.from_numpy(B)
    hip.trsm_right_upper(m, n, 2.0, dR, dB)
    want = 2.0 * sla.solve_triangular(R, B.T, trans="T", lower=False).T
    assert np.allclose(dB.numpy(), want, rtol=1e-11, atol=1e-12)


def test_trtri_ill_conditioned_within_kappa_bound(hip):
    """triu(N(0,1)) + 4 I at n = 257 has kappa ~ 7e5: the inverse is held to
    ||R X - I||_F <= n * u * kappa(R) * 10 (scipy's own inverse sits at ~6e-11)."""
    rnd = np.random.default_rng(5)
    n = 257
    R = np.triu(rnd.standard_normal((n, n))) + 4.0 * np.eye(n)
    dR, dX = hip.DeviceArray.from_numpy(R), hip.DeviceArray.zeros((n, n))
    hip.trtri_upper(n, dR, dX)
    bound = 10 * n * 2.0 ** -53 * np.linalg.cond(R)
    assert np.linalg.norm(R @ dX.numpy() - np.eye(n)) <= bound


@pytest.mark.parametrize("gram_slices", [None, 8])
@pytest.mark.parametrize("cond", [1e2, 1e5])
def test_cholesky_qr2(hip, cond, gram_slices):
    m, n = 40000, 128
    dA, _ = cholqr.test_matrix(m, n, cond, seed=11)
    dQ, dR = cholqr.cholesky_qr2(dA, m, n, gram_slices=gram_slices)
    orth, res = cholqr.qr_errors(dA, dQ, dR, m, n)
    assert orth < 1e-13, orth
    assert res < 1e-14, res
    R = dR.numpy()
    assert np.array_equal(np.tril(R, -1), np.zeros((n, n)))


def test_cholesky_breakdown_is_reported(hip):
    """A zero column makes pivot 6 of the first Gram exactly zero: the
    breakdown surfaces as an error with LAPACK's info, as the reference
    example prints it (examples/example_CholeskyQR.c:125-129)."""
    m, n = 1000, 16
    rnd = np.random.default_rng(9)
    A = rnd.standard_normal((m, n))
    A[:, 5] = 0.0
    dA = hip.DeviceArray.from_numpy(A)
    with pytest.raises(cholqr.CholeskyBreakdown) as ex:
        cholqr.cholesky_qr2(dA, m, n)
    assert ex.value.info == 6 and ex.value.step == 1


@pytest.mark.parametrize("variant", ["scholqr3", "auto"])
def test_shifted_cholesky_qr3_beyond_cholqr2(hip, variant):
    """kappa = 1e12 is beyond CholeskyQR2 (its Gram has kappa 1e24 > 1/u):
    plain CholeskyQR2 breaks down or loses orthogonality, while shifted
    CholeskyQR3 -- asked for, or chosen by the automatic variant from
    diag(R) -- keeps ||Q^T Q - I||_F and the residual at working precision."""
    m, n = 100000, 128
    dA, _ = cholqr.test_matrix(m, n, 1e12, seed=12)
    try:
        dQ, dR = cholqr.cholesky_qr2(dA, m, n)
        orth2, _ = cholqr.qr_errors(dA, dQ, dR, m, n)
        assert orth2 > 1e-10, orth2
    except cholqr.CholeskyBreakdown:
        pass
    v = hip.SCHOLQR3 if variant == "scholqr3" else hip.CHOLQR_AUTO
    dQ, dR, passes = cholqr.cholesky_qr(dA, m, n, v)
    orth, res = cholqr.qr_errors(dA, dQ, dR, m, n)
    assert passes >= 3, passes
    assert orth < 1e-13, orth
    assert res < 1e-14, res
    R = dR.numpy()
    assert np.array_equal(np.tril(R, -1), np.zeros((n, n)))
    assert np.all(np.diag(R) > 0)


def test_cholesky_qr_auto_keeps_cholqr2_when_well_conditioned(hip):
    m, n = 40000, 128
    dA, _ = cholqr.test_matrix(m, n, 1e3, seed=13)
    dQ, dR, passes = cholqr.cholesky_qr(dA, m, n, hip.CHOLQR_AUTO)
    assert passes == 2
    dQ2, dR2 = cholqr.cholesky_qr2(dA, m, n)
    assert np.array_equal(dQ.numpy(), dQ2.numpy())
    assert np.array_equal(dR.numpy(), dR2.numpy())


def test_cholesky_qr_rejects_bad_arguments(hip):
    from ca3dmm_amd import HipError
    dA = hip.DeviceArray.zeros((10, 20))
    dR = hip.DeviceArray.zeros((20, 20))
    with pytest.raises(HipError):
        hip.cholesky_qr(10, 20, dA, dR)            # m < n
    with pytest.raises(HipError):
        hip.cholesky_qr(20, 10, dA, dR, variant=5, lda=20)


def test_cholesky_qr_in_place_equals_out_of_place(hip, monkeypatch):
    """Passes written out of place through the m x n buffer (default) and
    solved in place (CA3DMM_CHOLQR_INPLACE=1) give the same bits, for an even
    (CholeskyQR2) and an odd (shifted CholeskyQR3) number of passes."""
    m, n = 30000, 200
    for variant in (hip.CHOLQR2, hip.SCHOLQR3):
        dA, _ = cholqr.test_matrix(m, n, 1e6, seed=21)
        monkeypatch.delenv("CA3DMM_CHOLQR_INPLACE", raising=False)
        dQ1, dR1, p1 = cholqr.cholesky_qr(dA, m, n, variant)
        monkeypatch.setenv("CA3DMM_CHOLQR_INPLACE", "1")
        dQ2, dR2, p2 = cholqr.cholesky_qr(dA, m, n, variant)
        assert p1 == p2
        assert np.array_equal(dQ1.numpy(), dQ2.numpy()), variant
        assert np.array_equal(dR1.numpy(), dR2.numpy()), variant



def test_cholesky_qr_random_shapes_and_padding(hip):
    """Seeded random m x n (n not a multiple of the 64-wide Cholesky blocks
    or of the 128-column solve blocks) with padded leading dimensions of A
    and R, each variant: Q orthonormal and A = Q R to working precision, R
    upper triangular, and the padding rows of both buffers untouched."""
    rnd = np.random.default_rng(1618)
    for case in range(9):
        n = int(rnd.integers(1, 300))
        m = int(rnd.integers(n, 20000))
        variant = (hip.CHOLQR2, hip.SCHOLQR3, hip.CHOLQR_AUTO)[case % 3]
        U, _ = np.linalg.qr(rnd.standard_normal((m, n)))
        V, _ = np.linalg.qr(rnd.standard_normal((n, n)))
        A = (U * np.logspace(0, -4, n)) @ V.T                 # kappa = 1e4
        lda, ldr = m + int(rnd.integers(1, 4)), n + int(rnd.integers(1, 4))
        bufA = np.full((lda, n), 7.0, order="F")
        bufA[:m] = A
        bufR = np.full((ldr, n), 7.0, order="F")
        dA, dR = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(bufR)
        print(f"[cholqr] random case {case}: m={m} n={n} variant={variant}", flush=True)
        info, passes = hip.cholesky_qr(m, n, dA, dR, variant=variant, lda=lda, ldr=ldr)
        assert info == 0, (case, info)
        Qb, Rb = dA.numpy(), dR.numpy()
        Q, R = Qb[:m], Rb[:n]
        assert np.all(Qb[m:] == 7.0) and np.all(Rb[n:] == 7.0), case
        assert np.array_equal(np.tril(R, -1), np.zeros((n, n))), case
        assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13 * max(1.0, np.sqrt(n)), case
        assert np.linalg.norm(A - Q @ R) / np.linalg.norm(A) < 1e-14 * max(1.0, np.sqrt(n)), case


def test_gram_random_cases(hip):
    """Seeded random n, k, alpha, beta, mirror and padded leading dimensions
    of A and C: the upper triangle is alpha A^T A + beta C, the lower
    triangle its mirror or untouched, the padding untouched; exact on
    integers."""
    rnd = np.random.default_rng(2718)
    for case in range(10):
        n, k = int(rnd.integers(1, 300)), int(rnd.integers(1, 5000))
        alpha, beta = float(rnd.integers(-2, 3)), float(rnd.integers(-2, 3))
        mirror = bool(rnd.integers(0, 2))
        lda, ldc = k + int(rnd.integers(0, 4)), n + int(rnd.integers(0, 4))
        A = rnd.integers(-3, 4, (k, n)).astype(np.float64)
        C0 = rnd.integers(-9, 10, (ldc, n)).astype(np.float64)
        bufA = np.full((lda, n), 5.0, order="F")
        bufA[:k] = A
        dA, dC = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(np.asfortranarray(C0))
        print(f"[gram] case {case}: n={n} k={k} alpha={alpha} beta={beta} mirror={mirror}", flush=True)
        hip.gram(n, k, dA, dC, alpha=alpha, beta=beta, mirror=mirror, lda=lda, ldc=ldc)
        got = dC.numpy()
        full = alpha * (A.T @ A) + beta * C0[:n]
        iu, il = np.triu_indices(n), np.tril_indices(n, -1)
        assert np.array_equal(got[:n][iu], full[iu]), case
        lower_want = full.T[il] if mirror else C0[:n][il]
        assert np.array_equal(got[:n][il], lower_want), case
        assert np.array_equal(got[n:], C0[n:]), case


def test_potrf_trtri_trsm_random_padded(hip):
    """Seeded random sizes with padded leading dimensions: the upper
    Cholesky factor satisfies R^T R = S (lower triangle and padding
    untouched), R X = I for X = R^{-1}, and the right solve gives X R =
    alpha B, each to working precision."""
    rnd = np.random.default_rng(3141)
    for case in range(10):
        n, m = int(rnd.integers(1, 700)), int(rnd.integers(1, 3000))
        X = rnd.standard_normal((2 * n + 3, n))
        S = X.T @ X + n * np.eye(n)
        ld = n + int(rnd.integers(1, 4))
        bufS = np.full((ld, n), 3.0, order="F")
        bufS[:n] = S
        dS = hip.DeviceArray.from_numpy(bufS)
        print(f"[solve] case {case}: n={n} m={m}", flush=True)
        assert hip.potrf_upper(n, dS) == 0, case
        got = dS.numpy()
        R = np.triu(got[:n])
        assert np.array_equal(np.tril(got[:n], -1), np.tril(S, -1)), case
        assert np.all(got[n:] == 3.0), case
        assert np.linalg.norm(R.T @ R - S) <= 1e-14 * n * np.linalg.norm(S), case
        dR = hip.DeviceArray.from_numpy(np.asfortranarray(np.vstack([R, np.zeros((ld - n, n))])))
        dX = hip.DeviceArray.from_numpy(np.full((ld, n), 3.0, order="F"))
        hip.trtri_upper(n, dR, dX)
        Xi = dX.numpy()
        assert np.all(Xi[n:] == 3.0), case
        assert np.linalg.norm(R @ np.triu(Xi[:n]) - np.eye(n)) <= 1e-13 * n, case
        alpha = float(rnd.choice([1.0, -0.5, 2.0]))
        B = rnd.standard_normal((m, n))
        ldb = m + int(rnd.integers(1, 4))
        bufB = np.full((ldb, n), 3.0, order="F")
        bufB[:m] = B
        dB = hip.DeviceArray.from_numpy(bufB)
        hip.trsm_right_upper(m, n, alpha, dR, dB)
        Y = dB.numpy()
        assert np.all(Y[m:] == 3.0), case
        assert np.linalg.norm(Y[:m] @ R - alpha * B) <= 1e-13 * n * np.linalg.norm(alpha * B), case
