"""GPU: the FP64 MFMA DGEMM (cuda_cublas_dgemm replacement, src/cuda_proxy.cu:211-230)
against the oracle.

Tolerance (componentwise, for random inputs): |C - C_ref| <= k * u * (|A||B|)_ij
with u = 2^-53, C_ref from double-double dot products; integer-valued inputs
must match exactly (every summation order is exact on them)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def _mats(rnd, m, n, k, ta, tb, kind):
    sa = (k, m) if ta else (m, k)
    sb = (n, k) if tb else (k, n)
    if kind == "int":
        return (rnd.integers(-16, 17, sa).astype(np.float64), rnd.integers(-16, 17, sb).astype(np.float64))
    return rnd.uniform(-1, 1, sa), rnd.uniform(-1, 1, sb)


SHAPES = [(1, 1, 1), (15, 17, 3), (16, 16, 4), (128, 128, 16), (129, 130, 17), (200, 300, 500),
          (257, 64, 1000), (64, 511, 33), (1000, 1000, 1000)]


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k", SHAPES)
def test_dgemm_integer_exact(hip, ta, tb, m, n, k):
    rnd = np.random.default_rng(m * 31 + n * 7 + k + ta * 2 + tb)
    A, B = _mats(rnd, m, n, k, ta, tb, "int")
    want = (A.T if ta else A) @ (B.T if tb else B)          # exact: integers < 2^53
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((m, n), 0xFF)                 # NaN garbage; beta = 0 ignores it
    hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dC)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_random_componentwise_bound(hip, ta, tb):
    m, n, k = 777, 555, 1999
    rnd = np.random.default_rng(42 + ta * 2 + tb)
    A, B = _mats(rnd, m, n, k, ta, tb, "f")
    Af, Bf = np.asfortranarray(A), np.asfortranarray(B)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((m, n))
    hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dC)
    got = dC.numpy()
    ii = rnd.integers(0, m, 3000); jj = rnd.integers(0, n, 3000)
    ref, ab = O.dd_sample(ta, tb, m, n, k, Af, Af.shape[0], Bf, Bf.shape[0], ii, jj)
    err = np.abs(got[ii, jj] - ref)
    assert np.all(err <= k * U * ab), float(np.max(err / ab))


def test_dgemm_alpha_beta_and_ld(hip):
    rnd = np.random.default_rng(3)
    m, n, k = 150, 140, 130
    lda, ldb, ldc = 157, 133, 151
    A = rnd.integers(-8, 9, (lda, k)).astype(np.float64)
    B = rnd.integers(-8, 9, (ldb, n)).astype(np.float64)
    Cm = rnd.integers(-8, 9, (ldc, n)).astype(np.float64)
    want = Cm.copy()
    want[:m] = 2.0 * (A[:m] @ B[:k]) - 3.0 * Cm[:m]
    dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (A, B, Cm))
    hip.dgemm(0, 0, m, n, k, 2.0, dA, dB, -3.0, dC, lda=lda, ldb=ldb, ldc=ldc)
    assert np.array_equal(dC.numpy(), want)          # rows m..ldc untouched


def test_dgemm_general_alpha_beta_rounded_separately(hip):
    """alpha * AB and beta * C are rounded separately, then added (the C
    oracle's order): with exact integer products, alpha = 1.1 and beta = -0.3
    a fused multiply-add in the epilogue would differ in the last bit."""
    rnd = np.random.default_rng(8)
    m, n, k = 133, 97, 70
    A = rnd.integers(-16, 17, (m, k)).astype(np.float64)
    B = rnd.integers(-16, 17, (k, n)).astype(np.float64)
    Cm = rnd.standard_normal((m, n))
    want = 1.1 * (A @ B) + (-0.3) * Cm
    dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (A, B, Cm))
    hip.dgemm(0, 0, m, n, k, 1.1, dA, dB, -0.3, dC)
    assert np.array_equal(dC.numpy(), want)


def test_dgemm_fold_random_cases(hip):
    """Seeded random shapes, trans cases, alpha, beta and 0..8 partials: the
    fused epilogue gives fold(P) + (alpha op(A) op(B) + beta C) bit for bit
    (as hip_fold_sum of the partials and a separate product), exact on
    integers."""
    rnd = np.random.default_rng(9191)
    for case in range(10):
        m, n = (int(v) for v in rnd.integers(1, 400, 2))
        k = int(rnd.integers(0, 500))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        nsrc = int(rnd.integers(0, 9))
        alpha, beta = float(rnd.integers(-2, 3)), float(rnd.integers(-2, 3))
        A = rnd.integers(-4, 5, (k, m) if ta else (m, k)).astype(np.float64)
        B = rnd.integers(-4, 5, (n, k) if tb else (k, n)).astype(np.float64)
        C0 = rnd.integers(-50, 51, (m, n)).astype(np.float64)
        P = [rnd.integers(-1000, 1001, (m, n)).astype(np.float64) for _ in range(nsrc)]
        dA = hip.DeviceArray.from_numpy(A if A.size else np.zeros((max(1, A.shape[0]), 1)))
        dB = hip.DeviceArray.from_numpy(B if B.size else np.zeros((max(1, B.shape[0]), 1)))
        dC = hip.DeviceArray.from_numpy(C0)
        dP = [hip.DeviceArray.from_numpy(x) for x in P]
        print(f"[fold] case {case}: m={m} n={n} k={k} ta={ta} tb={tb} nsrc={nsrc} alpha={alpha} beta={beta}",
              flush=True)
        hip.dgemm_fold(ta, tb, m, n, k, alpha, dA, dB, beta, dC, dP)
        prod = (A.T if ta else A) @ (B.T if tb else B) if k else np.zeros((m, n))
        want = alpha * prod + beta * C0
        if nsrc:
            s = P[0].copy()
            for t in range(1, nsrc):
                s = s + P[t]
            want = s + want
        assert np.array_equal(dC.numpy(), want), case


def test_dgemm_k_zero_scales_c(hip):
    Cm = np.arange(12.0).reshape(3, 4)
    dC = hip.DeviceArray.from_numpy(Cm)
    dA = hip.DeviceArray.zeros((3, 1)); dB = hip.DeviceArray.zeros((1, 4))
    hip.dgemm(0, 0, 3, 4, 0, 1.0, dA, dB, 0.5, dC)
    assert np.array_equal(dC.numpy(), 0.5 * Cm)


def test_dgemm_fold_epilogue_equals_fold_sum(hip):
    """C := (P0 + P1) + A*B equals hip_fold_sum over [P0, P1, A*B] bit for bit."""
    rnd = np.random.default_rng(8)
    m, n, k = 300, 260, 411
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    P = [rnd.uniform(-1, 1, (m, n)) * 100 for _ in range(2)]
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dP = [hip.DeviceArray.from_numpy(x) for x in P]
    prod = hip.DeviceArray.zeros((m, n))
    hip.dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, prod)
    want = hip.DeviceArray.zeros((m, n))
    hip.fold_sum(dP + [prod], m, n, want)
    fused = hip.DeviceArray.zeros((m, n))
    hip.dgemm_fold(0, 0, m, n, k, 1.0, dA, dB, 0.0, fused, dP)
    assert np.array_equal(fused.numpy(), want.numpy())
    assert np.array_equal(want.numpy(), (P[0] + P[1]) + prod.numpy())


def test_dgemm_rejects_bad_arguments(hip):
    from ca3dmm_amd import HipError
    dA = hip.DeviceArray.zeros((4, 4))
    with pytest.raises(HipError):
        hip.dgemm(0, 0, 4, 4, 4, 1.0, dA, dA, 0.0, dA, lda=2)


@pytest.mark.parametrize("variant", [("CA3DMM_DGEMM_LOADS", "reg"), ("CA3DMM_DGEMM_LOADS", "dword"),
                                     ("CA3DMM_DGEMM_LOADS", "aligned"),
                                     ("CA3DMM_DGEMM_TILE", "256x128"), ("CA3DMM_DGEMM_TILE", "128x256")])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_variants_same_bits(hip, monkeypatch, variant, ta, tb):
    """The register-staged 128 x 128 form (the default loads its stages
    direct-to-LDS, 16 bytes per lane), the 4-byte direct-to-LDS loads, the
    16-byte-aligned-only policy and the 8-wave tile shapes add the same
    products in the same order as the default kernel: the same bits on random inputs
    (interior and edge tiles, odd and even stage counts, a k tail), and
    exact integers."""
    rnd = np.random.default_rng(90 + 2 * ta + tb)
    for m, n, k in [(640, 520, 700), (1000, 1000, 1000), (256, 384, 32), (384, 256, 48), (130, 129, 17)]:
        A, B = _mats(rnd, m, n, k, ta, tb, "f")
        dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
        ref = hip.DeviceArray.zeros((m, n))
        monkeypatch.delenv("CA3DMM_DGEMM_LOADS", raising=False)
        monkeypatch.delenv("CA3DMM_DGEMM_TILE", raising=False)
        hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, ref)
        monkeypatch.setenv(*variant)
        got = hip.DeviceArray.filled((m, n), 0xFF)
        hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, got)
        assert np.array_equal(got.numpy(), ref.numpy()), (variant, m, n, k)
        Ai, Bi = _mats(rnd, m, n, k, ta, tb, "int")
        dA, dB = hip.DeviceArray.from_numpy(Ai), hip.DeviceArray.from_numpy(Bi)
        hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, got)
        assert np.array_equal(got.numpy(), (Ai.T if ta else Ai) @ (Bi.T if tb else Bi)), (variant, m, n, k)


class _At:
    """A device address inside a DeviceArray, for hip.dgemm's A / B / C."""
    def __init__(self, ptr):
        self.ptr = ptr


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_odd_leading_dimensions_and_offsets(hip, monkeypatch, ta, tb):
    """Operands with odd leading dimensions, or starting one double past a
    16-byte boundary, take 16-byte direct-to-LDS loads from 8-byte aligned
    addresses (with CA3DMM_DGEMM_LOADS=aligned: 4-byte loads where the
    operand is contiguous along m or n, register staging otherwise): the
    same bits as the register-staged kernel on random inputs, exact on
    integers."""
    rnd = np.random.default_rng(150 + 2 * ta + tb)
    for m, n, k, pa, pb in [(300, 260, 333, 1, 0), (257, 131, 160, 0, 1), (640, 384, 211, 1, 1), (129, 255, 48, 3, 5)]:
        ra, ca = (k, m) if ta else (m, k)
        rb, cb = (n, k) if tb else (k, n)
        lda, ldb = ra + pa, rb + pb                        # odd or even, per case
        for kind in ("f", "int"):
            bufA = np.zeros((lda * ca + 1,))
            bufB = np.zeros((ldb * cb + 1,))
            if kind == "f":
                A, B = rnd.uniform(-1, 1, (ra, ca)), rnd.uniform(-1, 1, (rb, cb))
            else:
                A, B = rnd.integers(-4, 5, (ra, ca)).astype(float), rnd.integers(-4, 5, (rb, cb)).astype(float)
            # column-major with leading dimension ld, starting one double in
            for c in range(ca):
                bufA[1 + c * lda:1 + c * lda + ra] = A[:, c]
            for c in range(cb):
                bufB[1 + c * ldb:1 + c * ldb + rb] = B[:, c]
            dA, dB = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(bufB)
            got, ref = hip.DeviceArray.filled((m, n), 0xFF), hip.DeviceArray.filled((m, n), 0xFF)
            monkeypatch.setenv("CA3DMM_DGEMM_LOADS", "reg")
            hip.dgemm(ta, tb, m, n, k, 1.0, _At(dA.at(1)), _At(dB.at(1)), 0.0, ref, lda=lda, ldb=ldb, ldc=m)
            for mode in (None, "aligned"):
                if mode is None:
                    monkeypatch.delenv("CA3DMM_DGEMM_LOADS", raising=False)
                else:
                    monkeypatch.setenv("CA3DMM_DGEMM_LOADS", mode)
                hip.dgemm(ta, tb, m, n, k, 1.0, _At(dA.at(1)), _At(dB.at(1)), 0.0, got, lda=lda, ldb=ldb, ldc=m)
                assert np.array_equal(got.numpy(), ref.numpy()), (m, n, k, lda, ldb, kind, mode)
            if kind == "int":
                want = (A.T if ta else A) @ (B.T if tb else B)
                assert np.array_equal(got.numpy(), want), (m, n, k, lda, ldb)


def test_dgemm_random_shapes_layouts_match_register_staged(hip, monkeypatch):
    """Seeded random shapes, leading dimensions, operand offsets and trans
    cases (16-byte, 4-byte and register-staged loads; interior, edge and
    one-tile products; k tails): the default kernel choice gives the bits of
    the register-staged kernel, and integer products are exact."""
    rnd = np.random.default_rng(2026)
    for case in range(40):
        m, n = (int(v) for v in rnd.integers(1, 700, 2))
        k = int(rnd.integers(1, 400))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        ra, ca = (k, m) if ta else (m, k)
        rb, cb = (n, k) if tb else (k, n)
        lda, ldb = ra + int(rnd.integers(0, 4)), rb + int(rnd.integers(0, 4))
        oa, ob = int(rnd.integers(0, 4)), int(rnd.integers(0, 4))
        A = rnd.integers(-8, 9, (ra, ca)).astype(np.float64)
        B = rnd.integers(-8, 9, (rb, cb)).astype(np.float64)
        bufA, bufB = np.zeros(oa + lda * ca), np.zeros(ob + ldb * cb)
        for c in range(ca):
            bufA[oa + c * lda:oa + c * lda + ra] = A[:, c]
        for c in range(cb):
            bufB[ob + c * ldb:ob + c * ldb + rb] = B[:, c]
        dA, dB = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(bufB)
        got, ref = hip.DeviceArray.filled((m, n), 0xFF), hip.DeviceArray.filled((m, n), 0xFF)
        monkeypatch.delenv("CA3DMM_DGEMM_LOADS", raising=False)
        hip.dgemm(ta, tb, m, n, k, 1.0, _At(dA.at(oa)), _At(dB.at(ob)), 0.0, got, lda=lda, ldb=ldb, ldc=m)
        monkeypatch.setenv("CA3DMM_DGEMM_LOADS", "reg")
        hip.dgemm(ta, tb, m, n, k, 1.0, _At(dA.at(oa)), _At(dB.at(ob)), 0.0, ref, lda=lda, ldb=ldb, ldc=m)
        want = (A.T if ta else A) @ (B.T if tb else B)
        assert np.array_equal(got.numpy(), want), (case, m, n, k, ta, tb, lda, ldb, oa, ob)
        assert np.array_equal(ref.numpy(), want), (case, m, n, k, ta, tb, lda, ldb, oa, ob)
