"""GPU: hip_ozaki2_dgemm (the CRT form of the Ozaki scheme) reproduces the
specification (oracle/ozaki.py, ozaki2_dgemm) bit for bit -- residue
products are exact and the Garner/Horner evaluation order is fixed -- and
reaches FP64 accuracy with 16 moduli."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z

pytestmark = pytest.mark.gpu


def _run(hip, ta, tb, opA, opB, N, alpha=1.0, beta=0.0, Cin=None):
    m, k = opA.shape
    n = opB.shape[1]
    dA = hip.DeviceArray.from_numpy(opA.T if ta else opA)
    dB = hip.DeviceArray.from_numpy(opB.T if tb else opB)
    dC = hip.DeviceArray.from_numpy(Cin) if Cin is not None else hip.DeviceArray.filled((m, n), 0xFF)
    hip.ozaki2_dgemm(ta, tb, m, n, k, alpha, dA, dB, beta, dC, N)
    return dC.numpy()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k,N", [(1, 1, 1, 4), (17, 33, 65, 6), (128, 128, 64, 16), (200, 130, 300, 14),
                                     (257, 129, 1000, 20), (64, 300, 129, 9), (300, 520, 257, 16),
                                     (513, 260, 3000, 18)])
def test_ozaki2_bit_exact_vs_spec(hip, ta, tb, m, n, k, N):
    rnd = np.random.default_rng(m * 7 + n + k + N + 2 * ta + tb)
    opA = Z.phi_matrix(rnd, (m, k), 2.0)
    opB = Z.phi_matrix(rnd, (k, n), 2.0)
    if m > 3:
        opA[3] = 0.0
    if n > 5:
        opB[:, 5] = 0.0
    want = Z.ozaki2_dgemm(opA, opB, N)
    assert np.array_equal(_run(hip, ta, tb, opA, opB, N), want)


@pytest.mark.parametrize("N", [4, 5, 7, 8, 10, 11, 12, 13, 15, 17, 19, 21, 22, 23, 24, 25])
def test_ozaki2_every_modulus_count(hip, N):
    rnd = np.random.default_rng(500 + N)
    opA, opB = Z.phi_matrix(rnd, (70, 300), 4.0), Z.phi_matrix(rnd, (300, 90), 4.0)
    assert np.array_equal(_run(hip, 0, 0, opA, opB, N), Z.ozaki2_dgemm(opA, opB, N))


def test_ozaki2_alpha_beta_and_wide_exponents(hip):
    rnd = np.random.default_rng(9)
    m, n, k = 150, 170, 400
    opA = Z.phi_matrix(rnd, (m, k), 4.0)
    opB = Z.phi_matrix(rnd, (k, n), 4.0)
    opA[7] *= 2.0 ** 500
    opA[8] *= 2.0 ** -600
    opB[:, 9] *= 2.0 ** -300
    opA[10, :5] = [1e-310, -5e-324, 0.0, 1.0, -2.0]
    Cin = rnd.standard_normal((m, n))
    want = Z.ozaki2_dgemm(opA, opB, 16, alpha=-1.1, beta=0.3, Cin=Cin)
    assert np.array_equal(_run(hip, 1, 0, opA, opB, 16, -1.1, 0.3, Cin), want)


def test_ozaki2_k_zero_and_zero_operands(hip):
    rnd = np.random.default_rng(1)
    Cin = rnd.standard_normal((40, 30))
    dA, dB = hip.DeviceArray.zeros((40, 1)), hip.DeviceArray.zeros((1, 30))
    dC = hip.DeviceArray.from_numpy(Cin)
    hip.ozaki2_dgemm(0, 0, 40, 30, 0, 1.0, dA, dB, 2.0, dC, 16)
    assert np.array_equal(dC.numpy(), 2.0 * Cin)
    opA, opB = np.zeros((40, 50)), Z.phi_matrix(rnd, (50, 30), 1.0)
    assert np.array_equal(_run(hip, 0, 0, opA, opB, 8), np.zeros((40, 30)))


def test_ozaki2_k_chunks(hip):
    """k above one INT32-exact chunk (65280; k = 140000 is three chunks):
    residues of the chunks are added modulo p in the epilogue; the result
    is still the exact residue."""
    rnd = np.random.default_rng(4)
    m, n, k = 40, 70, 140000
    opA = Z.phi_matrix(rnd, (m, k), 1.0)
    opB = Z.phi_matrix(rnd, (k, n), 1.0)
    want = Z.ozaki2_dgemm(opA, opB, 16)
    assert np.array_equal(_run(hip, 0, 1, opA, opB, 16), want)


def test_ozaki2_rejects_bad_arguments(hip):
    dA = hip.DeviceArray.zeros((8, 8))
    for N in (3, 26):
        with pytest.raises(RuntimeError):
            hip.ozaki2_dgemm(0, 0, 8, 8, 8, 1.0, dA, dA, 0.0, dA, N)


def test_ozaki2_accuracy_no_worse_than_native(hip):
    """1024^3, phi = 0.5: with 16 moduli the sampled max relative error and
    max |dC|/(|A||B|) are no worse than the FP64 MFMA DGEMM's."""
    rnd = np.random.default_rng(2)
    n = 1024
    A, B = Z.phi_matrix(rnd, (n, n), 0.5), Z.phi_matrix(rnd, (n, n), 0.5)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((n, n))
    ii, jj = rnd.integers(0, n, 2048), rnd.integers(0, n, 2048)
    ref, ab = O.dd_sample(0, 0, n, n, n, np.asfortranarray(A), n, np.asfortranarray(B), n, ii, jj)

    def errs():
        C = dC.numpy()[ii, jj]
        return np.max(np.abs(C - ref) / ab), np.max(np.abs(C - ref) / np.abs(ref))

    hip.dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC)
    nat = errs()
    hip.ozaki2_dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC, 16)
    got = errs()
    assert got[0] <= nat[0] and got[1] <= nat[1], (got, nat)


class _At:
    """A device address inside a DeviceArray, for the Ozaki entries' A / B / C."""
    def __init__(self, ptr):
        self.ptr = ptr


def test_both_forms_padded_layouts_and_offsets(hip):
    """Seeded random shapes and trans cases with padded leading dimensions
    and operands starting an odd number of doubles into their buffers (C
    too): both Ozaki forms reproduce the specification bit for bit."""
    rnd = np.random.default_rng(5150)
    for case in range(8):
        m, n = (int(v) for v in rnd.integers(1, 260, 2))
        k = int(rnd.integers(1, 600))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        N, s = int(rnd.integers(4, 26)), int(rnd.integers(2, 9))
        opA, opB = Z.phi_matrix(rnd, (m, k), 2.0), Z.phi_matrix(rnd, (k, n), 2.0)
        A, B = (opA.T if ta else opA), (opB.T if tb else opB)
        lda, ldb, ldc = A.shape[0] + int(rnd.integers(0, 4)), B.shape[0] + int(rnd.integers(0, 4)), m + 3
        oa, ob, oc = (int(v) for v in rnd.integers(0, 4, 3))

        def packed(X, ld, off):
            buf = np.zeros(off + ld * X.shape[1])
            for c in range(X.shape[1]):
                buf[off + c * ld:off + c * ld + X.shape[0]] = X[:, c]
            return hip.DeviceArray.from_numpy(buf)
        dA, dB = packed(A, lda, oa), packed(B, ldb, ob)
        print(f"[ozaki] padded case {case}: m={m} n={n} k={k} ta={ta} tb={tb} N={N} s={s}", flush=True)
        for form, cnt, spec in (("crt", N, Z.ozaki2_dgemm), ("slice", s, Z.ozaki_dgemm8)):
            dC = hip.DeviceArray.filled((oc + ldc * n,), 0xFF)
            f = hip.ozaki2_dgemm if form == "crt" else hip.ozaki_dgemm
            f(ta, tb, m, n, k, 1.0, _At(dA.at(oa)), _At(dB.at(ob)), 0.0, _At(dC.at(oc)), cnt,
              lda=lda, ldb=ldb, ldc=ldc)
            got = dC.numpy()[oc:].reshape(n, ldc).T[:m]
            assert np.array_equal(got, spec(opA, opB, cnt)), (case, form)
