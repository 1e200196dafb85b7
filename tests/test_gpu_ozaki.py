"""GPU: hip_ozaki_dgemm reproduces the Ozaki specification (oracle/ozaki.py,
ozaki_dgemm8) bit for bit -- INT32 slice products are exact and the FP64
recombination order is fixed -- and reaches FP64 accuracy."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("loads", ["pipe", "dtl", "reg"])
@pytest.mark.parametrize("tile", ["256x256", "256x128"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k,s", [(1, 1, 1, 1), (17, 33, 65, 3), (128, 128, 64, 8),
                                     (200, 130, 300, 8), (257, 129, 1000, 12), (64, 300, 129, 5),
                                     (300, 520, 257, 7)])
def test_ozaki_bit_exact_vs_spec(hip, monkeypatch, loads, tile, ta, tb, m, n, k, s):
    """Both workgroup tiles (CA3DMM_OZAKI_TILE) and both load paths
    (CA3DMM_OZAKI_LOADS: direct-to-LDS or register-staged) against the
    specification, at k from one pipeline round to several."""
    monkeypatch.setenv("CA3DMM_OZAKI_TILE", tile)
    monkeypatch.setenv("CA3DMM_OZAKI_LOADS", loads)
    rnd = np.random.default_rng(m * 7 + n + k + s + 2 * ta + tb)
    opA = Z.phi_matrix(rnd, (m, k), 2.0)
    opB = Z.phi_matrix(rnd, (k, n), 2.0)
    if m > 3:
        opA[3] = 0.0
    if n > 5:
        opB[:, 5] = 0.0
    want = Z.ozaki_dgemm8(opA, opB, s)
    A = opA.T if ta else opA
    B = opB.T if tb else opB
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((m, n), 0xFF)
    hip.ozaki_dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dC, s)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("waves", ["4", "8"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k,s", [(300, 520, 257, 7), (513, 260, 3000, 8), (40, 50, 20000, 16)])
def test_ozaki_pipelined_wave_counts(hip, monkeypatch, waves, ta, tb, m, n, k, s):
    """The pipelined form with 4 waves (fragment prefetch) and 8 waves per
    256 x 256 tile, per-segment (fused) and split-K, against the specification."""
    monkeypatch.setenv("CA3DMM_OZAKI_WAVES", waves)
    rnd = np.random.default_rng(m + n + k + s + ta + int(waves))
    opA, opB = Z.phi_matrix(rnd, (m, k), 2.0), Z.phi_matrix(rnd, (k, n), 2.0)
    want = Z.ozaki_dgemm8(opA, opB, s)
    dA = hip.DeviceArray.from_numpy(opA.T if ta else opA)
    dB = hip.DeviceArray.from_numpy(opB.T if tb else opB)
    dC = hip.DeviceArray.filled((m, n), 0xFF)
    for split in ("0", "1"):
        monkeypatch.setenv("CA3DMM_OZAKI_SPLIT", split)
        hip.ozaki_dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dC, s)
        assert np.array_equal(dC.numpy(), want), split


def test_ozaki_alpha_beta(hip):
    rnd = np.random.default_rng(5)
    m, n, k = 70, 90, 110
    A, B, Cin = Z.phi_matrix(rnd, (m, k), 0.5), Z.phi_matrix(rnd, (k, n), 0.5), rnd.standard_normal((m, n))
    want = Z.ozaki_dgemm8(A, B, 7, alpha=-2.0, beta=0.5, Cin=Cin)
    dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (A, B, Cin))
    hip.ozaki_dgemm(0, 0, m, n, k, -2.0, dA, dB, 0.5, dC, 7)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("split", ["0", "1"])
def test_ozaki_general_alpha_beta(hip, monkeypatch, split):
    """alpha = 1.1, beta = -0.3: the two products are rounded separately and
    then added, in both forms (a fused multiply-add would differ)."""
    monkeypatch.setenv("CA3DMM_OZAKI_SPLIT", split)
    rnd = np.random.default_rng(15)
    m, n, k = 90, 70, 300
    A, B, Cin = Z.phi_matrix(rnd, (m, k), 2.0), Z.phi_matrix(rnd, (k, n), 2.0), rnd.standard_normal((m, n))
    want = Z.ozaki_dgemm8(A, B, 8, alpha=1.1, beta=-0.3, Cin=Cin)
    dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (A, B, Cin))
    hip.ozaki_dgemm(0, 0, m, n, k, 1.1, dA, dB, -0.3, dC, 8)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("split", ["0", "1"])
def test_ozaki_k_chunks(hip, monkeypatch, split):
    """s = 16 limits an INT32-exact k-chunk to 7936: k = 20000 runs 3 chunks,
    through the per-diagonal launches (split = 0) and the split-K launch with
    INT32 partials and a separate fold (split = 1)."""
    monkeypatch.setenv("CA3DMM_OZAKI_SPLIT", split)
    rnd = np.random.default_rng(6)
    m, n, k = 40, 50, 20000
    A, B = Z.phi_matrix(rnd, (m, k), 1.0), Z.phi_matrix(rnd, (k, n), 1.0)
    want = Z.ozaki_dgemm8(A, B, 16)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((m, n))
    hip.ozaki_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, 16)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k,s", [(1, 1, 1, 1), (200, 130, 300, 8), (300, 70, 12000, 12)])
def test_ozaki_split_mode_bit_exact(hip, monkeypatch, ta, tb, m, n, k, s):
    monkeypatch.setenv("CA3DMM_OZAKI_SPLIT", "1")
    rnd = np.random.default_rng(m + n + k + s + ta)
    opA, opB = Z.phi_matrix(rnd, (m, k), 2.0), Z.phi_matrix(rnd, (k, n), 2.0)
    Cin = rnd.standard_normal((m, n))
    want = Z.ozaki_dgemm8(opA, opB, s, alpha=1.5, beta=-0.25, Cin=Cin)
    dA = hip.DeviceArray.from_numpy(opA.T if ta else opA)
    dB = hip.DeviceArray.from_numpy(opB.T if tb else opB)
    dC = hip.DeviceArray.from_numpy(Cin)
    hip.ozaki_dgemm(ta, tb, m, n, k, 1.5, dA, dB, -0.25, dC, s)
    assert np.array_equal(dC.numpy(), want)


@pytest.mark.parametrize("loads", ["pipe", "dtl", "reg"])
@pytest.mark.parametrize("tile", ["256x256", "256x128"])
@pytest.mark.parametrize("n,k,s,mirror", [(1, 5, 3, True), (100, 20000, 12, True), (300, 1000, 8, False),
                                          (513, 700, 8, True)])
def test_ozaki_gram_bit_exact(hip, monkeypatch, loads, tile, n, k, s, mirror):
    """hip_ozaki_gram equals ozaki_dgemm8(A^T, A) on the upper triangle; the
    lower triangle is its mirror, or C untouched when mirror = 0."""
    monkeypatch.setenv("CA3DMM_OZAKI_TILE", tile)
    monkeypatch.setenv("CA3DMM_OZAKI_LOADS", loads)
    rnd = np.random.default_rng(n + k + s)
    A = Z.phi_matrix(rnd, (k, n), 1.0)
    if n > 7:
        A[:, 7] = 0.0
    Cin = rnd.standard_normal((n, n))
    want = Z.ozaki_dgemm8(np.ascontiguousarray(A.T), A, s, alpha=0.5, beta=2.0, Cin=Cin)
    dA, dC = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(Cin)
    hip.ozaki_gram(n, k, dA, dC, s, alpha=0.5, beta=2.0, mirror=mirror)
    got = dC.numpy()
    iu = np.triu_indices(n)
    assert np.array_equal(got[iu], want[iu])
    il = np.tril_indices(n, -1)
    assert np.array_equal(got[il], (want.T if mirror else Cin)[il])


def test_ozaki_gram_accuracy(hip):
    """phi = 0.5, 256 columns, k = 100000, s = 8: on 500 sampled entries the
    error against double-double, over (|A|^T |A|), is no worse than the FP64
    MFMA Gram's and below 1e-15 (the FP64 recombination of s * 7 chunk
    partials costs a few ulp of |G|)."""
    rnd = np.random.default_rng(8)
    n, k = 256, 100000
    A = Z.phi_matrix(rnd, (k, n), 0.5)
    dA, dG = hip.DeviceArray.from_numpy(A), hip.DeviceArray.zeros((n, n))
    ii, jj = rnd.integers(0, n, 500), rnd.integers(0, n, 500)
    Af = np.asfortranarray(A)
    ref, ab = O.dd_sample(1, 0, n, n, k, Af, k, Af, k, ii, jj)
    hip.ozaki_gram(n, k, dA, dG, 8)
    oz = np.max(np.abs(dG.numpy()[ii, jj] - ref) / ab)
    hip.gram(n, k, dA, dG)
    native = np.max(np.abs(dG.numpy()[ii, jj] - ref) / ab)
    assert oz <= native and oz < 1e-15, (oz, native)


def test_ozaki_accuracy_matches_fp64(hip):
    """phi = 0.5, k = 8192, s = 8: max |C - C_dd| / (|A||B|) below 1e-16."""
    rnd = np.random.default_rng(7)
    m, n, k = 512, 384, 8192
    A, B = Z.phi_matrix(rnd, (m, k), 0.5), Z.phi_matrix(rnd, (k, n), 0.5)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((m, n))
    hip.ozaki_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, 8)
    got = dC.numpy()
    ii, jj = rnd.integers(0, m, 2000), rnd.integers(0, n, 2000)
    ref, ab = O.dd_sample(0, 0, m, n, k, np.asfortranarray(A), m, np.asfortranarray(B), k, ii, jj)
    assert np.max(np.abs(got[ii, jj] - ref) / ab) < 1e-16


def test_ozaki_rejects_bad_slices(hip):
    from ca3dmm_amd import HipError
    d = hip.DeviceArray.zeros((4, 4))
    with pytest.raises(HipError):
        hip.ozaki_dgemm(0, 0, 4, 4, 4, 1.0, d, d, 0.0, d, 0)
    with pytest.raises(HipError):
        hip.ozaki_dgemm(0, 0, 4, 4, 4, 1.0, d, d, 0.0, d, 17)


class _At:
    """A device address inside a DeviceArray, for hip_ozaki_gram's A / C."""
    def __init__(self, ptr):
        self.ptr = ptr


def test_ozaki_gram_padded_layouts(hip):
    """Seeded random n, k, s, scalars and mirror with A and C at padded
    leading dimensions and odd offsets into their buffers: the upper
    triangle is the specification's, the lower its mirror or untouched, the
    padding untouched."""
    rnd = np.random.default_rng(4711)
    for case in range(6):
        n, k, s = int(rnd.integers(1, 260)), int(rnd.integers(1, 3000)), int(rnd.integers(2, 10))
        mirror = bool(rnd.integers(0, 2))
        lda, ldc, oa, oc = k + int(rnd.integers(0, 4)), n + int(rnd.integers(1, 4)), int(rnd.integers(0, 4)), 1
        A = Z.phi_matrix(rnd, (k, n), 2.0)
        Cin = rnd.standard_normal((n, n))
        bufA = np.full(oa + lda * n, 9.0)
        for c in range(n):
            bufA[oa + c * lda:oa + c * lda + k] = A[:, c]
        bufC = np.full(oc + ldc * n, 9.0)
        for c in range(n):
            bufC[oc + c * ldc:oc + c * ldc + n] = Cin[:, c]
        dA, dC = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(bufC)
        print(f"[ozaki gram] padded case {case}: n={n} k={k} s={s} mirror={mirror}", flush=True)
        hip.ozaki_gram(n, k, _At(dA.at(oa)), _At(dC.at(oc)), s, alpha=-1.5, beta=0.5, mirror=mirror,
                       lda=lda, ldc=ldc)
        out = dC.numpy()
        got = out[oc:].reshape(n, ldc).T
        want = Z.ozaki_dgemm8(np.ascontiguousarray(A.T), A, s, alpha=-1.5, beta=0.5, Cin=Cin)
        iu, il = np.triu_indices(n), np.tril_indices(n, -1)
        assert np.array_equal(got[:n][iu], want[iu]), case
        assert np.array_equal(got[:n][il], (want.T if mirror else Cin)[il]), case
        assert np.all(got[n:] == 9.0) and out[0] == 9.0, case
