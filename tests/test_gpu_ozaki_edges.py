"""GPU: the Ozaki forms on non-finite inputs and on line scales near the ends
of the exponent range, bit for bit against the specifications
(oracle/ozaki.py; NaN positions compared as equal).

A line of op(A) or op(B) holding an Inf or a NaN gives a NaN row / column of
C (the rest of C is computed as usual); hip_dgemm's IEEE result has NaN or
Inf there too.  Scales are undone in one step, so a row near 2^1000 times a
column near 2^-1000 stays finite."""
import numpy as np
import pytest

from oracle import ozaki as Z
from tests.simutil import ozaki_grid_reference, run_sim

pytestmark = pytest.mark.gpu


def _with_nonfinite(rnd, shape_a, shape_b):
    A, B = rnd.uniform(-1, 1, shape_a), rnd.uniform(-1, 1, shape_b)
    A[3, 5] = np.nan
    A[min(10, shape_a[0] - 1), 0] = -np.inf
    B[7, 2] = np.inf
    B[0, min(9, shape_b[1] - 1)] = np.nan
    return A, B


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1), (0, 1)])
@pytest.mark.parametrize("form", ["slice", "crt"])
def test_nonfinite_inputs(hip, form, ta, tb):
    rnd = np.random.default_rng(70 + 2 * ta + tb)
    m, n, k = 300, 270, 333
    opA, opB = _with_nonfinite(rnd, (m, k), (k, n))
    A = np.asfortranarray(opA.T if ta else opA)
    B = np.asfortranarray(opB.T if tb else opB)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    Cin = rnd.standard_normal((m, n))
    dC = hip.DeviceArray.from_numpy(Cin)
    if form == "slice":
        want = Z.ozaki_dgemm8(opA, opB, 8, alpha=1.5, beta=-0.5, Cin=Cin)
        hip.ozaki_dgemm(ta, tb, m, n, k, 1.5, dA, dB, -0.5, dC, 8)
    else:
        want = Z.ozaki2_dgemm(opA, opB, 16, alpha=1.5, beta=-0.5, Cin=Cin)
        hip.ozaki2_dgemm(ta, tb, m, n, k, 1.5, dA, dB, -0.5, dC, 16)
    got = dC.numpy()
    assert np.array_equal(got, want, equal_nan=True)
    assert np.all(np.isnan(got[[3, 10], :])) and np.all(np.isnan(got[:, [2, 9]]))
    assert np.count_nonzero(np.isnan(got)) == 2 * n + 2 * m - 4
    # the IEEE product has non-finite entries wherever the Ozaki forms write NaN
    dN = hip.DeviceArray.zeros((m, n))
    hip.dgemm(ta, tb, m, n, k, 1.0, dA, dB, 0.0, dN)
    assert np.all(~np.isfinite(dN.numpy())[np.isnan(got)])


def test_nonfinite_split_and_gram(hip, monkeypatch):
    rnd = np.random.default_rng(75)
    n, k = 200, 5000
    A = rnd.uniform(-1, 1, (k, n))
    A[17, 4] = np.inf
    dA = hip.DeviceArray.from_numpy(A)
    dC = hip.DeviceArray.zeros((n, n))
    hip.ozaki_gram(n, k, dA, dC, 7)
    want = Z.ozaki_dgemm8(np.ascontiguousarray(A.T), A, 7)
    got = dC.numpy()
    iu = np.triu_indices(n)
    assert np.array_equal(got[iu], want[iu], equal_nan=True)
    assert np.all(np.isnan(got[4, :])) and np.all(np.isnan(got[:, 4]))
    monkeypatch.setenv("CA3DMM_OZAKI_SPLIT", "1")
    B = rnd.uniform(-1, 1, (k, 64))
    B[2, 60] = np.nan
    dB = hip.DeviceArray.from_numpy(B)
    dD = hip.DeviceArray.zeros((n, 64))
    hip.ozaki_dgemm(1, 0, n, 64, k, 1.0, dA, dB, 0.0, dD, 7)
    assert np.array_equal(dD.numpy(), Z.ozaki_dgemm8(np.ascontiguousarray(A.T), B, 7), equal_nan=True)


def test_nonfinite_in_the_grouped_grid(hip):
    """The Ozaki form of the 4x2x3 grid: a NaN in A makes its row of C NaN
    through the layer sum; every other entry as the restated grid."""
    rnd = np.random.default_rng(76)
    m, n, k = 500, 420, 610
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    A[123, 400] = np.nan
    want = ozaki_grid_reference(m, n, k, 24, [4, 2, 3], A, B, 15)
    got = run_sim(24, m, n, k, 0, 0, np.asfortranarray(A), np.asfortranarray(B), proc_grid=[4, 2, 3],
                  grouped=True, ozaki=15)
    assert np.array_equal(got, want, equal_nan=True)
    assert np.all(np.isnan(got[123, :])) and np.count_nonzero(np.isnan(got)) == n


@pytest.mark.parametrize("form", ["slice", "crt"])
def test_scales_near_the_exponent_range_ends(hip, form):
    rnd = np.random.default_rng(77)
    m, n, k = 130, 150, 400
    A = np.ldexp(rnd.uniform(-1, 1, (m, k)), 1000)
    B = np.ldexp(rnd.uniform(-1, 1, (k, n)), -1000)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((m, n))
    if form == "slice":
        hip.ozaki_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, 8)
        want = Z.ozaki_dgemm8(A, B, 8)
    else:
        hip.ozaki2_dgemm(0, 0, m, n, k, 1.0, dA, dB, 0.0, dC, 16)
        want = Z.ozaki2_dgemm(A, B, 16)
    got = dC.numpy()
    assert np.all(np.isfinite(got))
    assert np.array_equal(got, want)
