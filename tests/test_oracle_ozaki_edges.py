"""CPU: the Ozaki specifications (oracle/ozaki.py) on non-finite inputs and
on line scales near the ends of the exponent range.

* A line of op(A) or op(B) with an Inf or a NaN is sliced as a zero line and
  its whole row / column of C is NaN; every other entry is unchanged, bit
  for bit, from the same product with that line replaced by zeros (each C
  entry depends only on its row, its column and their scales).
* The scales are undone in one step, 2^(eA + eB), so a row near 2^1000 times
  a column near 2^-1000 stays finite: the result is within the forms'
  accuracy of the double-double product, as the native DGEMM's would be."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z

FORMS = [("slice", lambda A, B: Z.ozaki_dgemm8(A, B, 8)), ("crt", lambda A, B: Z.ozaki2_dgemm(A, B, 16))]


@pytest.mark.parametrize("name,f", FORMS)
def test_nonfinite_lines_give_nan_rows_and_columns(name, f):
    rnd = np.random.default_rng(3)
    m, n, k = 40, 30, 50
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    A[3, 5] = np.nan
    A[10, 0] = -np.inf
    B[7, 2] = np.inf
    B[0, 9] = np.nan
    C = f(A, B)
    bad_r, bad_c = [3, 10], [2, 9]
    assert np.all(np.isnan(C[bad_r, :])) and np.all(np.isnan(C[:, bad_c]))
    ok_r = [i for i in range(m) if i not in bad_r]
    ok_c = [j for j in range(n) if j not in bad_c]
    A0, B0 = A.copy(), B.copy()
    A0[bad_r, :] = 0.0
    B0[:, bad_c] = 0.0
    ref = f(A0, B0)
    sub = np.ix_(ok_r, ok_c)
    assert np.all(np.isfinite(C[sub]))
    assert np.array_equal(C[sub], ref[sub])
    assert np.allclose(C[sub], (A0 @ B0)[sub], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name,f", FORMS)
def test_scales_near_the_exponent_range_ends(name, f):
    """Rows of A scaled by 2^1000 and columns of B by 2^-1000: the product is
    O(1); undoing the two scales one after the other would overflow."""
    rnd = np.random.default_rng(4)
    m, n, k = 24, 20, 64
    A = np.ldexp(rnd.uniform(-1, 1, (m, k)), 1000)
    B = np.ldexp(rnd.uniform(-1, 1, (k, n)), -1000)
    C = f(A, B)
    assert np.all(np.isfinite(C))
    ii, jj = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    Af, Bf = np.asfortranarray(A), np.asfortranarray(B)
    ref, ab = O.dd_sample(0, 0, m, n, k, Af, m, Bf, k, ii.ravel().astype(np.int32), jj.ravel().astype(np.int32))
    err = np.abs(C[ii.ravel(), jj.ravel()] - ref)
    assert np.all(err <= k * 2.0 ** -53 * ab), float(np.max(err / ab))
    # the same product with both scales removed gives the same bits
    C0 = f(np.ldexp(A, -1000), np.ldexp(B, 1000))
    assert np.array_equal(C, C0)
