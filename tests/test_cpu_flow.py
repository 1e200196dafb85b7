"""CPU: the CPU restatement of the reference's exec sequence used as the
bench's CPU baseline (oracle/cpu_baseline.py) computes the product: exact on
integer-valued inputs over grids with allgathers, Cannon rounds (q = 2) and
kp-layer sums, for all four transA/transB cases, and the example_AB
linear-fill validation (examples/example_AB.c:208-305) finds 0 errors."""
import numpy as np
import pytest

from oracle import cpu_baseline as CB


@pytest.mark.parametrize("m,n,k,p,ta,tb,grid", [
    (300, 200, 250, 4, 0, 0, None),          # 2 x 1 x 2: B allgathered, kp = 2
    (301, 203, 257, 24, 0, 0, [4, 2, 3]),    # Cannon q = 2, kp = 3, uneven blocks
    (301, 203, 257, 12, 1, 1, [2, 2, 3]),
    (150, 220, 170, 6, 0, 1, None),
    (200, 90, 300, 8, 1, 0, [1, 4, 2]),      # A allgathered
    (64, 64, 64, 5, 0, 0, [2, 2, 1]),        # one idle rank
])
def test_flow_exact_on_integers(m, n, k, p, ta, tb, grid):
    fl = CB.Flow(m, n, k, ta, tb, p, proc_grid=grid)
    rnd = np.random.default_rng(m + n + k + p)
    A = np.asfortranarray(rnd.integers(-3, 4, (k, m) if ta else (m, k)).astype(float))
    B = np.asfortranarray(rnd.integers(-3, 4, (n, k) if tb else (k, n)).astype(float))
    Ab, Bb, Cb = fl.blocks(A, B)
    fl.exec(Ab, Bb, Cb)
    want = (A.T if ta else A) @ (B.T if tb else B)
    assert np.array_equal(fl.gather_C(Cb), want)
    ms = fl.stat_ms()
    assert ms["exec"] >= ms["cannon"] > 0


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_linear_fill_closed_form_check(ta, tb):
    """The closed form finds 0 errors on a correct product and counts a
    planted one-ulp-scale error of 1e-9 relative."""
    m, n, k = 120, 90, 70
    A = CB.linear_fill(k if ta else m, m if ta else k, 0.19, 0.24)
    B = CB.linear_fill(n if tb else k, k if tb else n, 0.11, 0.12)
    C = (A.T if ta else A) @ (B.T if tb else B)
    assert CB.linear_fill_check(m, n, k, ta, tb, C) == 0
    C[5, 7] *= 1 + 1e-9
    assert CB.linear_fill_check(m, n, k, ta, tb, C) == 1


def test_run_example_ab_small():
    r = CB.run_example_ab(200, 180, 160, 4, n_test=2)
    assert r["grid"] == [2, 2, 1]        # the grid partition.c picks for 200 x 180 x 160 on 4 ranks
    assert r["errors"] == 0 and r["tflops"] > 0
