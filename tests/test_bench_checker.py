"""CPU: the bench's own double-double yardstick (bench.dd_sample) against
the oracle's compiled double-double sampler, on ill-conditioned sums and all
four trans cases."""
import numpy as np
import pytest

import bench
from oracle import oracle as O


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_bench_dd_sample_matches_the_oracle(ta, tb):
    rnd = np.random.default_rng(40 + 2 * ta + tb)
    m, n, k = 37, 29, 1001
    sa, sb = ((k, m) if ta else (m, k)), ((n, k) if tb else (k, n))
    A = bench.phi_matrix(rnd, sa, 4.0)
    B = bench.phi_matrix(rnd, sb, 4.0)
    # cancellation: make op(A) row 0 . op(B) column 0 nearly zero
    opA = A.T if ta else A
    opB = B.T if tb else B
    col = opB[:, 0].copy()
    opA[0, -1] = 0.0
    opA[0, -1] = -(opA[0] @ col) / col[-1]
    ii = np.concatenate([[0], rnd.integers(0, m, 200)])
    jj = np.concatenate([[0], rnd.integers(0, n, 200)])
    ref, ab = bench.dd_sample(ta, tb, A, B, ii, jj, chunk=16)
    want, wab = O.dd_sample(ta, tb, m, n, k, np.asfortranarray(A), sa[0], np.asfortranarray(B), sb[0], ii, jj)
    assert np.all(np.abs(ref - want) <= 2.0 ** -100 * wab), np.max(np.abs(ref - want) / wab)
    assert np.allclose(ab, wab, rtol=1e-13, atol=0)
    # the sums differ from the plain double dot products where it matters
    plain = np.einsum("ij,ji->i", opA[ii], opB[:, jj])
    assert np.max(np.abs(plain - want) / wab) > 0.0


def test_suite_plans_are_the_oracles():
    """The suites take their redistribution workloads from the library's own
    engines (bench_suites.grid_redist); they equal the oracle's restatement
    of the reference's plans for the 4x2x3 grid on 8000^3, and col1d is the
    oracle's 1D column layout."""
    import bench_suites as S
    w = O.ca3dmm_world(8000, 8000, 8000, 0, 0, 24, proc_grid=[4, 2, 3])
    assert S.col1d(8000, 8000, 24) == [tuple(int(v) for v in r) for r in w["src_A"]]
    got = S.grid_redist(8000, 8000, 8000, 24, proc_grid=[4, 2, 3])
    for name in ("A", "B"):
        red = w["redist"][name]
        for r in range(24):
            assert tuple(got[name]["src"][r]) == tuple(int(v) for v in red["src"][r]), (name, r)
            assert tuple(got[name]["req"][r]) == tuple(int(v) for v in red["req"][r]), (name, r)
            sl, sc, rl, rc = red["plans"][r]
            gsl, gsc, grl, grc = got[name]["plans"][r]
            assert gsc == sc and grc == rc, (name, r)
            assert [list(map(int, e)) for e in sl] == gsl, (name, r)
            assert [list(map(int, e)) for e in rl] == grl, (name, r)
