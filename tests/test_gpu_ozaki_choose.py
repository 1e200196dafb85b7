"""GPU: hip_ozaki_choose picks the slice and modulus counts of the oracle's
rule (oracle/ozaki.py: spread, counts_from_spread) from the device's own
sampled statistics, and the counts it picks reach the FP64 MFMA DGEMM's
errors on phi = 0.5, 2 and 4 inputs."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z

pytestmark = pytest.mark.gpu


def _near_boundary(L, Ls, k):
    """True when a rounding difference in the statistics could move a count."""
    T = [Z.crt_T(N) for N in range(Z.CRT_MAX + 1)]
    off = 54.0 + (2.33 * np.log2(k / 1024.0) if k > 1024 else 0.0)
    c = 53.5 + (0.5 * np.log2(k / 1024.0) if k > 1024 else 0.0)
    return (np.isfinite(L) and min(abs(t - (L + c)) for t in T) < 1e-6) or \
           (np.isfinite(Ls) and min(abs(8 * s - (Ls + off)) for s in range(1, 17)) < 1e-6)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("m,n,k,phi", [(300, 200, 500, 2.0), (64, 4100, 96, 4.0), (5000, 7, 1500, 0.5),
                                     (1, 1, 3, 2.0), (129, 257, 2049, 4.0)])
def test_choose_matches_oracle_rule(hip, ta, tb, m, n, k, phi):
    rnd = np.random.default_rng(m + 3 * n + k + 7 * ta + 11 * tb)
    opA, opB = Z.phi_matrix(rnd, (m, k), phi), Z.phi_matrix(rnd, (k, n), phi)
    dA = hip.DeviceArray.from_numpy(opA.T if ta else opA)
    dB = hip.DeviceArray.from_numpy(opB.T if tb else opB)
    L, Ls = Z.spread(opA, opB)
    assert not _near_boundary(L, Ls, k)
    assert hip.ozaki_choose(ta, tb, m, n, k, dA, dB) == Z.counts_from_spread(L, Ls, k)


def test_choose_edge_cases(hip):
    m, k, n = 40, 64, 30
    A = np.zeros((m, k))
    B = np.zeros((k, n))
    A[:, :32] = 1.0
    B[32:, :] = 1.0                                   # disjoint supports: the largest counts
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    assert hip.ozaki_choose(0, 0, m, n, k, dA, dB) == (16, 25)
    dZ = hip.DeviceArray.from_numpy(np.zeros((m, k)))
    assert hip.ozaki_choose(0, 0, m, n, k, dZ, dB) == Z.counts_from_spread(0.0, 0.0, k)
    assert hip.ozaki_choose(0, 0, m, n, 0, dA, dB) == Z.counts_from_spread(0.0, 0.0, 0)
    with pytest.raises(Exception):
        hip.ozaki_choose(0, 0, m, n, k, dA, dB, lda=m - 1)


@pytest.mark.parametrize("phi", [0.5, 2.0, 4.0])
def test_chosen_counts_reach_native_accuracy(hip, phi):
    rnd = np.random.default_rng(int(phi * 10) + 1)
    n = 1024
    A, B = Z.phi_matrix(rnd, (n, n), phi), Z.phi_matrix(rnd, (n, n), phi)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.zeros((n, n))
    ii, jj = rnd.integers(0, n, 2048), rnd.integers(0, n, 2048)
    ref, ab = O.dd_sample(0, 0, n, n, n, np.asfortranarray(A), n, np.asfortranarray(B), n, ii, jj)

    def errs():
        C = dC.numpy()[ii, jj]
        return np.max(np.abs(C - ref) / ab), np.max(np.abs(C - ref) / np.abs(ref))

    hip.dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC)
    nat = errs()
    s, N = hip.ozaki_choose(0, 0, n, n, n, dA, dB)
    hip.ozaki2_dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC, N)
    crt = errs()
    hip.ozaki_dgemm(0, 0, n, n, n, 1.0, dA, dB, 0.0, dC, s)
    sl = errs()
    print(f"phi={phi}: slices {s}, moduli {N}, native {nat}, CRT {crt}, slices {sl}", flush=True)
    assert crt[0] <= nat[0] and crt[1] <= nat[1], (N, crt, nat)
    assert sl[0] <= nat[0] and sl[1] <= nat[1], (s, sl, nat)
