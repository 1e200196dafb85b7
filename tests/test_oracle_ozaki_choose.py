"""CPU: the slice / modulus count chooser.  The spread statistics of the
oracle (oracle/ozaki.py, spread) are pinned by inputs whose value is known
in closed form; the library's count rule (hip_ozaki_counts, host code, no
device work) equals the oracle's on a grid of statistics; the rule maps the
calibration runs' statistics to the counts the device measured."""
import ctypes

import numpy as np
import pytest

from oracle import ozaki as Z


def _lib_counts(L, Ls, k):
    from ca3dmm_amd import LIB_PATH
    lib = ctypes.CDLL(LIB_PATH)
    lib.hip_ozaki_counts.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    s, N = ctypes.c_int(0), ctypes.c_int(0)
    assert lib.hip_ozaki_counts(L, Ls, k, ctypes.byref(s), ctypes.byref(N)) == 0
    return s.value, N.value


def test_sample_indices_follow_the_formula():
    i, j = Z.spread_samples(1000, 777)
    for s in (0, 1, 2, 4095):
        assert i[s] == ((s * 2654435761 + 2135587861) % 2 ** 32) % 1000
        assert j[s] == ((s * 2246822519 + 3266489917) % 2 ** 32) % 777
    assert len(i) == Z.SPREAD_SAMPLES == 4096


def test_spread_closed_forms():
    m, k, n = 40, 64, 30
    # all-ones: a.b = k = ||a|| ||b||, (max(a)|b|_1 + max(b)|a|_1) / a.b = 2
    L, Ls = Z.spread(np.ones((m, k)), -np.ones((k, n)))
    assert L == pytest.approx(0.0, abs=1e-12) and Ls == pytest.approx(1.0, abs=1e-12)
    # one nonzero per line on the same k: L = 0, L_s = 1 whatever the values
    A = np.zeros((m, k))
    B = np.zeros((k, n))
    A[:, 5] = np.arange(1, m + 1) * 2.0 ** -30
    B[5, :] = -np.arange(1, n + 1) * 2.0 ** 40
    L, Ls = Z.spread(A, B)
    assert L == pytest.approx(0.0, abs=1e-12) and Ls == pytest.approx(1.0, abs=1e-12)
    # disjoint supports: every sampled a.b is 0 while a and b are not
    A = np.zeros((m, k))
    B = np.zeros((k, n))
    A[:, :32] = 1.0
    B[32:, :] = 1.0
    assert Z.spread(A, B) == (np.inf, np.inf)
    # empty operands carry no spread
    assert Z.spread(np.zeros((m, k)), np.ones((k, n))) == (0.0, 0.0)


def test_spread_scale_invariant_and_grows_with_phi():
    rnd = np.random.default_rng(5)
    A, B = Z.phi_matrix(rnd, (64, 512), 2.0), Z.phi_matrix(rnd, (512, 48), 2.0)
    L, Ls = Z.spread(A, B)
    L2, Ls2 = Z.spread(A * 2.0 ** 100, B * 2.0 ** -70)
    assert L2 == pytest.approx(L, abs=1e-9) and Ls2 == pytest.approx(Ls, abs=1e-9)
    lows = [Z.spread(Z.phi_matrix(rnd, (64, 512), p), Z.phi_matrix(rnd, (512, 48), p))[0] for p in (0.5, 2.0, 4.0)]
    assert lows[0] < 2.0 < lows[1] < lows[2]


def test_library_rule_equals_oracle_rule():
    rnd = np.random.default_rng(11)
    Ls_vals = list(rnd.uniform(-3, 90, 60)) + [0.0, np.inf, np.nan]
    L_vals = list(rnd.uniform(-3, 60, 60)) + [0.0, np.inf, np.nan]
    for k in (0, 1, 700, 1024, 1025, 4096, 8192, 65536, 1 << 23):
        for L, Ls in zip(L_vals, Ls_vals):
            want = Z.counts_from_spread(np.inf if np.isnan(L) else L, np.inf if np.isnan(Ls) else Ls, k)
            assert _lib_counts(L, Ls, k) == want, (L, Ls, k)


# Calibration runs: (k, phi, L, L_s) of 8192^3 and 65536 x 65536 x 1024
# inputs (two seeds), and the modulus and slice counts whose sampled errors
# first matched the FP64 MFMA DGEMM's on the device.
CALIBRATION = [
    (8192, 0.5, 0.84, 5.13, 15, 8), (8192, 0.5, 0.83, 5.16, 15, 8),
    (8192, 2.0, 9.56, 13.58, 16, 9), (8192, 2.0, 9.97, 13.67, 16, 9),
    (8192, 4.0, 23.70, 25.18, 20, 11), (8192, 4.0, 23.50, 24.55, 20, 11),
    (1024, 0.5, 0.97, 4.84, 14, 7), (1024, 0.5, 0.96, 4.78, 14, 7),
    (1024, 2.0, 11.35, 13.30, 15, 8), (1024, 2.0, 9.91, 12.75, 15, 8),
    (1024, 4.0, 25.98, 27.04, 19, 10), (1024, 4.0, 22.01, 23.08, 19, 10),
]


@pytest.mark.parametrize("k,phi,L,Ls,N_meas,s_meas", CALIBRATION)
def test_rule_reproduces_the_calibration(k, phi, L, Ls, N_meas, s_meas):
    s, N = Z.counts_from_spread(L, Ls, k)
    assert N_meas <= N <= N_meas + 2          # never fewer moduli than measured
    assert s_meas <= s <= s_meas + 1


def test_grid_spread_one_layer_is_spread():
    rnd = np.random.default_rng(21)
    A, B = Z.phi_matrix(rnd, (90, 300), 2.0), Z.phi_matrix(rnd, (300, 70), 2.0)
    L, kmax = Z.grid_spread(A, B, [(0, 300)])
    assert L == pytest.approx(Z.spread(A, B)[0], abs=1e-12) and kmax == 300


def test_grid_spread_brute_force():
    """Per-sample loops in plain Python on a tiny grid: sample s in layer
    s mod 3, statistics over that layer's k range only."""
    import math
    rnd = np.random.default_rng(22)
    m, n, k, S = 5, 4, 10, 12
    A, B = Z.phi_matrix(rnd, (m, k), 4.0), Z.phi_matrix(rnd, (k, n), 4.0)
    A[3, 4:7] = 0.0                          # row 3 empty in layer 1: its samples there are skipped
    layers = [(0, 4), (4, 3), (7, 3)]
    best = -math.inf
    for s in range(S):
        i = ((s * 2654435761 + 2135587861) % 2 ** 32) % m
        j = ((s * 2246822519 + 3266489917) % 2 ** 32) % n
        k0, kn = layers[s % 3]
        a = [abs(float(A[i, c])) for c in range(k0, k0 + kn)]
        b = [abs(float(B[c, j])) for c in range(k0, k0 + kn)]
        na, nb = math.sqrt(sum(x * x for x in a)), math.sqrt(sum(y * y for y in b))
        if na > 0 and nb > 0:
            best = max(best, math.log2(na * nb / sum(x * y for x, y in zip(a, b))))
    L, kmax = Z.grid_spread(A, B, layers, S=S)
    assert L == pytest.approx(best, abs=1e-12) and kmax == 4


def test_grid_spread_keeps_the_layers_apart():
    """Scaling one layer's k range changes the whole-k spread but not the
    per-layer one; operands whose supports meet only across layers carry no
    spread per layer while the whole-k statistic is infinite."""
    rnd = np.random.default_rng(23)
    m, n, k = 64, 48, 600
    layers = [(0, 200), (200, 200), (400, 200)]
    A, B = Z.phi_matrix(rnd, (m, k), 0.5), Z.phi_matrix(rnd, (k, n), 0.5)
    A2, B2 = A.copy(), B.copy()
    A2[:, 200:400] *= 2.0 ** 30
    B2[200:400, :] *= 2.0 ** -30
    L, _ = Z.grid_spread(A, B, layers)
    assert Z.grid_spread(A2, B2, layers)[0] == pytest.approx(L, abs=1e-9)
    assert Z.spread(A2, B2)[0] > Z.spread(A, B)[0] + 20
    A3, B3 = A.copy(), B.copy()
    A3[:, 200:] = 0.0                        # A lives in layer 0 only, B outside it
    B3[:200, :] = 0.0
    assert Z.spread(A3, B3)[0] == np.inf
    assert Z.grid_spread(A3, B3, layers) == (0.0, 200)
