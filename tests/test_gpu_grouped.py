"""GPU: the persistent grouped launch of the simulated grid -- all ranks'
Cannon rounds and the k-layer fold in one kernel from a device work queue,
optionally replayed from a hipGraph -- against the oracle."""
import numpy as np
import pytest

from tests.simutil import run_sim

pytestmark = pytest.mark.gpu


def _ints(rnd, shape):
    return np.asfortranarray(rnd.integers(-4, 5, shape).astype(np.float64))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_grouped_24_ranks_exact(hip, ta, tb):
    """4x2x3 grid on 1001^3 (uneven cuboids, q = 2, kp = 3): bit-exact."""
    rnd = np.random.default_rng(30 + ta)
    m = n = k = 1001
    A = _ints(rnd, (k, m) if ta else (m, k))
    B = _ints(rnd, (n, k) if tb else (k, n))
    got = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3], grouped=True)
    assert np.array_equal(got, (A.T if ta else A) @ (B.T if tb else B))


def test_grouped_equals_per_launch_on_integers(hip):
    rnd = np.random.default_rng(31)
    m, n, k = 700, 500, 900
    A, B = _ints(rnd, (m, k)), _ints(rnd, (k, n))
    a = run_sim(12, m, n, k, 0, 0, A, B, proc_grid=[2, 2, 3])
    b = run_sim(12, m, n, k, 0, 0, A, B, proc_grid=[2, 2, 3], grouped=True, autotune=True)
    assert np.array_equal(a, b)


def test_grouped_fp_within_bound(hip):
    rnd = np.random.default_rng(32)
    m, n, k = 640, 520, 1500
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    got = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True)
    bound = k * 2.0 ** -53 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(got - A @ B) <= bound + 2.0 ** -53 * np.abs(A @ B))


def test_grouped_graph_replays(hip):
    """The captured graph is replayed several times with eager work and
    synchronisation in between; every replay equals the eager result."""
    from ca3dmm_amd import device as dev
    rnd = np.random.default_rng(33)
    m, n, k = 520, 610, 730
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    eager = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, graph=False)
    got, sim, w, args = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, graph=True,
                                keep=True)
    Ap, ldA, Bp, ldB, Cp, ldC, dC, dA, dB = args     # dA, dB own the memory Ap, Bp point into
    assert np.array_equal(got, eager)
    junk = dev.DeviceArray.zeros((1000, 1000))
    for rep in range(4):
        dev.dgemm(0, 0, 1000, 1000, 1000, 1.0, junk, junk, 0.0, junk)   # eager work in between
        dev.sync()
        dev.fold_sum([junk, junk], 1000, 1000, junk)
        sim.exec(Ap, ldA, Bp, ldB, Cp, ldC)
        assert np.array_equal(dC.numpy(), eager), rep
    sim.free()
