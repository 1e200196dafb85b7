"""GPU: the simulated grid's automatic modulus count (set_ozaki(-1)).  The
count and the spread statistic it reads are the oracle's rule
(oracle/ozaki.py: grid_spread, counts_from_spread) on the same operands;
the product is then bit for bit the restated grid at that count; on one rank
the count is hip_ozaki_choose's; a count that changes with the data between
execs rebuilds the plan and the captured hipGraphs; and the chosen count
reaches the FP64 MFMA form's error."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ozaki as Z
from tests.simutil import ozaki_grid_reference, run_sim

pytestmark = pytest.mark.gpu


def _phi(rnd, shape, phi):
    return np.asfortranarray(Z.phi_matrix(rnd, shape, phi))


def _layers(m, n, k, p, grid):
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    kp = w["plans"][0].task_k_num
    return [O.block_spos_size(k, kp, l) for l in range(kp)]


def _oracle_count(opA, opB, layers):
    L, kmax = Z.grid_spread(opA, opB, layers)
    c = 53.5 + (0.5 * np.log2(kmax / 1024.0) if kmax > 1024 else 0.0)
    # a rounding difference in L must not be able to move the count
    assert not np.isfinite(L) or min(abs(Z.crt_T(N) - (L + c)) for N in range(Z.CRT_MIN, Z.CRT_MAX + 1)) > 1e-6
    return L, Z.counts_from_spread(L, 0.0, kmax)[1]


def _layer_spread(rnd, m, n, k, layers, l):
    """phi = 0.5 operands whose products in k layer l pair large A entries
    with small B entries and the reverse (a wide spread in that layer only)."""
    A, B = _phi(rnd, (m, k), 0.5), _phi(rnd, (k, n), 0.5)
    k0, kn = layers[l]
    h = k0 + kn // 2
    A[:, k0:h] *= 2.0 ** 20
    A[:, h:k0 + kn] *= 2.0 ** -20
    B[k0:h, :] *= 2.0 ** -20
    B[h:k0 + kn, :] *= 2.0 ** 20
    return A, B


@pytest.mark.parametrize("p,grid,ta,tb,phi", [(24, [4, 2, 3], 0, 0, 0.5), (24, [4, 2, 3], 1, 0, 2.0),
                                              (8, [2, 2, 2], 0, 1, 4.0), (12, [2, 3, 2], 1, 1, "layer"),
                                              (6, [2, 3, 1], 0, 0, 2.0)])
def test_auto_count_is_the_oracle_rule_and_bits(hip, p, grid, ta, tb, phi):
    rnd = np.random.default_rng(7000 + p + 2 * ta + tb)
    m, n, k = 700, 530, 900
    layers = _layers(m, n, k, p, grid)
    if phi == "layer":
        opA, opB = _layer_spread(rnd, m, n, k, layers, len(layers) - 1)
    else:
        opA, opB = _phi(rnd, (m, k), phi), _phi(rnd, (k, n), phi)
    A = np.asfortranarray(opA.T) if ta else opA
    B = np.asfortranarray(opB.T) if tb else opB
    print(f"[auto] p={p} grid={grid} ta={ta} tb={tb} phi={phi}", flush=True)
    got, sim, w, keep = run_sim(p, m, n, k, ta, tb, A, B, proc_grid=grid, grouped=True, ozaki=-1, keep=True)
    N, auto, L = sim.get_ozaki()
    sim.free()
    Lo, No = _oracle_count(opA, opB, layers)
    assert auto and L == pytest.approx(Lo, rel=1e-9, abs=1e-9), (L, Lo)
    assert N == No, (N, No, L)
    assert np.array_equal(got, ozaki_grid_reference(m, n, k, p, grid, opA, opB, N))


def test_auto_count_on_one_rank_is_hip_ozaki_choose(hip):
    rnd = np.random.default_rng(7100)
    m, n, k = 600, 700, 2500
    A, B = _phi(rnd, (m, k), 4.0), _phi(rnd, (k, n), 4.0)
    print("[auto] one rank, 600 x 700 x 2500, phi 4", flush=True)
    got, sim, w, keep = run_sim(1, m, n, k, 0, 0, A, B, proc_grid=[1, 1, 1], ozaki=-1, keep=True)
    N, auto, L = sim.get_ozaki()
    sim.free()
    dA, dB = keep[7], keep[8]
    assert hip.ozaki_choose(0, 0, m, n, k, dA, dB)[1] == N
    assert L == pytest.approx(Z.spread(A, B)[0], rel=1e-9, abs=1e-9)
    assert np.array_equal(got, ozaki_grid_reference(m, n, k, 1, [1, 1, 1], A, B, N))


@pytest.mark.parametrize("graph", [False, True])
def test_auto_count_follows_the_data(hip, graph):
    """Three execs of one grid: phi = 0.5, then the same shapes with a wide
    spread in one k layer (more moduli), then phi = 0.5 again (the first
    count back); each result is the restated grid at its count."""
    rnd = np.random.default_rng(7200 + graph)
    p, grid = 24, [4, 2, 3]
    m, n, k = 640, 520, 1500
    layers = _layers(m, n, k, p, grid)
    data = [(_phi(rnd, (m, k), 0.5), _phi(rnd, (k, n), 0.5)), _layer_spread(rnd, m, n, k, layers, 1)]
    data.append(data[0])
    print(f"[auto] data changes between execs, graph={graph}", flush=True)
    got, sim, w, keep = run_sim(p, m, n, k, 0, 0, data[0][0], data[0][1], proc_grid=grid, grouped=True,
                                graph=graph, ozaki=-1, keep=True)
    Ap, ldA, Bp, ldB, Cp, ldC, dC, dA, dB = keep
    counts = []
    for step, (A, B) in enumerate(data):
        if step:
            dA.upload(A)
            dB.upload(B)
            sim.exec(Ap, ldA, Bp, ldB, Cp, ldC)
            got = dC.numpy()
        N = sim.get_ozaki()[0]
        counts.append(N)
        assert N == _oracle_count(A, B, layers)[1], (step, N)
        assert np.array_equal(got, ozaki_grid_reference(m, n, k, p, grid, A, B, N)), step
    sim.free()
    assert counts[1] > counts[0] == counts[2], counts


def test_auto_count_reaches_the_fp64_form_error(hip):
    rnd = np.random.default_rng(7300)
    m, n, k = 640, 520, 1500
    A, B = _phi(rnd, (m, k), 2.0), _phi(rnd, (k, n), 2.0)
    ii, jj = rnd.integers(0, m, 2000), rnd.integers(0, n, 2000)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    print("[auto] accuracy against the FP64 form, 24 ranks, phi 2", flush=True)
    fp = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True)
    oz = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, ozaki=-1)
    for e_oz, e_fp in ((np.max(np.abs(oz[ii, jj] - ref) / ab), np.max(np.abs(fp[ii, jj] - ref) / ab)),
                       (np.max(np.abs(oz[ii, jj] - ref) / np.abs(ref)), np.max(np.abs(fp[ii, jj] - ref) / np.abs(ref)))):
        assert e_oz <= e_fp, (e_oz, e_fp)


def test_set_ozaki_arguments(hip):
    rnd = np.random.default_rng(7400)
    A, B = _phi(rnd, (300, 200), 0.5), _phi(rnd, (200, 250), 0.5)
    got, sim, w, keep = run_sim(4, 300, 250, 200, 0, 0, A, B, proc_grid=[2, 2, 1], keep=True)
    assert sim.get_ozaki()[0] == 0 and np.isnan(sim.get_ozaki()[2])
    for bad in (-2, 3, 26):
        with pytest.raises(Exception):
            sim.set_ozaki(bad)
    sim.set_ozaki(-1)
    assert sim.get_ozaki()[:2] == (0, True)
    with pytest.raises(Exception):
        sim.autotune_ozaki(1)                       # no count before the first exec
    sim.set_ozaki(15)
    assert sim.get_ozaki()[:2] == (15, False) and np.isnan(sim.get_ozaki()[2])
    sim.free()
