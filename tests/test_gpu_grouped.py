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


def _phi(rnd, shape, phi):
    from oracle import ozaki as Z
    return np.asfortranarray(Z.phi_matrix(rnd, shape, phi))


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_grouped_ozaki_24_ranks_bit_exact(hip, ta, tb):
    """The Ozaki (CRT) form of the 4x2x3 grid on 1001^3: every rank's
    product over its two Cannon rounds in one persistent launch, the kp = 3
    fold on the reconstructed values, bit-exact against the restated grid."""
    from tests.simutil import ozaki_grid_reference
    rnd = np.random.default_rng(40 + ta)
    m = n = k = 1001
    A = _phi(rnd, (k, m) if ta else (m, k), 2.0)
    B = _phi(rnd, (n, k) if tb else (k, n), 2.0)
    opA, opB = (A.T if ta else A), (B.T if tb else B)
    want = ozaki_grid_reference(m, n, k, 24, [4, 2, 3], opA, opB, 15)
    got = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3], grouped=True, ozaki=15)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("grid,p,N", [([2, 2, 3], 12, 8), ([2, 3, 1], 6, 16), ([1, 1, 1], 1, 20)])
def test_grouped_ozaki_per_rank_graph_and_grids(hip, grid, p, N):
    """Per-rank product launches, the grouped launch and its hipGraph replay
    (replayed three times) give the same bits, on grids with and without a
    k-layer fold."""
    from tests.simutil import ozaki_grid_reference
    rnd = np.random.default_rng(41 + p)
    m, n, k = 700, 530, 900
    A, B = _phi(rnd, (m, k), 0.5), _phi(rnd, (k, n), 0.5)
    want = ozaki_grid_reference(m, n, k, p, grid, A, B, N)
    a = run_sim(p, m, n, k, 0, 0, A, B, proc_grid=grid, ozaki=N)
    b = run_sim(p, m, n, k, 0, 0, A, B, proc_grid=grid, grouped=True, ozaki=N)
    c = run_sim(p, m, n, k, 0, 0, A, B, proc_grid=grid, grouped=True, graph=True, ozaki=N, repeat=3)
    assert np.array_equal(a, want) and np.array_equal(b, want) and np.array_equal(c, want)


def test_grouped_ozaki_accuracy_vs_fp64_form(hip):
    """16 moduli on the 24-rank grid: the error against the exact product
    is no larger than the FP64 MFMA form's (integer-free inputs, phi = 0.5)."""
    from oracle import oracle as O
    rnd = np.random.default_rng(44)
    m, n, k = 640, 520, 1500
    A, B = _phi(rnd, (m, k), 0.5), _phi(rnd, (k, n), 0.5)
    ii, jj = rnd.integers(0, m, 2000), rnd.integers(0, n, 2000)
    ref, ab = O.dd_sample(0, 0, m, n, k, A, m, B, k, ii, jj)
    fp = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True)
    oz = run_sim(24, m, n, k, 0, 0, A, B, proc_grid=[4, 2, 3], grouped=True, ozaki=16)
    e_fp = np.max(np.abs(fp[ii, jj] - ref) / ab)
    e_oz = np.max(np.abs(oz[ii, jj] - ref) / ab)
    assert e_oz <= e_fp, (e_oz, e_fp)


def test_grouped_ozaki_every_tuning_bit_exact(hip):
    """The Ozaki form's grouped launch under each tuning (256 x 256 / 256 x
    128 tiles, per-XCD / single queue) and after autotuning, per-rank and
    grouped, against the restated grid."""
    from ca3dmm_amd import engine as E
    from oracle import oracle as O
    from tests.simutil import ozaki_grid_reference
    rnd = np.random.default_rng(45)
    m, n, k, p, grid, N = 650, 590, 700, 12, [2, 2, 3], 15
    A, B = _phi(rnd, (m, k), 1.0), _phi(rnd, (k, n), 1.0)
    want = ozaki_grid_reference(m, n, k, p, grid, A, B, N)
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((m, n), 0xFF)
    sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
    Ap = [dA.at(r[2] * m) for r in w["src_A"]]
    Bp = [dB.at(r[2] * k) for r in w["src_B"]]
    Cp = [dC.at(r[2] * m) for r in w["dst_C"]]
    sim.set_ozaki(N)
    for tile_n in (256, 128):
        for xq in (1, 0):
            sim.set_tuning_ozaki(tile_n, xq)
            for grouped in (False, True):
                sim.set_mode(grouped=grouped, graph=False)
                sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
                assert np.array_equal(dC.numpy(), want), (tile_n, xq, grouped)
    sim.set_mode(grouped=True, graph=True)
    t, q = sim.autotune_ozaki(1)
    assert t in (256, 128) and q in (0, 1)
    for _ in range(2):
        sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
        assert np.array_equal(dC.numpy(), want)
    sim.free()


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_redistribution_one_pass_equals_staged(hip, ta, tb, monkeypatch):
    """The device-resident redistribution (each block straight from the
    sender's source block to the receiver's destination, A and B in one
    launch) gives the same bits as the staged pack / exchange / unpack of
    src/mat_redist.c:317-416 (CA3DMM_REDIST_STAGED=1), on the 24-rank grid
    and on the single-rank engine's forced redistribution."""
    rnd = np.random.default_rng(50 + 2 * ta + tb)
    m, n, k = 601, 433, 517
    A = rnd.uniform(-1, 1, (k, m) if ta else (m, k))
    B = rnd.uniform(-1, 1, (n, k) if tb else (k, n))
    monkeypatch.setenv("CA3DMM_REDIST_STAGED", "1")
    staged = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3])
    monkeypatch.setenv("CA3DMM_REDIST_STAGED", "0")
    direct = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3])
    assert np.array_equal(staged, direct)
    graph = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3], grouped=True, graph=True, repeat=2)
    bound = k * 2.0 ** -53 * (np.abs(A.T if ta else A) @ np.abs(B.T if tb else B))
    assert np.all(np.abs(graph - (A.T if ta else A) @ (B.T if tb else B)) <= bound + 1e-300)


def test_sequence_graph_recaptured_for_new_arguments(hip):
    """The whole exec sequence replayed from its hipGraph follows new input
    buffers (a new capture) and a new tuning, bit-exact on integers."""
    from ca3dmm_amd import device as dev
    from ca3dmm_amd import engine as E
    from oracle import oracle as O
    rnd = np.random.default_rng(60)
    m, n, k, p, grid = 500, 460, 700, 24, [4, 2, 3]
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
    sim.set_mode(grouped=True, graph=True)
    dC = dev.DeviceArray.filled((m, n), 0xFF)
    Cp = [dC.at(r[2] * m) for r in w["dst_C"]]
    for it in range(3):
        A, B = _ints(rnd, (m, k)), _ints(rnd, (k, n))
        dA, dB = dev.DeviceArray.from_numpy(A), dev.DeviceArray.from_numpy(B)     # new buffers each time
        Ap = [dA.at(r[2] * m) for r in w["src_A"]]
        Bp = [dB.at(r[2] * k) for r in w["src_B"]]
        if it == 2:
            sim.set_tuning(4, 1)
        for _ in range(2):
            sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
            assert np.array_equal(dC.numpy(), A @ B), it
        del dA, dB
    sim.free()


@pytest.mark.parametrize("tile,wg", [(0, 1), (0, 2), (1, 1), (2, 1), (3, 2)])
def test_grouped_every_tile_shape_exact(hip, tile, wg):
    """The grouped launch under each workgroup tile (128x128 with one or two
    workgroups per CU, 256x128, 128x256, the batched direct-to-LDS launch
    with fold launches): bit-exact on integers over a grid with kp = 3
    folds and uneven cuboids."""
    from ca3dmm_amd import device as dev
    from ca3dmm_amd import engine as E
    from oracle import oracle as O
    rnd = np.random.default_rng(80 + tile)
    m, n, k, p, grid = 1001, 777, 905, 24, [4, 2, 3]
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
    A, B = _ints(rnd, (m, k)), _ints(rnd, (k, n))
    dA, dB = dev.DeviceArray.from_numpy(A), dev.DeviceArray.from_numpy(B)
    dC = dev.DeviceArray.filled((m, n), 0xFF)
    Ap = [dA.at(r[2] * m) for r in w["src_A"]]
    Bp = [dB.at(r[2] * k) for r in w["src_B"]]
    Cp = [dC.at(r[2] * m) for r in w["dst_C"]]
    sim.set_mode(grouped=True, graph=False)
    sim.set_tuning(4, wg, tile)
    sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
    assert np.array_equal(dC.numpy(), A @ B)
    sim.free()


def test_autotune_table_persists(hip, tmp_path, monkeypatch):
    """The tuner's pick is written to the tuning table and a later plan of the
    same shape starts from it (here: a second simulated grid in the same
    process reads the table)."""
    monkeypatch.setenv("CA3DMM_TUNE_CACHE", str(tmp_path / "tuning.txt"))
    from ca3dmm_amd import engine as E
    from oracle import oracle as O
    m, n, k, p, grid = 640, 520, 700, 12, [2, 2, 3]
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
    g, wg, tile = sim.autotune(1)
    sim.free()
    lines = (tmp_path / "tuning.txt").read_text().splitlines()
    fp64 = [l for l in lines if l.startswith("fp64 ") and "640x" not in l[:5]]
    assert any(l.endswith(f"tile={tile} rows={g} wg={wg}") for l in fp64), lines


@pytest.mark.parametrize("grid,p,mnk", [([4, 2, 3], 24, (1001, 777, 1999)), ([2, 3, 2], 12, (640, 900, 1333)),
                                        ([2, 3, 1], 6, (700, 530, 900)), ([2, 1, 3], 7, (130, 77, 210))])
def test_grouped_batched_candidate_same_bits_as_persistent(hip, grid, p, mnk):
    """Random FP64 inputs: the batched direct-to-LDS candidate (tile 3,
    rounds in registers, fold launches) stores the bits of the persistent
    register-staged kernel (tile 0), eagerly and replayed from the exec's
    hipGraph, with and without k-layer folds, idle ranks included."""
    from ca3dmm_amd import device as dev
    from ca3dmm_amd import engine as E
    from oracle import oracle as O
    m, n, k = mnk
    rnd = np.random.default_rng(90 + p)
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    dA, dB = dev.DeviceArray.from_numpy(A), dev.DeviceArray.from_numpy(B)
    Ap = [dA.at(r[2] * m) for r in w["src_A"]]
    Bp = [dB.at(r[2] * k) for r in w["src_B"]]
    out = {}
    print(f"[grouped] batched candidate p={p} grid={grid} mnk={mnk}", flush=True)
    for tile, wg, graph in ((0, 2, False), (3, 2, False), (3, 2, True)):
        sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
        dC = dev.DeviceArray.filled((m, n), 0xFF)
        Cp = [dC.at(r[2] * m) for r in w["dst_C"]]
        sim.set_mode(grouped=True, graph=graph)
        sim.set_tuning(8, wg, tile)
        for _ in range(2 if graph else 1):
            sim.exec(Ap, [m] * p, Bp, [k] * p, Cp, [m] * p)
        out[(tile, graph)] = dC.numpy()
        sim.free()
    ref = out[(0, False)]
    assert np.array_equal(out[(3, False)], ref) and np.array_equal(out[(3, True)], ref)
    bound = k * 2.0 ** -53 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(ref - A @ B) <= bound + 2.0 ** -53 * np.abs(A @ B))
