"""GPU: the CA3DMM engine end to end (src/ca3dmm.c exec sequence) -- single
rank on the device, and simulated p-rank grids on one GPU -- against the
oracle.  Integer-valued inputs make every product exact, so results must be
bit-identical; the linear fill of examples/example_AB.c is checked with the
reference's own criterion (relative error <= 1e-12, zero failures)."""
import ctypes as C

import numpy as np
import pytest

from ca3dmm_amd import device as hip
from ca3dmm_amd import engine as E
from tests.simutil import linear_fill, run_sim, single_engine

pytestmark = pytest.mark.gpu


def _ints(rnd, shape):
    return np.asfortranarray(rnd.integers(-4, 5, shape).astype(np.float64))


@pytest.mark.parametrize("force", [0, 1])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_single_rank_engine_exact(hip, monkeypatch, ta, tb, force):
    monkeypatch.setenv("CA3DMM_FORCE_REDIST", str(force))
    rnd = np.random.default_rng(10 + 2 * ta + tb)
    m, n, k = 333, 257, 400
    A = _ints(rnd, (k, m) if ta else (m, k))
    B = _ints(rnd, (n, k) if tb else (k, n))
    want = (A.T if ta else A) @ (B.T if tb else B)
    eng = single_engine(m, n, k, ta, tb)
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((m, n), 0xFF)
    E.lib.ca3dmm_engine_exec(eng, dA.ptr, A.shape[0], dB.ptr, B.shape[0], dC.ptr, m)
    assert np.array_equal(dC.numpy(), want)
    e = eng.contents
    assert e.n_exec == 1 and e.mp == e.np == e.kp == 1
    E.lib.ca3dmm_engine_free(C.byref(eng))


def test_single_rank_sub_block_output(hip):
    """dst_C asks for rows 10..59, cols 5..24 only (redistribution of C)."""
    rnd = np.random.default_rng(3)
    m, n, k = 100, 40, 70
    A, B = _ints(rnd, (m, k)), _ints(rnd, (k, n))
    eng = single_engine(m, n, k, 0, 0, dst=(10, 50, 5, 20))
    dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((50, 20), 0xFF)
    E.lib.ca3dmm_engine_exec(eng, dA.ptr, m, dB.ptr, k, dC.ptr, 50)
    assert np.array_equal(dC.numpy(), (A @ B)[10:60, 5:25])
    E.lib.ca3dmm_engine_free(C.byref(eng))


def test_single_rank_btb(hip):
    rnd = np.random.default_rng(4)
    n, k = 120, 900
    B = _ints(rnd, (k, n))
    eng = single_engine(n, n, k, 0, 0, btb=True)
    dB = hip.DeviceArray.from_numpy(B)
    dC = hip.DeviceArray.filled((n, n), 0xFF)
    E.lib.ca3dmm_engine_exec(eng, None, 0, dB.ptr, k, dC.ptr, n)
    got = dC.numpy()
    assert np.array_equal(got, B.T @ B)
    assert np.array_equal(got, got.T)
    E.lib.ca3dmm_engine_free(C.byref(eng))


def test_mat_redist_single_rank(hip):
    """A requested window of the source block (row-major view)."""
    rnd = np.random.default_rng(5)
    src = rnd.standard_normal((100, 50))          # row-major rows x cols
    eng = C.POINTER(E.MatRedist)()
    E.lib.mat_redist_engine_init(0, 0, 100, 50, 20, 10, 50, 30, E.comm_self(), 0, 8, E.DEV_TYPE_HIP,
                                 C.byref(eng), None)
    assert eng and eng.contents.n_proc_send == 1 and eng.contents.recv_cnt == 1500
    ds = hip.DeviceArray.from_numpy(src.ravel())
    dd = hip.DeviceArray.zeros((1500,))
    E.lib.mat_redist_engine_exec(eng, ds.ptr, 50, dd.ptr, 30)
    assert np.array_equal(dd.numpy().reshape(50, 30), src[20:70, 10:40])
    E.lib.mat_redist_engine_free(C.byref(eng))


def test_cannon_engine_single(hip):
    rnd = np.random.default_rng(6)
    m, n, k = 90, 70, 50
    A, B, Cm = _ints(rnd, (m, k)), _ints(rnd, (k, n)), _ints(rnd, (m, n))
    ce = C.POINTER(E.Cannon)()
    E.lib.cannon_engine_init(m, n, k, E.comm_self(), E.DEV_TYPE_HIP, C.byref(ce), None)
    dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (A, B, Cm))
    E.lib.cannon_engine_exec(ce, 2.0, 1.0, dA.ptr, dB.ptr, dC.ptr)
    assert np.array_equal(dC.numpy(), 2.0 * (A @ B) + Cm)
    E.lib.cannon_engine_free(C.byref(ce))


def test_sim_example_ab_2000_p4_linear_fill(hip):
    """BASELINE config 0 on the GPU: 4 ranks (grid 2x1x2), 2000^3 NN, the
    example_AB linear fill, validated as example_AB.c:287-305 does."""
    m = n = k = 2000
    A = linear_fill(m, k, 0.19, 0.24)
    B = linear_fill(k, n, 0.11, 0.12)
    got, sim, w, _ = run_sim(4, m, n, k, 0, 0, A, B, keep=True)
    e = sim.engine(0)
    assert (e.mp, e.np, e.kp) == (2, 1, 2)
    ref = A @ B
    errors = int(np.sum(np.abs((got - ref) / ref) > 1e-12))
    assert errors == 0
    sim.free()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_sim_24_ranks_uneven_exact(hip, ta, tb):
    """Grid 4x2x3 (k split in 3 layers, q = 2 Cannon rounds, B replicated over
    2 tasks) on an uneven 1001^3 problem, integer inputs: bit-exact."""
    rnd = np.random.default_rng(20 + 2 * ta + tb)
    m = n = k = 1001
    A = _ints(rnd, (k, m) if ta else (m, k))
    B = _ints(rnd, (n, k) if tb else (k, n))
    got, sim, w, _ = run_sim(24, m, n, k, ta, tb, A, B, proc_grid=[4, 2, 3], keep=True)
    assert np.array_equal(got, (A.T if ta else A) @ (B.T if tb else B))
    sim.free()


def test_sim_random_fp_within_bound(hip):
    """Random FP64 inputs on a 2x2x3 grid: within the componentwise bound
    k*u*(|A||B|) of the exact product."""
    rnd = np.random.default_rng(7)
    m, n, k = 240, 300, 900
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    got = run_sim(12, m, n, k, 0, 0, A, B, proc_grid=[2, 2, 3])
    bound = k * 2.0 ** -53 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(got - A @ B) <= bound + 2.0 ** -53 * np.abs(A @ B))


@pytest.mark.parametrize("p,grid,ta,tb,mnk", [(24, [4, 2, 3], 0, 0, (1001, 777, 1999)), (24, [4, 2, 3], 1, 1, (1001, 777, 1999)),
                                              (12, [2, 3, 2], 0, 1, (640, 900, 1333)), (8, [2, 2, 2], 1, 0, (513, 515, 2050)),
                                              (6, [2, 3, 1], 0, 0, (700, 530, 900)), (1, [1, 1, 1], 1, 0, (300, 260, 333)),
                                              (24, [4, 2, 3], 0, 0, (300, 200, 20)), (7, [2, 1, 3], 0, 0, (130, 77, 210))])
def test_sim_batched_rounds_equal_per_round_launches(hip, monkeypatch, p, grid, ta, tb, mnk):
    """The per-rank Cannon rounds of every rank as one batched launch give
    the bits of one hip_dgemm launch per rank and round (CA3DMM_SIM_BATCH=0),
    eagerly and replayed from a hipGraph, on random FP64 inputs: uneven and
    odd block sizes, k blocks shorter than a stage, idle ranks."""
    m, n, k = mnk
    rnd = np.random.default_rng(m + n + k + p)
    A = rnd.uniform(-1, 1, (k, m) if ta else (m, k))
    B = rnd.uniform(-1, 1, (n, k) if tb else (k, n))
    print(f"[batch] p={p} grid={grid} ta={ta} tb={tb} mnk={mnk}", flush=True)
    monkeypatch.setenv("CA3DMM_SIM_BATCH", "0")
    want = run_sim(p, m, n, k, ta, tb, A, B, proc_grid=grid)
    monkeypatch.delenv("CA3DMM_SIM_BATCH")
    got = run_sim(p, m, n, k, ta, tb, A, B, proc_grid=grid)
    graph = run_sim(p, m, n, k, ta, tb, A, B, proc_grid=grid, graph=True, repeat=3)
    assert np.array_equal(got, want) and np.array_equal(graph, want)
    opA, opB = (A.T if ta else A), (B.T if tb else B)
    bound = k * 2.0 ** -53 * (np.abs(opA) @ np.abs(opB))
    assert np.all(np.abs(got - opA @ opB) <= bound + 2.0 ** -53 * np.abs(opA @ opB))


def test_sim_idle_ranks(hip):
    """p = 7 with grid 2x1x3: rank 6 idles but still takes part in the
    redistribution of its input columns."""
    rnd = np.random.default_rng(8)
    m, n, k = 130, 77, 210
    A, B = _ints(rnd, (m, k)), _ints(rnd, (k, n))
    got, sim, w, _ = run_sim(7, m, n, k, 0, 0, A, B, proc_grid=[2, 1, 3], keep=True)
    assert sim.engine(6).is_active == 0 and sim.engine(0).rp == 1
    assert np.array_equal(got, A @ B)
    sim.free()


@pytest.mark.parametrize("p,pg", [(4, None), (8, [2, 2, 2]), (12, [2, 2, 3])])
def test_sim_btb(hip, p, pg):
    rnd = np.random.default_rng(p)
    n, k = 150, 1200
    B = _ints(rnd, (k, n))
    got = run_sim(p, n, n, k, 0, 0, None, B, proc_grid=pg, is_btb=True)
    assert np.array_equal(got, B.T @ B)


def test_sim_random_grids_batched_paths_same_bits(hip, monkeypatch):
    """Seeded random rank counts, shapes and trans cases on the grid the
    library picks (idle ranks, k blocks shorter than a stage, uneven
    cuboids): the batched per-rank rounds give the bits of one launch per
    rank and round, and the grouped mode's batched candidate (tile 3) those
    of its persistent kernel (tile 0)."""
    from oracle import oracle as O
    rnd = np.random.default_rng(4242)
    for case in range(12):
        p = int(rnd.integers(1, 25))
        m, n, k = (int(v) for v in rnd.integers(1, 600, 3))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        A = rnd.uniform(-1, 1, (k, m) if ta else (m, k))
        B = rnd.uniform(-1, 1, (n, k) if tb else (k, n))
        print(f"[batch] random case {case}: p={p} m={m} n={n} k={k} ta={ta} tb={tb}", flush=True)
        monkeypatch.setenv("CA3DMM_SIM_BATCH", "0")
        per_round = run_sim(p, m, n, k, ta, tb, A, B)
        monkeypatch.delenv("CA3DMM_SIM_BATCH")
        batched = run_sim(p, m, n, k, ta, tb, A, B)
        assert np.array_equal(batched, per_round), case
        w = O.ca3dmm_world(m, n, k, ta, tb, p)
        if w["plans"][0].task_k_num > 16:
            continue                          # the grouped plans hold at most 16 k layers
        out = []
        for tile in (0, 3):
            dA, dB = hip.DeviceArray.from_numpy(A), hip.DeviceArray.from_numpy(B)
            dC = hip.DeviceArray.filled((m, n), 0xFF)
            sim = E.SimGrid(p, m, n, k, ta, tb, w["src_A"], w["src_B"], w["dst_C"], allocate=True)
            sim.set_mode(grouped=True, graph=False)
            sim.set_tuning(4, 2, tile)
            sim.exec([dA.at(r[2] * A.shape[0]) for r in w["src_A"]], [A.shape[0]] * p,
                     [dB.at(r[2] * B.shape[0]) for r in w["src_B"]], [B.shape[0]] * p,
                     [dC.at(r[2] * m) for r in w["dst_C"]], [m] * p)
            out.append(dC.numpy())
            sim.free()
        assert np.array_equal(out[0], out[1]), case
        opA, opB = (A.T if ta else A), (B.T if tb else B)
        bound = k * 2.0 ** -53 * (np.abs(opA) @ np.abs(opB))
        assert np.all(np.abs(out[0] - opA @ opB) <= bound + 2.0 ** -53 * np.abs(opA @ opB)), case


@pytest.mark.parametrize("q", [9, 17])
def test_sim_more_rounds_than_a_batch_item_holds(hip, monkeypatch, q):
    """A q x q Cannon group: with 9 rounds per rank the per-rank mode's
    batched launch holds them; with 17, more than a batched item holds, it
    falls back to one launch per rank and round.  Either way the bits of
    the launches per round asked for explicitly."""
    from oracle import oracle as O
    rnd = np.random.default_rng(99 + q)
    m, n, k, p = 300, 270, 333, q * q
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=[q, q, 1])
    assert max(pl.cannon.np_dim for pl in w["plans"] if pl.is_active) == q
    A, B = rnd.uniform(-1, 1, (m, k)), rnd.uniform(-1, 1, (k, n))
    print(f"[batch] {p} ranks on {q} x {q} x 1", flush=True)
    got = run_sim(p, m, n, k, 0, 0, A, B, proc_grid=[q, q, 1])
    monkeypatch.setenv("CA3DMM_SIM_BATCH", "0")
    want = run_sim(p, m, n, k, 0, 0, A, B, proc_grid=[q, q, 1])
    assert np.array_equal(got, want)
    bound = k * 2.0 ** -53 * (np.abs(A) @ np.abs(B))
    assert np.all(np.abs(got - A @ B) <= bound + 2.0 ** -53 * np.abs(A @ B))


@pytest.mark.parametrize("grid", [[9, 9, 1], [1, 1, 12], [2, 1, 16]])
def test_sim_grouped_and_ozaki_many_rounds_or_layers(hip, grid):
    """Grids with 9 Cannon rounds per rank or 12 and 16 k layers through the
    grouped forms (persistent kernel and batched candidate; exact on
    integers) and the Ozaki form (bit for bit the restated grid)."""
    from oracle import oracle as O
    from tests.simutil import ozaki_grid_reference
    p = grid[0] * grid[1] * grid[2]
    m, n, k = 300, 270, 900
    rnd = np.random.default_rng(p + grid[2])
    Ai, Bi = _ints(rnd, (m, k)), _ints(rnd, (k, n))
    Af, Bf = np.asfortranarray(rnd.uniform(-1, 1, (m, k))), np.asfortranarray(rnd.uniform(-1, 1, (k, n)))
    w = O.ca3dmm_world(m, n, k, 0, 0, p, proc_grid=grid)
    print(f"[grouped] grid {grid}", flush=True)
    for tile in (0, 3):
        dA, dB = hip.DeviceArray.from_numpy(Ai), hip.DeviceArray.from_numpy(Bi)
        dC = hip.DeviceArray.filled((m, n), 0xFF)
        sim = E.SimGrid(p, m, n, k, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=grid, allocate=True)
        sim.set_mode(grouped=True, graph=False)
        sim.set_tuning(4, 2, tile)
        sim.exec([dA.at(r[2] * m) for r in w["src_A"]], [m] * p, [dB.at(r[2] * k) for r in w["src_B"]], [k] * p,
                 [dC.at(r[2] * m) for r in w["dst_C"]], [m] * p)
        sim.free()
        assert np.array_equal(dC.numpy(), Ai @ Bi), tile
    got = run_sim(p, m, n, k, 0, 0, Af, Bf, proc_grid=grid, grouped=True, ozaki=14)
    assert np.array_equal(got, ozaki_grid_reference(m, n, k, p, grid, Af, Bf, 14))


def test_single_rank_engine_random_layouts(hip, monkeypatch):
    """The drop-in boundary on one rank with seeded random shapes, trans
    cases, padded leading dimensions, output windows (dst_C a sub-block or
    the whole C) and redistribution forced or not: exact on integers."""
    rnd = np.random.default_rng(777)
    for case in range(16):
        m, n, k = (int(v) for v in rnd.integers(1, 700, 3))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        force = int(rnd.integers(0, 2))
        if rnd.integers(0, 2):
            sr, sc = int(rnd.integers(0, m)), int(rnd.integers(0, n))
            dst = (sr, int(rnd.integers(1, m - sr + 1)), sc, int(rnd.integers(1, n - sc + 1)))
        else:
            dst = (0, m, 0, n)
        ra, ca = (k, m) if ta else (m, k)
        rb, cb = (n, k) if tb else (k, n)
        lda, ldb, ldc = ra + int(rnd.integers(0, 5)), rb + int(rnd.integers(0, 5)), dst[1] + int(rnd.integers(0, 5))
        A, B = _ints(rnd, (ra, ca)), _ints(rnd, (rb, cb))
        bufA, bufB = np.zeros((lda, ca)), np.zeros((ldb, cb))
        bufA[:ra], bufB[:rb] = A, B
        print(f"[engine] random case {case}: m={m} n={n} k={k} ta={ta} tb={tb} dst={dst} force={force}", flush=True)
        monkeypatch.setenv("CA3DMM_FORCE_REDIST", str(force))
        eng = single_engine(m, n, k, ta, tb, dst=dst)
        dA, dB = hip.DeviceArray.from_numpy(bufA), hip.DeviceArray.from_numpy(bufB)
        dC = hip.DeviceArray.filled((ldc, dst[3]), 0xFF)
        E.lib.ca3dmm_engine_exec(eng, dA.ptr, lda, dB.ptr, ldb, dC.ptr, ldc)
        want = ((A.T if ta else A) @ (B.T if tb else B))[dst[0]:dst[0] + dst[1], dst[2]:dst[2] + dst[3]]
        assert np.array_equal(dC.numpy()[:dst[1]], want), case
        E.lib.ca3dmm_engine_free(C.byref(eng))


def _cut(rnd, length, parts):
    """parts consecutive ranges covering [0, length) at random cut points
    (empty ranges possible)."""
    c = np.sort(rnd.integers(0, length + 1, parts - 1))
    return [0] + [int(v) for v in c] + [length]


def _random_rects(rnd, nrow, ncol, p):
    """A random pr x pc block partition of an nrow x ncol matrix (pr pc = p),
    blocks dealt to the ranks in random order: (srow, nrow, scol, ncol)."""
    pr = int(rnd.choice([d for d in range(1, p + 1) if p % d == 0]))
    rc, cc = _cut(rnd, nrow, pr), _cut(rnd, ncol, p // pr)
    rects = [(rc[i], rc[i + 1] - rc[i], cc[j], cc[j + 1] - cc[j]) for i in range(pr) for j in range(p // pr)]
    return [rects[o] for o in rnd.permutation(p)]


def test_sim_random_block_layouts(hip):
    """The grid's redistributions with seeded random 2-D block layouts of A,
    B and C (each rank's block in its own buffer with a padded leading
    dimension, empty blocks included), random trans cases and the grid the
    library picks: exact on integers, per-rank and grouped modes."""
    rnd = np.random.default_rng(31337)
    for case in range(8):
        p = int(rnd.choice([2, 3, 4, 6, 8, 12]))
        m, n, k = (int(v) for v in rnd.integers(1, 400, 3))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        A = _ints(rnd, (k, m) if ta else (m, k))
        B = _ints(rnd, (n, k) if tb else (k, n))
        rA, rB, rC = _random_rects(rnd, *A.shape, p), _random_rects(rnd, *B.shape, p), _random_rects(rnd, m, n, p)
        print(f"[layouts] case {case}: p={p} m={m} n={n} k={k} ta={ta} tb={tb}", flush=True)

        def blocks(X, rects):
            out = []
            for (sr, nr, sc, nc) in rects:
                ld = max(1, nr) + int(rnd.integers(0, 3))
                buf = np.zeros((ld, max(1, nc)), order="F")
                buf[:nr, :nc] = X[sr:sr + nr, sc:sc + nc]
                out.append((hip.DeviceArray.from_numpy(buf), ld))
            return out
        bA, bB = blocks(A, rA), blocks(B, rB)
        want = (A.T if ta else A) @ (B.T if tb else B)
        for grouped in (False, True):
            bC = [(hip.DeviceArray.filled((max(1, nr) + 1, max(1, nc)), 0xFF), max(1, nr) + 1) for (sr, nr, sc, nc) in rC]
            sim = E.SimGrid(p, m, n, k, ta, tb, rA, rB, rC, allocate=True)
            sim.set_mode(grouped=grouped, graph=False)
            sim.exec([b.ptr for b, _ in bA], [ld for _, ld in bA], [b.ptr for b, _ in bB], [ld for _, ld in bB],
                     [b.ptr for b, _ in bC], [ld for _, ld in bC])
            got = np.full((m, n), np.nan)
            for (sr, nr, sc, nc), (b, _) in zip(rC, bC):
                got[sr:sr + nr, sc:sc + nc] = b.numpy()[:nr, :nc]
            sim.free()
            assert np.array_equal(got, want), (case, grouped)


@pytest.mark.parametrize("mode", ["grouped-persistent", "grouped-batched", "ozaki-15", "ozaki-auto"])
def test_sim_btb_other_forms(hip, mode):
    """B^T B on a 12-rank grid through the grouped forms (persistent kernel,
    batched candidate; exact on integers) and the Ozaki form (15 moduli, and
    the automatic count; bit for bit the restated grid at that count, whose
    FP64 reconstruction need not be exact), eager and replayed from the
    exec's hipGraph."""
    from oracle import oracle as O
    from tests.simutil import ozaki_grid_reference
    rnd = np.random.default_rng(1212)
    p, pg, n, k = 12, [2, 2, 3], 150, 1200
    B = _ints(rnd, (k, n))
    w = O.ca3dmm_world(n, n, k, 0, 0, p, proc_grid=pg, is_btb=True)
    dB = hip.DeviceArray.from_numpy(B)
    print(f"[btb] {mode}", flush=True)
    for graph in (False, True):
        dC = hip.DeviceArray.filled((n, n), 0xFF)
        sim = E.SimGrid(p, n, n, k, 0, 0, None, w["src_B"], w["dst_C"], proc_grid=pg, is_btb=True, allocate=True)
        if mode.startswith("ozaki"):
            sim.set_ozaki(15 if mode == "ozaki-15" else -1)
        sim.set_mode(grouped=True, graph=graph)
        if mode.startswith("grouped"):
            sim.set_tuning(4, 2, 0 if mode == "grouped-persistent" else 3)
        for _ in range(2 if graph else 1):
            sim.exec(None, None, [dB.at(r[2] * k) for r in w["src_B"]], [k] * p,
                     [dC.at(r[2] * n) for r in w["dst_C"]], [n] * p)
        N, _, L = sim.get_ozaki()
        sim.free()
        if mode == "ozaki-auto":
            # the count is the oracle's rule on B^T and B over the plan's k layers
            from oracle import ozaki as Z
            kp = w["plans"][0].task_k_num
            layers = [O.block_spos_size(k, kp, l) for l in range(kp)]
            Lo, kmax = Z.grid_spread(np.ascontiguousarray(B.T), B, layers)
            assert L == pytest.approx(Lo, rel=1e-9, abs=1e-9), (L, Lo)
            assert N == Z.counts_from_spread(Lo, 0.0, kmax)[1], (N, Lo)
        if mode.startswith("ozaki"):
            assert (N == 15) if mode == "ozaki-15" else (4 <= N <= 25)
            want = ozaki_grid_reference(n, n, k, p, pg, np.ascontiguousarray(B.T), B, N, is_btb=True)
        else:
            want = B.T @ B
        assert np.array_equal(dC.numpy(), want), (mode, graph)
