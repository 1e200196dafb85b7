"""CPU: the product's planning (grid, task groups, Cannon blocks, redistribution
lists, workspace sizes -- src/ca3dmm.c:23-722, src/cannon.c:16-133,
src/mat_redist.c:44-236) equals the oracle's on every rank, integer for integer.
Plan-only simulated grids allocate nothing, so no GPU is needed."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from ca3dmm_amd import engine as E
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "grids.json")

FIELDS = ["mp", "np", "kp", "rp", "is_active", "rank_m", "rank_n", "rank_k",
          "task_m_num", "task_m_id", "task_n_num", "task_n_id", "task_k_num", "task_k_id",
          "use_ag", "use_rsb",
          "A_rd_srow", "A_rd_scol", "A_rd_nrow", "A_rd_ncol",
          "B_rd_srow", "B_rd_scol", "B_rd_nrow", "B_rd_ncol",
          "A_2dmm_srow", "A_2dmm_scol", "A_2dmm_nrow", "A_2dmm_ncol",
          "B_2dmm_srow", "B_2dmm_scol", "B_2dmm_nrow", "B_2dmm_ncol",
          "C_2dmm_srow", "C_2dmm_scol", "C_2dmm_nrow", "C_2dmm_ncol",
          "C_out_srow", "C_out_scol", "C_out_nrow", "C_out_ncol"]
CANNON = ["np_dim", "rank_row", "rank_col", "A_srow", "A_scol", "A_nrow", "A_ncol",
          "B_srow", "B_scol", "B_nrow", "B_ncol", "C_srow", "C_scol", "C_nrow", "C_ncol",
          "gemm_cycle", "max_A_blk_size", "max_B_blk_size", "max_C_blk_size"]

CASES = [  # p, m, n, k, transA, transB, btb, proc_grid
    (1, 64, 48, 80, 0, 0, False, None), (4, 2000, 2000, 2000, 0, 0, False, None),
    (24, 8000, 8000, 8000, 0, 0, False, None), (24, 10007, 10007, 10007, 0, 0, False, None),
    (24, 800, 800, 800, 1, 1, False, None), (6, 300, 200, 700, 1, 0, False, None),
    (8, 120, 530, 90, 0, 1, False, None), (12, 77, 91, 130, 1, 1, False, None),
    (16, 1000, 40, 40, 0, 0, False, None), (16, 40, 1000, 40, 1, 0, False, None),
    (7, 50, 60, 70, 0, 0, False, [2, 1, 3]), (9, 30, 30, 3000, 0, 0, False, None),
    (10, 500, 500, 50, 0, 0, False, [2, 4, 1]), (5, 33, 33, 33, 0, 0, False, [4, 4, 4]),
    (4, 0, 512, 2000, 0, 0, True, None), (16, 0, 512, 200000, 0, 0, True, None),
    (12, 0, 100, 100, 0, 0, True, [2, 2, 3]), (24, 0, 300, 30000, 0, 0, True, None),
]


def _world(p, m, n, k, ta, tb, btb, pg, layouts=None):
    kw = dict(zip(("src_A", "src_B", "dst_C"), layouts)) if layouts else {}
    w = O.ca3dmm_world(m, n, k, ta, tb, p, proc_grid=pg, is_btb=btb, **kw)
    srcA = w["src_A"] if not btb else None
    sim = E.SimGrid(p, m, n, k, ta, tb, srcA, w["src_B"], w["dst_C"], proc_grid=pg, is_btb=btb,
                    allocate=False)
    return w, sim


@pytest.mark.parametrize("case", CASES, ids=[str(c[:7]) for c in CASES])
def test_every_rank_matches_oracle(case):
    _check_against_oracle(*case)


def _rects(rnd, nrow, ncol, p):
    """A random pr x pc block partition (pr pc = p), blocks dealt to the
    ranks in random order, empty blocks possible: (srow, nrow, scol, ncol)."""
    pr = int(rnd.choice([d for d in range(1, p + 1) if p % d == 0]))
    cut = lambda length, parts: [0] + sorted(int(v) for v in rnd.integers(0, length + 1, parts - 1)) + [length]
    rc, cc = cut(nrow, pr), cut(ncol, p // pr)
    rects = [(rc[i], rc[i + 1] - rc[i], cc[j], cc[j + 1] - cc[j]) for i in range(pr) for j in range(p // pr)]
    return [rects[o] for o in rnd.permutation(p)]


def test_random_worlds_match_oracle():
    """Seeded random rank counts, shapes, trans cases, B^T B and random 2-D
    block layouts of the inputs and of C: every rank's plan, Cannon blocks,
    redistribution lists and workspace equal the oracle's."""
    rnd = np.random.default_rng(8128)
    for case in range(30):
        p = int(rnd.integers(1, 33))
        m, n, k = (int(v) for v in rnd.integers(1, 5000, 3))
        btb = bool(rnd.integers(0, 5) == 0)
        ta, tb = (0, 0) if btb else tuple(int(v) for v in rnd.integers(0, 2, 2))
        layouts = None
        if not btb and rnd.integers(0, 2):
            layouts = (_rects(rnd, *((k, m) if ta else (m, k)), p), _rects(rnd, *((n, k) if tb else (k, n)), p),
                       _rects(rnd, m, n, p))
        _check_against_oracle(p, 0 if btb else m, n, k, ta, tb, btb, None, layouts)


def _check_against_oracle(p, m, n, k, ta, tb, btb, pg, layouts=None):
    w, sim = _world(p, m, n, k, ta, tb, btb, pg, layouts)
    for r in range(p):
        e, pl = sim.engine(r), w["plans"][r]
        assert e.my_rank == r and e.n_proc == p
        for f in FIELDS[:5]:
            assert getattr(e, f) == getattr(pl, f), (r, f)
        if pl.is_active:
            for f in FIELDS:
                assert getattr(e, f) == getattr(pl, f), (r, f)
            ce = e.cannon_engine.contents
            for f in CANNON:
                assert getattr(ce, f) == getattr(pl.cannon, f), (r, "cannon." + f)
            assert e.cannon_workbuf_bytes == pl.cannon_workbuf_bytes
            assert e.self_workbuf_bytes == pl.self_workbuf_bytes
        for name, ptr in (("A", e.redist_A), ("B", e.redist_B), ("C", e.redist_C)):
            if name not in w["redist"]:
                assert not ptr
                continue
            red = ptr.contents
            sl, sc, rl, rc = w["redist"][name]["plans"][r]
            assert red.send_cnt == sc and red.recv_cnt == rc, (r, name)
            assert np.array_equal(red.send_list(), sl), (r, name, "send")
            assert np.array_equal(red.recv_list(), rl), (r, name, "recv")
        assert sim.workbuf_bytes(r) == w["workbuf"][r], r
    sim.free()


def test_readme_rank0_workspace_through_product():
    w, sim = _world(24, 8000, 8000, 8000, 0, 0, False, None)
    assert sim.workbuf_bytes(0) == 256149360
    e = sim.engine(0)
    assert (e.mp, e.np, e.kp, e.rp) == (4, 2, 3, 0)
    sim.free()


def test_product_grid_search_matches_fixture():
    g = json.load(open(GOLD))
    lib = E.lib
    for p, m, n, k, *want in g["cannon"]:
        v = [C.c_int() for _ in range(4)]
        lib.calc_3d_decomposition_cannon(p, m, n, k, *[C.byref(x) for x in v])
        assert [x.value for x in v] == want, (p, m, n, k)
    for p, n, k, *want in g["nk"]:
        v = [C.c_int() for _ in range(3)]
        lib.calc_3d_decomposition_nk(p, n, k, *[C.byref(x) for x in v])
        assert [x.value for x in v] == want, (p, n, k)


def test_product_block_partition_matches_oracle():
    s, z = C.c_int(), C.c_int()
    for n in range(0, 60):
        for b in range(1, 9):
            for i in range(-1, b + 2):
                E.lib.calc_block_spos_size(n, b, i, C.byref(s), C.byref(z))
                assert (s.value, z.value) == O.block_spos_size(n, b, i)


@pytest.mark.parametrize("p,n,want", [(12, 1000, (6, 2, 1)), (16, 2000, (4, 2, 2)), (24, 8000, (4, 2, 3))])
def test_reduce_kp_env(monkeypatch, p, n, want):
    """CA3DMM_REDUCE_KP=1 folds a factor 3 (else 2) of kp into the smaller of
    mp, np when divisibility allows (src/ca3dmm.c:75-124)."""
    monkeypatch.setenv("CA3DMM_REDUCE_KP", "1")
    w = O.ca3dmm_world(n, n, n, 0, 0, p, reduce_kp=1)
    sim = E.SimGrid(p, n, n, n, 0, 0, w["src_A"], w["src_B"], w["dst_C"], allocate=False)
    for r in range(p):
        e, pl = sim.engine(r), w["plans"][r]
        assert (e.mp, e.np, e.kp) == (pl.mp, pl.np, pl.kp) == want
        assert sim.workbuf_bytes(r) == w["workbuf"][r]
    sim.free()


# ---- failure paths a CPU test can force (the reference signals them with a
# NULL engine and a message on stderr) ----
def _init_single(dev_type=E.DEV_TYPE_HIP, comm=None, pg=None, m=64, n=64, k=64):
    eng = E.CA3P()
    wb = C.c_size_t()
    comm = comm or E.comm_self()
    E.lib.ca3dmm_engine_init(m, n, k, 0, 0, 0, m, 0, k, 0, k, 0, n, 0, m, 0, n,
                             E._ints(pg) if pg else None, comm, dev_type, C.byref(eng), C.byref(wb))
    return eng, wb.value


def test_host_dev_type_rejected():
    eng, _ = _init_single(dev_type=E.DEV_TYPE_HOST)
    assert not eng


def test_invalid_dev_type_rejected():
    eng, _ = _init_single(dev_type=7)
    assert not eng


def test_multi_rank_group_needs_simulated_grid():
    eng, _ = _init_single(comm=E.Comm(4, 1))
    assert not eng


def test_grid_larger_than_group_is_replaced():
    eng, wb = _init_single(pg=[2, 2, 2])          # 8 > p = 1: warning, automatic grid
    assert eng and (eng.contents.mp, eng.contents.np, eng.contents.kp) == (1, 1, 1)
    assert wb > 0
    E.lib.ca3dmm_engine_free(C.byref(eng))
    assert not eng


def test_exec_on_null_engine_is_harmless():
    E.lib.ca3dmm_engine_exec(None, None, 1, None, 1, None, 1)


def test_cannon_non_square_group_rejected():
    ce = C.POINTER(E.Cannon)()
    wb = C.c_size_t()
    E.lib.cannon_engine_init(10, 10, 10, E.Comm(3, 0), E.DEV_TYPE_HIP, C.byref(ce), C.byref(wb))
    assert not ce


def test_cannon_min_kblk_env(monkeypatch):
    """CANNON_MIN_KBLK_SIZE (cannon.c:86-99) drives gemm_cycle; out-of-range
    values fall back to the default 160."""
    for env, want in (("1000", 3), ("5", 2), (None, 2)):
        if env is None:
            monkeypatch.delenv("CANNON_MIN_KBLK_SIZE", raising=False)
        else:
            monkeypatch.setenv("CANNON_MIN_KBLK_SIZE", env)
        w = O.ca3dmm_world(60, 60, 300, 0, 0, 9, proc_grid=[3, 3, 1],
                           min_k_blk=int(env) if env and 16 <= int(env) <= 8192 else 160)
        sim = E.SimGrid(9, 60, 60, 300, 0, 0, w["src_A"], w["src_B"], w["dst_C"], proc_grid=[3, 3, 1],
                        allocate=False)
        assert sim.engine(0).cannon_engine.contents.gemm_cycle == w["plans"][0].cannon.gemm_cycle == want
        sim.free()


@pytest.mark.parametrize("p,m,n,k", [(1, 46341, 46341, 64), (1, 64, 46341, 46341), (4, 100000, 100000, 100)])
def test_counts_above_int_max_refused(p, m, n, k):
    """Redistribution counts, Cannon block sizes and reduce-scatter counts are
    int fields of the public structs (src/mat_redist.h:7-45, src/cannon.h:7-50,
    src/ca3dmm.h:10-82).  A shape whose counts pass INT_MAX is refused at
    planning time with an error instead of wrapping them (plan only, nothing
    is allocated)."""
    from ca3dmm_amd import HipError
    col = lambda nr, nc: [[0, nr, nc * r // p, nc * (r + 1) // p - nc * r // p] for r in range(p)]
    with pytest.raises(HipError):
        E.SimGrid(p, m, n, k, 0, 0, col(m, k), col(k, n), col(m, n), allocate=False)


def test_counts_just_below_int_max_planned():
    """The same planner accepts a rank whose largest count stays below
    INT_MAX: at m = n = 46339 the Cannon bound (m/q + 1)(n/q + 1) is
    46340^2 = 2,147,395,600."""
    m = n = 46339
    sim = E.SimGrid(1, m, n, 64, 0, 0, [[0, m, 0, 64]], [[0, 64, 0, n]], [[0, m, 0, n]], allocate=False)
    e = sim.engine(0)
    assert e.C_2dmm_nrow * e.C_2dmm_ncol == m * n
    assert e.redist_C.contents.send_cnt == m * n
    sim.free()
