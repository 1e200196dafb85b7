"""CPU: the oracle's planning arithmetic against the reference's own compiled
partition.c/utils.c (oracle/_ref), the committed grid fixture, and the
known answers in the reference README."""
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "grids.json")


def test_block_spos_closed_form():
    rnd = random.Random(7)
    for _ in range(2000):
        n, b = rnd.randint(0, 5000), rnd.randint(1, 64)
        starts, sizes = zip(*[O.block_spos_size(n, b, i) for i in range(b)])
        assert sum(sizes) == n                        # a partition
        assert max(sizes) - min(sizes) <= 1            # balanced
        assert list(sizes) == sorted(sizes, reverse=True)  # larger blocks first
        assert all(starts[i] + sizes[i] == (starts[i + 1] if i + 1 < b else n) for i in range(b))
        assert O.block_spos_size(n, b, b)[0] == n      # iblk == nblk gives len
        assert O.block_spos_size(n, b, b + 1) == (-1, 0)


def test_grids_match_committed_fixture():
    g = json.load(open(GOLD))
    for p, m, n, k, mp, np_, kp, rp in g["cannon"]:
        assert O.grid_cannon(p, m, n, k) == (mp, np_, kp, rp), (p, m, n, k)
    for p, n, k, np_, kp, rp in g["nk"]:
        assert O.grid_nk(p, n, k) == (np_, kp, rp), (p, n, k)


def test_grids_match_compiled_reference_sweep():
    r = O.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rnd = random.Random(3)
    for _ in range(3000):
        p = rnd.randint(1, 300)
        m, n, k = (rnd.choice([1, 5, 64, 333, 1000, 4096, 10007, 65536, 2000000]) for _ in range(3))
        assert O.grid_cannon(p, m, n, k) == O.ref_grid_cannon(r, p, m, n, k)
        assert O.grid_nk(p, n, k) == O.ref_grid_nk(r, p, n, k)
    for n in range(0, 200):
        for b in range(1, 12):
            for i in range(-1, b + 2):
                s1, z1 = C_ref_block(r, n, b, i)
                assert O.block_spos_size(n, b, i) == (s1, z1)


def C_ref_block(r, n, b, i):
    import ctypes as C
    s, z = C.c_int(), C.c_int()
    r.calc_block_spos_size(n, b, i, C.byref(s), C.byref(z))
    return s.value, z.value


def test_readme_known_answers_8000_p24():
    """README.md:58-70: grid 4x2x3, cuboid 2000x4000x2667, rank-0 workbuf
    244.28 MB = 256,149,360 B, comm volume / lower bound 1.04."""
    m = n = k = 8000
    w = O.ca3dmm_world(m, n, k, 0, 0, 24)
    pl = w["plans"][0]
    assert (pl.mp, pl.np, pl.kp, pl.rp) == (4, 2, 3, 0)
    mb, nb, kb = -(-m // pl.mp), -(-n // pl.np), -(-k // pl.kp)
    assert (mb, nb, kb) == (2000, 4000, 2667)
    assert w["workbuf"][0] == 256149360
    lower = 3.0 * (m * n * k / 24) ** (2.0 / 3.0) * 24
    cur = (mb * nb + mb * kb + nb * kb) * 24.0
    assert round(cur / lower, 2) == 1.04
    # the Cannon step of rank 0: q = 2, k-blocks 1334 then 1333
    assert pl.cannon.np_dim == 2 and pl.C_2dmm_nrow == 2000 and pl.C_2dmm_ncol == 4000
    assert (pl.A_2dmm_ncol, pl.cannon.k - pl.A_2dmm_ncol) == (1334, 1333)
    assert pl.C_out_ncol == 1334


def test_small_grids_hand_traced():
    assert O.grid_cannon(4, 2000, 2000, 2000) == (2, 1, 2, 0)
    assert O.grid_cannon(1, 8192, 8192, 8192) == (1, 1, 1, 0)
    assert O.grid_cannon(24, 10007, 10007, 10007) == (4, 2, 3, 0)


def test_world_cuboids_cover_problem_10007():
    """All 24 cuboids of the 4x2x3 grid on 10007^3 tile the m x n x k box."""
    m = n = k = 10007
    w = O.ca3dmm_world(m, n, k, 0, 0, 24)
    flops = 0
    cover = np.zeros((4, 4), dtype=np.int64)
    for pl in w["plans"]:
        c = pl.cannon
        flops += 2 * pl.C_2dmm_nrow * pl.C_2dmm_ncol * c.k
    assert flops == 2 * m * n * k == 2004202940686
    shapes = {(pl.C_2dmm_nrow, pl.C_2dmm_ncol, pl.cannon.k) for pl in w["plans"]}
    assert all(r in (2501, 2502) and c in (5003, 5004) for r, c, _ in shapes)


@pytest.mark.parametrize("p,m,n,k,ta,tb", [
    (4, 2000, 2000, 2000, 0, 0), (24, 800, 800, 800, 0, 0), (6, 300, 200, 700, 1, 0),
    (8, 120, 530, 90, 0, 1), (12, 77, 91, 130, 1, 1), (5, 64, 64, 64, 0, 0), (1, 50, 60, 70, 1, 1)])
def test_redistribution_lists_are_exact_covers(p, m, n, k, ta, tb):
    """Every rank's requested rectangle is covered exactly once by the
    received blocks, and every send matches the peer's receive."""
    w = O.ca3dmm_world(m, n, k, ta, tb, p)
    for name, red in w["redist"].items():
        for r in range(p):
            sl, sc, rl, rc = red["plans"][r]
            req = red["req"][r]
            assert rc == req[2] * req[3]
            assert sum(e[2] * e[3] for e in rl) == rc
            for e in sl:                     # each send has a matching receive
                peer = e[4]
                prl = red["plans"][peer][2]
                match = [f for f in prl if f[4] == r]
                assert len(match) == 1 and list(match[0][:4]) == list(e[:4])


def test_redistribution_moves_every_element_home():
    """Index identity: after the simulated exchange each rank holds exactly
    the global elements it requested (dst(i,j) == global(i,j))."""
    rnd = np.random.default_rng(1)
    for p, m, n, k, ta, tb in [(4, 40, 30, 50, 0, 0), (6, 33, 47, 29, 1, 1), (8, 16, 70, 20, 1, 0)]:
        w = O.ca3dmm_world(m, n, k, ta, tb, p)
        red = w["redist"]["A"]
        An, Ac = (k, m) if ta else (m, k)
        G = rnd.standard_normal((An, Ac))
        srcs, dsts, slds, dlds = [], [], [], []
        for r in range(p):
            s = red["src"][r]     # row-major: rows = global columns
            blk = np.ascontiguousarray(G[s[1]:s[1] + s[3], s[0]:s[0] + s[2]].T).ravel()
            srcs.append(blk if blk.size else np.zeros(1)); slds.append(max(1, s[3]))
            q = red["req"][r]
            dsts.append(np.zeros(max(1, q[2] * q[3]))); dlds.append(max(1, q[3]))
        O.redist_exec_all(red["src"], red["req"], srcs, slds, dsts, dlds)
        for r in range(p):
            q = red["req"][r]
            if q[2] * q[3] == 0:
                continue
            want = G[q[1]:q[1] + q[3], q[0]:q[0] + q[2]].T
            assert np.array_equal(dsts[r][:q[2] * q[3]].reshape(q[2], q[3]), want)
