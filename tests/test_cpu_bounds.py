"""CPU: every device address the data-movement and GEMM kernels form, checked
against its allocation, by restating the kernels' index arithmetic in numpy
for the shapes the tests and the measured configurations use.  A GPU fault
from an out-of-range access shows up here first.

* dgemm_kernel / grouped_kernel / gram_partial_kernel operand loaders
  (dgemm_tile.h OperandLoader<XC>): which operand is read as contiguous
  along m / n (XC = 1) or along k (XC = 0) is read from the kernel sources
  themselves, and each case's extreme element offsets are compared with the
  allocation its BLAS convention implies (A is lda x (transA ? m : k), B is
  ldb x (transB ? k : n)).  The full-stage (unpredicated) path must stay
  inside the operand for interior tiles, including the repeated last load.
* tile_coords (XCD remap + groups of 16 tile-rows) and grouped_coords visit
  every tile exactly once.
* copy_blocks_kernel: the row of a flattened index from the double
  reciprocal plus one correction is exact, and the tiles of a batch cover
  every row of every block once.
* the one-pass redistribution: each block lies inside the sender's source
  block and the receiver's destination block, for the grids of the tests
  and of the measured configurations (oracle plans).
"""
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ca3dmm_amd", "csrc")
BM = BN = 128
BK = 16


def _src(name):
    with open(os.path.join(CSRC, name)) as f:
        return f.read()


# --------------------------------------------------------------------------- loaders
def loader_xc(text, var, flag):
    """XC of `OperandLoader<...> var` in a kernel source: 1, 0, or for a
    template flag ('TA'/'TB') a function of it."""
    m = re.search(r"OperandLoader<\s*(!?)\s*(\w+)\s*(?:,[^>]*)?>\s*" + var + r"\s*;", text)
    assert m, f"no OperandLoader declaration for {var}"
    neg, arg = m.group(1) == "!", m.group(2)
    if arg in ("0", "1"):
        return lambda t: int(arg)
    assert arg == flag, (arg, flag)
    return (lambda t: 1 - t) if neg else (lambda t: t)


def operand_extremes(xc, xdim, K, ld, x0s, kspans, W=128):
    """Largest element offset a loader touches: XC = 1: x + k ld, XC = 0:
    k + x ld, for x in [x0, x0 + W) clipped to xdim (the predicate) and the
    k ranges of its stages."""
    hi = 0
    for x0, (k0, k1) in zip(x0s, kspans):
        xl = min(xdim, x0 + W) - 1
        kl = k1 - 1
        if xl < 0 or kl < 0:
            continue
        hi = max(hi, xl + kl * ld if xc else kl + xl * ld)
    return hi


def stage_spans(M, N, K, m0, n0, bm=BM, bn=BN):
    """k ranges the loaders read for tile (m0, n0) of mainloop: interior
    tiles read whole 16-deep stages unpredicated (the last one repeated),
    then the predicated tail; edge tiles are predicated throughout."""
    interior = m0 + bm <= M and n0 + bn <= N
    nfull = K // BK if interior else 0
    spans = []
    if nfull:
        spans.append((0, nfull * BK))                 # includes the repeated last stage
    if nfull * BK < K:
        spans.append((nfull * BK, K))                 # predicated: never past K
    return spans


DGEMM_SHAPES = [(1, 1, 1), (15, 17, 3), (129, 130, 17), (257, 64, 1000), (64, 511, 33), (777, 555, 1999),
                (1000, 1000, 1000), (2000, 4000, 1334), (2000, 4000, 1333), (2501, 5003, 1668),
                (8192, 8192, 8192), (8192, 128, 8192), (128, 8192, 8192)]


TILES = [(128, 128), (256, 128), (128, 256)]      # TileShape instances of the DGEMM and grouped kernels


def test_loader_thread_layouts_cover_each_stage_once():
    """OperandLoader<XC, W, T>: the (k, x) elements of one 16 x W stage the T
    threads load (P each; XC = 1 at k rows kk0 + KSTEP p, XC = 0 at columns
    x + XSTEP p) cover the stage exactly once, for every W / T in use."""
    for W, T in [(128, 256), (256, 512), (128, 512)]:
        P, KSTEP, XSTEP = W * BK // T, T // W, T // BK
        for xc in (0, 1):
            seen = np.zeros((BK, W), int)
            for t in range(T):
                for p in range(P):
                    if xc:
                        seen[t // W + KSTEP * p, t % W] += 1
                    else:
                        seen[t & 15, (t >> 4) + XSTEP * p] += 1
            assert np.all(seen == 1), (W, T, xc)


@pytest.mark.parametrize("bm,bn", TILES)
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_dgemm_loaders_stay_inside_the_operands(ta, tb, bm, bn):
    text = _src("dgemm.hip")
    xa, xb = loader_xc(text, "la", "TA")(ta), loader_xc(text, "lb", "TB")(tb)
    for M, N, K in DGEMM_SHAPES:
        lda, ldb = (K if ta else M), (N if tb else K)          # tight leading dims
        alloc_a = lda * (M if ta else K)
        alloc_b = ldb * (K if tb else N)
        tiles = [(tm * bm, tn * bn) for tm in range(-(-M // bm)) for tn in range(-(-N // bn))]
        for m0, n0 in tiles:
            sp = stage_spans(M, N, K, m0, n0, bm, bn)
            for k0, k1 in sp:
                # full stages read every element of the bm (bn) x 16 block:
                # they must be whole and in range on both axes
                if (k0, k1) == sp[0] and m0 + bm <= M and n0 + bn <= N and k1 - k0 >= BK:
                    assert k1 <= K and m0 + bm <= M and n0 + bn <= N
            ha = operand_extremes(xa, M, K, lda, [m0] * len(sp), sp, bm)
            hb = operand_extremes(xb, N, K, ldb, [n0] * len(sp), sp, bn)
            assert ha < alloc_a, (ta, tb, M, N, K, m0, ha, alloc_a)
            assert hb < alloc_b, (ta, tb, M, N, K, n0, hb, alloc_b)


def test_grouped_and_gram_loaders_read_their_blocks_as_stored():
    """grouped_kernel reads each Cannon round's A block (M x k, lda = M) along
    m and its B block (k x N, ldb = k) along k; the Gram reads A^T and A of
    a k x n A along k.  Any other XC would leave the blocks."""
    g = _src("grouped.hip")
    assert loader_xc(g, "la", "-")(0) == 1 and loader_xc(g, "lb", "-")(0) == 0
    for bm, bn in TILES:
        for M, N, K in [(2501, 5003, 1668), (2502, 5004, 1667), (1000, 500, 334), (3, 5, 7)]:
            for m0 in range(0, M, bm):
                for n0 in range(0, N, bn):
                    sp = stage_spans(M, N, K, m0, n0, bm, bn)
                    assert operand_extremes(1, M, K, M, [m0] * len(sp), sp, bm) < M * K
                    assert operand_extremes(0, N, K, K, [n0] * len(sp), sp, bn) < K * N
    gram = _src("gram.hip")
    m = re.search(r"OperandLoader<(\d)>\s+la,\s*lb;", gram)
    assert m and m.group(1) == "0"
    for n, k in [(512, 2000000), (130, 1000), (5, 3)]:
        for t0 in range(0, n, BM):
            assert operand_extremes(0, n, k, k, [t0], [(0, k)]) < k * n


# --------------------------------------------------------------------------- tile order
def tile_coords(bid, nb, M, N, bm=BM, bn=BN, G=16):
    """dgemm_tile.h tile_coords, restated."""
    tiles_m, tiles_n = -(-M // bm), -(-N // bn)
    if nb % 8 == 0:
        bid = (bid & 7) * (nb >> 3) + (bid >> 3)
    per = G * tiles_n
    grp = bid // per
    first = grp * G
    gsz = min(tiles_m - first, G)
    r = bid - grp * per
    return first + r % gsz, r // gsz


def batch_tiles(problems, G=16):
    """dgemm_batch_kernel (dgemm.hip): workgroup b of the padded grid ->
    (problem, tile row, tile col), or None for a padding workgroup; problems
    are (M, N) with their tiles at consecutive linear indices."""
    bases, total = [], 0
    for M, N in problems:
        bases.append(total)
        total += -(-M // BM) * -(-N // BN)
    nb = (total + 7) // 8 * 8
    out = []
    for b in range(nb):
        t = (b & 7) * (nb >> 3) + (b >> 3)
        if t >= total:
            out.append(None)
            continue
        lo, hi = 0, len(problems) - 1
        while lo < hi:
            mid = (lo + hi + 1) >> 1
            if bases[mid] <= t:
                lo = mid
            else:
                hi = mid - 1
        M, N = problems[lo]
        out.append((lo,) + tile_coords(t - bases[lo], 1, M, N, G=G))
    return out


@pytest.mark.parametrize("problems", [[(2000, 4000)] * 8 + [(2000, 4000)] * 16,
                                      [(2502, 5004), (2502, 5003), (2501, 5004), (2501, 5003)] * 6,
                                      [(1, 1)], [(130, 77), (65, 77), (130, 1), (1, 300)],
                                      [(128, 128)] * 7, [(8192, 8192)]])
def test_batch_tiles_cover_every_tile_once(problems):
    for G in (1, 16):
        got = [x for x in batch_tiles(problems, G) if x is not None]
        want = [(p, i, j) for p, (M, N) in enumerate(problems)
                for i in range(-(-M // BM)) for j in range(-(-N // BN))]
        assert sorted(got) == sorted(want) and len(got) == len(set(got))


@pytest.mark.parametrize("bm,bn", TILES)
@pytest.mark.parametrize("M,N", [(128, 128), (1000, 1000), (2000, 4000), (8192, 8192), (2501, 5003),
                                 (130, 9000), (9000, 130), (17 * 128, 3 * 128), (100, 100)])
def test_tile_order_is_a_bijection(M, N, bm, bn):
    tm_n, tn_n = -(-M // bm), -(-N // bn)
    nb = tm_n * tn_n
    for G in (1, 4, 8, 16, 32, 256):                  # CA3DMM_DGEMM_GROUP
        seen = {tile_coords(b, nb, M, N, bm, bn, G) for b in range(nb)}
        assert seen == {(i, j) for i in range(tm_n) for j in range(tn_n)}, G


@pytest.mark.parametrize("G", [1, 4, 8, 16])
def test_grouped_tile_order_is_a_bijection(G):
    for tiles_m, tiles_n in [(20, 40), (19, 39), (1, 7), (33, 2)]:
        seen = set()
        for t in range(tiles_m * tiles_n):
            per, grp = G * tiles_n, t // (G * tiles_n)
            first = grp * G
            gsz = min(tiles_m - first, G)
            r = t - grp * per
            seen.add((first + r % gsz, r // gsz))
        assert seen == {(i, j) for i in range(tiles_m) for j in range(tiles_n)}


# --------------------------------------------------------------------------- copies
def row_col(e, ncol):
    """copy_kernels.hip row_col, restated on uint32 values."""
    inv = 1.0 / np.float64(ncol)
    r = (e.astype(np.float64) * inv).astype(np.int64)
    c = e.astype(np.int64) - r * ncol
    low = c < 0
    r = np.where(low, r - 1, r)
    c = np.where(low, c + ncol, c)
    high = c >= ncol
    r = np.where(high, r + 1, r)
    c = np.where(high, c - ncol, c)
    return r, c


def test_copy_row_of_flattened_index_is_exact():
    rnd = np.random.default_rng(1)
    ncols = np.unique(np.concatenate([np.arange(1, 300), rnd.integers(1, 1 << 20, 400),
                                      [4000, 4001, 8000, 10007, 65535, 65536, (1 << 20) - 1]]))
    for ncol in ncols:
        ncol = int(ncol)
        rows = np.unique(np.concatenate([np.arange(0, 64), rnd.integers(0, (2 ** 31 - 1) // ncol, 200)]))
        rows = rows[rows * ncol < 2 ** 31]
        # every position around each row start
        e = (rows[:, None] * ncol + np.array([-1, 0, 1, ncol - 1])[None, :]).ravel()
        e = e[(e >= 0) & (e < 2 ** 31)]
        r, c = row_col(e, ncol)
        assert np.array_equal(r, e // ncol) and np.array_equal(c, e % ncol), ncol


def copy_tiles(descs, words=8):
    """Tile partition of one CopyBatch (launch_copy): per descriptor its first
    tile, rows per tile and pieces per row (a row longer than a tile is cut
    into tile-long pieces, one tile each)."""
    tile_words = 256 * 4 * 4
    out, tiles = [], 0
    for nrow, ncol in descs:
        ncw = ncol * 8 // words
        if ncw > tile_words:
            chunks = -(-ncw // tile_words)
            out.append((tiles, 1, nrow, chunks, ncw))
            tiles += nrow * chunks
        else:
            rpt = max(1, tile_words // ncw)
            out.append((tiles, rpt, nrow, 1, ncw))
            tiles += -(-nrow // rpt)
    return out, tiles


@pytest.mark.parametrize("long_rows", [False, True])
def test_copy_batch_tiles_cover_every_word_once(long_rows):
    """copy_blocks_kernel's tile -> (rows, words) mapping, restated: every word
    of every block is copied by exactly one tile, and a tile's flattened
    index stays inside 32 bits."""
    rnd = np.random.default_rng(2 + long_rows)
    tile_words = 4096
    for _ in range(30):
        if long_rows:
            descs = [(int(rnd.integers(1, 4)), int(rnd.integers(1, 40000))) for _ in range(int(rnd.integers(1, 48)))]
        else:
            descs = [(int(rnd.integers(1, 5000)), int(rnd.integers(1, 9000))) for _ in range(int(rnd.integers(1, 48)))]
        parts, total = copy_tiles(descs, words=16)
        cover = [np.zeros((nrow, p[4]), np.int8) for (nrow, _), p in zip(descs, parts)]
        starts = [p[0] for p in parts]
        for t in range(total):
            lo = int(np.searchsorted(starts, t, side="right") - 1)      # the kernel's binary search
            t0, rpt, nrow, chunks, ncw = parts[lo]
            if chunks > 1:
                q = t - t0
                r0, c0 = q // chunks, (q % chunks) * tile_words
                nc = min(tile_words, ncw - c0)
                assert r0 < nrow and nc > 0
                cover[lo][r0, c0:c0 + nc] += 1
            else:
                r0 = (t - t0) * rpt
                nr = min(rpt, nrow - r0)
                assert nr > 0 and r0 + nr <= nrow and nr * ncw < 2 ** 31
                cover[lo][r0:r0 + nr, :] += 1
        assert all(np.all(c == 1) for c in cover)


# --------------------------------------------------------------------------- redistribution
GRIDS = [(8000, 8000, 8000, 24, [4, 2, 3], 0, 0), (10007, 10007, 10007, 24, [4, 2, 3], 0, 0),
         (2000, 2000, 2000, 4, None, 0, 0), (1001, 1001, 1001, 24, [4, 2, 3], 1, 1),
         (601, 433, 517, 24, [4, 2, 3], 1, 0), (700, 500, 900, 12, [2, 2, 3], 0, 1),
         (300, 200, 250, 5, [2, 2, 1], 0, 0)]


@pytest.mark.parametrize("m,n,k,p,grid,ta,tb", GRIDS)
def test_one_pass_redistribution_blocks_inside_both_buffers(m, n, k, p, grid, ta, tb):
    """redist_direct copies block b (row-major rect of a receive entry) from
    the sender's source block to the receiver's destination block: it must
    lie inside both rects, with leading dimensions no smaller than the rows
    (row-major: ncol) of either buffer."""
    w = O.ca3dmm_world(m, n, k, ta, tb, p, proc_grid=grid)
    for name, red in w["redist"].items():
        src, req = red["src"], red["req"]
        moved = 0
        for r in range(p):
            _, _, rl, rc = red["plans"][r]
            for b in rl:
                s = src[int(b[4])]
                q = req[r]
                for rect in (s, q):
                    assert rect[0] <= b[0] and b[0] + b[2] <= rect[0] + rect[2], (name, r, b, rect)
                    assert rect[1] <= b[1] and b[1] + b[3] <= rect[1] + rect[3], (name, r, b, rect)
                moved += int(b[2]) * int(b[3])
            assert moved >= 0
        # every requested element arrives exactly once
        for r in range(p):
            q = req[r]
            area = int(q[2]) * int(q[3])
            got = sum(int(b[2]) * int(b[3]) for b in red["plans"][r][2])
            assert got == area, (name, r, got, area)


# --------------------------------------------------------------------------- direct-to-LDS stages
def dtl_index(xc, k, x):
    """dgemm_tile.h dtl_index, restated (in doubles)."""
    if xc:
        return k * 128 + 2 * ((x >> 1) ^ ((k & 1) << 3)) + (x & 1)
    return x * 16 + 2 * ((k >> 1) ^ ((x >> 1) & 7)) + (k & 1)


def dtl_writes(xc):
    """(LDS double index, k, x) of every double the 4 waves' DtlLoader issues
    write: lane l of wave w, block b = 4 w + i lands at b * 128 + 2 l (+1)."""
    out = []
    for w in range(4):
        for i in range(4):
            b = 4 * w + i
            for lane in range(64):
                for h in range(2):
                    lds = b * 128 + 2 * lane + h
                    if xc:
                        k, x = b, 2 * (lane ^ ((b & 1) << 3)) + h
                    else:
                        x = 8 * b + (lane >> 3)
                        k, x = 2 * ((lane & 7) ^ ((x >> 1) & 7)) + h, x
                    out.append((lds, k, x))
    return out


@pytest.mark.parametrize("xc", [0, 1])
def test_dtl_stage_layout_is_consistent_and_conflict_free(xc):
    """Every (k, x) of a 16 x 128 stage is written exactly once, at the slot
    the MFMA reads use (dtl_index); the global chunk of each lane is 16-byte
    aligned and inside the stage's 16 x 128 block; and each 32-lane half of
    an operand read (16 consecutive x at 2 k rows) touches 64 distinct banks
    (4-byte banks, 8 bytes per lane)."""
    wr = dtl_writes(xc)
    seen = {}
    for lds, k, x in wr:
        assert 0 <= k < 16 and 0 <= x < 128
        assert (k, x) not in seen
        seen[(k, x)] = lds
        assert dtl_index(xc, k, x) == lds
    assert len(seen) == 16 * 128
    for X0 in range(0, 128, 16):
        for ks in range(4):
            for half in range(2):
                banks = set()
                for lane in range(32 * half, 32 * half + 32):
                    k, x = 4 * ks + (lane >> 4), X0 + (lane & 15)
                    byte = 8 * dtl_index(xc, k, x)
                    banks.update({(byte // 4) % 64, (byte // 4 + 1) % 64})
                assert len(banks) == 64, (xc, X0, ks, half)


def dtl_loads(xc, xn, ld):
    """DtlLoader::init / issue (dgemm_tile.h), restated for one stage at
    koff = 0 of a tile at x0 = 0 with xn existing x: (LDS double index, k, x
    of the element loaded there, element offset from the operand origin) of
    both doubles of every lane's 16-byte load."""
    out = []
    for w in range(4):
        for i in range(4):
            b = 4 * w + i
            for lane in range(64):
                if xc:
                    base = b * ld
                    off = 16 * min(lane ^ ((b & 1) << 3), (xn >> 1) - 1)
                else:
                    xb, xl = min(8 * b, xn - 1), 8 * b + (lane >> 3)
                    x = min(xl, xn - 1)
                    base = xb * ld
                    off = (x - xb) * ld * 8 + 16 * ((lane & 7) ^ ((xl >> 1) & 7))
                    assert 0 <= off < 2 ** 32
                # 16-byte aligned for an even ld; 8-byte aligned (a double
                # pair, possibly split over two cache lines) for an odd one
                assert off % (16 if ld % 2 == 0 else 8) == 0
                for h in range(2):
                    e = base + off // 8 + h
                    k, x = (b, e - b * ld) if xc else (e % ld, e // ld)
                    out.append((b * 128 + 2 * lane + h, k, x, e))
    return out


@pytest.mark.parametrize("xc", [0, 1])
@pytest.mark.parametrize("xn", [128, 127, 120, 104, 64, 9, 8, 2, 1])
@pytest.mark.parametrize("odd_ld", [False, True])
def test_dtl_edge_tile_loads_stay_inside_and_fill_existing_slots(xc, xn, odd_ld):
    """Edge tiles through the direct-to-LDS stages: every load reads an
    element of the 16 x xn block that exists (x < xn, k < 16), and every LDS
    slot of an existing (k, x) holds exactly that element, as for an
    interior tile; the remaining slots hold copies of existing elements.
    XC = 1 needs xn even (odd counts take the checked path).  Odd leading
    dimensions (16-byte loads from 8-byte aligned addresses) map the same."""
    if xc and xn % 2:
        pytest.skip("XC = 1 edge tiles need an even count")
    ld = (1000 if xc else 2048) + (1 if odd_ld else 0)  # ld >= 128 (XC = 1) or >= 16 (XC = 0)
    slot = {dtl_index(xc, kk, xx): (kk, xx) for kk in range(16) for xx in range(128)}
    for lds, k, x, e in dtl_loads(xc, xn, ld):
        assert 0 <= k < 16 and 0 <= x < xn, (xc, xn, lds, k, x)
        want_k, want_x = slot[lds]
        if want_x < xn:
            assert (k, x) == (want_k, want_x), (xc, xn, lds)
        else:
            assert k == want_k, (xc, xn, lds)


def dtl_loads_dword(xc, xn, ld):
    """DtlLoader<XC, 1>::issue (dgemm_tile.h), restated for one stage at koff
    = 0 of a tile at x0 = 0: (LDS byte, k, x, byte offset from the operand
    origin) of every 4-byte load (lane l of load q of block b lands at byte
    1024 b + 256 q + 4 l)."""
    out = []
    for w in range(4):
        for i in range(4):
            b = 4 * w + i
            for q in range(4):
                for lane in range(64):
                    t, j, h = lane >> 2, (lane >> 1) & 1, lane & 1
                    if xc:
                        base = b * ld * 8
                        x = min(2 * ((16 * q + t) ^ ((b & 1) << 3)) + j, xn - 1)
                        off = 8 * x + 4 * h
                    else:
                        xb = min(8 * b + 2 * q, xn - 1)
                        base = xb * ld * 8
                        xl = 8 * b + 2 * q + (t >> 3)
                        x = min(xl, xn - 1)
                        off = (x - xb) * ld * 8 + 16 * ((t & 7) ^ ((xl >> 1) & 7)) + 8 * j + 4 * h
                    assert 0 <= off < 2 ** 32
                    e8 = base + off                            # byte offset
                    assert e8 % 4 == 0
                    e = e8 // 8
                    k, x = (b, e - b * ld) if xc else (e % ld, e // ld)
                    out.append((1024 * b + 256 * q + 4 * lane, k, x, e8))
    return out


@pytest.mark.parametrize("xc", [0, 1])
@pytest.mark.parametrize("xn", [128, 127, 104, 9, 1])
@pytest.mark.parametrize("ld", [1001, 2048])
def test_dtl_dword_loads_fill_the_same_slots(xc, xn, ld):
    """The 4-byte loads (operands with odd ld or 8-byte alignment): every
    load reads 4 bytes of an existing element of the stage's block, both
    halves of a double come from the same element, and every LDS slot of an
    existing (k, x) receives that element -- the same image as the 16-byte
    loads."""
    slot = {dtl_index(xc, kk, xx): (kk, xx) for kk in range(16) for xx in range(128)}
    halves = {}
    for byte, k, x, e8 in dtl_loads_dword(xc, xn, ld):
        assert 0 <= k < 16 and 0 <= x < xn, (xc, xn, byte, k, x)
        d = byte // 8
        halves.setdefault(d, set()).add((e8 // 8, (byte % 8) == (e8 % 8)))
        want_k, want_x = slot[d]
        if want_x < xn:
            assert (k, x) == (want_k, want_x), (xc, xn, ld, byte)
        else:
            assert k == want_k, (xc, xn, ld, byte)
    assert len(halves) == 16 * 128
    for d, hs in halves.items():
        assert len(hs) == 1 and all(same for _, same in hs), d   # one element, halves in place
