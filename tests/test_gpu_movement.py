"""GPU: pack/unpack copies, transposes and k-layer sums, bit-exact against the
oracle (src/mat_redist.c pack/unpack, src/utils.c copy/transpose,
src/mpi_op_omp.c sum)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("nrow,ncol,lds,ldd,off", [
    (1, 1, 1, 1, 0), (7, 5, 9, 6, 1), (64, 128, 128, 130, 0), (33, 1000, 1001, 1003, 3),
    (334, 2000, 8000, 2000, 0), (2, 70000, 70000, 70001, 0), (0, 5, 5, 5, 0), (5, 0, 5, 5, 0)])
def test_copy_matrix_block_bit_exact(hip, dt, nrow, ncol, lds, ldd, off):
    rnd = np.random.default_rng(nrow * 7 + ncol)
    src = rnd.standard_normal(max(1, nrow * lds + off)).astype(dt)
    dst0 = rnd.standard_normal(max(1, nrow * ldd + off)).astype(dt)
    want = dst0.copy()
    O.copy_block(src[off:], nrow, ncol, lds, want[off:], ldd)
    ds, dd = hip.DeviceArray.from_numpy(src), hip.DeviceArray.from_numpy(dst0)
    hip.copy_matrix_block(src.itemsize, nrow, ncol, ds.at(off), lds, dd.at(off), ldd)
    assert np.array_equal(dd.numpy().view(np.uint8), want.view(np.uint8))


def test_batched_copy_many_blocks_one_call(hip):
    """All blocks of a pack in one launch; > 48 blocks forces a split."""
    rnd = np.random.default_rng(11)
    src = rnd.standard_normal(300000)
    dst = np.zeros(400000)
    ds, dd = hip.DeviceArray.from_numpy(src), hip.DeviceArray.from_numpy(dst)
    blocks, want = [], dst.copy()
    so = do = 0
    for b in range(70):
        nr, nc = int(rnd.integers(1, 40)), int(rnd.integers(1, 90))
        lds, ldd = nc + int(rnd.integers(0, 5)), nc + int(rnd.integers(0, 5))
        O.copy_block(src[so:], nr, nc, lds, want[do:], ldd)
        blocks.append((ds.at(so), lds, dd.at(do), ldd, nr, nc))
        so += nr * lds + int(rnd.integers(0, 3)); do += nr * ldd + int(rnd.integers(0, 3))
    hip.copy_blocks(8, blocks)
    assert np.array_equal(dd.numpy(), want)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_batched_copy_random_blocks(hip, dt):
    """Seeded random batches (empty blocks, long and short rows, leading
    dimensions and offsets of any parity, so aligned and unaligned launches
    both run, more blocks than one launch takes): bit-exact against the
    oracle's block copy, bytes outside the blocks untouched."""
    rnd = np.random.default_rng(606 + np.dtype(dt).itemsize)
    for case in range(6):
        src = rnd.standard_normal(400000).astype(dt)
        dst = rnd.standard_normal(500000).astype(dt)
        want = dst.copy()
        ds, dd = hip.DeviceArray.from_numpy(src), hip.DeviceArray.from_numpy(dst)
        blocks = []
        so = int(rnd.integers(0, 4))
        do = int(rnd.integers(0, 4))
        for b in range(int(rnd.integers(1, 120))):
            nr = int(rnd.integers(0, 30))
            nc = int(rnd.choice([0, 1, 2, 3, int(rnd.integers(1, 64)), int(rnd.integers(64, 1500))]))
            lds, ldd = max(nc, 1) + int(rnd.integers(0, 4)), max(nc, 1) + int(rnd.integers(0, 4))
            if so + nr * lds >= src.size or do + nr * ldd >= dst.size:
                break
            O.copy_block(src[so:], nr, nc, lds, want[do:], ldd)
            blocks.append((ds.at(so), lds, dd.at(do), ldd, nr, nc))
            so += nr * lds + int(rnd.integers(0, 4))
            do += nr * ldd + int(rnd.integers(0, 4))
        hip.copy_blocks(np.dtype(dt).itemsize, blocks)
        assert np.array_equal(dd.numpy().view(np.uint8), want.view(np.uint8)), case


def test_redistribution_pack_unpack_matches_oracle(hip):
    """One rank's pack and unpack lists of a 6-rank world, run on the device,
    give the same send buffer and destination block as the oracle."""
    p, m, n, k = 6, 120, 90, 150
    w = O.ca3dmm_world(m, n, k, 0, 0, p)
    red = w["redist"]["A"]
    rnd = np.random.default_rng(5)
    r = 1
    s = red["src"][r]
    src = rnd.standard_normal(s[2] * s[3])        # row-major view: s[2] rows of s[3]
    sl, sc, rl, rc = red["plans"][r]
    want_send = np.zeros(max(1, sc))
    blocks = []
    dsrc = hip.DeviceArray.from_numpy(src)
    dsend = hip.DeviceArray.zeros((max(1, sc),))
    for e in sl:
        lr, lc = e[0] - s[0], e[1] - s[1]
        O.copy_block(src[lr * s[3] + lc:], e[2], e[3], s[3], want_send[e[5]:], e[3])
        blocks.append((dsrc.at(lr * s[3] + lc), s[3], dsend.at(e[5]), e[3], e[2], e[3]))
    hip.copy_blocks(8, blocks)
    assert np.array_equal(dsend.numpy()[:sc], want_send[:sc])
    # unpack from a synthetic receive buffer
    q = red["req"][r]
    recv = rnd.standard_normal(max(1, rc))
    want_dst = np.zeros(max(1, q[2] * q[3]))
    drecv = hip.DeviceArray.from_numpy(recv)
    ddst = hip.DeviceArray.zeros((max(1, q[2] * q[3]),))
    blocks = []
    for e in rl:
        lr, lc = e[0] - q[0], e[1] - q[1]
        O.copy_block(recv[e[5]:], e[2], e[3], e[3], want_dst[lr * q[3] + lc:], q[3])
        blocks.append((drecv.at(e[5]), e[3], ddst.at(lr * q[3] + lc), q[3], e[2], e[3]))
    hip.copy_blocks(8, blocks)
    assert np.array_equal(ddst.numpy(), want_dst)


@pytest.mark.parametrize("nsrc,nrow,ncol", [(1, 5, 3), (2, 64, 64), (3, 2000, 37), (3, 1001, 13), (8, 17, 129),
                                            (16, 40, 7), (17, 33, 9), (21, 64, 5), (40, 31, 11)])
def test_fold_sum_bit_exact(hip, nsrc, nrow, ncol):
    rnd = np.random.default_rng(nsrc * 100 + nrow)
    parts = [np.asfortranarray(rnd.standard_normal((nrow, ncol)) * 10.0 ** rnd.integers(-8, 8))
             for _ in range(nsrc)]
    want = np.zeros((nrow, ncol), order="F")
    wflat = want.ravel(order="F")
    assert np.shares_memory(wflat, want)       # the oracle writes through this view
    O.fold_sum([p.ravel(order="F") for p in parts], nrow, ncol, [nrow] * nsrc, wflat, nrow)
    dparts = [hip.DeviceArray.from_numpy(p) for p in parts]
    out = hip.DeviceArray.zeros((nrow, ncol))
    hip.fold_sum(dparts, nrow, ncol, out)
    got = out.numpy()
    assert np.array_equal(got, want)           # the oracle's fixed-order fold
    ref = parts[0].copy()
    for t in range(1, nsrc):
        ref = ref + parts[t]
    assert np.array_equal(want, ref)           # which is ((P0 + P1) + P2) + ...


@pytest.mark.parametrize("dt", [np.float64, np.float32, np.int32])
def test_op_sum_bit_exact(hip, dt):
    rnd = np.random.default_rng(9)
    a = (rnd.standard_normal(100003) * 1000).astype(dt)
    b = (rnd.standard_normal(100003) * 1000).astype(dt)
    want = b.copy()
    O.op_sum(a, want)
    da, db = hip.DeviceArray.from_numpy(a), hip.DeviceArray.from_numpy(b)
    hip.op_sum(da, db)
    assert np.array_equal(db.numpy(), want)


@pytest.mark.parametrize("nrow,ncol", [(1, 1), (5, 3), (64, 64), (130, 67), (1000, 333)])
def test_transpose_bit_exact(hip, nrow, ncol):
    rnd = np.random.default_rng(nrow + ncol)
    A = np.asfortranarray(rnd.standard_normal((ncol, nrow)))
    want = np.zeros(nrow * ncol)
    O.transpose_cm(nrow, ncol, A.ravel(order="F"), ncol, want, nrow)
    dA = hip.DeviceArray.from_numpy(A)
    dT = hip.DeviceArray.zeros((nrow, ncol))
    hip.transpose_cm(nrow, ncol, dA, dT)
    assert np.array_equal(dT.numpy().ravel(order="F"), want)


def test_memset_fill(hip):
    d = hip.DeviceArray.filled((1003,), 0xAB, np.uint8)
    assert np.all(d.numpy() == 0xAB)


def test_copy_long_rows_and_mixed_alignment(hip):
    """Rows longer than a copy tile (the allgather's single-row blocks) are
    cut into tile-long pieces, and a list mixing 16-byte aligned and
    unaligned blocks goes out with the widest word each block allows:
    bit-exact either way."""
    rnd = np.random.default_rng(21)
    src = rnd.standard_normal(700001)
    dst = np.zeros(700001)
    ds, dd = hip.DeviceArray.from_numpy(src), hip.DeviceArray.from_numpy(dst)
    blocks = [(ds.at(0), 300000, dd.at(0), 300000, 1, 300000),          # one long aligned row
              (ds.at(300001), 150000, dd.at(300001), 150000, 2, 149999),  # odd offsets: 8-byte words
              (ds.at(600002), 33333, dd.at(600002), 33333, 3, 33333)]
    hip.copy_blocks(8, blocks)
    got = dd.numpy()
    want = np.zeros(700001)
    want[:300000] = src[:300000]
    for r in range(2):
        want[300001 + r * 150000:300001 + r * 150000 + 149999] = src[300001 + r * 150000:300001 + r * 150000 + 149999]
    want[600002:600002 + 3 * 33333] = src[600002:600002 + 3 * 33333]
    assert np.array_equal(got, want)


def test_dgeam_random_cases(hip):
    """hip_dgeam (the reference's cuda_cublas_dgeam, used as a transpose):
    C := alpha op(A) + beta op(B) over seeded random shapes, trans cases,
    scalars (beta = 0 ignores B) and padded leading dimensions; exact on
    integers, the padding of C untouched."""
    from ca3dmm_amd._lib import lib
    rnd = np.random.default_rng(1414)
    for case in range(12):
        m, n = (int(v) for v in rnd.integers(1, 700, 2))
        ta, tb = (int(v) for v in rnd.integers(0, 2, 2))
        alpha, beta = float(rnd.integers(-3, 4)), float(rnd.integers(-3, 4))
        A = rnd.integers(-50, 51, (n, m) if ta else (m, n)).astype(np.float64)
        B = rnd.integers(-50, 51, (n, m) if tb else (m, n)).astype(np.float64)
        lda, ldb, ldc = A.shape[0] + int(rnd.integers(0, 3)), B.shape[0] + int(rnd.integers(0, 3)), m + 2
        bA = np.zeros((lda, A.shape[1]), order="F"); bA[:A.shape[0]] = A
        bB = np.zeros((ldb, B.shape[1]), order="F"); bB[:B.shape[0]] = B
        C0 = np.full((ldc, n), 7.0, order="F")
        dA, dB, dC = (hip.DeviceArray.from_numpy(x) for x in (bA, bB, C0))
        rc = lib.hip_dgeam(ta, tb, m, n, alpha, dA.ptr, lda, beta, dB.ptr, ldb, dC.ptr, ldc)
        assert rc == 0, case
        got = dC.numpy()
        want = alpha * (A.T if ta else A) + beta * (B.T if tb else B)
        assert np.array_equal(got[:m], want), case
        assert np.all(got[m:] == 7.0), case
