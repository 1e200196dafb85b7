"""CPU baseline of the CA3DMM exec path (TEST / BENCH INFRASTRUCTURE ONLY).

A CPU restatement of the reference's MPI + OpenMP exec sequence
(src/ca3dmm.c:926-1372), run as p simulated ranks in one process on the
example_AB layout (examples/example_AB.c:85-130: 1D column partition of A, B
and C, uniform[-1,1] inputs from a fixed seed; one warm-up then n_test timed
executions, :166-190):

  1. redistribute A and B        mat_redist pack / exchange / unpack
                                 (src/mat_redist.c:317-416) through persistent
                                 send and receive buffers, lists from
                                 or_redist_plan, blocks by or_copy_block
  2. allgather of the replicated operand (src/ca3dmm.c:1065-1101): the
                                 members' column slices concatenated in task order
  3. Cannon                      q rounds per rank, round s multiplies the A block
                                 of (row, (row+col+s) % q) by the B block of
                                 ((row+col+s) % q, col), beta 0 then 1
                                 (src/cannon.c:319-433); the DGEMM is the host
                                 BLAS behind numpy (cblas_dgemm in the reference)
  4. reduce-scatter of C         fixed-order kp-layer sum (mpi_op_omp_sum,
                                 src/mpi_op_omp.c:6-23)      -> or_fold_sum
  5. redistribute C              as in 1.

The ranks run one after another, each DGEMM on all BLAS threads, so the
per-phase times are the sums over ranks; with the reference's p concurrent
ranks of cores/p threads each, the products take about as long.  Phase times
are reported like ca3dmm_engine_print_stat (src/ca3dmm.c:1390-1410).
The result is validated as example_AB does (:208-305): with the reference's
linear fill, every entry of C within 1e-12 relative of the closed form
sum_l (a i + b l)(c l + d j).
"""
import os
import time

import numpy as np

from oracle import oracle as O


def blas_name():
    try:
        cfg = np.show_config(mode="dicts")
        return cfg["Build Dependencies"]["blas"]["name"]
    except Exception:  # pragma: no cover - numpy without config dicts
        return "numpy-blas"


def threads():
    for k in ("OPENBLAS_NUM_THREADS", "OMP_NUM_THREADS"):
        v = os.environ.get(k)
        if v and v.isdigit():
            return int(v)
    return os.cpu_count() or 1


class Flow:
    """All p ranks of one (m, n, k, transA, transB) problem, planned by the
    oracle as ca3dmm_engine_init would plan them."""

    def __init__(self, m, n, k, transA, transB, p, proc_grid=None):
        self.m, self.n, self.k, self.tA, self.tB, self.p = m, n, k, transA, transB, p
        self.w = O.ca3dmm_world(m, n, k, transA, transB, p, proc_grid=proc_grid)
        self.plans = self.w["plans"]
        pl0 = self.plans[0]
        self.mp, self.np_, self.kp = pl0.mp, pl0.np, pl0.kp
        self.An, self.Ac = (k, m) if transA else (m, k)
        self.Bn, self.Bc = (n, k) if transB else (k, n)
        self.stats = {"redist": 0.0, "allgather": 0.0, "cannon": 0.0, "reduce_scatter": 0.0, "exec": 0.0}
        self.n_exec = 0
        # per-rank buffers (row-major rects of mat_redist are column ranges here)
        self.A_rd = [np.zeros(max(1, self._area(self.w["redist"]["A"]["req"][r]))) for r in range(p)]
        self.B_rd = [np.zeros(max(1, self._area(self.w["redist"]["B"]["req"][r]))) for r in range(p)]
        self.C_out = [np.zeros(max(1, self._area(self.w["redist"]["C"]["src"][r]))) for r in range(p)]
        # mat_redist work buffers, attached once (src/mat_redist.c:239-267)
        self.bufs = {name: ([np.zeros(max(1, pl[1])) for pl in red["plans"]],
                            [np.zeros(max(1, pl[3])) for pl in red["plans"]])
                     for name, red in self.w["redist"].items()}

    @staticmethod
    def _area(rect):
        return int(rect[2]) * int(rect[3])

    def _redist(self, name, src, src_ld, dst, dst_ld):
        """Pack, exchange and unpack of one matrix for all ranks
        (src/mat_redist.c:317-416); rects and lists are row-major."""
        red = self.w["redist"][name]
        sbuf, rbuf = self.bufs[name]
        for r in range(self.p):
            s0, s1 = int(red["src"][r][0]), int(red["src"][r][1])
            for e in red["plans"][r][0]:
                off = (int(e[0]) - s0) * src_ld[r] + (int(e[1]) - s1)
                O.copy_block(src[r][off:], int(e[2]), int(e[3]), src_ld[r], sbuf[r][int(e[5]):], int(e[3]))
        for r in range(self.p):
            for e in red["plans"][r][2]:
                peer, n = int(e[4]), int(e[2]) * int(e[3])
                f = next(f for f in red["plans"][peer][0] if int(f[4]) == r)
                rbuf[r][int(e[5]):int(e[5]) + n] = sbuf[peer][int(f[5]):int(f[5]) + n]
        for r in range(self.p):
            q0, q1 = int(red["req"][r][0]), int(red["req"][r][1])
            for e in red["plans"][r][2]:
                off = (int(e[0]) - q0) * dst_ld[r] + (int(e[1]) - q1)
                O.copy_block(rbuf[r][int(e[5]):], int(e[2]), int(e[3]), int(e[3]), dst[r][off:], dst_ld[r])

    def _member(self, pl, task, row_col=None):
        """Global rank with pl's layer, the given task id and either pl's
        position in its Cannon group or (row, col) in it."""
        group = self.mp * self.np_ // max(pl.task_m_num * pl.task_n_num, 1)
        q = int(round(group ** 0.5))
        r2 = pl.rank_2dmm if row_col is None else row_col[0] * q + row_col[1]
        return pl.rank_k * self.mp * self.np_ + task * group + r2

    def _op_block(self, buf, nrow, ncol, trans):
        """The redistributed block as a column-major nrow x ncol op() block
        (transpose_cm_mat, src/utils.c:181-196, when op() transposes)."""
        if not trans:
            return buf[:nrow * ncol].reshape((ncol, nrow)).T
        return buf[:nrow * ncol].reshape((nrow, ncol))    # stored ncol x nrow: its transpose

    def exec(self, A_blocks, B_blocks, C_blocks):
        """A_blocks[r], B_blocks[r]: rank r's column block (flat, column-major,
        leading dim = rows of A / B as stored); C_blocks[r]: its C output."""
        p, plans, red = self.p, self.plans, self.w["redist"]
        t_exec = t0 = time.perf_counter()
        self._redist("A", A_blocks, [self.An] * p, self.A_rd, [max(1, int(r[3])) for r in red["A"]["req"]])
        self._redist("B", B_blocks, [self.Bn] * p, self.B_rd, [max(1, int(r[3])) for r in red["B"]["req"]])
        t1 = time.perf_counter(); self.stats["redist"] += t1 - t0; t0 = t1
        # op() blocks of the redistribution, then the allgather of the
        # replicated operand along the columns of its Cannon block
        Ard = [self._op_block(self.A_rd[r], pl.A_rd_nrow, pl.A_rd_ncol, self.tA) if pl.is_active else None
               for r, pl in enumerate(plans)]
        Brd = [self._op_block(self.B_rd[r], pl.B_rd_nrow, pl.B_rd_ncol, self.tB) if pl.is_active else None
               for r, pl in enumerate(plans)]
        A2, B2 = [None] * p, [None] * p
        for r, pl in enumerate(plans):
            if not pl.is_active:
                continue
            if pl.task_n_num > 1:
                A2[r] = np.hstack([Ard[self._member(pl, t)] for t in range(pl.task_n_num)])
            else:
                A2[r] = np.asfortranarray(Ard[r])
            if pl.task_m_num > 1:
                B2[r] = np.hstack([Brd[self._member(pl, t)] for t in range(pl.task_m_num)])
            else:
                B2[r] = np.asfortranarray(Brd[r])
        t1 = time.perf_counter(); self.stats["allgather"] += t1 - t0; t0 = t1
        # Cannon rounds
        C2 = [None] * p
        for r, pl in enumerate(plans):
            if not pl.is_active:
                continue
            c = pl.cannon
            q, row, col = c.np_dim, c.rank_row, c.rank_col
            task = pl.task_m_id + pl.task_n_id
            Cb = np.zeros((c.C_nrow, c.C_ncol), order="F")
            for s in range(q):
                off = (row + col + s) % q
                a = A2[self._member(pl, task, (row, off))]
                b = B2[self._member(pl, task, (off, col))]
                if a.shape[0] and b.shape[1] and a.shape[1]:
                    Cb += a @ b
            C2[r] = Cb
        t1 = time.perf_counter(); self.stats["cannon"] += t1 - t0; t0 = t1
        # reduce-scatter over the kp layers, fixed (layer) order
        mn = self.mp * self.np_
        for r, pl in enumerate(plans):
            if not pl.is_active or not pl.C_out_nrow or not pl.C_out_ncol:
                continue
            off = pl.C_out_scol - pl.C_2dmm_scol
            parts = [np.asfortranarray(C2[l * mn + r % mn][:, off:off + pl.C_out_ncol]).ravel(order="F")
                     for l in range(self.kp)]
            O.fold_sum(parts, pl.C_out_nrow, pl.C_out_ncol, [pl.C_out_nrow] * self.kp, self.C_out[r],
                       pl.C_out_nrow)
        t1 = time.perf_counter(); self.stats["reduce_scatter"] += t1 - t0; t0 = t1
        self._redist("C", self.C_out, [max(1, int(rr[3])) for rr in red["C"]["src"]], C_blocks, [self.m] * p)
        t1 = time.perf_counter()
        self.stats["redist"] += t1 - t0
        self.stats["exec"] += t1 - t_exec
        self.n_exec += 1

    def reset_stat(self):
        for key in self.stats:
            self.stats[key] = 0.0
        self.n_exec = 0

    def stat_ms(self):
        return {key: 1e3 * v / max(1, self.n_exec) for key, v in self.stats.items()}

    # --- the example_AB layout ------------------------------------------
    def blocks(self, A, B):
        """Rank r's column blocks of A and B (as stored) and empty C blocks."""
        Ab, Bb, Cb = [], [], []
        for r in range(self.p):
            _, _, sc, nc = self.w["src_A"][r]
            Ab.append(np.ascontiguousarray(A[:, sc:sc + nc].ravel(order="F")) if nc else np.zeros(1))
            _, _, sc, nc = self.w["src_B"][r]
            Bb.append(np.ascontiguousarray(B[:, sc:sc + nc].ravel(order="F")) if nc else np.zeros(1))
            _, _, sc, nc = self.w["dst_C"][r]
            Cb.append(np.zeros(max(1, self.m * nc)))
        return Ab, Bb, Cb

    def gather_C(self, Cb):
        out = np.empty((self.m, self.n), order="F")
        for r in range(self.p):
            _, _, sc, nc = self.w["dst_C"][r]
            if nc:
                out[:, sc:sc + nc] = Cb[r][:self.m * nc].reshape((nc, self.m)).T
        return out


def linear_fill_check(m, n, k, transA, transB, C):
    """example_AB's validation (examples/example_AB.c:208-305) against the
    closed form of its linear fill: the number of entries with relative
    error above 1e-12."""
    a, b = (0.24, 0.19) if transA else (0.19, 0.24)
    c, d = (0.12, 0.11) if transB else (0.11, 0.12)
    K = np.longdouble(k)
    S1, S2 = K * (K - 1) / 2, K * (K - 1) * (2 * K - 1) / 6
    ai = np.longdouble(a) * np.arange(m, dtype=np.longdouble)[:, None]
    dj = np.longdouble(d) * np.arange(n, dtype=np.longdouble)[None, :]
    want = K * ai * dj + (ai * np.longdouble(c) + np.longdouble(b) * dj) * S1 + np.longdouble(b) * np.longdouble(c) * S2
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.abs((C.astype(np.longdouble) - want) / want)
    return int(np.count_nonzero(rel > 1e-12))


def linear_fill(nrow, ncol, fr, fc):
    """X(i, j) = fr i + fc j on stored indices (examples/example_AB.c:107-128)."""
    return np.asfortranarray(fr * np.arange(nrow)[:, None] + fc * np.arange(ncol)[None, :])


def run_example_ab(m=2000, n=2000, k=2000, p=4, n_test=5, seed=20221113):
    """The example_AB CPU flow: p simulated ranks, NN, uniform[-1,1] inputs,
    one warm-up and n_test timed executions, then the linear-fill
    validation.  Returns a dict of averages (ms), TFLOP/s and errors."""
    fl = Flow(m, n, k, 0, 0, p)
    rnd = np.random.default_rng(seed)
    A = np.asfortranarray(rnd.uniform(-1, 1, (m, k)))
    B = np.asfortranarray(rnd.uniform(-1, 1, (k, n)))
    Ab, Bb, Cb = fl.blocks(A, B)
    fl.exec(Ab, Bb, Cb)                       # warm-up
    fl.reset_stat()
    for _ in range(n_test):
        fl.exec(Ab, Bb, Cb)
    ms = fl.stat_ms()
    # validation with the reference's linear fill
    Ab, Bb, Cb = fl.blocks(linear_fill(m, k, 0.19, 0.24), linear_fill(k, n, 0.11, 0.12))
    fl.exec(Ab, Bb, Cb)
    errors = linear_fill_check(m, n, k, 0, 0, fl.gather_C(Cb))
    return {"grid": [fl.mp, fl.np_, fl.kp], "ms": ms, "tflops": 2.0 * m * n * k / (ms["exec"] / 1e3) / 1e12,
            "errors": errors, "n_test": n_test}
