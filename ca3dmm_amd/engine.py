"""ctypes mirrors of the engine structs (include/mat_redist.h, cannon.h,
ca3dmm.h) and thin wrappers over the engine and simulated-grid entry points.
Field order matches the C headers exactly."""
import ctypes as C

import numpy as np

from ._lib import check, lib

I, D, P, SZ = C.c_int, C.c_double, C.c_void_p, C.c_size_t
IP = C.POINTER(C.c_int)

DEV_TYPE_HOST, DEV_TYPE_HIP, DEV_TYPE_HIP_MPI_DIRECT = 0, 1, 2


class Comm(C.Structure):
    _fields_ = [("nproc", I), ("rank", I)]


class MatRedist(C.Structure):
    _fields_ = [("comm", Comm), ("dtype", I), ("dt_size", SZ)] + [(f, I) for f in (
        "nproc", "rank", "src_srow", "src_scol", "src_nrow", "src_ncol",
        "req_srow", "req_scol", "req_nrow", "req_ncol", "n_proc_send", "n_proc_recv",
        "send_cnt", "recv_cnt", "alloc_workbuf")] + [(f, IP) for f in (
        "send_ranks", "send_sizes", "send_displs", "sblk_sizes",
        "recv_ranks", "recv_sizes", "recv_displs", "rblk_sizes", "send_info0", "recv_info0")] + [
        (f, P) for f in ("sendbuf_h", "recvbuf_h", "sendbuf_d", "recvbuf_d", "workbuf_h", "workbuf_d")] + [
        ("hd_trans_ms", D), ("dev_type", I)]

    def send_list(self):
        """rows of {srow, scol, nrow, ncol, peer, displ} (oracle list format)."""
        out = []
        for i in range(self.n_proc_send):
            b = [self.sblk_sizes[4 * i + j] for j in range(4)]
            out.append(b + [self.send_ranks[i], self.send_displs[i]])
        return np.array(out, dtype=np.int32).reshape(-1, 6)

    def recv_list(self):
        out = []
        for i in range(self.n_proc_recv):
            b = [self.rblk_sizes[4 * i + j] for j in range(4)]
            out.append(b + [self.recv_ranks[i], self.recv_displs[i]])
        return np.array(out, dtype=np.int32).reshape(-1, 6)


class Cannon(C.Structure):
    _fields_ = [(f, I) for f in (
        "m", "n", "k", "my_rank", "rank_row", "rank_col", "np_dim",
        "A_srow", "A_scol", "B_srow", "B_scol", "C_srow", "C_scol",
        "A_nrow", "A_ncol", "B_nrow", "B_ncol", "C_nrow", "C_ncol",
        "gemm_cycle", "max_A_blk_size", "max_B_blk_size", "max_C_blk_size", "alloc_workbuf")] + [
        ("m_displs", IP), ("n_displs", IP), ("k_displs", IP)] + [(f, P) for f in (
            "A_recv_h", "A_stack_h", "B_recv_h", "B_stack_h", "A_recv_d", "A_stack_d",
            "B_recv_d", "B_stack_d", "workbuf_h", "workbuf_d", "A_gemm_h", "B_gemm_h")] + [
        ("dev_type", I), ("comm", Comm)] + [(f, D) for f in (
            "init_ms", "shift0_ms", "lshift_ms", "gemm_ms", "exec_ms", "hd_trans_ms")] + [("n_exec", I)]


_CA3_INTS = ("m", "n", "k", "mp", "np", "kp", "rp", "my_rank", "n_proc", "is_active",
             "rank_m", "rank_n", "rank_k", "trans_A", "trans_B", "task_m_num", "task_m_id",
             "task_n_num", "task_n_id", "task_k_num", "task_k_id", "is_BTB", "use_ag", "use_rsb",
             "A_rd_srow", "A_rd_scol", "B_rd_srow", "B_rd_scol", "A_2dmm_srow", "A_2dmm_scol",
             "B_2dmm_srow", "B_2dmm_scol", "C_2dmm_srow", "C_2dmm_scol", "C_out_srow", "C_out_scol",
             "A_rd_nrow", "A_rd_ncol", "B_rd_nrow", "B_rd_ncol", "A_2dmm_nrow", "A_2dmm_ncol",
             "B_2dmm_nrow", "B_2dmm_ncol", "C_2dmm_nrow", "C_2dmm_ncol", "C_out_nrow", "C_out_ncol",
             "alloc_workbuf")


class Ca3dmm(C.Structure):
    _fields_ = [(f, I) for f in _CA3_INTS] + [
        ("AB_agv_recvcnts", IP), ("AB_agv_displs", IP), ("C_rs_recvcnts", IP)] + [(f, P) for f in (
            "A_rd_recv_h", "A_trans_h", "A_2dmm_h", "B_rd_recv_h", "B_trans_h", "B_2dmm_h", "C_2dmm_h",
            "C_out_h", "A_rd_recv_d", "A_trans_d", "A_2dmm_d", "B_rd_recv_d", "B_trans_d", "B_2dmm_d",
            "C_2dmm_d", "C_out_d", "workbuf_h", "workbuf_d")] + [(f, SZ) for f in (
            "rdA_workbuf_bytes", "rdB_workbuf_bytes", "rdC_workbuf_bytes", "cannon_workbuf_bytes",
            "self_workbuf_bytes")] + [
        ("comm_AB_agv", Comm), ("comm_C_rs", Comm), ("comm_2dmm", Comm),
        ("redist_A", C.POINTER(MatRedist)), ("redist_B", C.POINTER(MatRedist)),
        ("redist_C", C.POINTER(MatRedist)), ("cannon_engine", C.POINTER(Cannon)), ("dev_type", I),
        ("print_timing", I)] + [(f, D) for f in (
            "init_ms", "redist_ms", "agvAB_ms", "cannon_ms", "reduce_ms", "hd_trans_ms", "exec_ms")] + [
        ("n_exec", I)]


CA3P = C.POINTER(Ca3dmm)

_SIGS = {
    "ca3dmm_comm_self": (Comm, []),
    "calc_3d_decomposition_cannon": (None, [I, I, I, I, IP, IP, IP, IP]),
    "calc_3d_decomposition_nk": (None, [I, I, I, IP, IP, IP]),
    "calc_block_spos_size": (None, [I, I, I, IP, IP]),
    "mat_redist_engine_init": (None, [I, I, I, I, I, I, I, I, Comm, I, SZ, I,
                                      C.POINTER(C.POINTER(MatRedist)), C.POINTER(SZ)]),
    "mat_redist_engine_attach_workbuf": (None, [C.POINTER(MatRedist), P, P]),
    "mat_redist_engine_exec": (None, [C.POINTER(MatRedist), P, I, P, I]),
    "mat_redist_engine_free": (None, [C.POINTER(C.POINTER(MatRedist))]),
    "cannon_engine_init": (None, [I, I, I, Comm, I, C.POINTER(C.POINTER(Cannon)), C.POINTER(SZ)]),
    "cannon_engine_attach_workbuf": (None, [C.POINTER(Cannon), P, P]),
    "cannon_engine_free": (None, [C.POINTER(C.POINTER(Cannon))]),
    "cannon_engine_exec": (None, [C.POINTER(Cannon), D, D, P, P, P]),
    "cannon_engine_reset_stat": (None, [C.POINTER(Cannon)]),
    "cannon_engine_print_stat": (None, [C.POINTER(Cannon)]),
    "ca3dmm_engine_init": (None, [I, I, I, I, I] + [I] * 12 + [IP, Comm, I, C.POINTER(CA3P), C.POINTER(SZ)]),
    "ca3dmm_engine_init_BTB": (None, [I, I] + [I] * 8 + [IP, Comm, I, C.POINTER(CA3P), C.POINTER(SZ)]),
    "ca3dmm_engine_attach_workbuf": (None, [CA3P, P, P]),
    "ca3dmm_engine_free": (None, [C.POINTER(CA3P)]),
    "ca3dmm_engine_exec": (None, [CA3P, P, I, P, I, P, I]),
    "ca3dmm_engine_reset_stat": (None, [CA3P]),
    "ca3dmm_engine_print_stat": (None, [CA3P]),
    "ca3dmm_sim_init": (I, [I, I, I, I, I, I, IP, IP, IP, IP, I, I, I, C.POINTER(P)]),
    "ca3dmm_sim_workbuf_bytes": (SZ, [P, I]),
    "ca3dmm_sim_engine": (CA3P, [P, I]),
    "ca3dmm_sim_exec": (I, [P, C.POINTER(P), IP, C.POINTER(P), IP, C.POINTER(P), IP]),
    "ca3dmm_sim_c_out": (P, [P, I]),
    "ca3dmm_sim_free": (None, [C.POINTER(P)]),
    "ca3dmm_sim_set_mode": (I, [P, I, I]),
    "ca3dmm_sim_autotune": (I, [P, I, IP, IP, IP]),
    "ca3dmm_sim_set_tuning": (I, [P, I, I, I]),
    "ca3dmm_sim_set_ozaki": (I, [P, I]),
    "ca3dmm_sim_get_ozaki": (I, [P, IP, IP, C.POINTER(D)]),
    "ca3dmm_sim_autotune_ozaki": (I, [P, I, IP, IP]),
    "ca3dmm_sim_set_tuning_ozaki": (I, [P, I, I]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _ints(vals):
    vals = list(vals)
    return (C.c_int * max(1, len(vals)))(*vals)


def comm_self():
    return lib.ca3dmm_comm_self()


class SimGrid:
    """All ranks of a p-rank grid on one GPU (include/ca3dmm_sim.h)."""

    def __init__(self, p, m, n, k, transA, transB, src_A, src_B, dst_C=None, proc_grid=None,
                 is_btb=False, allocate=True, dev_type=DEV_TYPE_HIP):
        self.p = p
        flat = lambda rects: _ints([v for r in rects for v in r]) if rects is not None else None
        h = P()
        rc = lib.ca3dmm_sim_init(p, m, n, k, transA, transB, flat(src_A) if src_A is not None else None,
                                 flat(src_B), flat(dst_C) if dst_C is not None else None,
                                 _ints(proc_grid) if proc_grid is not None else None,
                                 1 if is_btb else 0, dev_type, 1 if allocate else 0, C.byref(h))
        check(rc, "ca3dmm_sim_init")
        self.h = h

    def engine(self, r):
        return lib.ca3dmm_sim_engine(self.h, r).contents

    def workbuf_bytes(self, r):
        return lib.ca3dmm_sim_workbuf_bytes(self.h, r)

    def exec(self, A_ptrs, ldA, B_ptrs, ldB, C_ptrs=None, ldC=None):
        p = self.p
        arr = lambda xs: (P * p)(*xs) if xs is not None else None
        return check(lib.ca3dmm_sim_exec(self.h, arr(A_ptrs), _ints(ldA) if ldA is not None else None,
                                         arr(B_ptrs), _ints(ldB), arr(C_ptrs),
                                         _ints(ldC) if ldC is not None else None), "ca3dmm_sim_exec")

    def set_mode(self, grouped=False, graph=False):
        return check(lib.ca3dmm_sim_set_mode(self.h, 1 if grouped else 0, 1 if graph else 0),
                     "ca3dmm_sim_set_mode")

    def autotune(self, reps=3):
        """Tune the FP64 grouped launch; returns (group_rows, wg_per_cu, tile)
        with tile 0 = 128x128, 1 = 256x128, 2 = 128x256 (persistent), 3 =
        the batched direct-to-LDS launch plus fold launches."""
        g, w, t = C.c_int(), C.c_int(), C.c_int()
        check(lib.ca3dmm_sim_autotune(self.h, reps, C.byref(g), C.byref(w), C.byref(t)), "ca3dmm_sim_autotune")
        return g.value, w.value, t.value

    def set_ozaki(self, nmoduli):
        """nmoduli > 0: the local products through the CRT form (0: FP64
        MFMA); -1: the CRT form with the count chosen at every exec."""
        return check(lib.ca3dmm_sim_set_ozaki(self.h, nmoduli), "ca3dmm_sim_set_ozaki")

    def get_ozaki(self):
        """(nmoduli, automatic, spread): the count the last exec used (or
        the one set), whether it is chosen per exec, and the spread
        statistic L of the last automatic choice (NaN before one)."""
        n, a, L = C.c_int(), C.c_int(), C.c_double()
        check(lib.ca3dmm_sim_get_ozaki(self.h, C.byref(n), C.byref(a), C.byref(L)), "ca3dmm_sim_get_ozaki")
        return n.value, bool(a.value), L.value

    def autotune_ozaki(self, reps=3):
        """Tune the Ozaki form's grouped launch; returns (tile_n, xcd_queue)."""
        t, q = C.c_int(), C.c_int()
        check(lib.ca3dmm_sim_autotune_ozaki(self.h, reps, C.byref(t), C.byref(q)), "ca3dmm_sim_autotune_ozaki")
        return t.value, q.value

    def set_tuning_ozaki(self, tile_n, xcd_queue):
        return check(lib.ca3dmm_sim_set_tuning_ozaki(self.h, tile_n, 1 if xcd_queue else 0),
                     "ca3dmm_sim_set_tuning_ozaki")

    def set_tuning(self, group_rows, wg_per_cu, tile=0):
        return check(lib.ca3dmm_sim_set_tuning(self.h, group_rows, wg_per_cu, tile), "ca3dmm_sim_set_tuning")

    def c_out(self, r):
        return lib.ca3dmm_sim_c_out(self.h, r)

    def free(self):
        if self.h:
            lib.ca3dmm_sim_free(C.byref(self.h))
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
