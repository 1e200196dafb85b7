"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP product is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.

Also exposes oracle/_ref/libca3dmm_ref.so -- the reference's own
src/partition.c and src/utils.c compiled unmodified -- when it has been built.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_I = C.c_int
_P = C.c_void_p
_IP = C.POINTER(C.c_int)


class CannonPlan(C.Structure):
    _fields_ = [(f, C.c_int) for f in (
        "m", "n", "k", "np_dim", "rank_row", "rank_col",
        "A_srow", "A_scol", "A_nrow", "A_ncol",
        "B_srow", "B_scol", "B_nrow", "B_ncol",
        "C_srow", "C_scol", "C_nrow", "C_ncol",
        "gemm_cycle", "max_A_blk_size", "max_B_blk_size", "max_C_blk_size")] + [
        ("workbuf_bytes", C.c_longlong)]


class Ca3dmmPlan(C.Structure):
    _fields_ = [(f, C.c_int) for f in (
        "mp", "np", "kp", "rp", "is_active", "rank_m", "rank_n", "rank_k",
        "task_m_num", "task_m_id", "task_n_num", "task_n_id", "task_k_num", "task_k_id",
        "rank_2dmm", "use_ag", "use_rsb",
        "A_rd_srow", "A_rd_scol", "A_rd_nrow", "A_rd_ncol",
        "B_rd_srow", "B_rd_scol", "B_rd_nrow", "B_rd_ncol",
        "A_2dmm_srow", "A_2dmm_scol", "A_2dmm_nrow", "A_2dmm_ncol",
        "B_2dmm_srow", "B_2dmm_scol", "B_2dmm_nrow", "B_2dmm_ncol",
        "C_2dmm_srow", "C_2dmm_scol", "C_2dmm_nrow", "C_2dmm_ncol",
        "C_out_srow", "C_out_scol", "C_out_nrow", "C_out_ncol")] + [
        ("self_workbuf_bytes", C.c_longlong), ("cannon_workbuf_bytes", C.c_longlong),
        ("cannon", CannonPlan)]


def _load(path):
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run `make -C oracle`")
    return C.CDLL(path)


lib = _load(os.path.join(_HERE, "liboracle.so"))
lib.or_block_spos_size.argtypes = [_I, _I, _I, _IP, _IP]
lib.or_grid_cannon.argtypes = [_I, _I, _I, _I, _IP, _IP, _IP, _IP]
lib.or_grid_nk.argtypes = [_I, _I, _I, _IP, _IP, _IP]
lib.or_copy_block.argtypes = [C.c_size_t, _I, _I, _P, _I, _P, _I]
lib.or_transpose_cm.argtypes = [_I, _I, _P, _I, _P, _I]
lib.or_op_sum.argtypes = [_I, _I, _P, _P]
lib.or_fold_sum.argtypes = [_I, C.POINTER(_P), _IP, _I, _I, _P, _I]
lib.or_redist_plan.argtypes = [_I, _IP, _IP, _I, _IP, _IP, _IP, _IP, _IP, _IP]
lib.or_redist_exec_all.argtypes = [_I, C.c_size_t, _IP, _IP, C.POINTER(_P), _IP,
                                   C.POINTER(_P), _IP]
lib.or_dgemm.argtypes = [_I, _I, _I, _I, _I, C.c_double, _P, _I, _P, _I, C.c_double, _P, _I]
lib.or_dd_sample.argtypes = [_I, _I, _I, _I, _I, _P, _I, _P, _I, _I, _P, _P, _P, _P]
lib.or_cannon_plan.argtypes = [_I, _I, _I, _I, _I, _I, C.POINTER(CannonPlan)]
lib.or_ca3dmm_plan.argtypes = [_I, _I, _I, _I, _I, _I, _I, _IP, _I, _I, _I,
                               C.POINTER(Ca3dmmPlan)]
lib.or_ca3dmm_plan.restype = _I


def _ptr(a):
    return a.ctypes.data_as(_P)


def block_spos_size(length, nblk, iblk):
    s, z = _I(), _I()
    lib.or_block_spos_size(length, nblk, iblk, C.byref(s), C.byref(z))
    return s.value, z.value


def grid_cannon(p, m, n, k):
    v = [_I() for _ in range(4)]
    lib.or_grid_cannon(p, m, n, k, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def grid_nk(p, n, k):
    v = [_I() for _ in range(3)]
    lib.or_grid_nk(p, n, k, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def copy_block(src, nrow, ncol, lds, dst, ldd):
    """Row-major block copy on flat numpy arrays (element size from dtype)."""
    lib.or_copy_block(src.itemsize, nrow, ncol, _ptr(src), lds, _ptr(dst), ldd)


def transpose_cm(nrow, ncol, A, ldA, AT, ldAT):
    lib.or_transpose_cm(nrow, ncol, _ptr(A), ldA, _ptr(AT), ldAT)


def op_sum(invec, inoutvec):
    code = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2}[invec.dtype]
    lib.or_op_sum(code, invec.size, _ptr(invec), _ptr(inoutvec))


def fold_sum(parts, nrow, ncol, lds, dst, ldd):
    arr = (_P * len(parts))(*[p.ctypes.data for p in parts])
    ld = (_I * len(parts))(*lds)
    lib.or_fold_sum(len(parts), arr, ld, nrow, ncol, _ptr(dst), ldd)


def redist_plan(all_src, all_req, rank):
    """all_src/all_req: (nproc, 4) int arrays of row-major rects."""
    nproc = len(all_src)
    s = np.ascontiguousarray(all_src, dtype=np.int32).ravel()
    r = np.ascontiguousarray(all_req, dtype=np.int32).ravel()
    sl = np.zeros(6 * nproc, np.int32)
    rl = np.zeros(6 * nproc, np.int32)
    ns, sc, nr, rc = _I(), _I(), _I(), _I()
    ip = lambda a: a.ctypes.data_as(_IP)
    lib.or_redist_plan(nproc, ip(s), ip(r), rank, ip(sl), C.byref(ns), C.byref(sc),
                       ip(rl), C.byref(nr), C.byref(rc))
    return (sl[:6 * ns.value].reshape(-1, 6), sc.value, rl[:6 * nr.value].reshape(-1, 6), rc.value)


def redist_exec_all(all_src, all_req, src_blks, src_lds, dst_blks, dst_lds):
    nproc = len(all_src)
    s = np.ascontiguousarray(all_src, dtype=np.int32).ravel()
    r = np.ascontiguousarray(all_req, dtype=np.int32).ravel()
    ip = lambda a: a.ctypes.data_as(_IP)
    sp = (_P * nproc)(*[b.ctypes.data for b in src_blks])
    dp = (_P * nproc)(*[b.ctypes.data for b in dst_blks])
    sld = (_I * nproc)(*src_lds)
    dld = (_I * nproc)(*dst_lds)
    lib.or_redist_exec_all(nproc, src_blks[0].itemsize, ip(s), ip(r), sp, sld, dp, dld)


def dgemm(transA, transB, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
    lib.or_dgemm(transA, transB, m, n, k, alpha, _ptr(A), lda, _ptr(B), ldb, beta, _ptr(Cm), ldc)


def dd_sample(transA, transB, m, n, k, A, lda, B, ldb, ii, jj):
    ii = np.ascontiguousarray(ii, np.int32)
    jj = np.ascontiguousarray(jj, np.int32)
    ref = np.zeros(len(ii))
    ab = np.zeros(len(ii))
    lib.or_dd_sample(transA, transB, m, n, k, _ptr(A), lda, _ptr(B), ldb, len(ii),
                     _ptr(ii), _ptr(jj), _ptr(ref), _ptr(ab))
    return ref, ab


def cannon_plan(m, n, k, q, rank2d, min_k_blk=160):
    pl = CannonPlan()
    lib.or_cannon_plan(m, n, k, q, rank2d, min_k_blk, C.byref(pl))
    return pl


def ca3dmm_plan(m, n, k, transA, transB, p, rank, proc_grid=None, reduce_kp=None,
                is_btb=False, min_k_blk=160):
    if reduce_kp is None:
        reduce_kp = 1 if is_btb else 0
    pl = Ca3dmmPlan()
    pg = (_I * 3)(*proc_grid) if proc_grid is not None else None
    rc = lib.or_ca3dmm_plan(m, n, k, transA, transB, p, rank, pg, reduce_kp,
                            1 if is_btb else 0, min_k_blk, C.byref(pl))
    if rc != 0:
        raise ValueError("invalid process grid")
    return pl


def plan_dict(pl):
    d = {f: getattr(pl, f) for f, _ in pl._fields_ if f != "cannon"}
    d["cannon"] = {f: getattr(pl.cannon, f) for f, _ in pl.cannon._fields_}
    return d


# --- the reference's own partition.c / utils.c, compiled unmodified -------
REF_PATH = os.path.join(_HERE, "_ref", "libca3dmm_ref.so")


def ref_lib():
    """Return the compiled reference library, or None if it is not built."""
    if not os.path.exists(REF_PATH):
        return None
    r = C.CDLL(REF_PATH)
    r.calc_3d_decomposition_cannon.argtypes = [_I, _I, _I, _I, _IP, _IP, _IP, _IP]
    r.calc_3d_decomposition_nk.argtypes = [_I, _I, _I, _IP, _IP, _IP]
    r.calc_block_spos_size.argtypes = [_I, _I, _I, _IP, _IP]
    r.copy_matrix_block.argtypes = [C.c_size_t, _I, _I, _P, _I, _P, _I, _I]
    r.transpose_cm_mat.argtypes = [_I, _I, _P, _I, _P, _I]
    return r


def ref_grid_cannon(r, p, m, n, k):
    v = [_I() for _ in range(4)]
    r.calc_3d_decomposition_cannon(p, m, n, k, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


def ref_grid_nk(r, p, n, k):
    v = [_I() for _ in range(3)]
    r.calc_3d_decomposition_nk(p, n, k, *[C.byref(x) for x in v])
    return tuple(x.value for x in v)


# --- whole-world planning for the example_AB / example_BTB layouts ---------
def col1d_rect(nrow, ncol, p, rank):
    """1D column partition of an nrow x ncol col-major matrix
    (examples/example_AB.c:86-100): (srow, nrow, scol, ncol)."""
    sc, nc = block_spos_size(ncol, p, rank)
    return (0, nrow, sc, nc)


def _rm(rect):
    """col-major (srow, nrow, scol, ncol) -> mat_redist row-major rect
    {srow, scol, nrow, ncol} with rows and columns swapped
    (src/ca3dmm.c:359-370)."""
    srow, nrow, scol, ncol = rect
    return (scol, srow, ncol, nrow)


def ca3dmm_world(m, n, k, transA, transB, p, proc_grid=None, reduce_kp=None, is_btb=False,
                 min_k_blk=160, src_A=None, src_B=None, dst_C=None):
    """Plan every rank as ca3dmm_engine_init would (src/ca3dmm.c:23-722).

    src_A/src_B/dst_C: per-rank col-major (srow, nrow, scol, ncol) lists;
    default is the example_AB 1D column layout.  dst_C=False skips the C
    redistribution.  Returns dict with 'plans', 'redist' and 'workbuf'."""
    An, Ac = (k, m) if transA else (m, k)
    Bn, Bc = (n, k) if transB else (k, n)
    if src_A is None:
        src_A = [col1d_rect(An, Ac, p, r) for r in range(p)]
    if src_B is None:
        src_B = [col1d_rect(Bn, Bc, p, r) for r in range(p)]
    if dst_C is None:
        dst_C = [col1d_rect(m if not is_btb else n, n, p, r) for r in range(p)]
    plans = [ca3dmm_plan(m, n, k, transA, transB, p, r, proc_grid, reduce_kp, is_btb, min_k_blk)
             for r in range(p)]

    def req_A(pl):
        if not pl.is_active:
            return (0, 0, 0, 0)
        rect = (pl.A_rd_srow, pl.A_rd_nrow, pl.A_rd_scol, pl.A_rd_ncol)
        if transA:
            rect = (rect[2], rect[3], rect[0], rect[1])
        return _rm(rect)

    def req_B(pl):
        if not pl.is_active:
            return (0, 0, 0, 0)
        rect = (pl.B_rd_srow, pl.B_rd_nrow, pl.B_rd_scol, pl.B_rd_ncol)
        if transB and not is_btb:
            rect = (rect[2], rect[3], rect[0], rect[1])
        return _rm(rect)

    def cout(pl):
        if not pl.is_active:
            return (0, 0, 0, 0)
        return _rm((pl.C_out_srow, pl.C_out_nrow, pl.C_out_scol, pl.C_out_ncol))

    red = {}
    mats = [("B", [_rm(r) for r in src_B], [req_B(pl) for pl in plans])]
    if not is_btb:
        mats.insert(0, ("A", [_rm(r) for r in src_A], [req_A(pl) for pl in plans]))
    if dst_C is not False:
        mats.append(("C", [cout(pl) for pl in plans], [_rm(r) for r in dst_C]))
    for name, srcs, reqs in mats:
        red[name] = {"src": srcs, "req": reqs,
                     "plans": [redist_plan(srcs, reqs, r) for r in range(p)]}
    workbuf = []
    for r, pl in enumerate(plans):
        child = [8 * (red[x]["plans"][r][1] + red[x]["plans"][r][3]) for x in red]
        if pl.is_active:
            child.append(pl.cannon_workbuf_bytes)
        workbuf.append(max(child + [0]) + (pl.self_workbuf_bytes if pl.is_active else 0))
    return {"plans": plans, "redist": red, "workbuf": workbuf,
            "src_A": src_A, "src_B": src_B, "dst_C": dst_C}
