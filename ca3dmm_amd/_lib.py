"""Loader for libca3dmm_amd.so (the C-ABI library built from ca3dmm_amd/csrc).

There is no fallback: if the library is not built, importing this module
raises ImportError.  Argument types are declared for every exported symbol
so that ctypes checks them at call time.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libca3dmm_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
        "(or `make -C ca3dmm_amd/csrc`)")

lib = C.CDLL(LIB_PATH)

I = C.c_int
D = C.c_double
P = C.c_void_p
SZ = C.c_size_t
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)
PP = C.POINTER(C.c_void_p)


class BlkCopy(C.Structure):
    """hip_blk_copy_t (include/ca3dmm_ext.h)."""
    _fields_ = [("src", P), ("dst", P), ("lds", C.c_int64), ("ldd", C.c_int64),
                ("nrow", C.c_int32), ("ncol", C.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    # hip_proxy.h
    "select_hip_device_by_mpi_local_rank": (I, []),
    "hip_set_rt_dev_id": (I, [I]),
    "hip_proxy_device_count": (I, []),
    "hip_proxy_set_sync": (I, [I]),
    "hip_device_sync": (I, []),
    "hip_stream_sync": (I, [P]),
    "hip_proxy_stream": (P, []),
    "hip_timer_start": (I, []),
    "hip_timer_stop": (I, [DP]),
    "hip_proxy_last_error": (C.c_char_p, []),
    "hip_malloc_dev": (I, [PP, SZ]),
    "hip_malloc_host": (I, [PP, SZ]),
    "hip_memset_dev": (I, [P, I, SZ]),
    "hip_free_dev": (I, [P]),
    "hip_free_host": (I, [P]),
    "hip_memcpy_h2d": (I, [P, P, SZ]),
    "hip_memcpy_d2h": (I, [P, P, SZ]),
    "hip_memcpy_d2d": (I, [P, P, SZ]),
    "hip_memcpy_auto": (I, [P, P, SZ]),
    "hip_copy_matrix_block": (I, [SZ, I, I, P, I, P, I]),
    "hip_dcopy": (I, [I, P, I, P, I]),
    "hip_dgemm": (I, [I, I, I, I, I, D, P, I, P, I, D, P, I]),
    "hip_dgeam": (I, [I, I, I, I, D, P, I, D, P, I, P, I]),
    "hip_dtrsm": (I, [I, I, I, I, I, I, D, P, I, P, I]),
    "hip_dpotrf": (I, [C.c_char, I, P, I]),
    # ca3dmm_ext.h
    "hip_copy_matrix_blocks": (I, [SZ, I, C.POINTER(BlkCopy)]),
    "hip_op_sum": (I, [I, I, P, P]),
    "hip_fold_sum": (I, [I, PP, IP, I, I, P, I]),
    "hip_dgemm_fold": (I, [I, I, I, I, I, D, P, I, P, I, D, P, I, I, PP, IP]),
    "hip_transpose_cm": (I, [I, I, P, I, P, I]),
    "hip_ozaki_dgemm": (I, [I, I, I, I, I, D, P, I, P, I, D, P, I, I]),
    "hip_ozaki_workspace_bytes": (SZ, [I, I, I, I]),
    "hip_ozaki2_dgemm": (I, [I, I, I, I, I, D, P, I, P, I, D, P, I, I]),
    "hip_ozaki2_workspace_bytes": (SZ, [I, I, I, I]),
    "hip_ozaki_choose": (I, [I, I, I, I, I, P, I, P, I, IP, IP]),
    "hip_ozaki_counts": (I, [D, D, I, IP, IP]),
    "hip_dsyrk_gram": (I, [I, I, D, P, I, D, P, I, I]),
    "hip_dsyrk_gram_noalloc": (I, [I, I, D, P, I, D, P, I, I]),
    "hip_ozaki_gram": (I, [I, I, D, P, I, D, P, I, I, I]),
    "hip_dtrtri_upper": (I, [I, P, I, P, I]),
    "hip_cholesky_qr": (I, [I, I, P, I, P, I, I, I, IP]),
    # dev_type.h
    "is_dev_type_valid": (I, [I]),
    "dev_type_malloc": (P, [SZ, I]),
    "dev_type_free": (None, [P, I]),
    "dev_type_realloc": (None, [C.POINTER(SZ), SZ, I, PP]),
    "dev_type_memset": (None, [P, I, SZ, I]),
    "dev_type_memcpy": (None, [P, P, SZ, I, I]),
    "dev_type_copy_mat_blk": (None, [SZ, I, I, P, I, P, I, I]),
    "dev_type_dgemm_cm": (None, [I, I, I, I, I, D, P, I, P, I, D, P, I, I]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class HipError(RuntimeError):
    pass


def check(rc, what="call"):
    """Raise HipError for a negative status from the library."""
    if rc is not None and rc < 0:
        msg = lib.hip_proxy_last_error()
        raise HipError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc
