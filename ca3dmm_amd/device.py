"""Device buffers and thin Python calls over the C-ABI.

The host side of tests and benchmarks: numpy arrays in, device buffers from
the library's own allocator (hip_malloc_dev), results copied back with
hip_memcpy_d2h.  Matrices are column-major (Fortran order), as in the
reference.
"""
import ctypes as C

import numpy as np

from ._lib import BlkCopy, check, lib

NOTRANS, TRANS = 0, 1


def device_count():
    return lib.hip_proxy_device_count()


class DeviceArray:
    """A device allocation holding a column-major 2-D array (or a vector)."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib.hip_malloc_dev(C.byref(p), self.nbytes), "hip_malloc_dev")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a):
        a = np.asarray(a)
        out = cls(a.shape, a.dtype)
        host = np.asfortranarray(a)
        if out.nbytes:
            check(lib.hip_memcpy_h2d(host.ctypes.data, out.ptr, out.nbytes), "hip_memcpy_h2d")
        return out

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        out = cls(shape, dtype)
        if out.nbytes:
            check(lib.hip_memset_dev(out.ptr, 0, out.nbytes), "hip_memset_dev")
        return out

    @classmethod
    def filled(cls, shape, byte, dtype=np.float64):
        out = cls(shape, dtype)
        if out.nbytes:
            check(lib.hip_memset_dev(out.ptr, byte, out.nbytes), "hip_memset_dev")
        return out

    def upload(self, a):
        """Overwrite the contents with host array a of the same shape."""
        host = np.asfortranarray(np.asarray(a, dtype=self.dtype))
        if host.shape != self.shape:
            raise ValueError(f"upload: shape {host.shape} into {self.shape}")
        if self.nbytes:
            check(lib.hip_memcpy_h2d(host.ctypes.data, self.ptr, self.nbytes), "hip_memcpy_h2d")
        return self

    def numpy(self):
        host = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.nbytes:
            check(lib.hip_stream_sync(None), "hip_stream_sync")
            check(lib.hip_memcpy_d2h(self.ptr, host.ctypes.data, self.nbytes), "hip_memcpy_d2h")
        return host

    def sample(self, ii, jj):
        """Entries (ii[t], jj[t]) of a 2-D array, read one by one (for arrays
        too large to copy back whole)."""
        assert len(self.shape) == 2
        ld = self.shape[0]
        out = np.empty(len(ii), dtype=self.dtype)
        check(lib.hip_stream_sync(None), "hip_stream_sync")
        for t, (i, j) in enumerate(zip(ii, jj)):
            if not (0 <= i < self.shape[0] and 0 <= j < self.shape[1]):
                raise IndexError((i, j))
            check(lib.hip_memcpy_d2h(self.at(int(i) + int(j) * ld), out[t:].ctypes.data,
                                     self.dtype.itemsize), "hip_memcpy_d2h")
        return out

    def at(self, offset_elems):
        """Raw device address of element `offset_elems` (flat, column-major)."""
        return self.ptr + int(offset_elems) * self.dtype.itemsize

    def free(self):
        if self.ptr:
            lib.hip_free_dev(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def sync():
    check(lib.hip_stream_sync(None), "hip_stream_sync")


def set_sync(flag):
    lib.hip_proxy_set_sync(1 if flag else 0)


class Timer:
    """HIP events on the library stream (where every kernel is launched)."""

    def start(self):
        check(lib.hip_timer_start(), "hip_timer_start")

    def stop(self):
        ms = C.c_double()
        check(lib.hip_timer_stop(C.byref(ms)), "hip_timer_stop")
        return ms.value


def _ld(a):
    return max(1, a.shape[0])


def dgemm(transA, transB, m, n, k, alpha, A, B, beta, Cm, lda=None, ldb=None, ldc=None):
    return check(lib.hip_dgemm(transA, transB, m, n, k, alpha, A.ptr, lda or _ld(A), B.ptr,
                               ldb or _ld(B), beta, Cm.ptr, ldc or _ld(Cm)), "hip_dgemm")


def ozaki_dgemm(transA, transB, m, n, k, alpha, A, B, beta, Cm, nslices, lda=None, ldb=None, ldc=None):
    return check(lib.hip_ozaki_dgemm(transA, transB, m, n, k, alpha, A.ptr, lda or _ld(A), B.ptr,
                                     ldb or _ld(B), beta, Cm.ptr, ldc or _ld(Cm), nslices),
                 "hip_ozaki_dgemm")


def ozaki2_dgemm(transA, transB, m, n, k, alpha, A, B, beta, Cm, nmoduli, lda=None, ldb=None, ldc=None):
    return check(lib.hip_ozaki2_dgemm(transA, transB, m, n, k, alpha, A.ptr, lda or _ld(A), B.ptr,
                                      ldb or _ld(B), beta, Cm.ptr, ldc or _ld(Cm), nmoduli),
                 "hip_ozaki2_dgemm")


def ozaki_choose(transA, transB, m, n, k, A, B, lda=None, ldb=None):
    """(slices, moduli) for hip_ozaki_dgemm / hip_ozaki2_dgemm chosen from
    the operands' spread on 4096 sampled entries (synchronous)."""
    s, N = C.c_int(0), C.c_int(0)
    check(lib.hip_ozaki_choose(transA, transB, m, n, k, A.ptr, lda or _ld(A), B.ptr, ldb or _ld(B),
                               C.byref(s), C.byref(N)), "hip_ozaki_choose")
    return s.value, N.value


def ozaki_counts(L, Ls, k):
    """The count rule of hip_ozaki_choose on given spread statistics."""
    s, N = C.c_int(0), C.c_int(0)
    check(lib.hip_ozaki_counts(float(L), float(Ls), k, C.byref(s), C.byref(N)), "hip_ozaki_counts")
    return s.value, N.value


def dgemm_fold(transA, transB, m, n, k, alpha, A, B, beta, Cm, parts):
    arr = (C.c_void_p * len(parts))(*[p.ptr for p in parts])
    lds = (C.c_int * len(parts))(*[_ld(p) for p in parts])
    return check(lib.hip_dgemm_fold(transA, transB, m, n, k, alpha, A.ptr, _ld(A), B.ptr, _ld(B),
                                    beta, Cm.ptr, _ld(Cm), len(parts), arr, lds), "hip_dgemm_fold")


def copy_blocks(dt_size, blocks):
    """blocks: list of (src_addr, lds, dst_addr, ldd, nrow, ncol) row-major copies."""
    arr = (BlkCopy * max(1, len(blocks)))()
    for i, (s, lds, d, ldd, nr, nc) in enumerate(blocks):
        arr[i] = BlkCopy(s, d, lds, ldd, nr, nc)
    return check(lib.hip_copy_matrix_blocks(dt_size, len(blocks), arr), "hip_copy_matrix_blocks")


def copy_matrix_block(dt_size, nrow, ncol, src_addr, lds, dst_addr, ldd):
    return check(lib.hip_copy_matrix_block(dt_size, nrow, ncol, src_addr, lds, dst_addr, ldd),
                 "hip_copy_matrix_block")


def fold_sum(parts, nrow, ncol, dst, ldd=None):
    arr = (C.c_void_p * len(parts))(*[p.ptr for p in parts])
    lds = (C.c_int * len(parts))(*[_ld(p) for p in parts])
    return check(lib.hip_fold_sum(len(parts), arr, lds, nrow, ncol, dst.ptr, ldd or _ld(dst)),
                 "hip_fold_sum")


def op_sum(invec, inoutvec):
    code = {np.dtype(np.float64): 0, np.dtype(np.float32): 1, np.dtype(np.int32): 2}[invec.dtype]
    n = int(np.prod(invec.shape))
    return check(lib.hip_op_sum(code, n, invec.ptr, inoutvec.ptr), "hip_op_sum")


def transpose_cm(nrow, ncol, A, AT):
    return check(lib.hip_transpose_cm(nrow, ncol, A.ptr, _ld(A), AT.ptr, _ld(AT)), "hip_transpose_cm")


def gram(n, k, A, Cm, alpha=1.0, beta=0.0, mirror=True, lda=None, ldc=None):
    return check(lib.hip_dsyrk_gram(n, k, alpha, A.ptr, lda or _ld(A), beta, Cm.ptr, ldc or _ld(Cm),
                                    1 if mirror else 0), "hip_dsyrk_gram")


def ozaki_gram(n, k, A, Cm, nslices, alpha=1.0, beta=0.0, mirror=True, lda=None, ldc=None):
    return check(lib.hip_ozaki_gram(n, k, alpha, A.ptr, lda or _ld(A), beta, Cm.ptr, ldc or _ld(Cm),
                                    1 if mirror else 0, nslices), "hip_ozaki_gram")


def potrf_upper(n, A, lda=None):
    """Returns LAPACK info (0 = success)."""
    return check(lib.hip_dpotrf(b"U", n, A.ptr, lda or _ld(A)), "hip_dpotrf")


def trtri_upper(n, R, X):
    return check(lib.hip_dtrtri_upper(n, R.ptr, _ld(R), X.ptr, _ld(X)), "hip_dtrtri_upper")


HIP_RIGHT, HIP_UPPER, HIP_NONUNIT = 4, 5, 8
CHOLQR_AUTO, CHOLQR2, SCHOLQR3 = 0, 2, 3


def cholesky_qr(m, n, A, R, variant=CHOLQR_AUTO, gram_slices=0, lda=None, ldr=None):
    """hip_cholesky_qr: A (m x n) := Q in place, R (n x n) := the upper
    triangular factor.  Returns (info, passes): info is 0 on success or
    LAPACK's info (> 0) of a Cholesky breakdown the variant did not recover."""
    passes = C.c_int(0)
    info = check(lib.hip_cholesky_qr(m, n, A.ptr, lda or _ld(A), R.ptr, ldr or _ld(R), variant, gram_slices,
                                     C.byref(passes)), "hip_cholesky_qr")
    return info, passes.value


def trsm_right_upper(m, n, alpha, R, B):
    """B := alpha * B * R^{-1}."""
    return check(lib.hip_dtrsm(HIP_RIGHT, HIP_UPPER, NOTRANS, HIP_NONUNIT, m, n, alpha, R.ptr, _ld(R),
                               B.ptr, _ld(B)), "hip_dtrsm")
