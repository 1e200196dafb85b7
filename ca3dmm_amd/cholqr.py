"""CholeskyQR2 on the device, through the C-ABI.

Follows examples/example_CholeskyQR.c (S = B^T B, R = chol(S) upper,
B_orth = B R^{-1}, check Q^T Q ~ I), run twice (CholeskyQR2) with every
step on the GPU: Gram (hip_dsyrk_gram, upper triangle + mirror), Cholesky
(hip_dpotrf 'U'), Q = A R^{-1} (hip_dtrsm), R = R2 R1 (hip_dgemm).
Norms come from Gram traces: ||X||_F^2 = trace(X^T X)."""
import numpy as np

from . import device as hip


class CholeskyBreakdown(RuntimeError):
    def __init__(self, info, step):
        super().__init__(f"Cholesky of the Gram matrix failed at pass {step}: leading minor {info} "
                         "is not positive definite")
        self.info = info
        self.step = step


def _zero_lower(R):
    """R (device, n x n): strictly lower part := 0 (examples/example_CholeskyQR.c:130-135)."""
    h = R.numpy()
    h = np.triu(h)
    return hip.DeviceArray.from_numpy(h)


def cholesky_qr2(dA, m, n, passes=2, gram_slices=None):
    """dA: device m x n (column-major).  Returns (dQ, dR) with Q m x n
    orthonormal and R n x n upper triangular, A = Q R.  gram_slices: compute
    each Gram through the Ozaki scheme with that many INT8 slices
    (hip_ozaki_gram) instead of the FP64 MFMA kernel."""
    dQ = hip.DeviceArray((m, n))
    hip.copy_matrix_block(8, n, m, dA.ptr, m, dQ.ptr, m)
    dR = None
    for step in range(1, passes + 1):
        dG = hip.DeviceArray.zeros((n, n))
        if gram_slices:
            hip.ozaki_gram(n, m, dQ, dG, gram_slices)
        else:
            hip.gram(n, m, dQ, dG)
        info = hip.potrf_upper(n, dG)
        if info != 0:
            raise CholeskyBreakdown(info, step)
        dRs = _zero_lower(dG)
        hip.trsm_right_upper(m, n, 1.0, dRs, dQ)
        if dR is None:
            dR = dRs
        else:
            dRn = hip.DeviceArray.zeros((n, n))
            hip.dgemm(0, 0, n, n, n, 1.0, dRs, dR, 0.0, dRn)
            dR = dRn
    return dQ, dR


def qr_errors(dA, dQ, dR, m, n):
    """(||Q^T Q - I||_F, ||A - QR||_F / ||A||_F), computed on the device."""
    dG = hip.DeviceArray.zeros((n, n))
    hip.gram(n, m, dQ, dG)
    G = dG.numpy()
    orth = float(np.linalg.norm(G - np.eye(n)))
    dD = hip.DeviceArray((m, n))
    hip.copy_matrix_block(8, n, m, dA.ptr, m, dD.ptr, m)
    hip.dgemm(0, 0, m, n, n, -1.0, dQ, dR, 1.0, dD)
    hip.gram(n, m, dD, dG)
    res2 = float(np.trace(dG.numpy()))
    hip.gram(n, m, dA, dG)
    a2 = float(np.trace(dG.numpy()))
    return orth, float(np.sqrt(max(res2, 0.0) / a2))


def test_matrix(m, n, cond, seed=2022, dU=None):
    """A = U diag(sigma) V^T with orthonormal U, V and sigma log-spaced from 1
    to 1/cond.  U comes from CholeskyQR2 of a seeded Gaussian matrix on the
    device (kappa ~ 1, so it is orthonormal to working precision), V from
    numpy's QR of an n x n Gaussian."""
    rnd = np.random.default_rng(seed)
    if dU is None:
        dG = hip.DeviceArray.from_numpy(rnd.standard_normal((m, n)))
        dU, _ = cholesky_qr2(dG, m, n)
    V, _ = np.linalg.qr(rnd.standard_normal((n, n)))
    sigma = np.logspace(0, -np.log10(cond), n)
    dSVt = hip.DeviceArray.from_numpy(np.asfortranarray(sigma[:, None] * V.T))
    dA = hip.DeviceArray.zeros((m, n))
    hip.dgemm(0, 0, m, n, n, 1.0, dU, dSVt, 0.0, dA)
    return dA, dU
