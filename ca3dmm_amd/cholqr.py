"""CholeskyQR2 and shifted CholeskyQR3 on the device, through the C-ABI.

Follows examples/example_CholeskyQR.c (S = B^T B, R = chol(S) upper,
B_orth = B R^{-1}, check Q^T Q ~ I), with every pass on the GPU inside
hip_cholesky_qr: Gram (hip_dsyrk_gram, or hip_ozaki_gram), optional
diagonal shift, Cholesky (hip_dpotrf 'U'), strictly lower part zeroed on the
device, Q = A R^{-1} (hip_dtrsm), R = R_p R (hip_dgemm).
Norms come from Gram traces: ||X||_F^2 = trace(X^T X)."""
import numpy as np

from . import device as hip


class CholeskyBreakdown(RuntimeError):
    def __init__(self, info, step):
        super().__init__(f"Cholesky of the Gram matrix failed at pass {step}: leading minor {info} "
                         "is not positive definite")
        self.info = info
        self.step = step


def cholesky_qr(dA, m, n, variant=hip.CHOLQR2, gram_slices=None):
    """dA: device m x n (column-major, left unchanged).  Returns (dQ, dR,
    passes) with Q m x n orthonormal, R n x n upper triangular, A = Q R.
    variant: hip.CHOLQR2, hip.SCHOLQR3 or hip.CHOLQR_AUTO.  gram_slices:
    compute each Gram through the Ozaki scheme with that many INT8 slices
    (hip_ozaki_gram) instead of the FP64 MFMA kernel."""
    dQ = hip.DeviceArray((m, n))
    hip.copy_matrix_block(8, n, m, dA.ptr, m, dQ.ptr, m)
    dR = hip.DeviceArray((n, n))
    info, passes = hip.cholesky_qr(m, n, dQ, dR, variant=variant, gram_slices=gram_slices or 0)
    if info > 0:
        raise CholeskyBreakdown(info, passes + 1)
    return dQ, dR, passes


def cholesky_qr2(dA, m, n, gram_slices=None):
    """CholeskyQR2 (two plain passes): (dQ, dR)."""
    dQ, dR, _ = cholesky_qr(dA, m, n, hip.CHOLQR2, gram_slices)
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
