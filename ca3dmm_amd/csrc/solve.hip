// Small dense factorizations for CholeskyQR on the device.
//
// Reference counterparts: cuda_cusolver_dpotrf (src/cuda_proxy.cu:315-339),
// cuda_cublas_dtrsm (:254-275), and the host LAPACKE_dpotrf / LAPACKE_dtrtri
// calls of examples/example_CholeskyQR.c:120-139.
//
//   hip_dpotrf(U)   blocked right-looking Cholesky A = R^T R over 64-wide
//                   block columns: the diagonal block factored in LDS by one
//                   workgroup, the block row R_jj^{-T} A_j,rest by forward
//                   substitution (one wave per column), the trailing upper
//                   triangle updated by the FP64 MFMA Gram (hip_dsyrk_gram,
//                   alpha = -1, beta = 1).  The strictly lower part of A is
//                   not touched.  Returns LAPACK's info (k+1 when the pivot k
//                   is not > 0).
//   hip_dtrtri_upper  R^{-1} for upper-triangular R: one workgroup per
//                   column j solves R x = e_j by a column sweep in LDS.
//   hip_dtrsm(R,U,N,N)  B := alpha * B * R^{-1}: R^{-1} (trtri) followed by
//                   one FP64 MFMA DGEMM per column block of the result, each
//                   reading only the non-zero rows of R^{-1}; blocks are
//                   produced from the last to the first through a workspace,
//                   so B can be overwritten in place.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "ca3dmm_ext.h"
#include "hip_common.h"

namespace {

constexpr int kMaxN = 8192;            // potrf / trtri keep one column in LDS

constexpr int kPb = 64;                // block width of the blocked Cholesky

// The jb x jb diagonal block at A(j0, j0) (its upper triangle) factored in
// LDS: R_jj with R_jj^T R_jj = A_jj.  info: 0 written by the first block;
// a later block does nothing once info is set, and writes j0 + k + 1 when
// its pivot k is not > 0.
__global__ __launch_bounds__(256) void potrf_diag_kernel(double *A, int64_t lda, int j0, int jb, int *info)
{
    __shared__ double L[kPb][kPb + 1];     // L[j][i] = A(j0 + i, j0 + j), i <= j
    __shared__ int bad;
    const int t = threadIdx.x;
    if (j0 > 0 && *info) return;
    if (t == 0) bad = 0;
    double *D = A + (int64_t) j0 * lda + j0;
    for (int e = t; e < jb * jb; e += 256) {
        const int i = e % jb, j = e / jb;
        if (i <= j) L[j][i] = D[(int64_t) j * lda + i];
    }
    __syncthreads();
    for (int k = 0; k < jb; ++k) {
        if (t == 0) {
            const double d = L[k][k];
            if (!(d > 0.0)) bad = k + 1;
            else L[k][k] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        const double piv = L[k][k];
        for (int j = k + 1 + t; j < jb; j += 256) L[j][k] = L[j][k] / piv;      // row k of R
        __syncthreads();
        // A(i, j) -= R(k, i) R(k, j), k < i <= j
        const int r = jb - k - 1;
        for (int e = t; e < r * r; e += 256) {
            const int i = k + 1 + e % r, j = k + 1 + e / r;
            if (i <= j) L[j][i] = fma(-L[i][k], L[j][k], L[j][i]);
        }
        __syncthreads();
    }
    for (int e = t; e < jb * jb; e += 256) {
        const int i = e % jb, j = e / jb;
        if (i <= j) D[(int64_t) j * lda + i] = L[j][i];
    }
    if (t == 0 && (j0 == 0 || bad)) *info = bad ? j0 + bad : 0;
}

// Block row: X = R_jj^{-T} A(j0:j0+jb, c) for the columns c >= j0 + jb, in
// place, by forward substitution; one wave per column, lane i holds row i.
__global__ __launch_bounds__(256) void potrf_panel_kernel(double *A, int64_t lda, int j0, int jb, int n,
                                                          const int *info)
{
    __shared__ double R[kPb][kPb + 1];     // R[j][i] = R_jj(i, j)
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (*info) return;
    const double *D = A + (int64_t) j0 * lda + j0;
    for (int e = t; e < jb * jb; e += 256) {
        const int i = e % jb, j = e / jb;
        R[j][i] = (i <= j) ? D[(int64_t) j * lda + i] : 0.0;
    }
    __syncthreads();
    const int c = j0 + jb + blockIdx.x * 4 + wave;
    if (c >= n) return;
    double *col = A + (int64_t) c * lda + j0;
    double x = lane < jb ? col[lane] : 0.0;
    for (int i = 0; i < jb; ++i) {
        // x_i is final once the rows above it are eliminated
        const double xi = __shfl(x, i) / R[i][i];
        if (lane == i) x = xi;
        if (lane > i) x = fma(-R[lane][i], xi, x);
    }
    if (lane < jb) col[lane] = x;
}

// X = R^{-1}, both n x n upper triangular (ldx, ldr); one block per column.
__global__ __launch_bounds__(256) void trtri_upper_kernel(const double *R, int64_t ldr, double *X, int64_t ldx, int n)
{
    __shared__ double x[kMaxN];
    const int j = blockIdx.x, t = threadIdx.x;
    for (int p = t; p <= j; p += 256) x[p] = (p == j) ? 1.0 : 0.0;
    __syncthreads();
    for (int i = j; i >= 0; --i) {
        const double xi = x[i] / R[(int64_t) i * ldr + i];
        __syncthreads();
        if (t == 0) x[i] = xi;
        for (int p = t; p < i; p += 256) x[p] = fma(-R[(int64_t) i * ldr + p], xi, x[p]);
        __syncthreads();
    }
    for (int p = t; p < n; p += 256) X[(int64_t) j * ldx + p] = (p <= j) ? x[p] : 0.0;
}

struct Ws {
    void *p = nullptr;
    size_t bytes = 0;
};
Ws g_ws[16];
std::mutex g_mu;

void *scratch(size_t bytes, int *rc)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    Ws &w = g_ws[dev & 15];
    std::lock_guard<std::mutex> lk(g_mu);
    *rc = 0;
    if (w.bytes < bytes) {
        if (w.p) {
            (void) hipStreamSynchronize(ca3::stream());
            (void) hipFree(w.p);
            w.p = nullptr;
            w.bytes = 0;
        }
        hipError_t e = hipMalloc(&w.p, bytes);
        if (e != hipSuccess) { *rc = ca3::fail(e, "solve workspace"); return nullptr; }
        w.bytes = bytes;
    }
    return w.p;
}

}  // namespace

extern "C" {

int hip_dpotrf(const char uplo, const int n, double *A, const int ldA)
{
    if (uplo != 'U' && uplo != 'u')
        return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dpotrf: only uplo = 'U' is implemented");
    if (n < 0 || ldA < std::max(1, n) || (n > 0 && !A)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dpotrf: bad arguments");
    if (n > kMaxN) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dpotrf: n > %d", kMaxN);
    if (n == 0) return 0;
    int rc = 0;
    int *dinfo = (int *) scratch(256, &rc);
    if (rc) return rc;
    const int save = ca3::sync_mode();
    hip_proxy_set_sync(0);
    for (int j0 = 0; j0 < n && !rc; j0 += kPb) {
        const int jb = std::min(kPb, n - j0), nr = n - j0 - jb;
        potrf_diag_kernel<<<1, 256, 0, ca3::stream()>>>(A, ldA, j0, jb, dinfo);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) { rc = ca3::fail(e, "potrf_diag_kernel"); break; }
        if (nr == 0) break;
        potrf_panel_kernel<<<(nr + 3) / 4, 256, 0, ca3::stream()>>>(A, ldA, j0, jb, n, dinfo);
        e = hipGetLastError();
        if (e != hipSuccess) { rc = ca3::fail(e, "potrf_panel_kernel"); break; }
        // trailing upper triangle: A(r, r) -= X^T X, X = A(j0:j0+jb, r)
        double *X = A + (int64_t) (j0 + jb) * ldA + j0;
        rc = hip_dsyrk_gram(nr, jb, -1.0, X, ldA, 1.0, A + (int64_t) (j0 + jb) * ldA + j0 + jb, ldA, 0);
    }
    hip_proxy_set_sync(save);
    if (rc) return rc;
    int info = 0;
    CA3_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, ca3::stream()));
    CA3_CHECK(hipStreamSynchronize(ca3::stream()));
    return info;
}

int hip_dtrtri_upper(const int n, const double *R, const int ldR, double *X, const int ldX)
{
    if (n < 0 || (n > 0 && (!R || !X)) || ldR < std::max(1, n) || ldX < std::max(1, n))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dtrtri: bad arguments");
    if (n > kMaxN) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dtrtri: n > %d", kMaxN);
    if (n == 0) return HIP_PROXY_OK;
    trtri_upper_kernel<<<n, 256, 0, ca3::stream()>>>(R, ldR, X, ldX, n);
    return ca3::finish();
}

int hip_dtrsm(hip_proxy_enum_t side, hip_proxy_enum_t uplo, hip_proxy_enum_t trans, hip_proxy_enum_t diag,
              const int m, const int n, const double alpha, const double *A, const int ldA, double *B,
              const int ldB)
{
    if (side != HipRight || uplo != HipUpper || trans != HipNoTrans || diag != HipNonUnit)
        return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dtrsm: only side=Right, uplo=Upper, trans=NoTrans, "
                                                   "diag=NonUnit is implemented");
    if (m < 0 || n < 0 || ldA < std::max(1, n) || ldB < std::max(1, m))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dtrsm: bad arguments");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (n > kMaxN) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dtrsm: n > %d", kMaxN);
    const int nb = 128;
    const int wcols = std::min(n, nb);
    int rc = 0;
    const size_t rinv_bytes = sizeof(double) * (size_t) n * n;
    char *ws = (char *) scratch(rinv_bytes + sizeof(double) * (size_t) m * wcols + 256, &rc);
    if (rc) return rc;
    double *Rinv = (double *) ws;
    double *T = (double *) (ws + ((rinv_bytes + 255) / 256) * 256);
    const int save = ca3::sync_mode();
    hip_proxy_set_sync(0);
    rc = hip_dtrtri_upper(n, A, ldA, Rinv, n);
    // column block jb of X = B R^{-1} needs rows 0 .. end(jb) of R^{-1}
    for (int j0 = ((n - 1) / nb) * nb; !rc && j0 >= 0; j0 -= nb) {
        const int w = std::min(nb, n - j0), kk = j0 + w;
        rc = hip_dgemm(HipNoTrans, HipNoTrans, m, w, kk, alpha, B, ldB, Rinv + (size_t) j0 * n, n, 0.0, T, m);
        if (!rc) rc = hip_copy_matrix_block(8, w, m, T, m, B + (size_t) j0 * ldB, ldB);
    }
    hip_proxy_set_sync(save);
    if (rc) return rc;
    return ca3::finish();
}

}  // extern "C"

int ca3::trsm_right_upper_out(int m, int n, const double *R, int ldR, const double *B, int ldB, double *X, int ldX)
{
    if (m < 0 || n < 0 || ldR < std::max(1, n) || ldB < std::max(1, m) || ldX < std::max(1, m))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "trsm_right_upper_out: bad arguments");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (n > kMaxN) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "trsm_right_upper_out: n > %d", kMaxN);
    int rc = 0;
    double *Rinv = (double *) scratch(sizeof(double) * (size_t) n * n, &rc);
    if (rc) return rc;
    const int save = ca3::sync_mode();
    hip_proxy_set_sync(0);
    rc = hip_dtrtri_upper(n, R, ldR, Rinv, n);
    // column block jb of X = B R^{-1} needs rows 0 .. end(jb) of R^{-1}
    const int nb = 128;
    for (int j0 = 0; !rc && j0 < n; j0 += nb) {
        const int w = std::min(nb, n - j0), kk = j0 + w;
        rc = hip_dgemm(HipNoTrans, HipNoTrans, m, w, kk, 1.0, B, ldB, Rinv + (size_t) j0 * n, n, 0.0,
                       X + (size_t) j0 * ldX, ldX);
    }
    hip_proxy_set_sync(save);
    if (rc) return rc;
    return ca3::finish();
}
