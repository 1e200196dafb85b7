// Small dense factorizations for CholeskyQR on the device.
//
// Reference counterparts: cuda_cusolver_dpotrf (src/cuda_proxy.cu:315-339),
// cuda_cublas_dtrsm (:254-275), and the host LAPACKE_dpotrf / LAPACKE_dtrtri
// calls of examples/example_CholeskyQR.c:120-139.
//
//   hip_dpotrf(U)   right-looking Cholesky A = R^T R in one workgroup: the
//                   n x n matrix (512 x 512 in CholeskyQR) stays L2 resident;
//                   row k of R is staged in LDS for each rank-1 update.
//                   Returns LAPACK's info (k+1 when the pivot k is not > 0).
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

constexpr int kPotrfThreads = 1024;
constexpr int kMaxN = 8192;            // potrf / trtri keep one column in LDS

__global__ __launch_bounds__(kPotrfThreads) void potrf_upper_kernel(double *A, int n, int64_t lda, int *info)
{
    __shared__ double rowk[kMaxN];
    __shared__ int bad;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = kPotrfThreads / 64;
    if (t == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (t == 0) {
            const double d = A[(int64_t) k * lda + k];
            if (!(d > 0.0)) {
                bad = k + 1;
            } else {
                A[(int64_t) k * lda + k] = sqrt(d);
            }
        }
        __syncthreads();
        if (bad) break;
        const double piv = A[(int64_t) k * lda + k];
        // row k of R: R[k][j] = A[k][j] / R[k][k], j > k
        for (int j = k + 1 + t; j < n; j += kPotrfThreads) {
            const double v = A[(int64_t) j * lda + k] / piv;
            A[(int64_t) j * lda + k] = v;
            rowk[j] = v;
        }
        __syncthreads();
        // trailing update of the upper triangle: A[i][j] -= R[k][i] R[k][j], k < i <= j
        for (int j = k + 1 + wave; j < n; j += nw) {
            const double rj = rowk[j];
            double *col = A + (int64_t) j * lda;
            for (int i = k + 1 + lane; i <= j; i += 64) col[i] = fma(-rowk[i], rj, col[i]);
        }
        __syncthreads();
    }
    if (t == 0) *info = bad;
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
    potrf_upper_kernel<<<1, kPotrfThreads, 0, ca3::stream()>>>(A, n, ldA, dinfo);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "potrf_upper_kernel");
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
