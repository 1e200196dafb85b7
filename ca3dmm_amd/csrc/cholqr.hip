// CholeskyQR on the device: CholeskyQR2 and shifted CholeskyQR3.
//
// Reference driver: examples/example_CholeskyQR.c:109-168 -- S0 = B^T B
// (BTB engine), R = chol(S0, 'U') (LAPACKE_dpotrf, :124), strictly lower part
// of R zeroed (:130-135), R^{-1} (LAPACKE_dtrtri, :137), B_orth = B R^{-1}
// (:145-154), and the check of S1 = B_orth^T B_orth against I (:160-193).
// The reference runs one pass and only prints a warning when the Cholesky
// breaks down (:125-129).
//
// Here every step runs on the GPU and a pass is repeated:
//   G   = Q^T Q                      hip_dsyrk_gram (or hip_ozaki_gram)
//   G  += s I  (shifted passes only) s = 11 (m n + n (n + 1)) u ||Q||_F^2,
//                                    ||Q||_F^2 = trace(G), u = 2^-53
//   R_p = chol(G) 'U'                hip_dpotrf, strictly lower part zeroed
//   Q   = Q R_p^{-1}                 hip_dtrsm
//   R   = R_p R                      hip_dgemm
// The shift is the one of Fukaya, Kannan, Nakatsukasa, Yamamoto, Yanagisawa,
// "Shifted Cholesky QR for computing the QR factorization of ill-conditioned
// matrices", SIAM J. Sci. Comput. 42(1), 2020 (with ||A||_2 bounded by
// ||A||_F): one shifted pass makes Q well enough conditioned (kappa(Q) of
// order u^-1/2) for two plain passes to reach orthogonality of order u.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "ca3dmm_ext.h"
#include "hip_common.h"

namespace {

constexpr int kMaxN = 8192;   // the limit of hip_dpotrf / hip_dtrsm

// G(i, j) := 0 for i > j (n x n, leading dim ld)
__global__ __launch_bounds__(256) void zero_lower_kernel(double *G, int n, int64_t ld)
{
    const int j = blockIdx.x;
    for (int i = j + 1 + (int) threadIdx.x; i < n; i += 256) G[(int64_t) j * ld + i] = 0.0;
}

// G(i, i) += 11 (m n + n (n + 1)) 2^-53 trace(G), trace summed in a fixed
// order (per-thread strided sums, then a tree in LDS), one workgroup.
__global__ __launch_bounds__(256) void shift_diag_kernel(double *G, int n, int64_t ld, double coef)
{
    __shared__ double part[256];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < n; i += 256) s = ca3::add_rn(s, G[(int64_t) i * ld + i]);
    part[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) part[t] = ca3::add_rn(part[t], part[t + w]);
        __syncthreads();
    }
    const double shift = ca3::mul_rn(coef, part[0]);
    for (int i = t; i < n; i += 256) G[(int64_t) i * ld + i] = ca3::add_rn(G[(int64_t) i * ld + i], shift);
}

// out[0] = max_i R(i, i), out[1] = min_i R(i, i), one workgroup
__global__ __launch_bounds__(256) void diag_range_kernel(const double *R, int n, int64_t ld, double *out)
{
    __shared__ double hi[256], lo[256];
    const int t = threadIdx.x;
    double a = 0.0, b = INFINITY;
    for (int i = t; i < n; i += 256) {
        const double d = R[(int64_t) i * ld + i];
        a = fmax(a, d);
        b = fmin(b, d);
    }
    hi[t] = a;
    lo[t] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            hi[t] = fmax(hi[t], hi[t + w]);
            lo[t] = fmin(lo[t], lo[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = hi[0];
        out[1] = lo[0];
    }
}

struct Ws {
    void *p = nullptr;
    size_t bytes = 0;
};
Ws g_ws[16];
std::mutex g_mu;

double *workspace(size_t bytes, int *rc)
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
        if (e != hipSuccess) { *rc = ca3::fail(e, "cholesky_qr workspace"); return nullptr; }
        w.bytes = bytes;
    }
    return (double *) w.p;
}

// Largest max/min ratio of diag(R_p) that the automatic variant accepts
// from a plain pass: kappa(A) >= that ratio, and plain CholeskyQR2 loses
// orthogonality once kappa(A) nears u^-1/2 = 2^26.5.  Above 2^23 the pass
// is redone with the shift.
constexpr double kAutoDiagRatio = 0x1p23;
constexpr int kAutoMaxPasses = 6;

// One pass on Q (m x n): G = Q^T Q (+ shift), R_p = chol(G), Q' = Q R_p^{-1}
// into Qo (out of place; Qo == Q: in place through hip_dtrsm).  Returns the
// Cholesky's info (> 0: breakdown, Q untouched) or a status < 0.  With
// check_ratio, max and min of diag(R_p) are read back and a ratio above
// kAutoDiagRatio stops the pass before Q' is formed (returns kIllConditioned).
constexpr int kIllConditioned = 1 << 30;

int one_pass(int m, int n, double *Q, int ldQ, double *Qo, int ldQo, double *G, double *dscal, bool shifted,
             int gram_slices, bool check_ratio)
{
    int rc = gram_slices > 0 ? hip_ozaki_gram(n, m, 1.0, Q, ldQ, 0.0, G, n, 0, gram_slices)
                             : hip_dsyrk_gram(n, m, 1.0, Q, ldQ, 0.0, G, n, 0);
    if (rc) return rc;
    if (shifted) {
        const double coef = 11.0 * ((double) m * n + (double) n * (n + 1)) * 0x1p-53;
        shift_diag_kernel<<<1, 256, 0, ca3::stream()>>>(G, n, n, coef);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ca3::fail(e, "shift_diag_kernel");
    }
    const int info = hip_dpotrf('U', n, G, n);          // synchronous: reads info back
    if (info != 0) return info;
    if (check_ratio) {
        diag_range_kernel<<<1, 256, 0, ca3::stream()>>>(G, n, n, dscal);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ca3::fail(e, "diag_range_kernel");
        double h[2] = {0.0, 0.0};
        CA3_CHECK(hipMemcpyAsync(h, dscal, sizeof(h), hipMemcpyDeviceToHost, ca3::stream()));
        CA3_CHECK(hipStreamSynchronize(ca3::stream()));
        if (!(h[0] <= kAutoDiagRatio * h[1])) return kIllConditioned;
    }
    zero_lower_kernel<<<n, 256, 0, ca3::stream()>>>(G, n, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "zero_lower_kernel");
    if (Qo == Q) return hip_dtrsm(HipRight, HipUpper, HipNoTrans, HipNonUnit, m, n, 1.0, G, n, Q, ldQ);
    return ca3::trsm_right_upper_out(m, n, G, n, Q, ldQ, Qo, ldQo);
}

// The out-of-place Q buffer: cached like the workspace; nullptr (and the
// in-place solve) when it cannot be allocated.
Ws g_qbuf[16];

double *qbuffer(size_t bytes)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    Ws &w = g_qbuf[dev & 15];
    std::lock_guard<std::mutex> lk(g_mu);
    if (w.bytes < bytes) {
        if (w.p) {
            (void) hipStreamSynchronize(ca3::stream());
            (void) hipFree(w.p);
            w.p = nullptr;
            w.bytes = 0;
        }
        if (hipMalloc(&w.p, bytes) != hipSuccess) {
            (void) hipGetLastError();          // the failed allocation is not an error of the call
            w.p = nullptr;
            return nullptr;
        }
        w.bytes = bytes;
    }
    return (double *) w.p;
}

}  // namespace

extern "C" {

int hip_cholesky_qr(const int m, const int n, double *A, const int ldA, double *R, const int ldR,
                    const int variant, const int gram_slices, int *passes_run)
{
    if (passes_run) *passes_run = 0;
    if (m < 0 || n < 0 || ldA < std::max(1, m) || ldR < std::max(1, n) || (m > 0 && n > 0 && (!A || !R)))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "cholesky_qr: bad arguments");
    if (variant != HIP_CHOLQR_AUTO && variant != HIP_CHOLQR2 && variant != HIP_SCHOLQR3)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "cholesky_qr: variant must be 0 (auto), 2 or 3");
    if (gram_slices < 0 || gram_slices > 16) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "cholesky_qr: gram_slices in 0..16");
    if (n > kMaxN) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "cholesky_qr: n > %d", kMaxN);
    if (m < n) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "cholesky_qr: needs m >= n (tall and skinny)");
    if (n == 0) return HIP_PROXY_OK;
    int rc = 0;
    const size_t nn = (size_t) n * n;
    double *G = workspace((2 * nn + 2) * sizeof(double), &rc);
    if (rc) return rc;
    double *T = G + nn, *dscal = T + nn;
    // an m x n buffer (when the device has room for it) lets each pass write
    // its Q out of place: the passes alternate between A and the buffer, and
    // an odd count ends with one copy back into A
    const char *ip = getenv("CA3DMM_CHOLQR_INPLACE");
    double *W = (ip && !strcmp(ip, "1")) ? nullptr : qbuffer((size_t) m * n * sizeof(double));
    double *Q = A, *Qo = W ? W : A;
    int ldQ = ldA, ldQo = W ? m : ldA;
    const int save = ca3::sync_mode();
    hip_proxy_set_sync(0);
    // plain passes still to do after the last shifted one; a plain pass
    // that breaks down is redone shifted when the variant allows it.  One
    // shifted pass divides kappa by about sqrt(11 (m n + n (n + 1)) u)
    // ||Q||_F / ||Q||_2 only, so near kappa = 1e12 at m = 2e6 the automatic
    // variant may need two or three of them before its two plain passes.
    const int max_passes = (variant == HIP_CHOLQR_AUTO) ? kAutoMaxPasses : 3;
    int plain_left = 2, passes = 0;
    bool shift_next = (variant == HIP_SCHOLQR3);
    while (plain_left > 0 && passes < max_passes) {
        const bool autoplain = (variant == HIP_CHOLQR_AUTO && !shift_next);
        rc = one_pass(m, n, Q, ldQ, Qo, ldQo, G, dscal, shift_next, gram_slices, autoplain);
        if (rc > 0 && autoplain) {
            shift_next = true;                          // Q untouched: redo this pass shifted
            continue;
        }
        if (rc) break;
        if (passes == 0) {
            rc = hip_copy_matrix_block(8, n, n, G, n, R, ldR);
        } else {
            rc = hip_dgemm(HipNoTrans, HipNoTrans, n, n, n, 1.0, G, n, R, ldR, 0.0, T, n);
            if (!rc) rc = hip_copy_matrix_block(8, n, n, T, n, R, ldR);
        }
        if (rc) break;
        std::swap(Q, Qo);
        std::swap(ldQ, ldQo);
        ++passes;
        plain_left = shift_next ? 2 : plain_left - 1;
        shift_next = false;
    }
    // Q holds the last pass's result: in the buffer after an odd count
    if (!rc && Q != A) rc = hip_copy_matrix_block(8, n, m, Q, ldQ, A, ldA);
    hip_proxy_set_sync(save);
    if (passes_run) *passes_run = passes;
    if (rc) return rc;
    return ca3::finish();
}

}  // extern "C"
