/*
 * Entry points of libca3dmm_amd with no one-to-one reference counterpart.
 * Each either batches or fuses a reference step (cited) or adds a path the
 * reference lacks.  Same conventions as hip_proxy.h: int status, plain
 * device pointers, work on the library stream.
 */
#ifndef CA3DMM_AMD_EXT_H
#define CA3DMM_AMD_EXT_H

#include <stddef.h>
#include <stdint.h>
#include "hip_proxy.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One row-major block copy: nrow rows of ncol elements, src leading dim lds,
 * dst leading dim ldd (elements). */
typedef struct {
    const void *src;
    void       *dst;
    int64_t     lds, ldd;
    int32_t     nrow, ncol;
} hip_blk_copy_t;

/* All blocks of one mat_redist pack (mat_redist.c:317-348) or unpack
 * (:389-416) -- or any list of block copies -- in ONE launch.  descs is a
 * host array; dt_size 4 or 8.  Bit-exact. */
int hip_copy_matrix_blocks(size_t dt_size, int nblk, const hip_blk_copy_t *descs);

/* The reduce operator of the C reduce-scatter, mpi_op_omp.c:6-23:
 * inoutvec[i] += invec[i]; dtype 0 = double, 1 = float, 2 = int. */
int hip_op_sum(int dtype, int len, const void *invec, void *inoutvec);

/* k-layer partial-C reduction (ca3dmm.c:1311-1324 with mpi_op_omp_sum) in
 * one pass and a fixed order: dst = ((P[0] + P[1]) + P[2]) + ...
 * Column-major nrow x ncol blocks; P[t] leading dim ldp[t].  dst may alias
 * P[0] (no other source).  Any nsrc >= 1: above 16 sources the fold runs
 * as several launches onto the running value, in the same order.
 * Bit-exact against the same fold on the CPU. */
int hip_fold_sum(int nsrc, const double *const *P, const int *ldp,
                 int nrow, int ncol, double *dst, int ldd);

/* hip_dgemm whose epilogue also absorbs k-layer partials:
 * C := fold(P[0..nsrc-1]) + (alpha*op(A)*op(B) + beta*C), left to right.
 * With beta == 0 and the last layer's product this equals hip_fold_sum of
 * all kp partials bit for bit.  0 <= nsrc <= 8 (HIP_PROXY_ERR_ARG above;
 * fold more layers with hip_fold_sum). */
int hip_dgemm_fold(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
                   const int m, const int n, const int k, const double alpha,
                   const double *A, const int ldA, const double *B, const int ldB,
                   const double beta, double *C, const int ldC,
                   int nsrc, const double *const *P, const int *ldp);

/* utils.c:181-196 (transpose_cm_mat) on the device: AT (nrow x ncol,
 * col-major, ldAT) := transpose of A (ncol x nrow, col-major, ldA). */
int hip_transpose_cm(int nrow, int ncol, const double *A, int ldA, double *AT, int ldAT);

/* ---- Ozaki-scheme split-precision DGEMM (no reference counterpart) ----
 * C := alpha*op(A)*op(B) + beta*C computed from nslices INT8 slices of each
 * operand (per-row / per-column power-of-two scaling, balanced base-256
 * digits) on the INT8 matrix cores with exact INT32 accumulation and FP64
 * recombination.  Accuracy grows by ~8 bits per slice; nslices in 1..16.
 * The result is reproducible bit for bit (fixed recombination order). */
int hip_ozaki_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
                    const int m, const int n, const int k, const double alpha,
                    const double *A, const int ldA, const double *B, const int ldB,
                    const double beta, double *C, const int ldC, int nslices);

/* Device workspace the library keeps for hip_ozaki_dgemm at this size. */
size_t hip_ozaki_workspace_bytes(int m, int n, int k, int nslices);

/* C := alpha * op(A) * op(B) + beta * C through the integer-modular (CRT)
 * form of the Ozaki scheme (Ozaki scheme II): lines of op(A) and op(B) are
 * scaled by powers of two to integers whose 2-norms fit the product of the
 * first nmoduli moduli (4 <= nmoduli <= 25), one exact INT8 product per
 * modulus (v_mfma_i32_16x16x64_i8), then Garner's mixed-radix digits are
 * evaluated in FP64 in a fixed order, so results are reproducible bit for
 * bit (oracle: ozaki2_dgemm).  k <= 2^23.  No reference counterpart: an
 * alternative to hip_dgemm (the local DGEMM of src/cannon.c:405-427). */
int hip_ozaki2_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
                     const int m, const int n, const int k, const double alpha,
                     const double *A, const int ldA, const double *B, const int ldB,
                     const double beta, double *C, const int ldC, int nmoduli);

/* Device workspace the library keeps for hip_ozaki2_dgemm at this size. */
size_t hip_ozaki2_workspace_bytes(int m, int n, int k, int nmoduli);

/* Slice count for hip_ozaki_dgemm and modulus count for hip_ozaki2_dgemm,
 * chosen per problem from the operands' spread (oracle: ozaki.spread and
 * ozaki.counts_from_spread).  4096 deterministic entries (i, j) of C are
 * sampled; with a = |row i of op(A)| and b = |column j of op(B)|,
 *   L   = max log2(||a||_2 ||b||_2 / a.b)                        (CRT form),
 *   L_s = max log2((max(a) ||b||_1 + max(b) ||a||_1) / a.b)      (slice form),
 * and the counts are the smallest with T(nmoduli) >= L + 53.5 +
 * 0.5 log2(k / 1024) and 8 nslices >= L_s + 54 + 2.33 log2(k / 1024) (the
 * log terms for k > 1024),
 * clamped to 25 and 16.  The offsets aim at errors no worse than
 * hip_dgemm's and were calibrated on uniform[-0.5, 0.5] exp(phi N(0,1))
 * inputs (phi 0.5, 2, 4; 8192^3 and 65536 x 65536 x 1024): an estimate,
 * not a bound.  Synchronous (reads the sample statistics back).  k = 0 or
 * an empty C gives the counts of L = L_s = 0.  No reference counterpart. */
int hip_ozaki_choose(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
                     const int m, const int n, const int k,
                     const double *A, const int ldA, const double *B, const int ldB,
                     int *nslices, int *nmoduli);

/* The count rule of hip_ozaki_choose on given spread statistics (host
 * only, no device work): nslices from L_s and k, nmoduli from L; NaN counts
 * as +inf. */
int hip_ozaki_counts(double L, double Ls, int k, int *nslices, int *nmoduli);

/* hip_dsyrk_gram (below) through the Ozaki scheme: C := alpha * A^T A +
 * beta * C for A k x n (column-major), upper triangle only, mirror != 0
 * copies it to the lower triangle.  A is sliced once and serves as both
 * operands; k is split across workgroups and the INT32 partials are added
 * in FP64 in the fixed order of hip_ozaki_dgemm, so the result equals
 * hip_ozaki_dgemm(T, N, n, n, k, ...) on (A, A) bit for bit on j >= i.
 * No reference counterpart (the Gram of examples/example_CholeskyQR.c:109-117). */
int hip_ozaki_gram(const int n, const int k, const double alpha, const double *A, const int ldA,
                   const double beta, double *C, const int ldC, int mirror, int nslices);

/* C := alpha * A^T A + beta * C for A k x n (column-major): the B^T B of
 * ca3dmm_engine_init_BTB (src/ca3dmm.c:1172-1265) and of CholeskyQR
 * (examples/example_CholeskyQR.c:109-117).  Only the upper triangle is
 * computed (n(n+1)k flops); mirror != 0 also writes the lower triangle as an
 * exact copy.  k is split across workgroups and the partials are added in a
 * fixed order, so results are reproducible. */
int hip_dsyrk_gram(const int n, const int k, const double alpha, const double *A, const int ldA,
                   const double beta, double *C, const int ldC, int mirror);

/* hip_dsyrk_gram without any allocation: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * kernels, k-split and result, bit for bit, as hip_dsyrk_gram; the k-split
 * partial tiles live in a fixed 64 MiB device array of the library (512
 * tiles of 128 x 128 doubles, one per GPU the library runs on, held until
 * the process ends) instead of a library allocation.  Shapes needing more
 * than 512 partial tiles (n > 3968 on a 256-CU MI355X) are refused with
 * HIP_PROXY_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that
 * overlap in time on one GPU (two streams, two host threads) race.  Each
 * call writes every partial tile it reads, so nothing needs clearing. */
int hip_dsyrk_gram_noalloc(const int n, const int k, const double alpha, const double *A, const int ldA,
                           const double beta, double *C, const int ldC, int mirror);

/* X := R^{-1} for upper-triangular R (n x n, n <= 8192); the strictly
 * lower part of X is set to zero (examples/example_CholeskyQR.c:137). */
int hip_dtrtri_upper(const int n, const double *R, const int ldR, double *X, const int ldX);

/* CholeskyQR of a tall-skinny A (m x n column-major, m >= n, n <= 8192),
 * every step on the device: the driver of examples/example_CholeskyQR.c:109-168
 * (Gram, Cholesky 'U' with the strictly lower part zeroed, Q = A R^{-1}),
 * repeated.  On return A holds Q and R (n x n, leading dim ldR) the upper
 * triangular factor with A = Q R (strictly lower part zero).
 *   variant HIP_CHOLQR2:  two plain passes (CholeskyQR2);
 *   variant HIP_SCHOLQR3: one shifted pass, G + s I with
 *                         s = 11 (m n + n (n + 1)) 2^-53 ||A||_F^2, then two
 *                         plain passes (shifted CholeskyQR3: the shifted
 *                         pass divides kappa by about
 *                         sqrt(11 (m n + n (n + 1)) 2^-53), and the plain
 *                         passes need what is left below about 2^26; at
 *                         m = 2e6, n = 512 that is kappa(A) up to about 1e11);
 *   variant HIP_CHOLQR_AUTO: CholeskyQR2, but a plain pass whose Cholesky
 *                         breaks down, or whose max / min of diag(R_p)
 *                         exceeds 2^23 (kappa(A) too close to u^-1/2 for
 *                         CholeskyQR2), is redone shifted, as often as that
 *                         happens, and followed by two plain passes (at most
 *                         six passes in all).
 * gram_slices > 0 computes each Gram through hip_ozaki_gram with that many
 * slices, 0 through hip_dsyrk_gram.  Returns 0 on success, LAPACK's info
 * (> 0, the order of the leading minor that is not positive definite) when
 * a Cholesky breaks down and the variant does not recover, < 0 on errors.
 * *passes_run (may be NULL) receives the number of completed passes.
 * When it can, the call holds a device buffer of m x n doubles (cached
 * across calls) so that each pass writes Q out of place; otherwise, or with
 * CA3DMM_CHOLQR_INPLACE=1, the passes solve in place (same result).
 * Synchronous (one read-back of the Cholesky's info per pass).  The
 * reference runs one pass and only prints a warning on breakdown
 * (examples/example_CholeskyQR.c:125-129). */
#define HIP_CHOLQR_AUTO 0
#define HIP_CHOLQR2     2
#define HIP_SCHOLQR3    3
int hip_cholesky_qr(const int m, const int n, double *A, const int ldA, double *R, const int ldR,
                    const int variant, const int gram_slices, int *passes_run);

#ifdef __cplusplus
}
#endif

#endif
