/*
 * hip_proxy: the C-ABI device layer of libca3dmm_amd on MI355X (gfx950).
 *
 * Replaces the reference's src/cuda_proxy.h (the set of entries the
 * reference library and examples call; see the per-function citations).
 * Differences from the reference, all deliberate:
 *   - every entry returns an int status (0 = success, negative = error;
 *     hip_dpotrf returns LAPACK-style info >= 0) instead of exit(-1)
 *     (reference: src/cuda_utils.h:23-35);
 *   - all kernels are this library's own HIP kernels; no vendor BLAS,
 *     solver or random-number library is called;
 *   - work runs on one library-owned stream per device (default priority).
 *     By default each entry synchronises that stream before returning, as
 *     the reference's BLAS proxies do (cuda_proxy.cu:184,229); after
 *     hip_proxy_set_sync(0) entries return as soon as the work is queued and
 *     hip_stream_sync(NULL) waits for it.
 * Pointers are plain device (or pinned host) addresses; sizes are in
 * elements unless named *bytes.
 */
#ifndef CA3DMM_AMD_HIP_PROXY_H
#define CA3DMM_AMD_HIP_PROXY_H

#include <stddef.h>
#include <stdint.h>

/* Same integer values as the reference's cublas_enum_t (cuda_proxy.h:7-19). */
typedef enum {
    HipNoTrans = 0,
    HipTrans,
    HipConjTrans,
    HipLeft,
    HipRight,
    HipUpper,
    HipLower,
    HipUnit,
    HipNonUnit
} hip_proxy_enum_t;

#define HIP_PROXY_OK            0
#define HIP_PROXY_ERR_ARG      -1   /* invalid argument                       */
#define HIP_PROXY_ERR_RUNTIME  -2   /* HIP runtime call failed                */
#define HIP_PROXY_ERR_NODEV    -3   /* no usable GPU                          */
#define HIP_PROXY_ERR_UNSUPP   -4   /* combination not implemented            */
#define HIP_PROXY_ERR_NOMEM    -5   /* device allocation failed               */

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- device and stream ---------------- */

/* cuda_proxy.h:26 (select_cuda_device_by_mpi_local_rank, cuda_proxy.cu:13-48):
 * device = local rank % device count, local rank read from MPI_LOCALRANKID,
 * MV2_COMM_WORLD_LOCAL_RANK, OMPI_COMM_WORLD_NODE_RANK, SLURM_LOCALID or
 * PBS_O_VNODENUM (first one set; 0 if none).  Callable before MPI_Init. */
int select_hip_device_by_mpi_local_rank(void);

/* cuda_proxy.h:30 */
int hip_set_rt_dev_id(const int dev_id);

/* Number of visible GPUs (0 when none); never initialises a context. */
int hip_proxy_device_count(void);

/* 1: entries synchronise before returning (default); 0: asynchronous. */
int hip_proxy_set_sync(int sync);

/* cuda_proxy.h:59 / :61.  stream == NULL means the library stream. */
int hip_device_sync(void);
int hip_stream_sync(void *stream);

/* The library stream of the current device, as a hipStream_t. */
void *hip_proxy_stream(void);

/* Event timer on the library stream: ms between start and stop. */
int hip_timer_start(void);
int hip_timer_stop(double *ms);

/* Last HIP error text (thread-local copy). */
const char *hip_proxy_last_error(void);

/* ---------------- memory ---------------- */

int hip_malloc_dev(void **dptr_, const size_t bytes);                        /* cuda_proxy.h:49 */
int hip_malloc_host(void **hptr_, const size_t bytes);                       /* cuda_proxy.h:51 */
int hip_memset_dev(void *dptr, const int value, const size_t bytes);         /* cuda_proxy.h:53 */
int hip_free_dev(void *dptr);                                                /* cuda_proxy.h:55 */
int hip_free_host(void *hptr);                                               /* cuda_proxy.h:57 */
int hip_memcpy_h2d(const void *hptr, void *dptr, const size_t bytes);        /* cuda_proxy.h:32 */
int hip_memcpy_d2h(const void *dptr, void *hptr, const size_t bytes);        /* cuda_proxy.h:34 */
int hip_memcpy_d2d(const void *dptr_src, void *dptr_dst, const size_t bytes);/* cuda_proxy.h:36 */
int hip_memcpy_auto(const void *src, void *dst, const size_t bytes);         /* cuda_proxy.h:38 */

/* ---------------- data movement kernels ---------------- */

/* cuda_proxy.h:63-75 (kernel cuda_proxy.cu:130-149): copy nrow rows of ncol
 * elements (row-major view) from src (leading dim lds) to dst (ldd);
 * dt_size must be 4 or 8.  Bit-exact. */
int hip_copy_matrix_block(size_t dt_size, const int nrow, const int ncol,
                          const void *src, const int lds, void *dst, const int ldd);

/* cuda_proxy.h:80: y(1:incy:n) := x(1:incx:n) */
int hip_dcopy(const int n, const double *x, const int incx, double *y, const int incy);

/* ---------------- BLAS-like kernels ---------------- */

/* cuda_proxy.h:92-97 (cuda_cublas_dgemm, cuda_proxy.cu:211-230):
 * column-major C := alpha*op(A)*op(B) + beta*C, FP64 MFMA, all four
 * transA/transB cases (HipNoTrans / HipTrans). */
int hip_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
              const int m, const int n, const int k, const double alpha,
              const double *A, const int ldA, const double *B, const int ldB,
              const double beta, double *C, const int ldC);

/* cuda_proxy.h:100-106 (cuda_cublas_dgeam): C := alpha*op(A) + beta*op(B).
 * B may be NULL when beta == 0 (the reference's transpose idiom,
 * ca3dmm.c:1039,1050,1252). */
int hip_dgeam(hip_proxy_enum_t transA, hip_proxy_enum_t transB,
              const int m, const int n,
              const double alpha, const double *A, const int ldA,
              const double beta, const double *B, const int ldB,
              double *C, const int ldC);

/* cuda_proxy.h:109-114 (cuda_cublas_dtrsm).  Implemented for
 * side=HipRight, uplo=HipUpper, trans=HipNoTrans, diag=HipNonUnit
 * (B := alpha * B * inv(A), the Q = A R^-1 step of CholeskyQR); other
 * combinations return HIP_PROXY_ERR_UNSUPP.  n <= 8192. */
int hip_dtrsm(hip_proxy_enum_t side, hip_proxy_enum_t uplo,
              hip_proxy_enum_t trans, hip_proxy_enum_t diag,
              const int m, const int n, const double alpha,
              const double *A, const int ldA, double *B, const int ldB);

/* cuda_proxy.h:117 (cuda_cusolver_dpotrf): Cholesky A = R^T R in place,
 * uplo 'U' (the strictly lower part is not referenced).  Returns LAPACK's
 * info: 0 on success, i > 0 when the leading minor of order i is not
 * positive definite, < 0 on an argument or runtime error.  n <= 8192. */
int hip_dpotrf(const char uplo, const int n, double *A, const int ldA);

#ifdef __cplusplus
}
#endif

#endif
