/*
 * Device-type seam of libca3dmm_amd.
 *
 * Replaces the reference's src/dev_type.h:9-70.  The integer values keep the
 * reference's meaning (0 = host, 1 = GPU with host-staged exchange,
 * 2 = GPU with device-direct exchange), so a caller's "dev_type" argument
 * (examples/example_AB.c:32, README.md:41) keeps working; the GPU values
 * select the HIP path on MI355X.  There are no CUDA values.
 */
#ifndef CA3DMM_AMD_DEV_TYPE_H
#define CA3DMM_AMD_DEV_TYPE_H

#include <stddef.h>
#include <stdint.h>

typedef enum {
    DEV_TYPE_HOST = 0,           /* host memory                                   */
    DEV_TYPE_HIP = 1,            /* HIP device memory, exchange staged via host   */
    DEV_TYPE_HIP_MPI_DIRECT = 2  /* HIP device memory, exchange device-direct     */
} dev_type_t;

#ifdef __cplusplus
extern "C" {
#endif

/* dev_type.c:14-23 */
int is_dev_type_valid(dev_type_t dev_type);

/* dev_type.c:26-52.  Host memory is pinned (hipHostMalloc) when the library
 * runs with a GPU, as the reference does under USE_CUDA (dev_type.c:39-46). */
void *dev_type_malloc(size_t bytes, dev_type_t dev_type);

/* dev_type.c:55-76 */
void dev_type_free(void *mem, dev_type_t dev_type);

/* dev_type.c:78-85 */
void dev_type_realloc(size_t *curr_bytes, size_t req_bytes, dev_type_t dev_type, void **mem);

/* dev_type.c:88-102 */
void dev_type_memset(void *mem, int value, size_t bytes, dev_type_t dev_type);

/* dev_type.c:105-130 */
void dev_type_memcpy(void *dst, const void *src, size_t bytes,
                     dev_type_t dst_dev_type, dev_type_t src_dev_type);

/* dev_type.c:133-157: row-major block copy (nrow rows of ncol elements of
 * dt_size bytes, leading dimensions lds/ldd in elements). */
void dev_type_copy_mat_blk(size_t dt_size, const int nrow, const int ncol,
                           const void *src, const int lds, void *dst, const int ldd,
                           dev_type_t dev_type);

/* dev_type.c:192-217: column-major C := alpha*op(A)*op(B) + beta*C. */
void dev_type_dgemm_cm(const int transA, const int transB, const int m, const int n, const int k,
                       const double alpha, const double *A, const int ldA,
                       const double *B, const int ldB, const double beta,
                       double *C, const int ldC, dev_type_t dev_type);

#ifdef __cplusplus
}
#endif

#endif
