// dev_type dispatch (reference: src/dev_type.c:14-217).
//
// DEV_TYPE_HOST is valid for memory operations on host buffers (allocation,
// copies), as the reference's host path needs them for staging.  All
// arithmetic is device-only: dev_type_dgemm_cm on host memory returns an
// error instead of falling back to a CPU BLAS, and the engines refuse
// DEV_TYPE_HOST at init.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "dev_type.h"
#include "hip_proxy.h"

static bool is_gpu(dev_type_t t) { return t == DEV_TYPE_HIP || t == DEV_TYPE_HIP_MPI_DIRECT; }

extern "C" {

int is_dev_type_valid(dev_type_t dev_type)
{
    return dev_type == DEV_TYPE_HOST || is_gpu(dev_type);
}

void *dev_type_malloc(size_t bytes, dev_type_t dev_type)
{
    void *mem = nullptr;
    if (!is_dev_type_valid(dev_type)) {
        fprintf(stderr, "[ca3dmm_amd] invalid device type %d\n", (int) dev_type);
        return nullptr;
    }
    if (bytes == 0) return nullptr;
    if (dev_type == DEV_TYPE_HOST) {
        if (posix_memalign(&mem, 256, bytes) != 0) mem = nullptr;
    } else {
        hip_malloc_dev(&mem, bytes);
    }
    if (!mem) fprintf(stderr, "[ca3dmm_amd] failed to allocate %zu bytes on device type %d\n", bytes, (int) dev_type);
    return mem;
}

void dev_type_free(void *mem, dev_type_t dev_type)
{
    if (!mem || !is_dev_type_valid(dev_type)) return;
    if (dev_type == DEV_TYPE_HOST) free(mem);
    else hip_free_dev(mem);
}

void dev_type_realloc(size_t *curr_bytes, size_t req_bytes, dev_type_t dev_type, void **mem)
{
    if (*curr_bytes >= req_bytes) return;
    dev_type_free(*mem, dev_type);
    *curr_bytes = 0;
    *mem = dev_type_malloc(req_bytes, dev_type);
    if (*mem) *curr_bytes = req_bytes;
}

void dev_type_memset(void *mem, int value, size_t bytes, dev_type_t dev_type)
{
    if (!is_dev_type_valid(dev_type) || bytes == 0) return;
    if (dev_type == DEV_TYPE_HOST) memset(mem, value, bytes);
    else hip_memset_dev(mem, value, bytes);
}

void dev_type_memcpy(void *dst, const void *src, size_t bytes, dev_type_t dst_dev_type,
                     dev_type_t src_dev_type)
{
    if (!is_dev_type_valid(dst_dev_type) || !is_dev_type_valid(src_dev_type) || bytes == 0) return;
    bool gd = is_gpu(dst_dev_type), gs = is_gpu(src_dev_type);
    if (!gd && !gs) memcpy(dst, src, bytes);
    else if (gd && !gs) hip_memcpy_h2d(src, dst, bytes);
    else if (!gd && gs) hip_memcpy_d2h(src, dst, bytes);
    else hip_memcpy_d2d(src, dst, bytes);
}

void dev_type_copy_mat_blk(size_t dt_size, const int nrow, const int ncol, const void *src,
                           const int lds, void *dst, const int ldd, dev_type_t dev_type)
{
    if (!is_dev_type_valid(dev_type)) return;
    if (dev_type == DEV_TYPE_HOST) {
        const char *s = (const char *) src;
        char *d = (char *) dst;
        for (int r = 0; r < nrow; ++r)
            memcpy(d + dt_size * (size_t) r * ldd, s + dt_size * (size_t) r * lds, dt_size * (size_t) ncol);
        return;
    }
    hip_copy_matrix_block(dt_size, nrow, ncol, src, lds, dst, ldd);
}

void dev_type_dgemm_cm(const int transA, const int transB, const int m, const int n, const int k,
                       const double alpha, const double *A, const int ldA, const double *B,
                       const int ldB, const double beta, double *C, const int ldC,
                       dev_type_t dev_type)
{
    if (!is_gpu(dev_type)) {
        fprintf(stderr, "[ca3dmm_amd] dev_type_dgemm_cm: host arithmetic is not provided; "
                        "use DEV_TYPE_HIP with device buffers\n");
        return;
    }
    hip_dgemm(transA ? HipTrans : HipNoTrans, transB ? HipTrans : HipNoTrans, m, n, k, alpha, A, ldA,
              B, ldB, beta, C, ldC);
}

}  // extern "C"
