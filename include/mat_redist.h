/*
 * Block redistribution of a 2D-partitioned matrix (row-major view).
 * Interface of the reference's src/mat_redist.h:7-100; the engine fields keep
 * the reference's names and order (callers read n_proc_send / n_proc_recv,
 * examples/example_CholeskyQR.c:115-116).  MPI handles are replaced by the
 * process-group handle; on the HIP device types the pack and unpack of all
 * blocks each run as ONE batched copy launch.
 */
#ifndef CA3DMM_AMD_MAT_REDIST_H
#define CA3DMM_AMD_MAT_REDIST_H

#include <stddef.h>
#include "ca3dmm_comm.h"
#include "dev_type.h"

struct mat_redist_engine {
    ca3dmm_comm_t comm;     /* group of the redistribution (was graph_comm)          */
    int     dtype;          /* CA3DMM_DTYPE_* (was an MPI_Datatype)                  */
    size_t  dt_size;        /* element size in bytes                                 */
    int     nproc;          /* ranks in the group                                    */
    int     rank;           /* this rank                                             */
    int     src_srow;       /* first row of this rank's source block                 */
    int     src_scol;       /* first column of this rank's source block              */
    int     src_nrow;       /* rows of this rank's source block                      */
    int     src_ncol;       /* columns of this rank's source block                   */
    int     req_srow;       /* first row this rank requires                          */
    int     req_scol;       /* first column this rank requires                       */
    int     req_nrow;       /* rows this rank requires                               */
    int     req_ncol;       /* columns this rank requires                            */
    int     n_proc_send;    /* ranks this rank sends a block to                      */
    int     n_proc_recv;    /* ranks this rank receives a block from                 */
    int     send_cnt;       /* elements sent                                         */
    int     recv_cnt;       /* elements received                                     */
    int     alloc_workbuf;  /* 1 if the engine owns its work buffer                  */
    int    *send_ranks;     /* [n_proc_send] destination ranks, ascending            */
    int    *send_sizes;     /* [n_proc_send] elements per send block                 */
    int    *send_displs;    /* [n_proc_send + 1] offsets in the send buffer          */
    int    *sblk_sizes;     /* [n_proc_send * 4] srow, scol, nrow, ncol per block     */
    int    *recv_ranks;     /* [n_proc_recv] source ranks, ascending                 */
    int    *recv_sizes;     /* [n_proc_recv] elements per receive block              */
    int    *recv_displs;    /* [n_proc_recv + 1] offsets in the receive buffer       */
    int    *rblk_sizes;     /* [n_proc_recv * 4] srow, scol, nrow, ncol per block     */
    int    *send_info0;     /* kept for layout compatibility; always NULL            */
    int    *recv_info0;     /* kept for layout compatibility; always NULL            */
    void   *sendbuf_h;      /* host send buffer (NULL on the HIP device types)       */
    void   *recvbuf_h;      /* host receive buffer (NULL on the HIP device types)    */
    void   *sendbuf_d;      /* device send buffer                                    */
    void   *recvbuf_d;      /* device receive buffer                                 */
    void   *workbuf_h;      /* host work buffer                                      */
    void   *workbuf_d;      /* device work buffer: send then receive buffer          */
    double  hd_trans_ms;    /* host <-> device copy time of the last exec            */
    dev_type_t dev_type;    /* where the data lives                                  */
};
typedef struct mat_redist_engine  mat_redist_engine_s;
typedef struct mat_redist_engine *mat_redist_engine_p;

#ifdef __cplusplus
extern "C" {
#endif

/* mat_redist.h:69-74.  For comm.nproc > 1 the other ranks' rectangles come
 * from the simulated grid (ca3dmm_sim.h); called directly, only the
 * single-process group is accepted.  *engine_ is NULL on failure. */
void mat_redist_engine_init(const int src_srow, const int src_scol, const int src_nrow,
                            const int src_ncol, const int req_srow, const int req_scol,
                            const int req_nrow, const int req_ncol, ca3dmm_comm_t comm,
                            int dtype, const size_t dt_size, dev_type_t dev_type,
                            mat_redist_engine_p *engine_, size_t *workbuf_bytes);

/* mat_redist.h:83 */
void mat_redist_engine_attach_workbuf(mat_redist_engine_p engine, void *workbuf_h, void *workbuf_d);

/* mat_redist.h:92-95: pack (one batched launch) -> exchange -> unpack (one
 * batched launch); returns with the destination complete. */
void mat_redist_engine_exec(mat_redist_engine_p engine, const void *src_blk, const int src_ld,
                            void *dst_blk, const int dst_ld);

/* mat_redist.h:98 */
void mat_redist_engine_free(mat_redist_engine_p *engine_);

#ifdef __cplusplus
}
#endif

/* Callers built with -DCA3DMM_MPI get the reference's MPI-typed
 * constructors (include/ca3dmm_mpi.h, libca3dmm_amd_mpi.so). */
#if defined(CA3DMM_MPI) && !defined(CA3DMM_AMD_CA3DMM_H)
#include "ca3dmm_mpi.h"
#endif

#endif
