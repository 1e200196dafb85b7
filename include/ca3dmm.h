/*
 * CA3DMM engine: C := op(A) * op(B) and C := B^T * B.
 * Interface of the reference's src/ca3dmm.h:10-200.  The engine fields keep
 * the reference's names and order (callers read mp/np/kp/rp/my_rank,
 * examples/example_AB.c:148-162, and B_rd_*, examples/example_CholeskyQR.c:58-61);
 * MPI communicators are replaced by process-group handles.
 *
 * Device data flow of one exec on the HIP device types:
 *   redistribute A, B (batched pack/unpack)  ->  op() fused into the GEMM
 *   loads  ->  local FP64 MFMA DGEMM (Cannon rounds)  ->  k-layer fold sum
 *   ->  redistribute C.  On the single-process group whose redistribution is
 *   the identity, exec is exactly one DGEMM from the caller's buffers into
 *   dst_C.
 */
#ifndef CA3DMM_AMD_CA3DMM_H
#define CA3DMM_AMD_CA3DMM_H

#include <stddef.h>
#include "ca3dmm_comm.h"
#include "cannon.h"
#include "dev_type.h"
#include "mat_redist.h"
#include "partition.h"

struct ca3dmm_engine {
    int  m, n, k;                   /* op(A) m x k, op(B) k x n, C m x n                  */
    int  mp, np, kp, rp;            /* process grid and idle ranks                        */
    int  my_rank;                   /* rank in the group                                  */
    int  n_proc;                    /* ranks in the group                                 */
    int  is_active;                 /* 1 if this rank computes                            */
    int  rank_m, rank_n, rank_k;    /* grid coordinates                                   */
    int  trans_A, trans_B;          /* op() of A and B                                    */
    int  task_m_num, task_m_id;     /* Cannon tasks along m (B replicated)                */
    int  task_n_num, task_n_id;     /* Cannon tasks along n (A replicated)                */
    int  task_k_num, task_k_id;     /* k layers                                           */
    int  is_BTB;                    /* 1 for C = B^T B                                    */
    int  use_ag, use_rsb;           /* equal-size allgather / reduce-scatter blocks       */
    int  A_rd_srow,   A_rd_scol;    /* op(A) block fetched by redistribution              */
    int  B_rd_srow,   B_rd_scol;    /* op(B) block fetched by redistribution              */
    int  A_2dmm_srow, A_2dmm_scol;  /* op(A) block used by the Cannon step                */
    int  B_2dmm_srow, B_2dmm_scol;  /* op(B) block used by the Cannon step                */
    int  C_2dmm_srow, C_2dmm_scol;  /* C block computed by the Cannon step                */
    int  C_out_srow,  C_out_scol;   /* C block owned after the k-layer reduction          */
    int  A_rd_nrow,   A_rd_ncol;
    int  B_rd_nrow,   B_rd_ncol;
    int  A_2dmm_nrow, A_2dmm_ncol;
    int  B_2dmm_nrow, B_2dmm_ncol;
    int  C_2dmm_nrow, C_2dmm_ncol;
    int  C_out_nrow,  C_out_ncol;
    int  alloc_workbuf;             /* 1 if the engine owns its work buffer               */
    int  *AB_agv_recvcnts;          /* allgather counts of the replicated operand         */
    int  *AB_agv_displs;            /* allgather displacements                            */
    int  *C_rs_recvcnts;            /* [task_k_num] reduce-scatter counts                 */
    void *A_rd_recv_h;              /* host buffers: unused on the HIP types (NULL)       */
    void *A_trans_h;
    void *A_2dmm_h;
    void *B_rd_recv_h;
    void *B_trans_h;
    void *B_2dmm_h;
    void *C_2dmm_h;
    void *C_out_h;
    void *A_rd_recv_d;              /* device aliases into workbuf_d                      */
    void *A_trans_d;
    void *A_2dmm_d;
    void *B_rd_recv_d;
    void *B_trans_d;
    void *B_2dmm_d;
    void *C_2dmm_d;
    void *C_out_d;
    void *workbuf_h;
    void *workbuf_d;
    size_t   rdA_workbuf_bytes;
    size_t   rdB_workbuf_bytes;
    size_t   rdC_workbuf_bytes;
    size_t   cannon_workbuf_bytes;
    size_t   self_workbuf_bytes;
    ca3dmm_comm_t comm_AB_agv;      /* allgather group                                    */
    ca3dmm_comm_t comm_C_rs;        /* reduce-scatter group (the kp layers)               */
    ca3dmm_comm_t comm_2dmm;        /* Cannon group                                       */
    mat_redist_engine_p redist_A;
    mat_redist_engine_p redist_B;
    mat_redist_engine_p redist_C;
    cannon_engine_p cannon_engine;
    dev_type_t dev_type;

    /* statistics (ca3dmm.h:72-81) */
    int    print_timing;
    double init_ms;
    double redist_ms;
    double agvAB_ms;
    double cannon_ms;
    double reduce_ms;
    double hd_trans_ms;
    double exec_ms;
    int    n_exec;
};
typedef struct ca3dmm_engine  ca3dmm_engine_s;
typedef struct ca3dmm_engine *ca3dmm_engine_p;

#ifdef __cplusplus
extern "C" {
#endif

/* ca3dmm.h:114-124.  Block arguments are column-major (srow, nrow, scol,
 * ncol) of A and B as stored; all dst_C_* == -1 skips the C redistribution.
 * proc_grid may be NULL (automatic, env CA3DMM_REDUCE_KP honoured).
 * workbuf_bytes == NULL: the engine allocates its own device work buffer.
 * Called directly, only the single-process group is accepted; a p-rank grid
 * is built by ca3dmm_sim_init (ca3dmm_sim.h).  *engine_ is NULL on failure. */
void ca3dmm_engine_init(const int m, const int n, const int k, const int trans_A, const int trans_B,
                        const int src_A_srow, const int src_A_nrow, const int src_A_scol,
                        const int src_A_ncol, const int src_B_srow, const int src_B_nrow,
                        const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                        const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                        const int *proc_grid, ca3dmm_comm_t comm, dev_type_t dev_type,
                        ca3dmm_engine_p *engine_, size_t *workbuf_bytes);

/* ca3dmm.h:149-157: C := B^T B, B k x n. */
void ca3dmm_engine_init_BTB(const int n, const int k, const int src_B_srow, const int src_B_nrow,
                            const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                            const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                            const int *proc_grid, ca3dmm_comm_t comm, dev_type_t dev_type,
                            ca3dmm_engine_p *engine_, size_t *workbuf_bytes);

/* ca3dmm.h:167: workbuf_d >= *workbuf_bytes of device memory; workbuf_h is
 * not used on the HIP device types (may be NULL). */
void ca3dmm_engine_attach_workbuf(ca3dmm_engine_p engine, void *workbuf_h, void *workbuf_d);

/* ca3dmm.h:170 */
void ca3dmm_engine_free(ca3dmm_engine_p *engine_);

/* ca3dmm.h:191-194: device pointers (column-major).  Returns with dst_C
 * complete. */
void ca3dmm_engine_exec(ca3dmm_engine_p engine, const void *src_A, const int ldA,
                        const void *src_B, const int ldB, void *dst_C, const int ldC);

/* ca3dmm.h:197 / :200 */
void ca3dmm_engine_reset_stat(ca3dmm_engine_p engine);
void ca3dmm_engine_print_stat(ca3dmm_engine_p engine);

#ifdef __cplusplus
}
#endif

/* dev_type.h:76-101: allocate and attach an engine's work buffer (device
 * memory on the HIP device types). */
#define MALLOC_ATTACH_WORKBUF(attach_func, free_func, engine, dev_type, workbuf_bytes, workbuf_h, \
                              workbuf_d)                                                        \
    do {                                                                                        \
        workbuf_h = NULL;                                                                       \
        workbuf_d = NULL;                                                                       \
        if ((dev_type) == DEV_TYPE_HOST) {                                                      \
            workbuf_h = dev_type_malloc(workbuf_bytes, DEV_TYPE_HOST);                          \
        } else {                                                                                \
            workbuf_d = dev_type_malloc(workbuf_bytes, dev_type);                               \
            if (workbuf_d == NULL && (workbuf_bytes) > 0) {                                     \
                free_func(&engine);                                                             \
                break;                                                                          \
            }                                                                                   \
        }                                                                                       \
        attach_func(engine, workbuf_h, workbuf_d);                                              \
    } while (0)

/* Callers built with -DCA3DMM_MPI get the reference's MPI-typed
 * constructors (include/ca3dmm_mpi.h, libca3dmm_amd_mpi.so). */
#ifdef CA3DMM_MPI
#include "ca3dmm_mpi.h"
#endif

#endif
