/*
 * 2D Cannon engine: C := alpha * A * B + beta * C on a q x q grid.
 * Interface of the reference's src/cannon.h:7-109 with the same field names
 * and order (ca3dmm reads rank_row / rank_col, src/ca3dmm.c:1196).  On one
 * GPU the q shift rounds need no data movement: the round-s operands are
 * read in place from the ranks that hold them (simulated grid), and each
 * local product is the FP64 MFMA DGEMM.
 */
#ifndef CA3DMM_AMD_CANNON_H
#define CA3DMM_AMD_CANNON_H

#include <stddef.h>
#include "ca3dmm_comm.h"
#include "dev_type.h"

struct cannon_engine {
    int  m, n, k;               /* A is m x k, B k x n, C m x n                        */
    int  my_rank;               /* rank in the Cannon group                            */
    int  rank_row;              /* grid row of this rank                               */
    int  rank_col;              /* grid column of this rank                            */
    int  np_dim;                /* q: ranks per grid dimension                         */
    int  A_srow, A_scol;        /* first row / column of this rank's A block           */
    int  B_srow, B_scol;        /* first row / column of this rank's B block           */
    int  C_srow, C_scol;        /* first row / column of this rank's C block           */
    int  A_nrow, A_ncol;        /* rows / columns of this rank's A block               */
    int  B_nrow, B_ncol;        /* rows / columns of this rank's B block               */
    int  C_nrow, C_ncol;        /* rows / columns of this rank's C block               */
    int  gemm_cycle;            /* shift steps per local GEMM (k-stacking)             */
    int  max_A_blk_size;        /* (m/q + 1) * (k/q + 1)                               */
    int  max_B_blk_size;        /* (k/q + 1) * (n/q + 1)                               */
    int  max_C_blk_size;        /* (m/q + 1) * (n/q + 1)                               */
    int  alloc_workbuf;         /* 1 if the engine owns its work buffer                */
    int  *m_displs;             /* [q + 1] block starts along m                        */
    int  *n_displs;             /* [q + 1] block starts along n                        */
    int  *k_displs;             /* [q + 1] block starts along k                        */
    void *A_recv_h;             /* host shift buffers: unused on one GPU (NULL)        */
    void *A_stack_h;
    void *B_recv_h;
    void *B_stack_h;
    void *A_recv_d;             /* device receive buffer of A (workspace)              */
    void *A_stack_d;            /* device k-stack of A (gemm_cycle > 1)                */
    void *B_recv_d;             /* device receive buffer of B                          */
    void *B_stack_d;            /* device k-stack of B                                 */
    void *workbuf_h;
    void *workbuf_d;
    void *A_gemm_h;             /* unused on one GPU (NULL)                            */
    void *B_gemm_h;
    dev_type_t dev_type;
    ca3dmm_comm_t comm;         /* the q*q Cannon group (was an MPI communicator)       */

    /* statistics (cannon.h:43-50) */
    double init_ms;
    double shift0_ms;
    double lshift_ms;
    double gemm_ms;
    double exec_ms;
    double hd_trans_ms;
    int    n_exec;
};
typedef struct cannon_engine  cannon_engine_s;
typedef struct cannon_engine *cannon_engine_p;

#ifdef __cplusplus
extern "C" {
#endif

/* cannon.h:69-72.  comm.nproc must be a square q*q; called directly only the
 * 1 x 1 grid is accepted (the simulated grid builds larger ones). */
void cannon_engine_init(const int m, const int n, const int k, ca3dmm_comm_t comm,
                        dev_type_t dev_type, cannon_engine_p *engine_, size_t *workbuf_bytes);

/* cannon.h:81 */
void cannon_engine_attach_workbuf(cannon_engine_p engine, void *workbuf_h, void *workbuf_d);

/* cannon.h:84 */
void cannon_engine_free(cannon_engine_p *engine_);

/* cannon.h:100-103: device pointers, column-major blocks with leading
 * dimensions A_nrow, B_nrow (= local k) and C_nrow.  Returns with C complete.
 * As in the reference, a zero m, n or k leaves C untouched. */
void cannon_engine_exec(cannon_engine_p engine, const double alpha, const double beta,
                        double *A_blk, double *B_blk, double *C_blk);

/* cannon.h:106 / :109 */
void cannon_engine_reset_stat(cannon_engine_p engine);
void cannon_engine_print_stat(cannon_engine_p engine);

#ifdef __cplusplus
}
#endif

/* Callers built with -DCA3DMM_MPI get the reference's MPI-typed
 * constructors (include/ca3dmm_mpi.h, libca3dmm_amd_mpi.so). */
#if defined(CA3DMM_MPI) && !defined(CA3DMM_AMD_CA3DMM_H)
#include "ca3dmm_mpi.h"
#endif

#endif
