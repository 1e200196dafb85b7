/*
 * CPU oracle for the CA3DMM local compute path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against; it is never linked into, loaded by or called
 * from the product library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may use it.
 *
 * Every function restates one piece of the reference's algorithm and cites
 * the reference file:line it follows (paths relative to the reference root).
 */
#ifndef CA3DMM_ORACLE_H
#define CA3DMM_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/utils.c:26-49 */
void or_block_spos_size(int len, int nblk, int iblk, int *spos, int *size);

/* src/partition.c:98-194 (+ fixed_pair_scale_grid :21-96, max_min_ratio :9-19) */
void or_grid_cannon(int p, int m, int n, int k, int *mp, int *np, int *kp, int *rp);

/* src/partition.c:196-232 */
void or_grid_nk(int p, int n, int k, int *np, int *kp, int *rp);

/* src/utils.c:92-120: row-major block copy, element size dt bytes */
void or_copy_block(size_t dt, int nrow, int ncol, const void *src, int lds, void *dst, int ldd);

/* src/utils.c:181-196: AT(nrow x ncol, col-major) := transpose of A (ncol x nrow, col-major) */
void or_transpose_cm(int nrow, int ncol, const double *A, int ldA, double *AT, int ldAT);

/* src/mpi_op_omp.c:6-23: inout[i] += in[i]; dtype 0 = double, 1 = float, 2 = int */
void or_op_sum(int dtype, int len, const void *in, void *inout);

/* The k-layer reduction in canonical order: dst = ((P0 + P1) + P2) + ... ,
 * each Pt a col-major nrow x ncol block with leading dimension ldp[t]. */
void or_fold_sum(int nsrc, const double *const *P, const int *ldp, int nrow, int ncol,
                 double *dst, int ldd);

/*
 * Redistribution plan (src/mat_redist.c:9-41 intersections, :44-236 lists).
 * Rectangles are row-major {srow, scol, nrow, ncol}, one per rank.
 * send/recv lists get 6 ints per entry {srow, scol, nrow, ncol, peer, displ},
 * in ascending peer order; capacity nproc entries each.
 */
void or_redist_plan(int nproc, const int *all_src, const int *all_req, int rank,
                    int *send_list, int *n_send, int *send_cnt,
                    int *recv_list, int *n_recv, int *recv_cnt);

/*
 * All ranks' pack -> neighbour all-to-all -> unpack, simulated in one process
 * (src/mat_redist.c:298-419).  src_blk[r] / dst_blk[r] are rank r's buffers.
 */
void or_redist_exec_all(int nproc, size_t dt, const int *all_src, const int *all_req,
                        const void *const *src_blk, const int *src_ld,
                        void *const *dst_blk, const int *dst_ld);

/* C := alpha*op(A)*op(B) + beta*C, col-major, sequential k order, no FMA. */
void or_dgemm(int transA, int transB, int m, int n, int k, double alpha,
              const double *A, int lda, const double *B, int ldb, double beta,
              double *C, int ldc);

/* Double-double dot products for sampled entries of op(A)*op(B):
 * ref[s] = sum_l op(A)[ii[s],l]*op(B)[l,jj[s]] rounded from a double-double
 * accumulator; absab[s] = sum_l |op(A)[ii[s],l]|*|op(B)[l,jj[s]]|. */
void or_dd_sample(int transA, int transB, int m, int n, int k,
                  const double *A, int lda, const double *B, int ldb,
                  int nsamp, const int *ii, const int *jj, double *ref, double *absab);

/* Cannon per-rank plan (src/cannon.c:16-133, :196-250) */
typedef struct {
    int m, n, k, np_dim, rank_row, rank_col;
    int A_srow, A_scol, A_nrow, A_ncol;
    int B_srow, B_scol, B_nrow, B_ncol;
    int C_srow, C_scol, C_nrow, C_ncol;
    int gemm_cycle, max_A_blk_size, max_B_blk_size, max_C_blk_size;
    long long workbuf_bytes;
} or_cannon_plan_t;

void or_cannon_plan(int m, int n, int k, int np_dim, int rank2d, int min_k_blk,
                    or_cannon_plan_t *pl);

/* CA3DMM per-rank plan (src/ca3dmm.c:23-444 general, :447-722 BTB) */
typedef struct {
    int mp, np, kp, rp, is_active;
    int rank_m, rank_n, rank_k;
    int task_m_num, task_m_id, task_n_num, task_n_id, task_k_num, task_k_id;
    int rank_2dmm, use_ag, use_rsb;
    int A_rd_srow, A_rd_scol, A_rd_nrow, A_rd_ncol;
    int B_rd_srow, B_rd_scol, B_rd_nrow, B_rd_ncol;
    int A_2dmm_srow, A_2dmm_scol, A_2dmm_nrow, A_2dmm_ncol;
    int B_2dmm_srow, B_2dmm_scol, B_2dmm_nrow, B_2dmm_ncol;
    int C_2dmm_srow, C_2dmm_scol, C_2dmm_nrow, C_2dmm_ncol;
    int C_out_srow, C_out_scol, C_out_nrow, C_out_ncol;
    long long self_workbuf_bytes, cannon_workbuf_bytes;
    or_cannon_plan_t cannon;
} or_ca3dmm_plan_t;

/* proc_grid may be NULL (automatic); reduce_kp follows CA3DMM_REDUCE_KP. */
int or_ca3dmm_plan(int m, int n, int k, int transA, int transB, int p, int rank,
                   const int *proc_grid, int reduce_kp, int is_btb, int min_k_blk,
                   or_ca3dmm_plan_t *pl);

#ifdef __cplusplus
}
#endif
#endif
