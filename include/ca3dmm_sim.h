/*
 * Simulated process grid: every rank of a p-rank CA3DMM grid planned and run
 * in ONE process on ONE GPU.  No reference counterpart as an API; it stands
 * in for `mpirun -np p` on one node (the reference's only multi-rank test
 * method, README.md:31-33) and shows the 3D decomposition at its real
 * per-rank shapes.
 *
 * Each rank's engine is built exactly as ca3dmm_engine_init(_BTB) would on
 * that rank (same grid, blocks, redistribution lists and workspace size) and
 * is readable through ca3dmm_sim_engine().  ca3dmm_sim_exec runs the
 * reference's exec sequence for all ranks at once, with every exchange
 * replaced by device-to-device copies:
 *   redistribute A, B   pack (all ranks, one launch) -> exchange -> unpack
 *   transpose, allgather of the replicated operand
 *   Cannon              q rounds per rank, operands read in place
 *   reduce-scatter      fixed-order k-layer fold (rank order)
 *   redistribute C
 */
#ifndef CA3DMM_AMD_SIM_H
#define CA3DMM_AMD_SIM_H

#include "ca3dmm.h"

typedef struct ca3dmm_sim *ca3dmm_sim_p;

#ifdef __cplusplus
extern "C" {
#endif

/* Plan all p ranks; with allocate != 0 also allocate each rank's device work
 * buffer (allocate == 0 plans only: fields are readable, exec refuses).
 * src_A, src_B, dst_C: 4*p ints, rank r's column-major block
 * {srow, nrow, scol, ncol} of A, B (as stored) and C; dst_C == NULL skips
 * the C redistribution.  For is_btb != 0, C := B^T B with B k x n (m, transA,
 * transB and src_A are ignored).  Returns 0 or a negative status. */
int ca3dmm_sim_init(int p, int m, int n, int k, int trans_A, int trans_B, const int *src_A,
                    const int *src_B, const int *dst_C, const int *proc_grid, int is_btb,
                    dev_type_t dev_type, int allocate, ca3dmm_sim_p *sim);

/* Work-buffer bytes rank r's ca3dmm_engine_init would report
 * (src/ca3dmm.c:398-423: largest child buffer + own buffers). */
size_t ca3dmm_sim_workbuf_bytes(ca3dmm_sim_p sim, int rank);

/* Rank r's engine (fields as the reference's ca3dmm_engine on that rank). */
ca3dmm_engine_p ca3dmm_sim_engine(ca3dmm_sim_p sim, int rank);

/* Run one exec of every rank.  Per-rank device pointers and leading
 * dimensions (column-major).  For is_btb src_A / ldA are ignored. */
int ca3dmm_sim_exec(ca3dmm_sim_p sim, const void *const *src_A, const int *ldA,
                    const void *const *src_B, const int *ldB, void *const *dst_C, const int *ldC);

/* Device address of rank r's C_out block (C_out_nrow x C_out_ncol,
 * leading dimension C_out_nrow) after exec. */
void *ca3dmm_sim_c_out(ca3dmm_sim_p sim, int rank);

/* grouped != 0: run the Cannon rounds of all ranks and the k-layer fold as
 * ONE persistent launch pulling tiles of every rank's C from a device work
 * queue (rounds accumulated in registers; the last layer to finish a tile
 * adds the kp layer tiles in layer order), or as the tuner's batched
 * candidate (see ca3dmm_sim_autotune).  The grouped mode and the Ozaki form
 * hold at most 16 k layers and 16 Cannon rounds per rank; exec returns
 * HIP_PROXY_ERR_UNSUPP on larger grids (the default mode has no such
 * limit: beyond 16 rounds it launches per round).  use_graph != 0: capture the whole
 * exec -- redistributions, transposes, allgathers, products, fold -- into one
 * hipGraph on the first call (again when an argument or setting changes)
 * and replay it.  Default: every rank's rounds in one batched DGEMM launch
 * (the bits of one launch per rank and round; CA3DMM_SIM_BATCH=0 makes
 * those launches), eager. */
int ca3dmm_sim_set_mode(ca3dmm_sim_p sim, int grouped, int use_graph);

/* Time the grouped launch over its tunings -- tile 0, 1, 2: the persistent
 * register-staged kernel with workgroup tile 128 x 128 (4 waves), 256 x 128
 * or 128 x 256 (8 waves); tile 3: one batched direct-to-LDS launch of
 * 128 x 128 tiles with the rounds accumulated in registers, then the fold
 * launches (the same bits); tile rows walked together: 1, 4, 8, 16;
 * persistent workgroups per CU: 1, 2 (tile 0 only) -- keep the fastest for
 * this grid's shape signature, in memory and in the tuning table on disk
 * (CA3DMM_TUNE_CACHE, default ~/.cache/ca3dmm_amd/tuning.txt; a later plan
 * of the same signature starts from it), and report it.  tile (may be
 * NULL): 0 to 3 as above. */
int ca3dmm_sim_autotune(ca3dmm_sim_p sim, int reps, int *group_rows, int *wg_per_cu, int *tile);

/* Fix the grouped launch's tuning by hand (e.g. the untuned baseline). */
int ca3dmm_sim_set_tuning(ca3dmm_sim_p sim, int group_rows, int wg_per_cu, int tile);

/* nmoduli > 0 (4..25): every rank's local product -- all its Cannon rounds
 * -- runs through the CRT form of the Ozaki scheme (hip_ozaki2_dgemm's
 * arithmetic on the concatenated k-segments), and the k-layer fold adds the
 * reconstructed layer products in layer order.  With set_mode(grouped = 1)
 * the products of all ranks run as one persistent launch fed by a device
 * queue (else one launch per rank); use_graph replays the whole sequence
 * (line scales, residues, products, reconstruction and fold) from a
 * captured hipGraph.  0 returns to the FP64 MFMA form.
 * nmoduli = -1: the CRT form with the count chosen at every exec, after the
 * operands are in place, by hip_ozaki_choose's rule on the local products:
 * its 4096 sampled entries (i_s, j_s) of C (same hashes over the whole
 * C), sample s taken in k layer s mod kp from the rank whose C block holds
 * it (a and b over that layer's k range), k = the longest layer.  The exec
 * then synchronises once before the products (the statistics are read
 * back), rebuilds the plan when the count changes, and with use_graph
 * replays the operand and product halves from one hipGraph each.
 * No reference counterpart (the reference's local GEMM is
 * src/cannon.c:405-427). */
int ca3dmm_sim_set_ozaki(ca3dmm_sim_p sim, int nmoduli);

/* The modulus count of the last exec (or the one set; 0: FP64 MFMA form),
 * whether it is chosen per exec (set_ozaki(-1)), and the spread statistic
 * L = max log2(||a||_2 ||b||_2 / a.b) of the last automatic choice (NaN
 * before one).  Any pointer may be NULL.  No reference counterpart. */
int ca3dmm_sim_get_ozaki(ca3dmm_sim_p sim, int *nmoduli, int *automatic, double *spread);

/* The Ozaki form's grouped launch: time its tunings (workgroup tile 256 x
 * 256 with 8 waves or 256 x 128 with 4 waves; queue items in 8 per-XCD
 * ranges or from one counter), keep the fastest for this grid's shape
 * signature and report it.  Needs set_ozaki(nmoduli > 0) first, or
 * set_ozaki(-1) and one exec.  No reference counterpart. */
int ca3dmm_sim_autotune_ozaki(ca3dmm_sim_p sim, int reps, int *tile_n, int *xcd_queue);

/* Fix the Ozaki form's tuning by hand (tile_n 256 or 128). */
int ca3dmm_sim_set_tuning_ozaki(ca3dmm_sim_p sim, int tile_n, int xcd_queue);

void ca3dmm_sim_free(ca3dmm_sim_p *sim);

#ifdef __cplusplus
}
#endif

#endif
