/*
 * 3D process-grid choice and block partition arithmetic (host only).
 * Same functions and integer results as the reference's src/partition.h:16-31
 * and src/utils.h (calc_block_spos_size); the grid search is restated from
 * src/partition.c:9-232 with the same floating-point evaluation order so that
 * every tie breaks the same way.
 */
#ifndef CA3DMM_AMD_PARTITION_H
#define CA3DMM_AMD_PARTITION_H

#ifdef __cplusplus
extern "C" {
#endif

/* partition.c:98-194: near-optimal mp x np x kp (+ rp idle) for an
 * m x n x k cuboid on p processes with max(mp,np) % min(mp,np) == 0. */
void calc_3d_decomposition_cannon(const int p, const int m, const int n, const int k,
                                  int *mp, int *np, int *kp, int *rp);

/* partition.c:196-232: np x np x kp (+ rp) for C = B^T B, B k x n. */
void calc_3d_decomposition_nk(const int p, const int n, const int k, int *np, int *kp, int *rp);

/* utils.c:26-49: start and size of block iblk of nblk near-equal blocks of
 * len (the first len % nblk blocks are one longer); iblk == nblk gives
 * (len, 0); iblk outside [0, nblk] gives (-1, 0). */
void calc_block_spos_size(const int len, const int nblk, const int iblk, int *blk_spos, int *blk_size);

#ifdef __cplusplus
}
#endif

#endif
