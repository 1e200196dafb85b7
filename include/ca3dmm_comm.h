/*
 * Process-group handle taken where the reference takes an MPI communicator
 * (src/ca3dmm.h:114-124, src/cannon.h:69-72, src/mat_redist.h:69-74).
 *
 * This library runs the local compute path on one GPU and builds no
 * inter-process exchange.  A group is therefore described by its size and
 * the caller's rank in it:
 *   - nproc == 1 (ca3dmm_comm_self()): the engine computes the whole product
 *     on the device -- the drop-in single-rank path;
 *   - nproc > 1: the engines of all ranks live in one process and are built
 *     and run together by the simulated-grid API (ca3dmm_sim.h), which
 *     replaces the exchange steps by device-to-device copies.
 * An MPI caller obtains the same values from MPI_Comm_size / MPI_Comm_rank.
 */
#ifndef CA3DMM_AMD_COMM_H
#define CA3DMM_AMD_COMM_H

typedef struct ca3dmm_comm {
    int nproc;   /* number of ranks in the group */
    int rank;    /* this rank, 0 <= rank < nproc */
} ca3dmm_comm_t;

#ifdef __cplusplus
extern "C" {
#endif

/* The single-process group {1, 0} (the role of MPI_COMM_SELF). */
ca3dmm_comm_t ca3dmm_comm_self(void);

#ifdef __cplusplus
}
#endif

/* Element type codes used where the reference passes an MPI_Datatype
 * (mat_redist.h:71-73); only the element size matters to the copies. */
#define CA3DMM_DTYPE_DOUBLE 0
#define CA3DMM_DTYPE_FLOAT  1
#define CA3DMM_DTYPE_INT    2

#endif
