/*
 * The reference's MPI-typed entry points, for callers compiled against MPI.
 *
 * The reference declares its engines with MPI types by value:
 *   ca3dmm_engine_init(..., MPI_Comm comm, ...)            src/ca3dmm.h:114-124
 *   ca3dmm_engine_init_BTB(..., MPI_Comm comm, ...)        src/ca3dmm.h:149-157
 *   cannon_engine_init(m, n, k, MPI_Comm comm, ...)        src/cannon.h:69-72
 *   mat_redist_engine_init(..., MPI_Comm comm,
 *                          MPI_Datatype dtype, ...)        src/mat_redist.h:69-74
 * libca3dmm_amd takes a ca3dmm_comm_t {nproc, rank} there and links no MPI.
 * This header and libca3dmm_amd_mpi.so (ca3dmm_amd/csrc/mpi_binding.c, built
 * with the caller's mpicc) restore the MPI signatures:
 *
 *   - a caller builds with -DCA3DMM_MPI and includes "ca3dmm.h" (or
 *     "cannon.h" / "mat_redist.h") as it includes the reference's headers;
 *     this header is then pulled in, and the four names above resolve to the
 *     *_mpi functions below, so calls such as
 *         ca3dmm_engine_init(..., proc_grid, MPI_COMM_WORLD, dev_type, &ce, &bytes);
 *     compile unchanged;
 *   - each *_mpi function reads MPI_Comm_size / MPI_Comm_rank and forwards to
 *     the ca3dmm_comm_t entry.  The library runs one GPU per process and
 *     builds no inter-process exchange, so a communicator of more than one
 *     rank is refused: *engine_ = NULL and a message on stderr, the
 *     reference's failure convention (src/ca3dmm.c:36-40).  A p-rank grid on
 *     one GPU is built by ca3dmm_sim_init (ca3dmm_sim.h) instead;
 *   - MPI_DOUBLE, MPI_FLOAT and MPI_INT map to the CA3DMM_DTYPE_* codes;
 *     any other datatype maps by dt_size (8 -> double, 4 -> float).
 *
 * Everything else (attach_workbuf, exec, free, stat calls) has no MPI type
 * in its signature and is called as declared in ca3dmm.h.  The comm fields
 * of the public structs hold ca3dmm_comm_t values.
 */
#ifndef CA3DMM_AMD_MPI_H
#define CA3DMM_AMD_MPI_H

#include <mpi.h>

#include "ca3dmm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* {MPI_Comm_size, MPI_Comm_rank} of comm. */
ca3dmm_comm_t ca3dmm_comm_from_mpi(MPI_Comm comm);

/* CA3DMM_DTYPE_* code of an MPI datatype (see above). */
int ca3dmm_dtype_from_mpi(MPI_Datatype dtype, size_t dt_size);

/* src/ca3dmm.h:114-124 */
void ca3dmm_engine_init_mpi(const int m, const int n, const int k, const int trans_A, const int trans_B,
                            const int src_A_srow, const int src_A_nrow, const int src_A_scol,
                            const int src_A_ncol, const int src_B_srow, const int src_B_nrow,
                            const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                            const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                            const int *proc_grid, MPI_Comm comm, dev_type_t dev_type,
                            ca3dmm_engine_p *engine_, size_t *workbuf_bytes);

/* src/ca3dmm.h:149-157 */
void ca3dmm_engine_init_BTB_mpi(const int n, const int k, const int src_B_srow, const int src_B_nrow,
                                const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                                const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                                const int *proc_grid, MPI_Comm comm, dev_type_t dev_type,
                                ca3dmm_engine_p *engine_, size_t *workbuf_bytes);

/* src/cannon.h:69-72 */
void cannon_engine_init_mpi(const int m, const int n, const int k, MPI_Comm comm, dev_type_t dev_type,
                            cannon_engine_p *engine_, size_t *workbuf_bytes);

/* src/mat_redist.h:69-74 */
void mat_redist_engine_init_mpi(const int src_srow, const int src_scol, const int src_nrow,
                                const int src_ncol, const int req_srow, const int req_scol,
                                const int req_nrow, const int req_ncol, MPI_Comm comm,
                                MPI_Datatype dtype, const size_t dt_size, dev_type_t dev_type,
                                mat_redist_engine_p *engine_, size_t *workbuf_bytes);

#ifdef __cplusplus
}
#endif

#ifndef CA3DMM_MPI_BINDING_SOURCE
#define ca3dmm_engine_init     ca3dmm_engine_init_mpi
#define ca3dmm_engine_init_BTB ca3dmm_engine_init_BTB_mpi
#define cannon_engine_init     cannon_engine_init_mpi
#define mat_redist_engine_init mat_redist_engine_init_mpi
#endif

#endif
