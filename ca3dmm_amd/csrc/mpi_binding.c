/*
 * libca3dmm_amd_mpi.so: the reference's MPI-typed engine constructors
 * (src/ca3dmm.h:114-124, :149-157, src/cannon.h:69-72,
 * src/mat_redist.h:69-74) over libca3dmm_amd's ca3dmm_comm_t entries.
 * Built with the caller's mpicc; declared in include/ca3dmm_mpi.h.
 */
#define CA3DMM_MPI_BINDING_SOURCE
#include <stdio.h>

#include "ca3dmm_mpi.h"

ca3dmm_comm_t ca3dmm_comm_from_mpi(MPI_Comm comm)
{
    ca3dmm_comm_t c = {0, 0};
    int inited = 0;
    MPI_Initialized(&inited);
    if (!inited || comm == MPI_COMM_NULL) return c;
    MPI_Comm_size(comm, &c.nproc);
    MPI_Comm_rank(comm, &c.rank);
    return c;
}

int ca3dmm_dtype_from_mpi(MPI_Datatype dtype, size_t dt_size)
{
    if (dtype == MPI_DOUBLE) return CA3DMM_DTYPE_DOUBLE;
    if (dtype == MPI_FLOAT) return CA3DMM_DTYPE_FLOAT;
    if (dtype == MPI_INT) return CA3DMM_DTYPE_INT;
    return dt_size == 4 ? CA3DMM_DTYPE_FLOAT : CA3DMM_DTYPE_DOUBLE;
}

/* The comm of a one-GPU engine: one rank.  Anything else is refused with
 * the reference's convention (NULL engine, message on stderr). */
static int single_rank(MPI_Comm comm, const char *who, ca3dmm_comm_t *out)
{
    *out = ca3dmm_comm_from_mpi(comm);
    if (out->nproc == 1) return 1;
    if (out->nproc == 0)
        fprintf(stderr, "[ca3dmm_amd] %s: MPI is not initialised or the communicator is MPI_COMM_NULL\n", who);
    else
        fprintf(stderr,
                "[ca3dmm_amd] %s: the communicator has %d ranks; the HIP engine runs one rank per process and "
                "builds no inter-process exchange (a p-rank grid on one GPU: ca3dmm_sim_init)\n",
                who, out->nproc);
    return 0;
}

void ca3dmm_engine_init_mpi(const int m, const int n, const int k, const int trans_A, const int trans_B,
                            const int src_A_srow, const int src_A_nrow, const int src_A_scol,
                            const int src_A_ncol, const int src_B_srow, const int src_B_nrow,
                            const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                            const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                            const int *proc_grid, MPI_Comm comm, dev_type_t dev_type,
                            ca3dmm_engine_p *engine_, size_t *workbuf_bytes)
{
    ca3dmm_comm_t c;
    *engine_ = NULL;
    if (!single_rank(comm, "ca3dmm_engine_init", &c)) return;
    ca3dmm_engine_init(m, n, k, trans_A, trans_B, src_A_srow, src_A_nrow, src_A_scol, src_A_ncol, src_B_srow,
                       src_B_nrow, src_B_scol, src_B_ncol, dst_C_srow, dst_C_nrow, dst_C_scol, dst_C_ncol,
                       proc_grid, c, dev_type, engine_, workbuf_bytes);
}

void ca3dmm_engine_init_BTB_mpi(const int n, const int k, const int src_B_srow, const int src_B_nrow,
                                const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                                const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                                const int *proc_grid, MPI_Comm comm, dev_type_t dev_type,
                                ca3dmm_engine_p *engine_, size_t *workbuf_bytes)
{
    ca3dmm_comm_t c;
    *engine_ = NULL;
    if (!single_rank(comm, "ca3dmm_engine_init_BTB", &c)) return;
    ca3dmm_engine_init_BTB(n, k, src_B_srow, src_B_nrow, src_B_scol, src_B_ncol, dst_C_srow, dst_C_nrow,
                           dst_C_scol, dst_C_ncol, proc_grid, c, dev_type, engine_, workbuf_bytes);
}

void cannon_engine_init_mpi(const int m, const int n, const int k, MPI_Comm comm, dev_type_t dev_type,
                            cannon_engine_p *engine_, size_t *workbuf_bytes)
{
    ca3dmm_comm_t c;
    *engine_ = NULL;
    if (!single_rank(comm, "cannon_engine_init", &c)) return;
    cannon_engine_init(m, n, k, c, dev_type, engine_, workbuf_bytes);
}

void mat_redist_engine_init_mpi(const int src_srow, const int src_scol, const int src_nrow,
                                const int src_ncol, const int req_srow, const int req_scol,
                                const int req_nrow, const int req_ncol, MPI_Comm comm,
                                MPI_Datatype dtype, const size_t dt_size, dev_type_t dev_type,
                                mat_redist_engine_p *engine_, size_t *workbuf_bytes)
{
    ca3dmm_comm_t c;
    *engine_ = NULL;
    if (!single_rank(comm, "mat_redist_engine_init", &c)) return;
    mat_redist_engine_init(src_srow, src_scol, src_nrow, src_ncol, req_srow, req_scol, req_nrow, req_ncol, c,
                           ca3dmm_dtype_from_mpi(dtype, dt_size), dt_size, dev_type, engine_, workbuf_bytes);
}
