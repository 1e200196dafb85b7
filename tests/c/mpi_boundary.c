/*
 * Link check of the MPI-typed constructors (include/ca3dmm_mpi.h): a caller
 * written against the reference's signatures (src/ca3dmm.h:114-124,
 * :149-157, src/cannon.h:69-72, src/mat_redist.h:69-74), built with
 * -DCA3DMM_MPI.  Plans only (workbuf_bytes given), so no GPU is needed.
 * Prints one line per constructor: "<name> <rank> <nproc> ok <bytes>" or
 * "<name> <rank> <nproc> null".
 */
#include <stdio.h>

#include "ca3dmm.h"

static void report(const char *name, const void *engine, int rank, int nproc, size_t bytes)
{
    if (engine) printf("%s %d %d ok %zu\n", name, rank, nproc, bytes);
    else printf("%s %d %d null\n", name, rank, nproc);
    fflush(stdout);
}

int main(int argc, char **argv)
{
    int rank, nproc;
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &rank);
    MPI_Comm_size(MPI_COMM_WORLD, &nproc);
    const int m = 300, n = 200, k = 500;
    size_t bytes = 0;

    ca3dmm_engine_p ce = NULL;
    ca3dmm_engine_init(m, n, k, 1, 0, 0, k, 0, m, 0, k, 0, n, 0, m, 0, n, NULL, MPI_COMM_WORLD, DEV_TYPE_HIP,
                       &ce, &bytes);
    report("ca3dmm_engine_init", ce, rank, nproc, bytes);
    if (ce) printf("grid %d %d %d rank %d\n", ce->mp, ce->np, ce->kp, ce->my_rank);
    ca3dmm_engine_free(&ce);

    ca3dmm_engine_p cb = NULL;
    ca3dmm_engine_init_BTB(n, k, 0, k, 0, n, 0, n, 0, n, NULL, MPI_COMM_WORLD, DEV_TYPE_HIP, &cb, &bytes);
    report("ca3dmm_engine_init_BTB", cb, rank, nproc, bytes);
    ca3dmm_engine_free(&cb);

    cannon_engine_p cn = NULL;
    cannon_engine_init(m, n, k, MPI_COMM_WORLD, DEV_TYPE_HIP, &cn, &bytes);
    report("cannon_engine_init", cn, rank, nproc, bytes);
    cannon_engine_free(&cn);

    mat_redist_engine_p rd = NULL;
    mat_redist_engine_init(0, 0, 40, 30, 10, 5, 20, 20, MPI_COMM_WORLD, MPI_DOUBLE, sizeof(double),
                           DEV_TYPE_HIP, &rd, &bytes);
    report("mat_redist_engine_init", rd, rank, nproc, bytes);
    if (rd) printf("dtype %d dt_size %zu send %d recv %d\n", rd->dtype, rd->dt_size, rd->send_cnt, rd->recv_cnt);
    mat_redist_engine_free(&rd);

    MPI_Finalize();
    return 0;
}
