/*
 * C := op(A) * op(B) through the CA3DMM engine on the HIP device type: the
 * command line, fill, timing report and validation of the reference's
 * examples/example_AB.c (CLI :19-41, fill :107-128, report :148-205,
 * validation :208-305), written against this library's headers.
 *
 *   example_AB m n k transA transB check_correct n_test dev_type [mp np kp]
 *
 * Built two ways from this one file:
 *   example_AB      plain C, the single-process group ca3dmm_comm_self();
 *   example_AB_mpi  -DCA3DMM_MPI with mpicc: MPI_Init / MPI_COMM_WORLD and
 *                   the reference's MPI-typed ca3dmm_engine_init
 *                   (include/ca3dmm_mpi.h, libca3dmm_amd_mpi.so); it runs as
 *                   one rank per process.
 * A, B and C use the reference's 1D column partition over the ranks.  The
 * check compares every entry of this rank's C block with the closed form of
 * the linear fill, sum_l (a i + b l)(c l + d j), evaluated in long double,
 * where the reference compares with a host cblas_dgemm of the same fill; an
 * entry with relative error above 1e-12 counts as an error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ca3dmm.h"
#include "hip_proxy.h"

#ifdef CA3DMM_MPI
#include <mpi.h>
#endif

static int int_arg(int argc, char **argv, int idx, int def, int lo, int hi)
{
    if (idx >= argc) return def;
    int v = atoi(argv[idx]);
    if (v < lo || v > hi) {
        printf("argument %d = %d is outside [%d, %d], using %d\n", idx, v, lo, hi, def);
        return def;
    }
    return v;
}

/* op(A)(i, l) = a i + b l and op(B)(l, j) = c l + d j for the reference's
 * fill A(r, s) = 0.19 r + 0.24 s, B(r, s) = 0.11 r + 0.12 s on stored indices. */
static long double closed_form(int i, int j, int K, int tA, int tB)
{
    const long double a = tA ? 0.24 : 0.19, b = tA ? 0.19 : 0.24;
    const long double c = tB ? 0.12 : 0.11, d = tB ? 0.11 : 0.12;
    const long double Kl = K, S1 = Kl * (Kl - 1) / 2, S2 = Kl * (Kl - 1) * (2 * Kl - 1) / 6;
    const long double ai = a * i, dj = d * j;
    return Kl * ai * dj + (ai * c + b * dj) * S1 + b * c * S2;
}

static void fill(double *X, int nrow, int ncol, int srow, int scol, double fr, double fc)
{
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nrow; i++)
            X[(size_t) j * nrow + i] = fr * (double) (i + srow) + fc * (double) (j + scol);
}

static void print_row(const char *name, const double *v, int n)
{
    printf("%s", name);
    for (int i = 0; i < n; i++) printf("%.0lf ", v[i]);
    printf("\n");
}

int main(int argc, char **argv)
{
    if (argc == 2 && (!strcmp(argv[1], "--help") || !strcmp(argv[1], "-h"))) {
        printf("Usage: %s m n k transA(0 or 1) transB(0 or 1) check_correct(0 or 1) n_test dev_type "
               "[mp np kp]\n", argv[0]);
        return 0;
    }
    int m = int_arg(argc, argv, 1, 4096, 1, 8388608);
    int n = int_arg(argc, argv, 2, 4096, 1, 8388608);
    int k = int_arg(argc, argv, 3, 4096, 1, 8388608);
    int trans_A = int_arg(argc, argv, 4, 0, 0, 1);
    int trans_B = int_arg(argc, argv, 5, 0, 0, 1);
    int chk_res = int_arg(argc, argv, 6, 1, 0, 1);
    int n_test = int_arg(argc, argv, 7, 10, 1, 100);
    /* host compute is not provided by this library: default to the HIP type */
    dev_type_t dev_type = (dev_type_t) int_arg(argc, argv, 8, DEV_TYPE_HIP, DEV_TYPE_HOST, DEV_TYPE_HIP_MPI_DIRECT);
    int grid[3], *proc_grid = NULL;
    if (argc >= 12) {
        grid[0] = atoi(argv[9]);
        grid[1] = atoi(argv[10]);
        grid[2] = atoi(argv[11]);
        proc_grid = grid;
    }
    if (dev_type != DEV_TYPE_HOST) select_hip_device_by_mpi_local_rank();

    int my_rank = 0, n_proc = 1;
#ifdef CA3DMM_MPI
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &my_rank);
    MPI_Comm_size(MPI_COMM_WORLD, &n_proc);
#endif
    if (my_rank == 0) {
        printf("Test problem size m * n * k : %d * %d * %d\n", m, n, k);
        printf("Transpose A / B             : %d / %d\n", trans_A, trans_B);
        printf("Number of tests             : %d\n", n_test);
        printf("Check result correctness    : %d\n", chk_res);
        printf("Device type                 : %d\n\n", dev_type);
        fflush(stdout);
    }

    const int A_nrow = trans_A ? k : m, A_ncol = trans_A ? m : k;
    const int B_nrow = trans_B ? n : k, B_ncol = trans_B ? k : n;
    /* 1D column partition of A, B and C */
    int A_scol, A_ncolb, B_scol, B_ncolb, C_scol, C_ncolb;
    calc_block_spos_size(A_ncol, n_proc, my_rank, &A_scol, &A_ncolb);
    calc_block_spos_size(B_ncol, n_proc, my_rank, &B_scol, &B_ncolb);
    calc_block_spos_size(n, n_proc, my_rank, &C_scol, &C_ncolb);
    const size_t A_bytes = sizeof(double) * (size_t) A_nrow * A_ncolb;
    const size_t B_bytes = sizeof(double) * (size_t) B_nrow * B_ncolb;
    const size_t C_bytes = sizeof(double) * (size_t) m * C_ncolb;
    double *A_h = (double *) malloc(A_bytes ? A_bytes : 1);
    double *B_h = (double *) malloc(B_bytes ? B_bytes : 1);
    double *A_d = (double *) dev_type_malloc(A_bytes, dev_type);
    double *B_d = (double *) dev_type_malloc(B_bytes, dev_type);
    double *C_d = (double *) dev_type_malloc(C_bytes, dev_type);
    if (!A_h || !B_h || (A_bytes && !A_d) || (B_bytes && !B_d) || (C_bytes && !C_d)) {
        fprintf(stderr, "allocation failed\n");
        return 1;
    }
    fill(A_h, A_nrow, A_ncolb, 0, A_scol, 0.19, 0.24);
    fill(B_h, B_nrow, B_ncolb, 0, B_scol, 0.11, 0.12);
    dev_type_memcpy(A_d, A_h, A_bytes, dev_type, DEV_TYPE_HOST);
    dev_type_memcpy(B_d, B_h, B_bytes, dev_type, DEV_TYPE_HOST);

    ca3dmm_engine_p ce = NULL;
    size_t ce_workbuf_bytes = 0;
    ca3dmm_engine_init(m, n, k, trans_A, trans_B, 0, A_nrow, A_scol, A_ncolb, 0, B_nrow, B_scol, B_ncolb,
                       0, m, C_scol, C_ncolb, proc_grid,
#ifdef CA3DMM_MPI
                       MPI_COMM_WORLD,
#else
                       ca3dmm_comm_self(),
#endif
                       dev_type, &ce, &ce_workbuf_bytes);
    if (ce == NULL) {
        fprintf(stderr, "ca3dmm_engine_init failed\n");
        return 1;
    }
    void *workbuf_h, *workbuf_d;
    MALLOC_ATTACH_WORKBUF(ca3dmm_engine_attach_workbuf, ca3dmm_engine_free, ce, dev_type, ce_workbuf_bytes,
                          workbuf_h, workbuf_d);
    if (ce == NULL) {
        fprintf(stderr, "work buffer allocation failed\n");
        return 1;
    }
    if (ce->my_rank == 0) {
        const int mb = (m + ce->mp - 1) / ce->mp, nb = (n + ce->np - 1) / ce->np, kb = (k + ce->kp - 1) / ce->kp;
        double lower = 3.0 * pow((double) m * n * k / n_proc, 2.0 / 3.0) * n_proc;
        double vol = ((double) mb * nb + (double) mb * kb + (double) nb * kb) * n_proc;
        printf("CA3DMM partition info:\n");
        printf("Process grid mp * np * kp  : %d * %d * %d\n", ce->mp, ce->np, ce->kp);
        printf("Work cuboid  mb * nb * kb  : %d * %d * %d\n", mb, nb, kb);
        printf("Process utilization        : %.2f %% \n", 100.0 * (1.0 - (double) ce->rp / n_proc));
        printf("Comm. volume / lower bound : %.2f\n", vol / lower);
        printf("Rank 0 work buffer size    : %.2f MBytes\n\n", (double) ce_workbuf_bytes / 1048576.0);
        fflush(stdout);
    }

    /* one warm-up, then n_test timed executions */
    ca3dmm_engine_exec(ce, A_d, A_nrow, B_d, B_nrow, C_d, m);
    ca3dmm_engine_reset_stat(ce);
    double *t = (double *) calloc(6 * (size_t) n_test, sizeof(double));
    double prev[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < n_test; i++) {
#ifdef CA3DMM_MPI
        MPI_Barrier(MPI_COMM_WORLD);
#endif
        ca3dmm_engine_exec(ce, A_d, A_nrow, B_d, B_nrow, C_d, m);
        const double now[5] = {ce->redist_ms, ce->agvAB_ms, ce->cannon_ms, ce->reduce_ms, ce->exec_ms};
        for (int p = 0; p < 5; p++) t[(size_t) p * n_test + i] = now[p] - prev[p];
        t[(size_t) 5 * n_test + i] = t[(size_t) 1 * n_test + i] + t[(size_t) 2 * n_test + i] +
                                     t[(size_t) 3 * n_test + i];
        memcpy(prev, now, sizeof(prev));
    }
    if (my_rank == 0) {
        printf("\n");
        print_row("A, B, C redist   : ", t, n_test);
        print_row("A / B allgather  : ", t + n_test, n_test);
        print_row("2D Cannon        : ", t + 2 * (size_t) n_test, n_test);
        print_row("C reduce-scatter : ", t + 3 * (size_t) n_test, n_test);
        print_row("matmul only      : ", t + 5 * (size_t) n_test, n_test);
        print_row("total execution  : ", t + 4 * (size_t) n_test, n_test);
        printf("\n");
        ca3dmm_engine_print_stat(ce);
    }
    free(t);

    int status = 0;
    if (chk_res) {
        double *C_h = (double *) malloc(C_bytes ? C_bytes : 1);
        dev_type_memcpy(C_h, C_d, C_bytes, DEV_TYPE_HOST, dev_type);
        int local_error = 0, total_error = 0;
        for (int j = 0; j < C_ncolb; j++)
            for (int i = 0; i < m; i++) {
                const long double want = closed_form(i, C_scol + j, k, trans_A, trans_B);
                const double relerr = (double) fabsl(((long double) C_h[(size_t) j * m + i] - want) / want);
                if (relerr > 1e-12) local_error++;   /* NaN (0 / 0 at want == 0) is not counted */
            }
        total_error = local_error;
#ifdef CA3DMM_MPI
        MPI_Reduce(&local_error, &total_error, 1, MPI_INT, MPI_SUM, 0, MPI_COMM_WORLD);
#endif
        if (my_rank == 0) printf("CA3DMM output : %d error(s)\n", total_error);
        status = total_error != 0;
        free(C_h);
    }

    free(A_h);
    free(B_h);
    dev_type_free(A_d, dev_type);
    dev_type_free(B_d, dev_type);
    dev_type_free(C_d, dev_type);
    dev_type_free(workbuf_d, dev_type);
    ca3dmm_engine_free(&ce);
#ifdef CA3DMM_MPI
    MPI_Finalize();
#endif
    return status;
}
