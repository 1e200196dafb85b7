/*
 * CPU oracle for the CA3DMM local compute path -- see oracle.h.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference's
 * algorithm, used by tests/ and bench.py's cpu_baseline leg as the checker.
 * It is never part of the product path.
 *
 * Pinning status of each function (how a plausible mistake would be caught):
 *   or_block_spos_size   - compiled reference (oracle/_ref) sweep + closed form
 *   or_grid_cannon/_nk   - compiled reference (oracle/_ref) sweep, committed
 *                          grid fixture, README known answer (8000^3, p=24)
 *   or_copy_block        - compiled reference copy_matrix_block on random data
 *   or_transpose_cm      - compiled reference transpose_cm_mat + involution
 *   or_op_sum / fold     - IEEE identities (kp=2 commutes, integers exact)
 *   or_redist_*          - index identity dst(i,j) == global(i,j) on every
 *                          rank, brute force on tiny grids
 *   or_dgemm             - integer-valued inputs (exact), linear-fill closed form
 *   or_dd_sample         - exact rational check on integer inputs, error-free
 *                          transformation identities
 *   or_cannon/ca3dmm plan- README known answers (grid 4x2x3, cuboid
 *                          2000x4000x2667, rank-0 workspace 256,149,360 B)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ------------------------------------------------------------------ */
/* Block partition arithmetic: src/utils.c:26-49                        */
/* ------------------------------------------------------------------ */
void or_block_spos_size(int len, int nblk, int iblk, int *spos, int *size)
{
    if (iblk < 0 || iblk > nblk) { *spos = -1; *size = 0; return; }
    int small = len / nblk;        /* size of the trailing blocks        */
    int nbig  = len % nblk;        /* the first nbig blocks get one more */
    if (iblk < nbig) {
        *spos = iblk * (small + 1);
        *size = small + 1;
    } else {
        *spos = nbig + iblk * small;
        *size = small;
    }
}

/* ------------------------------------------------------------------ */
/* Grid search: src/partition.c                                         */
/* The floating-point expressions keep the reference's evaluation order */
/* so that every comparison takes the same branch.                      */
/* ------------------------------------------------------------------ */
static double spread3(const double *v)      /* partition.c:9-19 */
{
    double hi = v[0], lo = v[0];
    for (int i = 1; i < 3; i++) {
        if (v[i] > hi) hi = v[i];
        if (v[i] < lo) lo = v[i];
    }
    return hi / lo;
}

/* partition.c:21-96: search (pg[a], pg[b] = ratio * pg[a], rest) around
 * the ideal real-valued grid, minimising cuboid volume then surface sum. */
static void scan_pair_grid(int p, int a, int b, const double *dims,
                           const double *ideal, int *pg)
{
    int c = 3 - a - b;
    double ratio_ab = ideal[b] / ideal[a];
    double pd = (double) p;
    double ideal_vol = 1.0;
    for (int i = 0; i < 3; i++) ideal_vol *= dims[i] / ideal[i];

    pg[a] = (int) floor(ideal[a]);
    pg[b] = (int) floor(ratio_ab) * pg[a];
    pg[c] = (int) floor(pd / (double) (pg[a] * pg[b]));
    double side[3], best_vol = 1.0;
    for (int i = 0; i < 3; i++) { side[i] = dims[i] / (double) pg[i]; best_vol *= side[i]; }
    double best_surf = side[0] * side[1] + side[0] * side[2] + side[1] * side[2];
    best_surf = best_surf * (double) pg[0] * (double) pg[1] * (double) pg[2];

    int a_lo = (int) ceil(ideal[a] * 0.3), a_hi = (int) floor(ideal[a] * 3.3);
    int r_lo = (int) ceil(ratio_ab * 0.3), r_hi = (int) floor(ratio_ab * 3.3);
    for (int ga = a_lo; ga <= a_hi; ga++) {
        for (int r = r_lo; r <= r_hi; r++) {
            int t[3];
            t[a] = ga;
            t[b] = r * ga;
            t[c] = (int) floor(pd / (double) (t[a] * t[b]));
            double s[3], vol = 1.0, cnt = 1.0;
            for (int i = 0; i < 3; i++) {
                s[i] = dims[i] / (double) t[i];
                vol *= s[i];
                cnt *= (double) t[i];
            }
            double surf = s[0] * s[1] + s[0] * s[2] + s[1] * s[2];
            double surfsum = cnt * surf;
            if (t[0] < 1 || t[1] < 1 || t[2] < 1) continue;
            if (vol / ideal_vol > 1.25) continue;
            if (spread3(s) >= 4.0) continue;
            if (vol < best_vol) {
                best_vol = vol; best_surf = surfsum;
                pg[0] = t[0]; pg[1] = t[1]; pg[2] = t[2];
            }
            if (vol < best_vol * 1.05 && surfsum < best_surf) {
                best_vol = vol; best_surf = surfsum;
                pg[0] = t[0]; pg[1] = t[1]; pg[2] = t[2];
            }
        }
    }
}

void or_grid_cannon(int p, int m, int n, int k, int *mp, int *np, int *kp, int *rp)
{
    /* partition.c:103-138: order the dimensions ascending.  The reference
     * tests the six orderings in sequence and keeps the LAST one that
     * matches, which decides ties; pos[d] is where dimension d (m,n,k) lands. */
    static const int perms[6][3] = {
        {0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}
    };
    const int len[3] = {m, n, k};
    double dims[3] = {0, 0, 0};
    int pos[3] = {0, 1, 2};
    for (int q = 0; q < 6; q++) {
        const int *o = perms[q];   /* o[0] <= o[1] <= o[2] must hold */
        if (len[o[0]] <= len[o[1]] && len[o[1]] <= len[o[2]]) {
            for (int i = 0; i < 3; i++) { dims[i] = len[o[i]]; pos[o[i]] = i; }
        }
    }
    int pg[3] = {1, 1, 1};
    double pd = p, d0 = dims[0], d1 = dims[1], d2 = dims[2];

    double r21 = d2 / d1;                      /* one large dimension */
    if (r21 > pd) pg[2] = p;

    double two_large = d1 * d2 / (d0 * d0);    /* two large dimensions */
    if (r21 <= pd && pd <= two_large) {
        double best = 1.0;
        int lo = (int) ceil(r21 * 0.5), hi = (int) floor(r21 * 2);
        for (int r = lo; r <= hi; r++) {
            int g1 = (int) floor(sqrt(pd / (double) r));
            int g2 = ((double) k == d0) ? (g1 * r) : (int) floor(pd / (double) g1);
            double prod = (double) g1 * g2;
            if (prod > best) { best = prod; pg[1] = g1; pg[2] = g2; }
        }
    }

    if (pd > two_large) {                      /* three large dimensions */
        double r20 = d2 / d0, r10 = d1 / d0;
        double g0 = cbrt(pd / (r20 * r10));
        double ideal[3] = {g0, g0 * r10, g0 * r20};
        /* the pair that must keep max % min == 0 is (m, n): the two
         * sorted slots that are not k's */
        if (pos[2] == 0) scan_pair_grid(p, 1, 2, dims, ideal, pg);
        if (pos[2] == 1) scan_pair_grid(p, 0, 2, dims, ideal, pg);
        if (pos[2] == 2) scan_pair_grid(p, 0, 1, dims, ideal, pg);
    }
    *mp = pg[pos[0]];
    *np = pg[pos[1]];
    *kp = pg[pos[2]];
    *rp = p - (*mp) * (*np) * (*kp);
}

void or_grid_nk(int p, int n, int k, int *np, int *kp, int *rp)
{
    /* partition.c:196-232: start from the cube-root guess and shrink the
     * square mn-grid while it lowers the idle count, until the k/n aspect
     * of the per-rank cuboid exceeds 4x the global one. */
    double aspect = (double) k / (double) n;
    int g = (int) ceil(cbrt((double) p / aspect));
    int gk = p / (g * g);
    *np = g; *kp = gk; *rp = p - g * g * gk;
    while (g > 1) {
        g--;
        gk = p / (g * g);
        int idle = p - g * g * gk;
        if ((double) gk / (double) g >= 4.0 * aspect) break;
        if (idle < *rp) { *np = g; *kp = gk; *rp = idle; }
    }
}

/* ------------------------------------------------------------------ */
/* Data movement                                                        */
/* ------------------------------------------------------------------ */
void or_copy_block(size_t dt, int nrow, int ncol, const void *src, int lds, void *dst, int ldd)
{   /* utils.c:92-120 */
    const char *s = (const char *) src;
    char *d = (char *) dst;
    for (int r = 0; r < nrow; r++)
        memcpy(d + dt * (size_t) r * (size_t) ldd, s + dt * (size_t) r * (size_t) lds, dt * (size_t) ncol);
}

void or_transpose_cm(int nrow, int ncol, const double *A, int ldA, double *AT, int ldAT)
{   /* utils.c:181-196: AT(i,j) = A(j,i) */
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nrow; i++)
            AT[(size_t) j * ldAT + i] = A[(size_t) i * ldA + j];
}

void or_op_sum(int dtype, int len, const void *in, void *inout)
{   /* mpi_op_omp.c:6-23 */
    if (dtype == 0) {
        const double *a = (const double *) in; double *b = (double *) inout;
        for (int i = 0; i < len; i++) b[i] += a[i];
    } else if (dtype == 1) {
        const float *a = (const float *) in; float *b = (float *) inout;
        for (int i = 0; i < len; i++) b[i] += a[i];
    } else if (dtype == 2) {
        const int *a = (const int *) in; int *b = (int *) inout;
        for (int i = 0; i < len; i++) b[i] += a[i];
    }
}

void or_fold_sum(int nsrc, const double *const *P, const int *ldp, int nrow, int ncol,
                 double *dst, int ldd)
{
    for (int j = 0; j < ncol; j++)
        for (int i = 0; i < nrow; i++) {
            double s = P[0][(size_t) j * ldp[0] + i];
            for (int t = 1; t < nsrc; t++) s += P[t][(size_t) j * ldp[t] + i];
            dst[(size_t) j * ldd + i] = s;
        }
}

/* ------------------------------------------------------------------ */
/* Redistribution: src/mat_redist.c                                     */
/* ------------------------------------------------------------------ */
/* Closed-interval overlap [s0,e0] with [s1,e1] (mat_redist.c:9-30). */
static int seg_overlap(int s0, int e0, int s1, int e1, int *is, int *ie)
{
    if (s0 > s1) { int t = s0; s0 = s1; s1 = t; t = e0; e0 = e1; e1 = t; }
    if (s1 > e0 || s1 > e1 || s0 > e0) { *is = -1; *ie = -1; return 0; }
    *is = s1;
    *ie = (e0 < e1) ? e0 : e1;
    return 1;
}

static int rect_overlap(const int *a, const int *b, int *out)
{   /* a, b = {srow, scol, nrow, ncol}; out = same form */
    int rs, re, cs, ce;
    if (!seg_overlap(a[0], a[0] + a[2] - 1, b[0], b[0] + b[2] - 1, &rs, &re)) return 0;
    if (!seg_overlap(a[1], a[1] + a[3] - 1, b[1], b[1] + b[3] - 1, &cs, &ce)) return 0;
    out[0] = rs; out[1] = cs; out[2] = re - rs + 1; out[3] = ce - cs + 1;
    return 1;
}

void or_redist_plan(int nproc, const int *all_src, const int *all_req, int rank,
                    int *send_list, int *n_send, int *send_cnt,
                    int *recv_list, int *n_recv, int *recv_cnt)
{
    const int *mysrc = all_src + 4 * rank, *myreq = all_req + 4 * rank;
    int ns = 0, nr = 0, cs = 0, cr = 0, ov[4];
    for (int peer = 0; peer < nproc; peer++) {
        if (rect_overlap(mysrc, all_req + 4 * peer, ov)) {
            int *e = send_list + 6 * ns++;
            e[0] = ov[0]; e[1] = ov[1]; e[2] = ov[2]; e[3] = ov[3]; e[4] = peer; e[5] = cs;
            cs += ov[2] * ov[3];
        }
    }
    for (int peer = 0; peer < nproc; peer++) {
        if (rect_overlap(myreq, all_src + 4 * peer, ov)) {
            int *e = recv_list + 6 * nr++;
            e[0] = ov[0]; e[1] = ov[1]; e[2] = ov[2]; e[3] = ov[3]; e[4] = peer; e[5] = cr;
            cr += ov[2] * ov[3];
        }
    }
    *n_send = ns; *send_cnt = cs; *n_recv = nr; *recv_cnt = cr;
}

void or_redist_exec_all(int nproc, size_t dt, const int *all_src, const int *all_req,
                        const void *const *src_blk, const int *src_ld,
                        void *const *dst_blk, const int *dst_ld)
{
    int *sl = (int *) malloc(sizeof(int) * 6 * nproc * nproc);
    int *rl = (int *) malloc(sizeof(int) * 6 * nproc * nproc);
    int *ns = (int *) malloc(sizeof(int) * nproc), *nr = (int *) malloc(sizeof(int) * nproc);
    int *sc = (int *) malloc(sizeof(int) * nproc), *rc = (int *) malloc(sizeof(int) * nproc);
    char **sbuf = (char **) malloc(sizeof(char *) * nproc);
    char **rbuf = (char **) malloc(sizeof(char *) * nproc);
    for (int r = 0; r < nproc; r++) {
        or_redist_plan(nproc, all_src, all_req, r, sl + 6 * nproc * r, &ns[r], &sc[r],
                       rl + 6 * nproc * r, &nr[r], &rc[r]);
        sbuf[r] = (char *) malloc(dt * (size_t) sc[r] + 1);
        rbuf[r] = (char *) malloc(dt * (size_t) rc[r] + 1);
        /* pack (mat_redist.c:317-348) */
        for (int i = 0; i < ns[r]; i++) {
            const int *e = sl + 6 * nproc * r + 6 * i;
            int lr = e[0] - all_src[4 * r], lc = e[1] - all_src[4 * r + 1];
            const char *s = (const char *) src_blk[r] + dt * ((size_t) lr * src_ld[r] + lc);
            or_copy_block(dt, e[2], e[3], s, src_ld[r], sbuf[r] + dt * (size_t) e[5], e[3]);
        }
    }
    /* exchange: rank r's recv entry from peer s is s's send entry to r */
    for (int r = 0; r < nproc; r++) {
        for (int i = 0; i < nr[r]; i++) {
            const int *e = rl + 6 * nproc * r + 6 * i;
            int s = e[4];
            for (int j = 0; j < ns[s]; j++) {
                const int *f = sl + 6 * nproc * s + 6 * j;
                if (f[4] == r) {
                    memcpy(rbuf[r] + dt * (size_t) e[5], sbuf[s] + dt * (size_t) f[5],
                           dt * (size_t) f[2] * (size_t) f[3]);
                    break;
                }
            }
        }
    }
    /* unpack (mat_redist.c:389-416) */
    for (int r = 0; r < nproc; r++) {
        for (int i = 0; i < nr[r]; i++) {
            const int *e = rl + 6 * nproc * r + 6 * i;
            int lr = e[0] - all_req[4 * r], lc = e[1] - all_req[4 * r + 1];
            char *d = (char *) dst_blk[r] + dt * ((size_t) lr * dst_ld[r] + lc);
            or_copy_block(dt, e[2], e[3], rbuf[r] + dt * (size_t) e[5], e[3], d, dst_ld[r]);
        }
        free(sbuf[r]); free(rbuf[r]);
    }
    free(sl); free(rl); free(ns); free(nr); free(sc); free(rc); free(sbuf); free(rbuf);
}

/* ------------------------------------------------------------------ */
/* Reference products                                                   */
/* ------------------------------------------------------------------ */
static inline double opA(int t, const double *A, int lda, int i, int l)
{ return t ? A[(size_t) i * lda + l] : A[(size_t) l * lda + i]; }
static inline double opB(int t, const double *B, int ldb, int l, int j)
{ return t ? B[(size_t) l * ldb + j] : B[(size_t) j * ldb + l]; }

void or_dgemm(int transA, int transB, int m, int n, int k, double alpha,
              const double *A, int lda, const double *B, int ldb, double beta,
              double *C, int ldc)
{
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double s = 0.0;
            for (int l = 0; l < k; l++) {
                double prod = opA(transA, A, lda, i, l) * opB(transB, B, ldb, l, j);
                s = s + prod;
            }
            double *c = &C[(size_t) j * ldc + i];
            *c = (beta == 0.0) ? alpha * s : alpha * s + beta * (*c);
        }
}

/* Error-free transformations (Knuth TwoSum, FMA-based TwoProd). */
static inline void two_sum(double a, double b, double *s, double *e)
{
    double x = a + b, z = x - a;
    *e = (a - (x - z)) + (b - z);
    *s = x;
}

void or_dd_sample(int transA, int transB, int m, int n, int k,
                  const double *A, int lda, const double *B, int ldb,
                  int nsamp, const int *ii, const int *jj, double *ref, double *absab)
{
    (void) m; (void) n;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int s = 0; s < nsamp; s++) {
        double hi = 0.0, lo = 0.0, ab = 0.0;
        for (int l = 0; l < k; l++) {
            double a = opA(transA, A, lda, ii[s], l), b = opB(transB, B, ldb, l, jj[s]);
            double p = a * b, pe = fma(a, b, -p);
            double t, te;
            two_sum(hi, p, &t, &te);
            lo += te + pe;
            two_sum(t, lo, &hi, &lo);
            ab += fabs(a) * fabs(b);
        }
        ref[s] = hi + lo;
        absab[s] = ab;
    }
}

/* ------------------------------------------------------------------ */
/* Cannon and CA3DMM planning arithmetic                                */
/* ------------------------------------------------------------------ */
void or_cannon_plan(int m, int n, int k, int q, int rank2d, int min_k_blk, or_cannon_plan_t *pl)
{   /* cannon.c:16-133 */
    memset(pl, 0, sizeof(*pl));
    pl->m = m; pl->n = n; pl->k = k; pl->np_dim = q;
    pl->rank_row = rank2d / q;        /* MPI_Cart_coords, row-major, no reorder */
    pl->rank_col = rank2d % q;
    or_block_spos_size(m, q, pl->rank_row, &pl->A_srow, &pl->A_nrow);
    or_block_spos_size(k, q, pl->rank_col, &pl->A_scol, &pl->A_ncol);
    or_block_spos_size(k, q, pl->rank_row, &pl->B_srow, &pl->B_nrow);
    or_block_spos_size(n, q, pl->rank_col, &pl->B_scol, &pl->B_ncol);
    or_block_spos_size(m, q, pl->rank_row, &pl->C_srow, &pl->C_nrow);
    or_block_spos_size(n, q, pl->rank_col, &pl->C_scol, &pl->C_ncol);
    pl->max_A_blk_size = (m / q + 1) * (k / q + 1);
    pl->max_B_blk_size = (k / q + 1) * (n / q + 1);
    pl->max_C_blk_size = (m / q + 1) * (n / q + 1);
    long long wb = 8LL * pl->max_A_blk_size + 8LL * pl->max_B_blk_size;
    int gc = 1;
    if (pl->A_ncol > 0 && pl->A_ncol < min_k_blk) {
        gc = (min_k_blk + pl->A_ncol - 1) / pl->A_ncol;
        if (gc > q) gc = q;
        wb += 8LL * pl->max_A_blk_size * gc + 8LL * pl->max_B_blk_size * gc;
    }
    pl->gemm_cycle = gc;
    pl->workbuf_bytes = wb;
}

/* ca3dmm.c:75-124: fold a factor 3 (else 2) of kp into the smaller of mp, np
 * when the Cannon divisibility still holds. */
static void shrink_kp(int *mp, int *np, int *kp)
{
    int f = 0;
    if (*kp >= 3 && *kp % 3 == 0) f = 3;
    else if (*kp >= 2 && *kp % 2 == 0) f = 2;
    if (!f) return;
    int *grow = (*mp <= *np) ? mp : np;
    int other = (*mp <= *np) ? *np : *mp;
    int g = *grow * f;
    int hi = (g > other) ? g : other, lo = (g < other) ? g : other;
    if (hi % lo == 0) { *kp /= f; *grow = g; }
}

int or_ca3dmm_plan(int m, int n, int k, int transA, int transB, int p, int rank,
                   const int *proc_grid, int reduce_kp, int is_btb, int min_k_blk,
                   or_ca3dmm_plan_t *pl)
{
    memset(pl, 0, sizeof(*pl));
    if (is_btb) m = n;
    int mp = 0, np = 0, kp = 0, rp = 0, auto_grid = 1;
    if (proc_grid) {
        mp = proc_grid[0]; np = proc_grid[1]; kp = proc_grid[2];
        rp = p - mp * np * kp;
        int hi = mp > np ? mp : np, lo = mp < np ? mp : np;
        int ok = is_btb ? (mp == np && rp >= 0) : (lo > 0 && hi % lo == 0 && rp >= 0);
        if (ok) auto_grid = 0;
    }
    if (auto_grid) {
        if (is_btb) {
            or_grid_nk(p, n, k, &np, &kp, &rp);
            if (reduce_kp && kp >= 4 && kp % 4 == 0) { kp /= 4; np *= 2; }
        } else {
            or_grid_cannon(p, m, n, k, &mp, &np, &kp, &rp);
            if (reduce_kp) shrink_kp(&mp, &np, &kp);
        }
    }
    if (is_btb) mp = np;
    if (mp < 1 || np < 1 || kp < 1 || mp * np * kp > p) return -1;
    pl->mp = mp; pl->np = np; pl->kp = kp; pl->rp = rp;

    /* ca3dmm.c:146-185: coordinates and task groups */
    int active = rank < mp * np * kp;
    pl->is_active = active;
    if (!active) return 0;
    int rank_k = rank / (mp * np), rank_mn = rank % (mp * np);
    pl->rank_k = rank_k; pl->rank_n = rank_mn / mp; pl->rank_m = rank_mn % mp;
    int tmn = 1, tmi = 0, tnn = 1, tni = 0, group = mp * np;
    if (!is_btb) {
        if (mp > np) { tmn = mp / np; tmi = rank_mn / (np * np); group = np * np; }
        if (np > mp) { tnn = np / mp; tni = rank_mn / (mp * mp); group = mp * mp; }
    }
    pl->task_m_num = tmn; pl->task_m_id = tmi;
    pl->task_n_num = tnn; pl->task_n_id = tni;
    pl->task_k_num = kp;  pl->task_k_id = rank_k;
    /* rank inside the Cannon sub-communicator: the split keys by global
     * rank and each task group is a contiguous run of rank_mn values */
    pl->rank_2dmm = rank_mn % group;

    /* ca3dmm.c:212-261: blocks of this rank */
    int msp, msz, nsp, nsz, ksp, ksz;
    or_block_spos_size(m, tmn, tmi, &msp, &msz);
    or_block_spos_size(n, tnn, tni, &nsp, &nsz);
    or_block_spos_size(k, kp, rank_k, &ksp, &ksz);
    int q = (int) floor(sqrt((double) group) + 0.5);
    or_cannon_plan(msz, nsz, ksz, q, pl->rank_2dmm, min_k_blk, &pl->cannon);
    const or_cannon_plan_t *c = &pl->cannon;
    pl->A_2dmm_srow = msp + c->A_srow; pl->A_2dmm_scol = ksp + c->A_scol;
    pl->A_2dmm_nrow = c->A_nrow;       pl->A_2dmm_ncol = c->A_ncol;
    pl->B_2dmm_srow = ksp + c->B_srow; pl->B_2dmm_scol = nsp + c->B_scol;
    pl->B_2dmm_nrow = c->B_nrow;       pl->B_2dmm_ncol = c->B_ncol;
    pl->C_2dmm_srow = msp + c->C_srow; pl->C_2dmm_scol = nsp + c->C_scol;
    pl->C_2dmm_nrow = c->C_nrow;       pl->C_2dmm_ncol = c->C_ncol;
    pl->C_out_srow = pl->C_2dmm_srow;  pl->C_out_nrow = pl->C_2dmm_nrow;
    int first = -1;
    pl->use_rsb = 1;
    for (int t = 0; t < kp; t++) {
        int sc, nc;
        or_block_spos_size(pl->C_2dmm_ncol, kp, t, &sc, &nc);
        if (first < 0) first = nc;
        if (nc != first) pl->use_rsb = 0;
        if (t == rank_k) { pl->C_out_scol = pl->C_2dmm_scol + sc; pl->C_out_ncol = nc; }
    }

    /* ca3dmm.c:298-356: the slice fetched in redistribution before the
     * allgather that replicates A (tnn > 1) or B (tmn > 1) */
    pl->A_rd_srow = pl->A_2dmm_srow; pl->A_rd_scol = pl->A_2dmm_scol;
    pl->A_rd_nrow = pl->A_2dmm_nrow; pl->A_rd_ncol = pl->A_2dmm_ncol;
    pl->B_rd_srow = pl->B_2dmm_srow; pl->B_rd_scol = pl->B_2dmm_scol;
    pl->B_rd_nrow = pl->B_2dmm_nrow; pl->B_rd_ncol = pl->B_2dmm_ncol;
    if (tnn > 1 || tmn > 1) {
        int ntask = (tnn > 1) ? tnn : tmn, id = (tnn > 1) ? tni : tmi;
        int full = (tnn > 1) ? pl->A_2dmm_ncol : pl->B_2dmm_ncol;
        int first_nc = -1;
        pl->use_ag = 1;
        for (int t = 0; t < ntask; t++) {
            int sc, nc;
            or_block_spos_size(full, ntask, t, &sc, &nc);
            if (first_nc < 0) first_nc = nc;
            if (nc != first_nc) pl->use_ag = 0;
            if (t == id) {
                if (tnn > 1) { pl->A_rd_scol = pl->A_2dmm_scol + sc; pl->A_rd_ncol = nc; }
                else         { pl->B_rd_scol = pl->B_2dmm_scol + sc; pl->B_rd_ncol = nc; }
            }
        }
    }

    /* ca3dmm.c:398-412 (general) and :680-690 (BTB): own buffers */
    long long Ab = 8LL * c->max_A_blk_size, Bb = 8LL * c->max_B_blk_size, self = 0;
    if (is_btb) {
        self = Ab + Ab + Bb;
    } else {
        self = Ab + Bb;
        if (transA) self += Ab;
        if (transB) self += Bb;
        if (tnn > 1) self += Ab;
        if (tmn > 1) self += Bb;
    }
    self += 8LL * pl->C_2dmm_nrow * pl->C_2dmm_ncol;
    self += 8LL * pl->C_out_nrow * pl->C_out_ncol;
    pl->self_workbuf_bytes = self;
    pl->cannon_workbuf_bytes = c->workbuf_bytes;
    return 0;
}
