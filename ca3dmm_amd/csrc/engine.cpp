// Host engines of libca3dmm_amd: grid choice, Cannon, redistribution and the
// CA3DMM engine, plus the simulated p-rank grid on one GPU.
//
// Reference counterparts: src/partition.c, src/utils.c:26-49,
// src/mat_redist.c, src/cannon.c, src/ca3dmm.c.  The planning arithmetic
// (grid, task groups, blocks, redistribution lists, workspace sizes) gives
// the same integers as the reference; the data path is rebuilt for one
// MI355X: every copy step is one batched launch, the Cannon rounds read
// their operands in place, the k-layer reduction is a fixed-order fold, and
// the single-rank exec with identity redistribution is a single DGEMM.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ca3dmm.h"
#include "ca3dmm_ext.h"
#include "ca3dmm_sim.h"
#include "grouped.h"
#include "hip_proxy.h"

#define ENG_ERR(...) fprintf(stderr, "[ca3dmm_amd] " __VA_ARGS__)

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// utils.h:71-83 semantics: unset -> default; out of range -> default.
int env_int(const char *name, int def, int lo, int hi)
{
    const char *s = getenv(name);
    if (!s) return def;
    int v = atoi(s);
    if (v < lo || v > hi) v = def;
    return v;
}

bool gpu_type(dev_type_t t) { return t == DEV_TYPE_HIP || t == DEV_TYPE_HIP_MPI_DIRECT; }

}  // namespace

// =====================================================================
// Partition arithmetic (partition.c, utils.c:26-49)
// =====================================================================
extern "C" {

ca3dmm_comm_t ca3dmm_comm_self(void)
{
    ca3dmm_comm_t c;
    c.nproc = 1;
    c.rank = 0;
    return c;
}

void calc_block_spos_size(const int len, const int nblk, const int iblk, int *blk_spos, int *blk_size)
{
    if (iblk < 0 || iblk > nblk) { *blk_spos = -1; *blk_size = 0; return; }
    const int base = len / nblk, extra = len % nblk;
    if (iblk < extra) { *blk_spos = iblk * (base + 1); *blk_size = base + 1; }
    else              { *blk_spos = extra + iblk * base; *blk_size = base; }
}

}  // extern "C"

namespace {

double ratio_of_extremes(const double *v)
{
    double hi = v[0], lo = v[0];
    for (int i = 1; i < 3; ++i) { hi = std::max(hi, v[i]); lo = std::min(lo, v[i]); }
    return hi / lo;
}

// partition.c:21-96.  Candidate grids keep g[b] a multiple of g[a].
void pair_scaled_search(int p, int a, int b, const double *dims, const double *ideal, int *g)
{
    const int c = 3 - a - b;
    const double ratio = ideal[b] / ideal[a], pd = (double) p;
    double ideal_vol = 1.0;
    for (int i = 0; i < 3; ++i) ideal_vol *= dims[i] / ideal[i];

    g[a] = (int) floor(ideal[a]);
    g[b] = (int) floor(ratio) * g[a];
    g[c] = (int) floor(pd / (double) (g[a] * g[b]));
    double side[3], vol_best = 1.0;
    for (int i = 0; i < 3; ++i) { side[i] = dims[i] / (double) g[i]; vol_best *= side[i]; }
    double surf_best = side[0] * side[1] + side[0] * side[2] + side[1] * side[2];
    surf_best = surf_best * (double) g[0] * (double) g[1] * (double) g[2];

    const int ga_lo = (int) ceil(ideal[a] * 0.3), ga_hi = (int) floor(ideal[a] * 3.3);
    const int r_lo = (int) ceil(ratio * 0.3), r_hi = (int) floor(ratio * 3.3);
    for (int ga = ga_lo; ga <= ga_hi; ++ga) {
        for (int r = r_lo; r <= r_hi; ++r) {
            int t[3];
            t[a] = ga;
            t[b] = r * ga;
            t[c] = (int) floor(pd / (double) (t[a] * t[b]));
            double s[3], vol = 1.0, count = 1.0;
            for (int i = 0; i < 3; ++i) {
                s[i] = dims[i] / (double) t[i];
                vol *= s[i];
                count *= (double) t[i];
            }
            const double surfsum = count * (s[0] * s[1] + s[0] * s[2] + s[1] * s[2]);
            if (t[0] < 1 || t[1] < 1 || t[2] < 1) continue;
            if (vol / ideal_vol > 1.25) continue;
            if (ratio_of_extremes(s) >= 4.0) continue;
            if (vol < vol_best) { vol_best = vol; surf_best = surfsum; std::copy(t, t + 3, g); }
            if (vol < vol_best * 1.05 && surfsum < surf_best) {
                vol_best = vol; surf_best = surfsum; std::copy(t, t + 3, g);
            }
        }
    }
}

}  // namespace

extern "C" {

void calc_3d_decomposition_cannon(const int p, const int m, const int n, const int k, int *mp,
                                  int *np, int *kp, int *rp)
{
    // Ascending order of (m, n, k); among the six orderings the last one that
    // fits wins, which fixes how ties are broken.
    static const int orders[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    const int len[3] = {m, n, k};
    double d[3] = {0, 0, 0};
    int slot[3] = {0, 1, 2};          // slot[dim] = position of dim in d
    for (const auto &o : orders) {
        if (len[o[0]] <= len[o[1]] && len[o[1]] <= len[o[2]])
            for (int i = 0; i < 3; ++i) { d[i] = len[o[i]]; slot[o[i]] = i; }
    }
    int g[3] = {1, 1, 1};
    const double pd = p;
    const double r21 = d[2] / d[1];
    if (r21 > pd) g[2] = p;                                  // one large dimension
    const double two_large = d[1] * d[2] / (d[0] * d[0]);
    if (r21 <= pd && pd <= two_large) {                     // two large dimensions
        double best = 1.0;
        for (int r = (int) ceil(r21 * 0.5); r <= (int) floor(r21 * 2); ++r) {
            const int g1 = (int) floor(sqrt(pd / (double) r));
            const int g2 = ((double) k == d[0]) ? (g1 * r) : (int) floor(pd / (double) g1);
            const double prod = (double) g1 * g2;
            if (prod > best) { best = prod; g[1] = g1; g[2] = g2; }
        }
    }
    if (pd > two_large) {                                   // three large dimensions
        const double r20 = d[2] / d[0], r10 = d[1] / d[0];
        const double g0 = cbrt(pd / (r20 * r10));
        const double ideal[3] = {g0, g0 * r10, g0 * r20};
        if (slot[2] == 0) pair_scaled_search(p, 1, 2, d, ideal, g);
        if (slot[2] == 1) pair_scaled_search(p, 0, 2, d, ideal, g);
        if (slot[2] == 2) pair_scaled_search(p, 0, 1, d, ideal, g);
    }
    *mp = g[slot[0]];
    *np = g[slot[1]];
    *kp = g[slot[2]];
    *rp = p - (*mp) * (*np) * (*kp);
    if (*rp * 10 > p)
        ENG_ERR("warning: poor grid for p=%d: %d x %d x %d + %d idle\n", p, *mp, *np, *kp, *rp);
}

void calc_3d_decomposition_nk(const int p, const int n, const int k, int *np, int *kp, int *rp)
{
    const double aspect = (double) k / (double) n;
    int g = (int) ceil(cbrt((double) p / aspect));
    int gk = p / (g * g);
    *np = g; *kp = gk; *rp = p - g * g * gk;
    while (g > 1) {
        --g;
        gk = p / (g * g);
        const int idle = p - g * g * gk;
        if ((double) gk / (double) g >= 4.0 * aspect) break;
        if (idle < *rp) { *np = g; *kp = gk; *rp = idle; }
    }
    if (*rp * 10 > p)
        ENG_ERR("warning: poor grid for p=%d: %d x %d x %d + %d idle\n", p, *np, *np, *kp, *rp);
}

}  // extern "C"

// =====================================================================
// Per-rank planning (ca3dmm.c:23-444, :447-722; cannon.c:16-133)
// =====================================================================
namespace {

struct Rect {     // column-major block of a matrix
    int srow, nrow, scol, ncol;
};

struct RankPlan {
    int active = 0, rank_m = 0, rank_n = 0, rank_k = 0;
    int tmn = 0, tmi = 0, tnn = 0, tni = 0, tkn = 0, tki = 0;
    int group = 1, q = 1, rank2d = 0;
    int msp = 0, msz = 0, nsp = 0, nsz = 0, ksp = 0, ksz = 0;
    Rect A_rd{}, B_rd{}, A_2dmm{}, B_2dmm{}, C_2dmm{}, C_out{};
    int use_ag = 0, use_rsb = 0;
    int cA_srow = 0, cA_scol = 0, cA_nrow = 0, cA_ncol = 0;   // Cannon-local blocks
    int cB_srow = 0, cB_scol = 0, cB_nrow = 0, cB_ncol = 0;
    int cC_srow = 0, cC_scol = 0, cC_nrow = 0, cC_ncol = 0;
};

struct Grid {
    int mp = 0, np = 0, kp = 0, rp = 0;
};

// ca3dmm.c:47-137 (general) and :470-510 (BTB).
int choose_grid(int p, int m, int n, int k, const int *proc_grid, bool btb, Grid *g)
{
    int mp = 0, np = 0, kp = 0, rp = 0;
    bool automatic = true;
    if (proc_grid) {
        mp = proc_grid[0]; np = proc_grid[1]; kp = proc_grid[2];
        rp = p - mp * np * kp;
        const int hi = std::max(mp, np), lo = std::min(mp, np);
        const bool ok = btb ? (mp == np && rp >= 0) : (lo > 0 && hi % lo == 0 && rp >= 0);
        if (ok) automatic = false;
        else ENG_ERR("warning: invalid process grid %d x %d x %d, choosing one\n", mp, np, kp);
    }
    if (automatic) {
        if (btb) {
            calc_3d_decomposition_nk(p, n, k, &np, &kp, &rp);
            if (env_int("CA3DMM_REDUCE_KP", 1, 0, 1) && kp >= 4 && kp % 4 == 0) { kp /= 4; np *= 2; }
        } else {
            calc_3d_decomposition_cannon(p, m, n, k, &mp, &np, &kp, &rp);
            if (env_int("CA3DMM_REDUCE_KP", 0, 0, 1)) {
                // fold a factor 3 (else 2) of kp into the smaller of mp, np
                const int f = (kp >= 3 && kp % 3 == 0) ? 3 : ((kp >= 2 && kp % 2 == 0) ? 2 : 0);
                if (f) {
                    int &grow = (mp <= np) ? mp : np;
                    const int other = (mp <= np) ? np : mp;
                    const int gnew = grow * f;
                    if (std::max(gnew, other) % std::min(gnew, other) == 0) { kp /= f; grow = gnew; }
                }
            }
        }
    }
    if (btb) mp = np;
    if (mp < 1 || np < 1 || kp < 1 || mp * np * kp > p) {
        ENG_ERR("invalid process grid: p = %d, mp, np, kp = %d, %d, %d\n", p, mp, np, kp);
        return -1;
    }
    g->mp = mp; g->np = np; g->kp = kp; g->rp = rp;
    return 0;
}

RankPlan plan_rank(const Grid &g, int m, int n, int k, bool btb, int rank)
{
    RankPlan r;
    const int mp = g.mp, np = g.np, kp = g.kp;
    r.active = rank < mp * np * kp;
    if (!r.active) return r;
    const int rank_mn = rank % (mp * np);
    r.rank_k = rank / (mp * np);
    r.rank_n = rank_mn / mp;
    r.rank_m = rank_mn % mp;
    r.tmn = 1; r.tnn = 1; r.tkn = kp; r.tki = r.rank_k;
    r.group = mp * np;
    if (!btb && mp > np) { r.tmn = mp / np; r.tmi = rank_mn / (np * np); r.group = np * np; }
    if (!btb && np > mp) { r.tnn = np / mp; r.tni = rank_mn / (mp * mp); r.group = mp * mp; }
    r.rank2d = rank_mn % r.group;
    r.q = (int) floor(sqrt((double) r.group) + 0.5);
    calc_block_spos_size(m, r.tmn, r.tmi, &r.msp, &r.msz);
    calc_block_spos_size(n, r.tnn, r.tni, &r.nsp, &r.nsz);
    calc_block_spos_size(k, kp, r.rank_k, &r.ksp, &r.ksz);
    const int row = r.rank2d / r.q, col = r.rank2d % r.q;
    calc_block_spos_size(r.msz, r.q, row, &r.cA_srow, &r.cA_nrow);
    calc_block_spos_size(r.ksz, r.q, col, &r.cA_scol, &r.cA_ncol);
    calc_block_spos_size(r.ksz, r.q, row, &r.cB_srow, &r.cB_nrow);
    calc_block_spos_size(r.nsz, r.q, col, &r.cB_scol, &r.cB_ncol);
    calc_block_spos_size(r.msz, r.q, row, &r.cC_srow, &r.cC_nrow);
    calc_block_spos_size(r.nsz, r.q, col, &r.cC_scol, &r.cC_ncol);
    r.A_2dmm = {r.msp + r.cA_srow, r.cA_nrow, r.ksp + r.cA_scol, r.cA_ncol};
    r.B_2dmm = {r.ksp + r.cB_srow, r.cB_nrow, r.nsp + r.cB_scol, r.cB_ncol};
    r.C_2dmm = {r.msp + r.cC_srow, r.cC_nrow, r.nsp + r.cC_scol, r.cC_ncol};
    r.C_out = {r.C_2dmm.srow, r.C_2dmm.nrow, 0, 0};
    r.use_rsb = 1;
    int first = -1;
    for (int t = 0; t < kp; ++t) {
        int sc, nc;
        calc_block_spos_size(r.C_2dmm.ncol, kp, t, &sc, &nc);
        if (first < 0) first = nc;
        if (nc != first) r.use_rsb = 0;
        if (t == r.rank_k) { r.C_out.scol = r.C_2dmm.scol + sc; r.C_out.ncol = nc; }
    }
    r.A_rd = r.A_2dmm;
    r.B_rd = r.B_2dmm;
    if (r.tnn > 1 || r.tmn > 1) {
        const bool repA = r.tnn > 1;
        const int ntask = repA ? r.tnn : r.tmn, id = repA ? r.tni : r.tmi;
        const int full = repA ? r.A_2dmm.ncol : r.B_2dmm.ncol;
        int f = -1;
        r.use_ag = 1;
        for (int t = 0; t < ntask; ++t) {
            int sc, nc;
            calc_block_spos_size(full, ntask, t, &sc, &nc);
            if (f < 0) f = nc;
            if (nc != f) r.use_ag = 0;
            if (t == id) {
                if (repA) { r.A_rd.scol = r.A_2dmm.scol + sc; r.A_rd.ncol = nc; }
                else      { r.B_rd.scol = r.B_2dmm.scol + sc; r.B_rd.ncol = nc; }
            }
        }
    }
    return r;
}

// mat_redist row-major rectangle {srow, scol, nrow, ncol} of a column-major
// block: rows and columns swap (ca3dmm.c:359-370).
void to_rm(const Rect &c, int *out) { out[0] = c.scol; out[1] = c.srow; out[2] = c.ncol; out[3] = c.nrow; }

// ---------------------------------------------------------------------
// mat_redist engine construction (mat_redist.c:9-236)
// ---------------------------------------------------------------------
bool seg_overlap(int s0, int e0, int s1, int e1, int *is, int *ie)
{
    if (s0 > s1) { std::swap(s0, s1); std::swap(e0, e1); }
    if (s1 > e0 || s1 > e1 || s0 > e0) { *is = *ie = -1; return false; }
    *is = s1;
    *ie = std::min(e0, e1);
    return true;
}

bool rect_overlap(const int *a, const int *b, int *o)
{
    int rs, re, cs, ce;
    if (!seg_overlap(a[0], a[0] + a[2] - 1, b[0], b[0] + b[2] - 1, &rs, &re)) return false;
    if (!seg_overlap(a[1], a[1] + a[3] - 1, b[1], b[1] + b[3] - 1, &cs, &ce)) return false;
    o[0] = rs; o[1] = cs; o[2] = re - rs + 1; o[3] = ce - cs + 1;
    return true;
}

mat_redist_engine_p redist_build(int nproc, int rank, const int *all_src, const int *all_req,
                                 int dtype, size_t dt_size, dev_type_t dev_type)
{
    auto *e = (mat_redist_engine_p) calloc(1, sizeof(mat_redist_engine_s));
    if (!e) return nullptr;
    e->comm.nproc = nproc;
    e->comm.rank = rank;
    e->dtype = dtype;
    e->dt_size = dt_size;
    e->nproc = nproc;
    e->rank = rank;
    const int *s = all_src + 4 * rank, *q = all_req + 4 * rank;
    e->src_srow = s[0]; e->src_scol = s[1]; e->src_nrow = s[2]; e->src_ncol = s[3];
    e->req_srow = q[0]; e->req_scol = q[1]; e->req_nrow = q[2]; e->req_ncol = q[3];
    e->dev_type = dev_type;
    std::vector<int> sl, rl;
    int o[4];
    for (int peer = 0; peer < nproc; ++peer) {
        if (rect_overlap(s, all_req + 4 * peer, o)) sl.insert(sl.end(), {o[0], o[1], o[2], o[3], peer});
        if (rect_overlap(q, all_src + 4 * peer, o)) rl.insert(rl.end(), {o[0], o[1], o[2], o[3], peer});
    }
    const int ns = (int) sl.size() / 5, nr = (int) rl.size() / 5;
    // Block sizes, displacements and totals are int fields of the public
    // struct (mat_redist.h:7-45); a list whose running total passes INT_MAX
    // would wrap them, so it is refused here instead.
    int64_t stot = 0, rtot = 0;
    for (int i = 0; i < ns; ++i) stot += (int64_t) sl[5 * i + 2] * sl[5 * i + 3];
    for (int i = 0; i < nr; ++i) rtot += (int64_t) rl[5 * i + 2] * rl[5 * i + 3];
    if (stot > INT_MAX || rtot > INT_MAX) {
        ENG_ERR("mat_redist: rank %d would send %lld / receive %lld elements; counts above INT_MAX "
                "(%d) are not supported\n", rank, (long long) stot, (long long) rtot, INT_MAX);
        free(e);
        return nullptr;
    }
    e->n_proc_send = ns;
    e->n_proc_recv = nr;
    e->send_ranks = (int *) malloc(sizeof(int) * std::max(1, ns));
    e->send_sizes = (int *) malloc(sizeof(int) * std::max(1, ns));
    e->send_displs = (int *) malloc(sizeof(int) * (ns + 1));
    e->sblk_sizes = (int *) malloc(sizeof(int) * std::max(1, 4 * ns));
    e->recv_ranks = (int *) malloc(sizeof(int) * std::max(1, nr));
    e->recv_sizes = (int *) malloc(sizeof(int) * std::max(1, nr));
    e->recv_displs = (int *) malloc(sizeof(int) * (nr + 1));
    e->rblk_sizes = (int *) malloc(sizeof(int) * std::max(1, 4 * nr));
    int cnt = 0;
    for (int i = 0; i < ns; ++i) {
        const int *b = &sl[5 * i];
        std::copy(b, b + 4, e->sblk_sizes + 4 * i);
        e->send_ranks[i] = b[4];
        e->send_sizes[i] = b[2] * b[3];
        e->send_displs[i] = cnt;
        cnt += b[2] * b[3];
    }
    e->send_displs[ns] = cnt;
    e->send_cnt = cnt;
    cnt = 0;
    for (int i = 0; i < nr; ++i) {
        const int *b = &rl[5 * i];
        std::copy(b, b + 4, e->rblk_sizes + 4 * i);
        e->recv_ranks[i] = b[4];
        e->recv_sizes[i] = b[2] * b[3];
        e->recv_displs[i] = cnt;
        cnt += b[2] * b[3];
    }
    e->recv_displs[nr] = cnt;
    e->recv_cnt = cnt;
    return e;
}

size_t redist_bytes(const mat_redist_engine_p e)
{
    return e ? e->dt_size * ((size_t) e->send_cnt + (size_t) e->recv_cnt) : 0;
}

// 1: the mat_redist steps run staged through the engines' send and receive
// buffers (pack, exchange, unpack: three passes over the data); 0 (default):
// device-resident, one pass.
bool redist_staged() { return env_int("CA3DMM_REDIST_STAGED", 0, 0, 1) != 0; }

// The device-resident redistribution of a set of engines that together form
// one group (all ranks of the simulated grid, or the single rank).  Every
// rank's memory is on this GPU, so each intersection block goes straight
// from the sender's source block to the receiver's destination block: the
// receiver reads the sender's data where it lies, as a P2P read over xGMI
// would, and the pack, exchange and unpack of src/mat_redist.c:317-416
// collapse into one pass.  The blocks, their order and the bytes are those
// of the staged path.  Descriptors are appended to d (one launch for
// several matrices).
void redist_direct(int n, const mat_redist_engine_p *eng, const void *const *src, const int *src_ld,
                   void *const *dst, const int *dst_ld, std::vector<hip_blk_copy_t> &d)
{
    (void) n;
    for (int r = 0; r < n; ++r) {
        const mat_redist_engine_p e = eng[r];
        const size_t dt = e->dt_size;
        for (int i = 0; i < e->n_proc_recv; ++i) {
            const int *b = e->rblk_sizes + 4 * i;
            const int sr = e->recv_ranks[i];
            const mat_redist_engine_p s = eng[sr];
            hip_blk_copy_t c;
            c.src = (const char *) src[sr] + dt * ((size_t) (b[0] - s->src_srow) * src_ld[sr] + (b[1] - s->src_scol));
            c.dst = (char *) dst[r] + dt * ((size_t) (b[0] - e->req_srow) * dst_ld[r] + (b[1] - e->req_scol));
            c.lds = src_ld[sr]; c.ldd = dst_ld[r]; c.nrow = b[2]; c.ncol = b[3];
            d.push_back(c);
        }
    }
}

// Pack, exchange and unpack of a set of engines that together form one group
// (mat_redist.c:317-416), each step one batched copy launch, through the
// engines' work buffers.  For a single rank the self exchange is the identity
// and the unpack reads the send buffer directly.  The default path is
// redist_direct (one pass); this one runs with CA3DMM_REDIST_STAGED=1.
int redist_staged_run(int n, const mat_redist_engine_p *eng, const void *const *src, const int *src_ld,
                      void *const *dst, const int *dst_ld)
{
    std::vector<hip_blk_copy_t> d;
    const size_t dt = eng[0]->dt_size;
    for (int r = 0; r < n; ++r) {           // pack (mat_redist.c:317-348)
        const mat_redist_engine_p e = eng[r];
        for (int i = 0; i < e->n_proc_send; ++i) {
            const int *b = e->sblk_sizes + 4 * i;
            hip_blk_copy_t c;
            c.src = (const char *) src[r] + dt * ((size_t) (b[0] - e->src_srow) * src_ld[r] + (b[1] - e->src_scol));
            c.dst = (char *) e->sendbuf_d + dt * (size_t) e->send_displs[i];
            c.lds = src_ld[r]; c.ldd = b[3]; c.nrow = b[2]; c.ncol = b[3];
            d.push_back(c);
        }
    }
    int rc = hip_copy_matrix_blocks(dt, (int) d.size(), d.data());
    if (rc) return rc;
    const bool self_only = (n == 1);
    if (!self_only) {                      // exchange (mat_redist.c:350-387)
        d.clear();
        for (int r = 0; r < n; ++r) {
            const mat_redist_engine_p e = eng[r];
            for (int i = 0; i < e->n_proc_recv; ++i) {
                const mat_redist_engine_p s = eng[e->recv_ranks[i]];
                int j = 0;
                while (j < s->n_proc_send && s->send_ranks[j] != r) ++j;
                if (j == s->n_proc_send || s->send_sizes[j] != e->recv_sizes[i]) return HIP_PROXY_ERR_ARG;
                hip_blk_copy_t c;
                c.src = (const char *) s->sendbuf_d + dt * (size_t) s->send_displs[j];
                c.dst = (char *) e->recvbuf_d + dt * (size_t) e->recv_displs[i];
                c.lds = c.ldd = e->recv_sizes[i]; c.nrow = 1; c.ncol = e->recv_sizes[i];
                d.push_back(c);
            }
        }
        rc = hip_copy_matrix_blocks(dt, (int) d.size(), d.data());
        if (rc) return rc;
    }
    d.clear();
    for (int r = 0; r < n; ++r) {           // unpack (mat_redist.c:389-416)
        const mat_redist_engine_p e = eng[r];
        const char *rbuf = (const char *) (self_only ? e->sendbuf_d : e->recvbuf_d);
        for (int i = 0; i < e->n_proc_recv; ++i) {
            const int *b = e->rblk_sizes + 4 * i;
            hip_blk_copy_t c;
            c.src = rbuf + dt * (size_t) e->recv_displs[i];
            c.dst = (char *) dst[r] + dt * ((size_t) (b[0] - e->req_srow) * dst_ld[r] + (b[1] - e->req_scol));
            c.lds = b[3]; c.ldd = dst_ld[r]; c.nrow = b[2]; c.ncol = b[3];
            d.push_back(c);
        }
    }
    return hip_copy_matrix_blocks(dt, (int) d.size(), d.data());
}

int redist_run(int n, const mat_redist_engine_p *eng, const void *const *src, const int *src_ld,
               void *const *dst, const int *dst_ld)
{
    if (redist_staged()) return redist_staged_run(n, eng, src, src_ld, dst, dst_ld);
    std::vector<hip_blk_copy_t> d;
    redist_direct(n, eng, src, src_ld, dst, dst_ld, d);
    return d.empty() ? 0 : hip_copy_matrix_blocks(eng[0]->dt_size, (int) d.size(), d.data());
}

// ---------------------------------------------------------------------
// Cannon engine construction (cannon.c:16-133)
// ---------------------------------------------------------------------
// The public struct plus what the library keeps privately; the public part
// comes first so a cannon_engine_p points at both.
struct CannonPriv {
    cannon_engine_s pub;
    size_t workbuf_bytes;
};

cannon_engine_p cannon_build(int m, int n, int k, int q, int rank2d, dev_type_t dev_type)
{
    auto *priv = (CannonPriv *) calloc(1, sizeof(CannonPriv));
    if (!priv) return nullptr;
    cannon_engine_p e = &priv->pub;
    e->m = m; e->n = n; e->k = k;
    e->np_dim = q;
    e->my_rank = rank2d;
    e->rank_row = rank2d / q;
    e->rank_col = rank2d % q;
    e->comm.nproc = q * q;
    e->comm.rank = rank2d;
    e->dev_type = dev_type;
    calc_block_spos_size(m, q, e->rank_row, &e->A_srow, &e->A_nrow);
    calc_block_spos_size(k, q, e->rank_col, &e->A_scol, &e->A_ncol);
    calc_block_spos_size(k, q, e->rank_row, &e->B_srow, &e->B_nrow);
    calc_block_spos_size(n, q, e->rank_col, &e->B_scol, &e->B_ncol);
    calc_block_spos_size(m, q, e->rank_row, &e->C_srow, &e->C_nrow);
    calc_block_spos_size(n, q, e->rank_col, &e->C_scol, &e->C_ncol);
    // int fields of the public struct (cannon.h:7-50): refuse shapes whose
    // block sizes would wrap them
    const int64_t mA = (int64_t) (m / q + 1) * (k / q + 1), mB = (int64_t) (k / q + 1) * (n / q + 1),
                  mC = (int64_t) (m / q + 1) * (n / q + 1);
    if (std::max(mA, std::max(mB, mC)) > INT_MAX) {
        ENG_ERR("cannon: a block of %lld elements exceeds INT_MAX (%d)\n",
                (long long) std::max(mA, std::max(mB, mC)), INT_MAX);
        free(priv);
        return nullptr;
    }
    e->max_A_blk_size = (int) mA;
    e->max_B_blk_size = (int) mB;
    e->max_C_blk_size = (int) mC;
    // The reference's GPU default (256) is overwritten by the env macro's
    // default (160) whenever the variable is unset (cannon.c:89-91); the
    // observable default is 160 on every device type.
    const int min_k = env_int("CANNON_MIN_KBLK_SIZE", 160, 16, 8192);
    e->gemm_cycle = 1;
    priv->workbuf_bytes = sizeof(double) * ((size_t) e->max_A_blk_size + (size_t) e->max_B_blk_size);
    if (e->A_ncol > 0 && e->A_ncol < min_k) {
        e->gemm_cycle = std::min(q, (min_k + e->A_ncol - 1) / e->A_ncol);
        // k-stack buffers are sized even when the cycle is clamped to 1
        // (cannon.c:92-98)
        priv->workbuf_bytes += sizeof(double) * ((size_t) e->max_A_blk_size + (size_t) e->max_B_blk_size) *
                               e->gemm_cycle;
    }
    e->m_displs = (int *) malloc(sizeof(int) * (q + 1));
    e->n_displs = (int *) malloc(sizeof(int) * (q + 1));
    e->k_displs = (int *) malloc(sizeof(int) * (q + 1));
    for (int i = 0; i <= q; ++i) {
        int dummy;
        calc_block_spos_size(m, q, i, e->m_displs + i, &dummy);
        calc_block_spos_size(n, q, i, e->n_displs + i, &dummy);
        calc_block_spos_size(k, q, i, e->k_displs + i, &dummy);
    }
    return e;
}

size_t cannon_bytes(const cannon_engine_p e) { return reinterpret_cast<const CannonPriv *>(e)->workbuf_bytes; }

// ---------------------------------------------------------------------
// CA3DMM engine construction for one rank, given every rank's plan
// ---------------------------------------------------------------------
struct Problem {
    int p, m, n, k, tA, tB;
    bool btb;
    Grid grid;
    std::vector<RankPlan> plans;
    std::vector<Rect> srcA, srcB, dstC;   // per rank (column-major blocks)
    bool have_dstC;
    dev_type_t dev_type;
};

Rect op_rect(const Rect &r, bool swap)   // rectangle of op(X) -> rectangle of X as stored
{
    return swap ? Rect{r.scol, r.ncol, r.srow, r.nrow} : r;
}

ca3dmm_engine_p build_rank(const Problem &P, int rank)
{
    auto *e = (ca3dmm_engine_p) calloc(1, sizeof(ca3dmm_engine_s));
    if (!e) return nullptr;
    const RankPlan &r = P.plans[rank];
    e->m = P.btb ? P.n : P.m; e->n = P.n; e->k = P.k;
    e->mp = P.grid.mp; e->np = P.grid.np; e->kp = P.grid.kp; e->rp = P.grid.rp;
    e->my_rank = rank; e->n_proc = P.p;
    e->is_active = r.active;
    e->rank_m = r.rank_m; e->rank_n = r.rank_n; e->rank_k = r.rank_k;
    e->trans_A = P.btb ? 0 : P.tA; e->trans_B = P.btb ? 0 : P.tB;
    e->task_m_num = r.tmn; e->task_m_id = r.tmi;
    e->task_n_num = r.tnn; e->task_n_id = r.tni;
    e->task_k_num = r.tkn; e->task_k_id = r.tki;
    e->is_BTB = P.btb;
    e->use_ag = r.use_ag; e->use_rsb = r.use_rsb;
    e->A_rd_srow = r.A_rd.srow; e->A_rd_scol = r.A_rd.scol; e->A_rd_nrow = r.A_rd.nrow; e->A_rd_ncol = r.A_rd.ncol;
    e->B_rd_srow = r.B_rd.srow; e->B_rd_scol = r.B_rd.scol; e->B_rd_nrow = r.B_rd.nrow; e->B_rd_ncol = r.B_rd.ncol;
    e->A_2dmm_srow = r.A_2dmm.srow; e->A_2dmm_scol = r.A_2dmm.scol;
    e->A_2dmm_nrow = r.A_2dmm.nrow; e->A_2dmm_ncol = r.A_2dmm.ncol;
    e->B_2dmm_srow = r.B_2dmm.srow; e->B_2dmm_scol = r.B_2dmm.scol;
    e->B_2dmm_nrow = r.B_2dmm.nrow; e->B_2dmm_ncol = r.B_2dmm.ncol;
    e->C_2dmm_srow = r.C_2dmm.srow; e->C_2dmm_scol = r.C_2dmm.scol;
    e->C_2dmm_nrow = r.C_2dmm.nrow; e->C_2dmm_ncol = r.C_2dmm.ncol;
    e->C_out_srow = r.C_out.srow; e->C_out_scol = r.C_out.scol;
    e->C_out_nrow = r.C_out.nrow; e->C_out_ncol = r.C_out.ncol;
    e->dev_type = P.dev_type;
    // sub-groups (ca3dmm.c:194-210): Cannon group, k-layer group, allgather group
    e->comm_2dmm.nproc = r.active ? r.group : 0;  e->comm_2dmm.rank = r.rank2d;
    e->comm_C_rs.nproc = r.active ? P.grid.kp : 0; e->comm_C_rs.rank = r.rank_k;
    e->comm_AB_agv.nproc = r.active ? r.tmn * r.tnn : 0; e->comm_AB_agv.rank = r.tmi + r.tni;
    // the C_2dmm block and the allgathered operand are counted in int fields
    // (C_rs_recvcnts, AB_agv_*; ca3dmm.h:10-82)
    if (r.active && ((int64_t) r.C_2dmm.nrow * r.C_2dmm.ncol > INT_MAX ||
                     (int64_t) r.A_2dmm.nrow * r.A_2dmm.ncol > INT_MAX ||
                     (int64_t) r.B_2dmm.nrow * r.B_2dmm.ncol > INT_MAX)) {
        ENG_ERR("ca3dmm: rank %d's blocks exceed INT_MAX (%d) elements\n", rank, INT_MAX);
        free(e);
        return nullptr;
    }
    const int tk = r.active ? P.grid.kp : 0;
    e->C_rs_recvcnts = (int *) malloc(sizeof(int) * std::max(1, tk));
    for (int t = 0; t < tk; ++t) {
        int sc, nc;
        calc_block_spos_size(r.C_2dmm.ncol, tk, t, &sc, &nc);
        e->C_rs_recvcnts[t] = nc * r.C_2dmm.nrow;
    }
    const int tmn = r.tmn * r.tnn;
    e->AB_agv_recvcnts = (int *) malloc(sizeof(int) * std::max(1, tmn));
    e->AB_agv_displs = (int *) calloc(std::max(1, tmn) + 1, sizeof(int));
    if (r.tnn > 1 || r.tmn > 1) {
        const bool repA = r.tnn > 1;
        for (int t = 0; t < tmn; ++t) {
            int sc, nc;
            calc_block_spos_size(repA ? r.A_2dmm.ncol : r.B_2dmm.ncol, tmn, t, &sc, &nc);
            e->AB_agv_recvcnts[t] = nc * (repA ? r.A_2dmm.nrow : r.B_2dmm.nrow);
            e->AB_agv_displs[t + 1] = e->AB_agv_displs[t] + e->AB_agv_recvcnts[t];
        }
    }
    if (r.active) {
        e->cannon_engine = cannon_build(r.msz, r.nsz, r.ksz, r.q, r.rank2d, P.dev_type);
        if (!e->cannon_engine) { ca3dmm_engine_free(&e); return nullptr; }
        e->cannon_workbuf_bytes = cannon_bytes(e->cannon_engine);
    }
    // redistributions (ca3dmm.c:357-397, :750-779)
    const int p = P.p;
    std::vector<int> src(4 * p), req(4 * p);
    auto active_rect = [&](int i, const Rect &x) { return P.plans[i].active ? x : Rect{0, 0, 0, 0}; };
    if (!P.btb) {
        for (int i = 0; i < p; ++i) {
            to_rm(P.srcA[i], &src[4 * i]);
            to_rm(active_rect(i, op_rect(P.plans[i].A_rd, P.tA)), &req[4 * i]);
        }
        e->redist_A = redist_build(p, rank, src.data(), req.data(), CA3DMM_DTYPE_DOUBLE, 8, P.dev_type);
        e->rdA_workbuf_bytes = redist_bytes(e->redist_A);
    }
    for (int i = 0; i < p; ++i) {
        to_rm(P.srcB[i], &src[4 * i]);
        to_rm(active_rect(i, op_rect(P.plans[i].B_rd, P.btb ? 0 : P.tB)), &req[4 * i]);
    }
    e->redist_B = redist_build(p, rank, src.data(), req.data(), CA3DMM_DTYPE_DOUBLE, 8, P.dev_type);
    e->rdB_workbuf_bytes = redist_bytes(e->redist_B);
    if (P.have_dstC) {
        for (int i = 0; i < p; ++i) {
            to_rm(active_rect(i, P.plans[i].C_out), &src[4 * i]);
            to_rm(P.dstC[i], &req[4 * i]);
        }
        e->redist_C = redist_build(p, rank, src.data(), req.data(), CA3DMM_DTYPE_DOUBLE, 8, P.dev_type);
        e->rdC_workbuf_bytes = redist_bytes(e->redist_C);
    }
    if ((!P.btb && !e->redist_A) || !e->redist_B || (P.have_dstC && !e->redist_C)) {
        ca3dmm_engine_free(&e);
        return nullptr;
    }
    // own buffers (ca3dmm.c:398-412, :680-690)
    size_t self = 0;
    if (r.active) {
        const size_t Ab = 8 * (size_t) e->cannon_engine->max_A_blk_size;
        const size_t Bb = 8 * (size_t) e->cannon_engine->max_B_blk_size;
        if (P.btb) {
            self = Ab + Ab + Bb;
        } else {
            self = Ab + Bb;
            if (P.tA) self += Ab;
            if (P.tB) self += Bb;
            if (r.tnn > 1) self += Ab;
            if (r.tmn > 1) self += Bb;
        }
        self += 8 * (size_t) r.C_2dmm.nrow * r.C_2dmm.ncol;
        self += 8 * (size_t) r.C_out.nrow * r.C_out.ncol;
    }
    e->self_workbuf_bytes = self;
    e->print_timing = env_int("CA3DMM_PRINT_TIMING", 0, 0, 1) && rank == 0;
    return e;
}

size_t total_bytes(const ca3dmm_engine_p e)
{
    size_t c = std::max(std::max(e->rdA_workbuf_bytes, e->rdB_workbuf_bytes),
                        std::max(e->rdC_workbuf_bytes, e->cannon_workbuf_bytes));
    return c + e->self_workbuf_bytes;
}

// Engines accept only device-resident data: the library computes on the GPU.
bool check_dev(dev_type_t t)
{
    if (!is_dev_type_valid(t)) { ENG_ERR("invalid device type %d\n", (int) t); return false; }
    if (!gpu_type(t)) {
        ENG_ERR("device type %d: this library computes on the GPU only; use DEV_TYPE_HIP\n", (int) t);
        return false;
    }
    return true;
}

}  // namespace

// =====================================================================
// Public engine API
// =====================================================================
extern "C" {

void mat_redist_engine_init(const int src_srow, const int src_scol, const int src_nrow,
                            const int src_ncol, const int req_srow, const int req_scol,
                            const int req_nrow, const int req_ncol, ca3dmm_comm_t comm, int dtype,
                            const size_t dt_size, dev_type_t dev_type, mat_redist_engine_p *engine_,
                            size_t *workbuf_bytes)
{
    *engine_ = nullptr;
    if (!check_dev(dev_type)) return;
    if (comm.nproc != 1 || comm.rank != 0) {
        ENG_ERR("mat_redist_engine_init: multi-rank groups are built by ca3dmm_sim_init\n");
        return;
    }
    if (dt_size != 4 && dt_size != 8) { ENG_ERR("mat_redist_engine_init: dt_size must be 4 or 8\n"); return; }
    const int s[4] = {src_srow, src_scol, src_nrow, src_ncol};
    const int q[4] = {req_srow, req_scol, req_nrow, req_ncol};
    mat_redist_engine_p e = redist_build(1, 0, s, q, dtype, dt_size, dev_type);
    if (!e) return;
    const size_t bytes = redist_bytes(e);
    if (workbuf_bytes) {
        e->alloc_workbuf = 0;
        *workbuf_bytes = bytes;
    } else {
        e->alloc_workbuf = 1;
        void *wb = dev_type_malloc(bytes, dev_type);
        if (!wb && bytes) { mat_redist_engine_free(&e); return; }
        mat_redist_engine_attach_workbuf(e, nullptr, wb);
    }
    *engine_ = e;
}

void mat_redist_engine_attach_workbuf(mat_redist_engine_p e, void *workbuf_h, void *workbuf_d)
{
    if (!e) return;
    e->workbuf_h = workbuf_h;
    e->workbuf_d = workbuf_d;
    e->sendbuf_d = workbuf_d;
    e->recvbuf_d = workbuf_d ? (char *) workbuf_d + e->dt_size * (size_t) e->send_cnt : nullptr;
}

void mat_redist_engine_exec(mat_redist_engine_p e, const void *src_blk, const int src_ld, void *dst_blk,
                            const int dst_ld)
{
    if (!e) { ENG_ERR("mat_redist_engine not initialized\n"); return; }
    if (e->nproc != 1) { ENG_ERR("mat_redist_engine_exec: multi-rank groups run via ca3dmm_sim_exec\n"); return; }
    const void *s[1] = {src_blk};
    void *d[1] = {dst_blk};
    const int sl[1] = {src_ld}, dl[1] = {dst_ld};
    mat_redist_engine_p one[1] = {e};
    if (redist_run(1, one, s, sl, d, dl)) ENG_ERR("mat_redist_engine_exec failed\n");
    hip_stream_sync(nullptr);
}

void mat_redist_engine_free(mat_redist_engine_p *engine_)
{
    mat_redist_engine_p e = engine_ ? *engine_ : nullptr;
    if (!e) return;
    if (e->alloc_workbuf) dev_type_free(e->workbuf_d, e->dev_type);
    free(e->send_ranks); free(e->send_sizes); free(e->send_displs); free(e->sblk_sizes);
    free(e->recv_ranks); free(e->recv_sizes); free(e->recv_displs); free(e->rblk_sizes);
    free(e);
    *engine_ = nullptr;
}

void cannon_engine_init(const int m, const int n, const int k, ca3dmm_comm_t comm, dev_type_t dev_type,
                        cannon_engine_p *engine_, size_t *workbuf_bytes)
{
    *engine_ = nullptr;
    if (!check_dev(dev_type)) return;
    const int q = (int) floor(sqrt((double) comm.nproc) + 0.5);
    if (q * q != comm.nproc) { ENG_ERR("group size %d is not a square number\n", comm.nproc); return; }
    if (comm.nproc != 1) { ENG_ERR("cannon_engine_init: q > 1 grids are built by ca3dmm_sim_init\n"); return; }
    const double t0 = now_ms();
    cannon_engine_p e = cannon_build(m, n, k, 1, 0, dev_type);
    if (!e) return;
    const size_t bytes = cannon_bytes(e);
    if (workbuf_bytes) {
        *workbuf_bytes = bytes;
    } else {
        e->alloc_workbuf = 1;
        void *wb = dev_type_malloc(bytes, dev_type);
        if (!wb && bytes) { cannon_engine_free(&e); return; }
        cannon_engine_attach_workbuf(e, nullptr, wb);
    }
    e->init_ms = now_ms() - t0;
    *engine_ = e;
}

void cannon_engine_attach_workbuf(cannon_engine_p e, void *workbuf_h, void *workbuf_d)
{
    if (!e) return;
    e->workbuf_h = workbuf_h;
    e->workbuf_d = workbuf_d;
    if (!workbuf_d) return;
    e->A_recv_d = workbuf_d;
    e->B_recv_d = (double *) e->A_recv_d + e->max_A_blk_size;
    if (e->gemm_cycle > 1) {
        e->A_stack_d = (double *) e->B_recv_d + e->max_B_blk_size;
        e->B_stack_d = (double *) e->A_stack_d + (size_t) e->max_A_blk_size * e->gemm_cycle;
    }
}

void cannon_engine_free(cannon_engine_p *engine_)
{
    cannon_engine_p e = engine_ ? *engine_ : nullptr;
    if (!e) return;
    if (e->alloc_workbuf) dev_type_free(e->workbuf_d, e->dev_type);
    free(e->m_displs); free(e->n_displs); free(e->k_displs);
    free(reinterpret_cast<CannonPriv *>(e));
    *engine_ = nullptr;
}

void cannon_engine_exec(cannon_engine_p e, const double alpha, const double beta, double *A_blk,
                        double *B_blk, double *C_blk)
{
    if (!e) { ENG_ERR("cannon_engine not initialized\n"); return; }
    if (e->m == 0 || e->n == 0 || e->k == 0) return;
    if (e->np_dim != 1) { ENG_ERR("cannon_engine_exec: q > 1 grids run via ca3dmm_sim_exec\n"); return; }
    const double t0 = now_ms();
    if (hip_dgemm(HipNoTrans, HipNoTrans, e->m, e->n, e->k, alpha, A_blk, e->m, B_blk, e->k, beta, C_blk, e->m))
        ENG_ERR("cannon_engine_exec: local DGEMM failed\n");
    hip_stream_sync(nullptr);
    const double dt = now_ms() - t0;
    e->gemm_ms += dt;
    e->exec_ms += dt;
    e->n_exec++;
}

void cannon_engine_reset_stat(cannon_engine_p e)
{
    if (!e) return;
    e->shift0_ms = e->lshift_ms = e->gemm_ms = e->hd_trans_ms = e->exec_ms = 0.0;
    e->n_exec = 0;
}

void cannon_engine_print_stat(cannon_engine_p e)
{
    if (!e) return;
    if (e->n_exec == 0) { ENG_ERR("no cannon_engine statistics to print\n"); return; }
    const double gf = 2.0 * (double) e->C_nrow * (double) e->C_ncol * (double) e->k * e->n_exec /
                      e->exec_ms * 1e3 / 1e9;
    printf("--------------- 2D Cannon engine (HIP) ---------------\n");
    printf("* Initialization        : %.2f ms\n", e->init_ms);
    printf("* Number of executions  : %d\n", e->n_exec);
    printf("* Execution time (avg)  : %.2f ms\n", e->exec_ms / e->n_exec);
    printf("  * Initial shift       : %.2f ms\n", e->shift0_ms / e->n_exec);
    printf("  * Loop shift wait     : %.2f ms\n", e->lshift_ms / e->n_exec);
    printf("  * Local DGEMM         : %.2f ms\n", e->gemm_ms / e->n_exec);
    printf("* Per-rank performance  : %.2f GFlops\n", gf);
    printf("------------------------------------------------------\n");
}

}  // extern "C"

// =====================================================================
// CA3DMM engine
// =====================================================================
namespace {

ca3dmm_engine_p init_single(int m, int n, int k, int tA, int tB, const Rect &sA, const Rect &sB,
                            const Rect *dC, const int *proc_grid, bool btb, ca3dmm_comm_t comm,
                            dev_type_t dev_type, size_t *workbuf_bytes)
{
    if (!check_dev(dev_type)) return nullptr;
    if (comm.nproc != 1 || comm.rank != 0) {
        ENG_ERR("ca3dmm_engine_init: multi-rank groups are planned and run by ca3dmm_sim_init / "
                "ca3dmm_sim_exec (no inter-process exchange is built)\n");
        return nullptr;
    }
    const double t0 = now_ms();
    Problem P;
    P.p = 1; P.m = m; P.n = n; P.k = k; P.tA = tA; P.tB = tB; P.btb = btb;
    P.dev_type = dev_type;
    if (choose_grid(1, btb ? n : m, n, k, proc_grid, btb, &P.grid)) return nullptr;
    P.plans = {plan_rank(P.grid, btb ? n : m, n, k, btb, 0)};
    P.srcA = {sA}; P.srcB = {sB};
    P.have_dstC = dC != nullptr;
    if (dC) P.dstC = {*dC};
    ca3dmm_engine_p e = build_rank(P, 0);
    if (!e) return nullptr;
    const size_t bytes = total_bytes(e);
    if (workbuf_bytes) {
        e->alloc_workbuf = 0;
        *workbuf_bytes = bytes;
    } else {
        e->alloc_workbuf = 1;
        void *wb = dev_type_malloc(bytes, dev_type);
        if (!wb && bytes) { ca3dmm_engine_free(&e); return nullptr; }
        ca3dmm_engine_attach_workbuf(e, nullptr, wb);
    }
    e->init_ms = now_ms() - t0;
    return e;
}

bool all_minus_one(int a, int b, int c, int d) { return a == -1 || b == -1 || c == -1 || d == -1; }

// The op(A) block received by redistribution, as stored: rows x cols and
// its leading dimension (ca3dmm.c:535-536).
int recv_ld(int rd_nrow, int rd_ncol, int trans) { return trans ? rd_ncol : rd_nrow; }

}  // namespace

extern "C" {

void ca3dmm_engine_init(const int m, const int n, const int k, const int trans_A, const int trans_B,
                        const int src_A_srow, const int src_A_nrow, const int src_A_scol,
                        const int src_A_ncol, const int src_B_srow, const int src_B_nrow,
                        const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                        const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                        const int *proc_grid, ca3dmm_comm_t comm, dev_type_t dev_type,
                        ca3dmm_engine_p *engine_, size_t *workbuf_bytes)
{
    *engine_ = nullptr;
    const Rect sA{src_A_srow, src_A_nrow, src_A_scol, src_A_ncol};
    const Rect sB{src_B_srow, src_B_nrow, src_B_scol, src_B_ncol};
    const Rect dC{dst_C_srow, dst_C_nrow, dst_C_scol, dst_C_ncol};
    const bool noC = all_minus_one(dst_C_srow, dst_C_nrow, dst_C_scol, dst_C_ncol);
    *engine_ = init_single(m, n, k, trans_A ? 1 : 0, trans_B ? 1 : 0, sA, sB, noC ? nullptr : &dC,
                           proc_grid, false, comm, dev_type, workbuf_bytes);
}

void ca3dmm_engine_init_BTB(const int n, const int k, const int src_B_srow, const int src_B_nrow,
                            const int src_B_scol, const int src_B_ncol, const int dst_C_srow,
                            const int dst_C_nrow, const int dst_C_scol, const int dst_C_ncol,
                            const int *proc_grid, ca3dmm_comm_t comm, dev_type_t dev_type,
                            ca3dmm_engine_p *engine_, size_t *workbuf_bytes)
{
    *engine_ = nullptr;
    const Rect sB{src_B_srow, src_B_nrow, src_B_scol, src_B_ncol};
    const Rect dC{dst_C_srow, dst_C_nrow, dst_C_scol, dst_C_ncol};
    const bool noC = all_minus_one(dst_C_srow, dst_C_nrow, dst_C_scol, dst_C_ncol);
    *engine_ = init_single(n, n, k, 0, 0, Rect{0, 0, 0, 0}, sB, noC ? nullptr : &dC, proc_grid, true,
                           comm, dev_type, workbuf_bytes);
}

void ca3dmm_engine_attach_workbuf(ca3dmm_engine_p e, void *workbuf_h, void *workbuf_d)
{
    if (!e) return;
    const size_t child = std::max(std::max(e->rdA_workbuf_bytes, e->rdB_workbuf_bytes),
                                  std::max(e->rdC_workbuf_bytes, e->cannon_workbuf_bytes));
    e->workbuf_h = workbuf_h;
    e->workbuf_d = workbuf_d;
    char *cw = (char *) workbuf_d;
    if (e->redist_A) mat_redist_engine_attach_workbuf(e->redist_A, nullptr, cw);
    if (e->redist_B) mat_redist_engine_attach_workbuf(e->redist_B, nullptr, cw);
    if (e->redist_C) mat_redist_engine_attach_workbuf(e->redist_C, nullptr, cw);
    if (e->cannon_engine) cannon_engine_attach_workbuf(e->cannon_engine, nullptr, cw);
    if (!e->is_active || !workbuf_d) return;
    char *s = cw + child;                      // own buffers (ca3dmm.c:725-894)
    const size_t Ab = 8 * (size_t) e->cannon_engine->max_A_blk_size;
    const size_t Bb = 8 * (size_t) e->cannon_engine->max_B_blk_size;
    if (!e->is_BTB) {
        e->A_rd_recv_d = s; s += Ab;
        e->B_rd_recv_d = s; s += Bb;
        if (e->trans_A) { e->A_trans_d = s; s += Ab; } else e->A_trans_d = e->A_rd_recv_d;
        if (e->trans_B) { e->B_trans_d = s; s += Bb; } else e->B_trans_d = e->B_rd_recv_d;
        if (e->task_n_num > 1) { e->A_2dmm_d = s; s += Ab; } else e->A_2dmm_d = e->A_trans_d;
        if (e->task_m_num > 1) { e->B_2dmm_d = s; s += Bb; } else e->B_2dmm_d = e->B_trans_d;
    } else {
        e->A_2dmm_d = s; s += Ab;
        e->A_trans_d = s; s += Ab;
        e->B_2dmm_d = s; s += Bb;
    }
    e->C_2dmm_d = s; s += 8 * (size_t) e->C_2dmm_nrow * e->C_2dmm_ncol;
    e->C_out_d = s;
}

void ca3dmm_engine_free(ca3dmm_engine_p *engine_)
{
    ca3dmm_engine_p e = engine_ ? *engine_ : nullptr;
    if (!e) return;
    free(e->AB_agv_recvcnts); free(e->AB_agv_displs); free(e->C_rs_recvcnts);
    if (e->alloc_workbuf) dev_type_free(e->workbuf_d, e->dev_type);
    mat_redist_engine_free(&e->redist_A);
    mat_redist_engine_free(&e->redist_B);
    mat_redist_engine_free(&e->redist_C);
    cannon_engine_free(&e->cannon_engine);
    free(e);
    *engine_ = nullptr;
}

void ca3dmm_engine_exec(ca3dmm_engine_p e, const void *src_A, const int ldA, const void *src_B,
                        const int ldB, void *dst_C, const int ldC)
{
    if (!e) { ENG_ERR("ca3dmm_engine not initialized\n"); return; }
    if (e->n_proc != 1) { ENG_ERR("ca3dmm_engine_exec: multi-rank grids run via ca3dmm_sim_exec\n"); return; }
    if (!e->workbuf_d && (e->self_workbuf_bytes || e->rdB_workbuf_bytes)) {
        ENG_ERR("ca3dmm_engine_exec: no work buffer attached\n");
        return;
    }
    const double t_exec = now_ms();
    double t0 = t_exec;
    const int m = e->m, n = e->n, k = e->k;
    int rc = 0;
    // Identity redistribution (one rank holding the whole matrices, C wanted
    // in full): the exec is one DGEMM from the caller's buffers, with op()
    // fused into the operand loads.
    const bool force = env_int("CA3DMM_FORCE_REDIST", 0, 0, 1) != 0;
    auto whole = [](const mat_redist_engine_p r) {
        return r && r->n_proc_send == 1 && r->n_proc_recv == 1 && r->src_srow == r->req_srow &&
               r->src_scol == r->req_scol && r->src_nrow == r->req_nrow && r->src_ncol == r->req_ncol;
    };
    const bool directC = e->redist_C ? whole(e->redist_C) : false;
    if (!e->is_BTB && !force && whole(e->redist_A) && whole(e->redist_B) && directC) {
        rc = hip_dgemm(e->trans_A ? HipTrans : HipNoTrans, e->trans_B ? HipTrans : HipNoTrans, m, n, k,
                       1.0, (const double *) src_A, ldA, (const double *) src_B, ldB, 0.0,
                       (double *) dst_C, ldC);
        hip_stream_sync(nullptr);
        const double t1 = now_ms();
        e->cannon_ms += t1 - t0;
        if (e->cannon_engine) {
            e->cannon_engine->gemm_ms += t1 - t0;
            e->cannon_engine->exec_ms += t1 - t0;
            e->cannon_engine->n_exec++;
        }
    } else {
        // redistribute A and B (ca3dmm.c:986-999)
        const mat_redist_engine_p rb[1] = {e->redist_B};
        if (!e->is_BTB) {
            const mat_redist_engine_p ra[1] = {e->redist_A};
            const void *sa[1] = {src_A};
            void *da[1] = {e->A_rd_recv_d};
            const int sla[1] = {ldA}, dla[1] = {recv_ld(e->A_rd_nrow, e->A_rd_ncol, e->trans_A)};
            rc = redist_run(1, ra, sa, sla, da, dla);
            const void *sb[1] = {src_B};
            void *db[1] = {e->B_rd_recv_d};
            const int slb[1] = {ldB}, dlb[1] = {recv_ld(e->B_rd_nrow, e->B_rd_ncol, e->trans_B)};
            if (!rc) rc = redist_run(1, rb, sb, slb, db, dlb);
        } else {
            const void *sb[1] = {src_B};
            void *db[1] = {e->B_2dmm_d};
            const int slb[1] = {ldB}, dlb[1] = {e->B_2dmm_nrow};
            rc = redist_run(1, rb, sb, slb, db, dlb);
        }
        hip_stream_sync(nullptr);
        double t1 = now_ms();
        e->redist_ms += t1 - t0;
        t0 = t1;
        // local product on the 1 x 1 grid; op() fused into the loads
        double *Cdst = (double *) (e->task_k_num == 1 ? e->C_out_d : e->C_2dmm_d);
        if (!rc && e->is_active && m && n && k) {
            if (!e->is_BTB) {
                rc = hip_dgemm(e->trans_A ? HipTrans : HipNoTrans, e->trans_B ? HipTrans : HipNoTrans, m,
                               n, k, 1.0, (const double *) e->A_rd_recv_d,
                               recv_ld(e->A_rd_nrow, e->A_rd_ncol, e->trans_A), (const double *) e->B_rd_recv_d,
                               recv_ld(e->B_rd_nrow, e->B_rd_ncol, e->trans_B), 0.0, Cdst, e->C_2dmm_nrow);
            } else {
                // C = B^T B from the redistributed B block (ca3dmm.c:1172-1265,
                // pair transpose fused into the loads)
                rc = hip_dgemm(HipTrans, HipNoTrans, n, n, k, 1.0, (const double *) e->B_2dmm_d, e->B_2dmm_nrow,
                               (const double *) e->B_2dmm_d, e->B_2dmm_nrow, 0.0, Cdst, e->C_2dmm_nrow);
            }
        }
        hip_stream_sync(nullptr);
        t1 = now_ms();
        e->cannon_ms += t1 - t0;
        if (e->cannon_engine) {
            e->cannon_engine->gemm_ms += t1 - t0;
            e->cannon_engine->exec_ms += t1 - t0;
            e->cannon_engine->n_exec++;
        }
        t0 = t1;
        if (!rc && e->redist_C) {
            const mat_redist_engine_p rcn[1] = {e->redist_C};
            const void *sc[1] = {e->C_out_d};
            void *dc[1] = {dst_C};
            const int slc[1] = {e->C_out_nrow}, dlc[1] = {ldC};
            rc = redist_run(1, rcn, sc, slc, dc, dlc);
        }
        hip_stream_sync(nullptr);
        e->redist_ms += now_ms() - t0;
    }
    if (rc) ENG_ERR("ca3dmm_engine_exec failed (%d): %s\n", rc, hip_proxy_last_error());
    e->exec_ms += now_ms() - t_exec;
    e->n_exec++;
}

void ca3dmm_engine_reset_stat(ca3dmm_engine_p e)
{
    if (!e) return;
    cannon_engine_reset_stat(e->cannon_engine);
    e->redist_ms = e->agvAB_ms = e->cannon_ms = e->reduce_ms = e->hd_trans_ms = e->exec_ms = 0.0;
    e->n_exec = 0;
}

void ca3dmm_engine_print_stat(ca3dmm_engine_p e)
{
    if (!e) return;
    if (e->n_exec == 0) { ENG_ERR("no ca3dmm_engine statistics to print\n"); return; }
    printf("================ CA3DMM engine (HIP) =================\n");
    printf("* Initialization         : %.2f ms\n", e->init_ms);
    printf("* Number of executions   : %d\n", e->n_exec);
    printf("* Execution time (avg)   : %.2f ms\n", e->exec_ms / e->n_exec);
    printf("  * Redistribute A, B, C : %.2f ms\n", e->redist_ms / e->n_exec);
    printf("  * Allgather A or B     : %.2f ms\n", e->agvAB_ms / e->n_exec);
    printf("  * 2D Cannon execution  : %.2f ms\n", e->cannon_ms / e->n_exec);
    printf("  * Reduce-scatter C     : %.2f ms\n", e->reduce_ms / e->n_exec);
    cannon_engine_print_stat(e->cannon_engine);
    printf("======================================================\n");
}

}  // extern "C"

// =====================================================================
// Simulated p-rank grid on one GPU
// =====================================================================
struct ca3dmm_sim {
    Problem P;
    std::vector<ca3dmm_engine_p> eng;
    int grouped = 0;              // Cannon rounds + k fold as one persistent launch
    int use_graph = 0;            // replay that launch from a captured hipGraph
    int ozaki = 0;                // > 0: the local products through the CRT form with this many moduli
    int ozaki_auto = 0;           // choose that count at every exec from the operands in place
    double ozaki_spread = NAN;    // the spread statistic L of the last automatic choice
    ca3::grouped::Plan *plan = nullptr;
    ca3::grouped::OzPlan *ozplan = nullptr;
    ca3::grouped::BatchPlan *batch = nullptr;  // the per-rank rounds as one launch
    bool batch_unsupported = false;            // more rounds per rank than a batch item holds
    uint64_t plan_gen = 0;        // plans built so far: a captured sequence holds their buffers
    // captured exec sequences (use_graph) and the arguments and settings each
    // was captured for: [0] the whole exec, [1] and [2] its operand and
    // product halves (automatic modulus count)
    ca3::grouped::SeqGraph *seq[3] = {nullptr, nullptr, nullptr};
    std::vector<uintptr_t> seq_key[3];
};

namespace {

// Rank of the member at Cannon coordinates (row, col) of rank r's group.
int cannon_member(const Problem &P, int r, int row, int col)
{
    const RankPlan &pl = P.plans[r];
    const int task = pl.tmi + pl.tni;
    return pl.rank_k * P.grid.mp * P.grid.np + task * pl.group + row * pl.q + col;
}

// Problems and fold groups of the persistent grouped launch: one problem
// per active rank (its q Cannon rounds as k-segments read in place), one
// fold group per mn-position when kp > 1.
int collect_grouped(ca3dmm_sim *S, std::vector<ca3::grouped::Problem> &probs,
                    std::vector<ca3::grouped::Fold> &folds, std::vector<int> *ranks = nullptr)
{
    const Problem &P = S->P;
    namespace G = ca3::grouped;
    const int mn = P.grid.mp * P.grid.np, kp = P.grid.kp;
    std::vector<int> fold_of(mn, -1);
    if (kp > 1) {
        if (kp > G::kMaxKp) return HIP_PROXY_ERR_UNSUPP;
        for (int r = 0; r < mn; ++r) {
            ca3dmm_engine_p e = S->eng[r];
            if (!e->is_active || !e->C_2dmm_nrow || !e->C_2dmm_ncol) continue;
            G::Fold f;
            memset(&f, 0, sizeof(f));
            f.kp = kp;
            f.M = e->C_2dmm_nrow; f.N = e->C_2dmm_ncol;
            f.ldl = e->C_2dmm_nrow;
            f.ldo = std::max(1, e->C_out_nrow);
            for (int l = 0; l < kp; ++l) {
                ca3dmm_engine_p el = S->eng[l * mn + r];
                f.Cl[l] = (double *) el->C_2dmm_d;
                f.Cout[l] = (double *) el->C_out_d;
                f.cs[l] = el->C_out_scol - e->C_2dmm_scol;
            }
            f.cs[kp] = f.N;
            fold_of[r] = (int) folds.size();
            folds.push_back(f);
        }
    }
    for (int r = 0; r < P.p; ++r) {
        ca3dmm_engine_p e = S->eng[r];
        if (!e->is_active) continue;
        const cannon_engine_p ce = e->cannon_engine;
        const int q = ce->np_dim, row = ce->rank_row, col = ce->rank_col;
        const int Am = ce->m_displs[row + 1] - ce->m_displs[row];
        const int Bn = ce->n_displs[col + 1] - ce->n_displs[col];
        if (!Am || !Bn) continue;
        if (q > G::kMaxSeg) return HIP_PROXY_ERR_UNSUPP;
        G::Problem pr;
        memset(&pr, 0, sizeof(pr));
        pr.M = Am; pr.N = Bn;
        for (int s2 = 0; s2 < q; ++s2) {
            const int off = (row + col + s2) % q;
            const int lk = ce->k_displs[off + 1] - ce->k_displs[off];
            if (!lk) continue;
            G::Seg &sg = pr.seg[pr.nseg++];
            sg.A = (const double *) S->eng[cannon_member(P, r, row, off)]->A_2dmm_d;
            sg.B = (const double *) S->eng[cannon_member(P, r, off, col)]->B_2dmm_d;
            sg.lda = Am; sg.ldb = lk; sg.k = lk;
        }
        pr.C = (double *) (kp == 1 ? e->C_out_d : e->C_2dmm_d);
        pr.ldc = Am;
        pr.fold = (kp > 1) ? fold_of[r % mn] : -1;
        probs.push_back(pr);
        if (ranks) ranks->push_back(r);
    }
    return 0;
}

int build_grouped(ca3dmm_sim *S)
{
    std::vector<ca3::grouped::Problem> probs;
    std::vector<ca3::grouped::Fold> folds;
    int rc = collect_grouped(S, probs, folds);
    ++S->plan_gen;
    return rc ? rc : ca3::grouped::build(probs, folds, &S->plan);
}

int build_ozaki(ca3dmm_sim *S)
{
    std::vector<ca3::grouped::Problem> probs;
    std::vector<ca3::grouped::Fold> folds;
    int rc = collect_grouped(S, probs, folds);
    ++S->plan_gen;
    return rc ? rc : ca3::grouped::build_ozaki(probs, folds, S->ozaki, &S->ozplan);
}

// The per-rank Cannon rounds (step 3 below) as one batched launch unless
// CA3DMM_SIM_BATCH=0 asks for one hip_dgemm launch per rank and round.
bool batch_on()
{
    const char *v = getenv("CA3DMM_SIM_BATCH");
    return !(v && !strcmp(v, "0"));
}

// The problems of the batched launch: exactly the hip_dgemm calls of the
// per-rank rounds, one problem per rank with its q rounds as segments in
// round order (rounds with k = 0 included).
int build_batch(ca3dmm_sim *S)
{
    const Problem &P = S->P;
    namespace G = ca3::grouped;
    std::vector<G::Problem> probs;
    for (int r = 0; r < P.p; ++r) {
        ca3dmm_engine_p e = S->eng[r];
        if (!e->is_active) continue;
        const cannon_engine_p ce = e->cannon_engine;
        if (ce->m == 0 || ce->n == 0 || ce->k == 0) continue;
        const int q = ce->np_dim, row = ce->rank_row, col = ce->rank_col;
        const int Am = ce->m_displs[row + 1] - ce->m_displs[row];
        const int Bn = ce->n_displs[col + 1] - ce->n_displs[col];
        if (!Am || !Bn) continue;
        if (q > G::kMaxSeg) return HIP_PROXY_ERR_UNSUPP;
        G::Problem pr;
        memset(&pr, 0, sizeof(pr));
        pr.M = Am; pr.N = Bn;
        pr.C = (double *) (e->task_k_num == 1 ? e->C_out_d : e->C_2dmm_d);
        pr.ldc = std::max(1, Am);
        pr.fold = -1;
        for (int s2 = 0; s2 < q; ++s2) {
            const int off = (row + col + s2) % q;
            const int lk = ce->k_displs[off + 1] - ce->k_displs[off];
            G::Seg &sg = pr.seg[pr.nseg++];
            sg.A = (const double *) S->eng[cannon_member(P, r, row, off)]->A_2dmm_d;
            sg.B = (const double *) S->eng[cannon_member(P, r, off, col)]->B_2dmm_d;
            sg.lda = std::max(1, Am); sg.ldb = std::max(1, lk); sg.k = lk;
        }
        probs.push_back(pr);
    }
    ++S->plan_gen;
    return G::build_batch(probs, true, &S->batch);
}

// The automatic modulus count, on the operands in place: sample s of
// hip_ozaki_choose's 4096 (entry (i_s, j_s) of the whole C, same hashes)
// is taken in k layer l = s mod kp, from the local problem of the layer-l
// rank whose C block holds it, so a and b run over layer l's k range; the
// count rule then takes the longest layer's k.  Sets *L to the spread
// statistic it used.  Synchronous.
int choose_ozaki_count(ca3dmm_sim *S, int *nmod, double *L)
{
    const Problem &P = S->P;
    std::vector<ca3::grouped::Problem> probs;
    std::vector<ca3::grouped::Fold> folds;
    std::vector<int> ranks;
    int rc = collect_grouped(S, probs, folds, &ranks);
    if (rc) return rc;
    const int mC = P.btb ? P.n : P.m, nC = P.n, kp = P.grid.kp;
    int kmax = 0;
    for (int l = 0; l < kp; ++l) {
        int s0 = 0, len = 0;
        calc_block_spos_size(P.k, kp, l, &s0, &len);
        kmax = std::max(kmax, len);
    }
    std::vector<int> samples;
    if (mC > 0 && nC > 0) {
        for (uint64_t s = 0; s < 4096; ++s) {
            const int i = (int) (((s * 2654435761ull + 2135587861ull) & 0xffffffffull) % (uint64_t) mC);
            const int j = (int) (((s * 2246822519ull + 3266489917ull) & 0xffffffffull) % (uint64_t) nC);
            const int layer = (int) (s % (uint64_t) kp);
            for (size_t t = 0; t < probs.size(); ++t) {
                const ca3dmm_engine_p e = S->eng[ranks[t]];
                if (P.plans[ranks[t]].rank_k != layer) continue;
                const int li = i - e->C_2dmm_srow, lj = j - e->C_2dmm_scol;
                if (li < 0 || li >= probs[t].M || lj < 0 || lj >= probs[t].N) continue;
                samples.push_back((int) t);
                samples.push_back(li);
                samples.push_back(lj);
                break;
            }
        }
    }
    return ca3::grouped::choose_moduli(probs, samples, kmax, nmod, L);
}

// Steps 1 and 2 of the exec: every rank's operand blocks in place for its
// Cannon rounds.
int sim_operands(ca3dmm_sim *S, const void *const *src_A, const int *ldA, const void *const *src_B,
                 const int *ldB)
{
    const Problem &P = S->P;
    const int p = P.p;
    std::vector<mat_redist_engine_p> re(p);
    std::vector<const void *> sp(p);
    std::vector<void *> dp(p);
    std::vector<int> sl(p), dl(p);
    int rc = 0;
    // 1. redistribute A and B: one pass, both matrices in one batched launch
    //    (staged: pack, exchange and unpack per matrix)
    const bool staged = redist_staged();
    std::vector<hip_blk_copy_t> rd;
    if (!P.btb) {
        for (int r = 0; r < p; ++r) {
            re[r] = S->eng[r]->redist_A; sp[r] = src_A[r]; sl[r] = ldA[r];
            dp[r] = S->eng[r]->A_rd_recv_d;
            dl[r] = std::max(1, recv_ld(S->eng[r]->A_rd_nrow, S->eng[r]->A_rd_ncol, P.tA));
        }
        if (staged) {
            if ((rc = redist_staged_run(p, re.data(), sp.data(), sl.data(), dp.data(), dl.data()))) return rc;
        } else {
            redist_direct(p, re.data(), sp.data(), sl.data(), dp.data(), dl.data(), rd);
        }
    }
    for (int r = 0; r < p; ++r) {
        ca3dmm_engine_p e = S->eng[r];
        re[r] = e->redist_B; sp[r] = src_B[r]; sl[r] = ldB[r];
        dp[r] = P.btb ? e->B_2dmm_d : e->B_rd_recv_d;
        dl[r] = std::max(1, P.btb ? e->B_2dmm_nrow : recv_ld(e->B_rd_nrow, e->B_rd_ncol, P.tB));
    }
    if (staged) {
        if ((rc = redist_staged_run(p, re.data(), sp.data(), sl.data(), dp.data(), dl.data()))) return rc;
    } else {
        redist_direct(p, re.data(), sp.data(), sl.data(), dp.data(), dl.data(), rd);
        if (!rd.empty() && (rc = hip_copy_matrix_blocks(8, (int) rd.size(), rd.data()))) return rc;
    }

    // 2. transposes and the allgather of the replicated operand
    if (!P.btb) {
        for (int r = 0; r < p && !rc; ++r) {
            ca3dmm_engine_p e = S->eng[r];
            if (!e->is_active) continue;
            const int Atn = e->task_n_num > 1 ? e->A_rd_nrow : e->A_2dmm_nrow;
            const int Atc = e->task_n_num > 1 ? e->A_rd_ncol : e->A_2dmm_ncol;
            const int Btn = e->task_m_num > 1 ? e->B_rd_nrow : e->B_2dmm_nrow;
            const int Btc = e->task_m_num > 1 ? e->B_rd_ncol : e->B_2dmm_ncol;
            if (P.tA && Atn && Atc)
                rc = hip_transpose_cm(Atn, Atc, (const double *) e->A_rd_recv_d, Atc, (double *) e->A_trans_d,
                                      std::max(1, e->A_2dmm_nrow));
            if (!rc && P.tB && Btn && Btc)
                rc = hip_transpose_cm(Btn, Btc, (const double *) e->B_rd_recv_d, Btc, (double *) e->B_trans_d,
                                      std::max(1, Btn));
        }
        if (rc) return rc;
        std::vector<hip_blk_copy_t> d;
        for (int r = 0; r < p; ++r) {
            ca3dmm_engine_p e = S->eng[r];
            if (!e->is_active || (e->task_m_num == 1 && e->task_n_num == 1)) continue;
            const bool repA = e->task_n_num > 1;
            const RankPlan &pl = P.plans[r];
            const int ntask = repA ? e->task_n_num : e->task_m_num;
            for (int t = 0; t < ntask; ++t) {
                // member of the allgather group with task id t: same layer,
                // same position inside its Cannon group
                const int g = pl.rank_k * P.grid.mp * P.grid.np + t * pl.group + pl.rank2d;
                ca3dmm_engine_p src = S->eng[g];
                hip_blk_copy_t c;
                c.src = repA ? src->A_trans_d : src->B_trans_d;
                c.dst = (double *) (repA ? e->A_2dmm_d : e->B_2dmm_d) + e->AB_agv_displs[t];
                c.nrow = 1; c.ncol = e->AB_agv_recvcnts[t];
                c.lds = c.ldd = std::max(1, c.ncol);
                d.push_back(c);
            }
        }
        if ((rc = hip_copy_matrix_blocks(8, (int) d.size(), d.data()))) return rc;
    } else {
        // BTB: rank (row, col) needs (B block of its partner (col, row))^T
        for (int r = 0; r < p && !rc; ++r) {
            ca3dmm_engine_p e = S->eng[r];
            if (!e->is_active || !e->A_2dmm_nrow || !e->A_2dmm_ncol) continue;
            const cannon_engine_p ce = e->cannon_engine;
            const int partner = cannon_member(P, r, ce->rank_col, ce->rank_row);
            ca3dmm_engine_p pe = S->eng[partner];
            rc = hip_transpose_cm(e->A_2dmm_nrow, e->A_2dmm_ncol, (const double *) pe->B_2dmm_d,
                                  std::max(1, pe->B_2dmm_nrow), (double *) e->A_2dmm_d,
                                  std::max(1, e->A_2dmm_nrow));
        }
        if (rc) return rc;
    }
    return 0;
}

// Steps 3 to 5: the local products, the k-layer fold and the redistribution
// of C.
int sim_products(ca3dmm_sim *S, void *const *dst_C, const int *ldC)
{
    const Problem &P = S->P;
    const int p = P.p;
    int rc = 0;

    // 3.+4. through the CRT form of the Ozaki scheme: every rank's product
    //    over all its rounds, the k-layer fold in layer order on the
    //    reconstructed values (products in one persistent launch when grouped)
    const bool fp64_rounds = !S->ozaki;
    if (S->ozaki) {
        if ((rc = ca3::grouped::run_ozaki(S->ozplan, S->grouped != 0, false))) return rc;
    } else if (S->grouped) {
        // 3.+4. as one persistent grouped launch (rounds accumulated in
        //    registers, fold done by the last layer of each tile)
        if ((rc = ca3::grouped::run(S->plan, false))) return rc;
    }

    // 3. Cannon rounds: round s of rank (row, col) multiplies the A block of
    //    (row, (row+col+s) % q) by the B block of ((row+col+s) % q, col),
    //    beta = 0 in round 0 and 1 afterwards (cannon.c:319-433); all ranks'
    //    rounds in one batched launch (the same bits as the launches below)
    const bool batched = fp64_rounds && !S->grouped && S->batch && batch_on();
    if (batched && (rc = ca3::grouped::run_batch(S->batch))) return rc;
    for (int r = 0; r < p && !rc && !S->grouped && fp64_rounds && !batched; ++r) {
        ca3dmm_engine_p e = S->eng[r];
        if (!e->is_active) continue;
        const cannon_engine_p ce = e->cannon_engine;
        const int q = ce->np_dim, row = ce->rank_row, col = ce->rank_col;
        double *C = (double *) (e->task_k_num == 1 ? e->C_out_d : e->C_2dmm_d);
        if (ce->m == 0 || ce->n == 0 || ce->k == 0) continue;
        const int Am = ce->m_displs[row + 1] - ce->m_displs[row];
        const int Bn = ce->n_displs[col + 1] - ce->n_displs[col];
        for (int s = 0; s < q && !rc; ++s) {
            const int off = (row + col + s) % q;
            const int lk = ce->k_displs[off + 1] - ce->k_displs[off];
            const ca3dmm_engine_p ea = S->eng[cannon_member(P, r, row, off)];
            const ca3dmm_engine_p eb = S->eng[cannon_member(P, r, off, col)];
            rc = hip_dgemm(HipNoTrans, HipNoTrans, Am, Bn, lk, 1.0, (const double *) ea->A_2dmm_d,
                           std::max(1, Am), (const double *) eb->B_2dmm_d, std::max(1, lk), s == 0 ? 0.0 : 1.0,
                           C, std::max(1, Am));
        }
    }
    if (rc) return rc;

    // 4. k-layer reduction in rank (layer) order (ca3dmm.c:1287-1335)
    for (int r = 0; r < p && !rc && !S->grouped && fp64_rounds; ++r) {
        ca3dmm_engine_p e = S->eng[r];
        if (!e->is_active || e->task_k_num == 1 || !e->C_out_nrow || !e->C_out_ncol) continue;
        const int kp = e->task_k_num, mn = P.grid.mp * P.grid.np;
        const int rank_mn = r % mn, off = e->C_out_scol - e->C_2dmm_scol;
        std::vector<const double *> parts(kp);
        std::vector<int> lds(kp, e->C_2dmm_nrow);
        for (int l = 0; l < kp; ++l)
            parts[l] = (const double *) S->eng[l * mn + rank_mn]->C_2dmm_d + (size_t) off * e->C_2dmm_nrow;
        rc = hip_fold_sum(kp, parts.data(), lds.data(), e->C_out_nrow, e->C_out_ncol, (double *) e->C_out_d,
                          e->C_out_nrow);
    }
    if (rc) return rc;

    // 5. redistribute C
    if (P.have_dstC) {
        std::vector<mat_redist_engine_p> re(p);
        std::vector<const void *> sp(p);
        std::vector<void *> dp(p);
        std::vector<int> sl(p), dl(p);
        for (int r = 0; r < p; ++r) {
            ca3dmm_engine_p e = S->eng[r];
            re[r] = e->redist_C; sp[r] = e->C_out_d; sl[r] = std::max(1, e->C_out_nrow);
            dp[r] = dst_C ? dst_C[r] : nullptr; dl[r] = ldC ? ldC[r] : 1;
        }
        if ((rc = redist_run(p, re.data(), sp.data(), sl.data(), dp.data(), dl.data()))) return rc;
    }
    return 0;
}

// The exec sequence of all ranks, enqueued on the library stream without
// synchronising (so that it can be captured whole into a hipGraph).
int sim_sequence(ca3dmm_sim *S, const void *const *src_A, const int *ldA, const void *const *src_B,
                 const int *ldB, void *const *dst_C, const int *ldC)
{
    const int rc = sim_operands(S, src_A, ldA, src_B, ldB);
    return rc ? rc : sim_products(S, dst_C, ldC);
}

// What a captured sequence depends on: every argument of the exec, the
// mode, the tunings and the redistribution path.
std::vector<uintptr_t> sequence_key(ca3dmm_sim *S, const void *const *src_A, const int *ldA,
                                    const void *const *src_B, const int *ldB, void *const *dst_C,
                                    const int *ldC)
{
    std::vector<uintptr_t> k;
    const int p = S->P.p;
    k.push_back((uintptr_t) S->grouped);
    k.push_back((uintptr_t) S->ozaki);
    k.push_back((uintptr_t) S->plan_gen);
    k.push_back((uintptr_t) redist_staged());
    k.push_back((uintptr_t) batch_on());
    if (S->plan) {
        const ca3::grouped::Tuning &t = ca3::grouped::tuning(S->plan);
        k.push_back((uintptr_t) t.tile);
        k.push_back((uintptr_t) t.group_rows);
        k.push_back((uintptr_t) t.wg_per_cu);
    }
    if (S->ozplan) {
        const ca3::grouped::OzTuning &t = ca3::grouped::tuning_ozaki(S->ozplan);
        k.push_back((uintptr_t) t.tile_n);
        k.push_back((uintptr_t) t.xcd_queue);
    }
    for (int r = 0; r < p; ++r) {
        k.push_back(src_A ? (uintptr_t) src_A[r] : 0);
        k.push_back(ldA ? (uintptr_t) ldA[r] : 0);
        k.push_back((uintptr_t) src_B[r]);
        k.push_back((uintptr_t) ldB[r]);
        k.push_back(dst_C ? (uintptr_t) dst_C[r] : 0);
        k.push_back(ldC ? (uintptr_t) ldC[r] : 0);
    }
    return k;
}

// Run body on the library stream, or with use_graph replay it from the
// hipGraph of slot q (captured again when its key changes).
int run_sequence(ca3dmm_sim *S, int q, const std::vector<uintptr_t> &key, const std::function<int()> &body)
{
    if (!S->use_graph) return body();
    int rc = 0;
    if (!S->seq[q] || key != S->seq_key[q]) {
        (void) hip_stream_sync(nullptr);
        ca3::grouped::destroy_sequence(S->seq[q]);
        S->seq[q] = nullptr;
        S->seq_key[q].clear();
        rc = ca3::grouped::capture_sequence(&S->seq[q], body);
        if (!rc) S->seq_key[q] = key;
    }
    return rc ? rc : ca3::grouped::replay_sequence(S->seq[q]);
}

void drop_sequences(ca3dmm_sim *S)
{
    (void) hip_stream_sync(nullptr);
    for (int q = 0; q < 3; ++q) {
        ca3::grouped::destroy_sequence(S->seq[q]);
        S->seq[q] = nullptr;
        S->seq_key[q].clear();
    }
}

// One exec of the whole grid.  With use_graph the sequence -- the
// redistributions of A and B, transposes, allgathers, the products and the
// fold, the redistribution of C -- is captured into one hipGraph on the first
// call (again when an argument or setting changes) and replayed after that.
// With the automatic modulus count the operands are put in place first, the
// count is chosen on them (synchronous), the Ozaki plan is rebuilt when it
// changes, then the products run; use_graph then replays each half from a
// hipGraph of its own.
int sim_exec(ca3dmm_sim *S, const void *const *src_A, const int *ldA, const void *const *src_B,
             const int *ldB, void *const *dst_C, const int *ldC)
{
    int rc = 0;
    // the plans allocate and synchronise: build them outside any capture
    if (!S->ozaki_auto && S->ozaki && !S->ozplan && (rc = build_ozaki(S))) return rc;
    if (!S->ozaki && !S->ozaki_auto && S->grouped && !S->plan && (rc = build_grouped(S))) return rc;
    if (!S->ozaki && !S->ozaki_auto && !S->grouped && !S->batch && !S->batch_unsupported && batch_on()) {
        rc = build_batch(S);
        if (rc == HIP_PROXY_ERR_UNSUPP) {      // more Cannon rounds than a batch item holds: one launch per round
            S->batch_unsupported = true;
            rc = 0;
        }
        if (rc) return rc;
    }
    const bool save = ca3::grouped::sync_mode();
    hip_proxy_set_sync(0);
    const std::vector<uintptr_t> key = sequence_key(S, src_A, ldA, src_B, ldB, dst_C, ldC);
    if (!S->ozaki_auto) {
        rc = run_sequence(S, 0, key, [&]() { return sim_sequence(S, src_A, ldA, src_B, ldB, dst_C, ldC); });
    } else {
        rc = run_sequence(S, 1, key, [&]() { return sim_operands(S, src_A, ldA, src_B, ldB); });
        int N = 0;
        double L = NAN;
        if (!rc) rc = choose_ozaki_count(S, &N, &L);
        if (!rc && N != S->ozaki && S->ozplan) {
            (void) hip_stream_sync(nullptr);
            ca3::grouped::destroy_ozaki(S->ozplan);
            S->ozplan = nullptr;
        }
        if (!rc) {
            S->ozaki = N;
            S->ozaki_spread = L;
        }
        if (!rc && !S->ozplan) rc = build_ozaki(S);
        if (!rc)
            rc = run_sequence(S, 2, sequence_key(S, src_A, ldA, src_B, ldB, dst_C, ldC),
                              [&]() { return sim_products(S, dst_C, ldC); });
    }
    hip_proxy_set_sync(save ? 1 : 0);
    if (rc) return rc;
    return hip_stream_sync(nullptr);
}

}  // namespace

extern "C" {

int ca3dmm_sim_init(int p, int m, int n, int k, int trans_A, int trans_B, const int *src_A,
                    const int *src_B, const int *dst_C, const int *proc_grid, int is_btb,
                    dev_type_t dev_type, int allocate, ca3dmm_sim_p *sim)
{
    *sim = nullptr;
    if (p < 1 || !src_B || (!is_btb && !src_A)) return HIP_PROXY_ERR_ARG;
    if (!check_dev(dev_type)) return HIP_PROXY_ERR_ARG;
    auto *S = new ca3dmm_sim;
    Problem &P = S->P;
    P.p = p; P.btb = is_btb != 0;
    P.m = P.btb ? n : m; P.n = n; P.k = k;
    P.tA = P.btb ? 0 : (trans_A ? 1 : 0);
    P.tB = P.btb ? 0 : (trans_B ? 1 : 0);
    P.dev_type = dev_type;
    if (choose_grid(p, P.m, n, k, proc_grid, P.btb, &P.grid)) { delete S; return HIP_PROXY_ERR_ARG; }
    for (int r = 0; r < p; ++r) {
        P.plans.push_back(plan_rank(P.grid, P.m, n, k, P.btb, r));
        if (!P.btb) P.srcA.push_back(Rect{src_A[4 * r], src_A[4 * r + 1], src_A[4 * r + 2], src_A[4 * r + 3]});
        P.srcB.push_back(Rect{src_B[4 * r], src_B[4 * r + 1], src_B[4 * r + 2], src_B[4 * r + 3]});
        if (dst_C) P.dstC.push_back(Rect{dst_C[4 * r], dst_C[4 * r + 1], dst_C[4 * r + 2], dst_C[4 * r + 3]});
    }
    P.have_dstC = dst_C != nullptr;
    for (int r = 0; r < p; ++r) {
        ca3dmm_engine_p e = build_rank(P, r);
        if (!e) { ca3dmm_sim_p s = S; ca3dmm_sim_free(&s); return HIP_PROXY_ERR_ARG; }
        S->eng.push_back(e);
        if (!allocate) continue;
        const size_t bytes = total_bytes(e);
        e->alloc_workbuf = 1;
        void *wb = bytes ? dev_type_malloc(bytes, dev_type) : nullptr;
        if (!wb && bytes) { ca3dmm_sim_p s = S; ca3dmm_sim_free(&s); return HIP_PROXY_ERR_NOMEM; }
        ca3dmm_engine_attach_workbuf(e, nullptr, wb);
    }
    *sim = S;
    return HIP_PROXY_OK;
}

size_t ca3dmm_sim_workbuf_bytes(ca3dmm_sim_p sim, int rank)
{
    ca3dmm_engine_p e = ca3dmm_sim_engine(sim, rank);
    return e ? total_bytes(e) : 0;
}

ca3dmm_engine_p ca3dmm_sim_engine(ca3dmm_sim_p sim, int rank)
{
    if (!sim || rank < 0 || rank >= (int) sim->eng.size()) return nullptr;
    return sim->eng[rank];
}

void *ca3dmm_sim_c_out(ca3dmm_sim_p sim, int rank)
{
    ca3dmm_engine_p e = ca3dmm_sim_engine(sim, rank);
    return e ? e->C_out_d : nullptr;
}

int ca3dmm_sim_exec(ca3dmm_sim_p sim, const void *const *src_A, const int *ldA, const void *const *src_B,
                    const int *ldB, void *const *dst_C, const int *ldC)
{
    if (!sim) return HIP_PROXY_ERR_ARG;
    for (auto e : sim->eng)
        if (!e->workbuf_d && total_bytes(e)) {
            ENG_ERR("ca3dmm_sim_exec: grid was planned without work buffers\n");
            return HIP_PROXY_ERR_ARG;
        }
    const double t0 = now_ms();
    int rc = sim_exec(sim, src_A, ldA, src_B, ldB, dst_C, ldC);
    const double dt = now_ms() - t0;
    for (auto e : sim->eng) { e->exec_ms += dt; e->n_exec++; }
    if (rc) ENG_ERR("ca3dmm_sim_exec failed (%d): %s\n", rc, hip_proxy_last_error());
    return rc;
}

int ca3dmm_sim_set_mode(ca3dmm_sim_p sim, int grouped, int use_graph)
{
    if (!sim) return HIP_PROXY_ERR_ARG;
    sim->grouped = grouped ? 1 : 0;
    sim->use_graph = use_graph ? 1 : 0;
    return HIP_PROXY_OK;
}

int ca3dmm_sim_set_ozaki(ca3dmm_sim_p sim, int nmoduli)
{
    if (!sim || nmoduli < -1 || (nmoduli > 0 && (nmoduli < 4 || nmoduli > 25))) return HIP_PROXY_ERR_ARG;
    const int auto_count = nmoduli == -1;
    if (auto_count != sim->ozaki_auto) drop_sequences(sim);
    sim->ozaki_auto = auto_count;
    if (auto_count) return HIP_PROXY_OK;      // the count and plan of the last exec stay until the next
    sim->ozaki_spread = NAN;
    if (nmoduli != sim->ozaki && sim->ozplan) {
        drop_sequences(sim);
        ca3::grouped::destroy_ozaki(sim->ozplan);
        sim->ozplan = nullptr;
    }
    sim->ozaki = nmoduli;
    return HIP_PROXY_OK;
}

int ca3dmm_sim_get_ozaki(ca3dmm_sim_p sim, int *nmoduli, int *automatic, double *spread)
{
    if (!sim) return HIP_PROXY_ERR_ARG;
    if (nmoduli) *nmoduli = sim->ozaki;
    if (automatic) *automatic = sim->ozaki_auto;
    if (spread) *spread = sim->ozaki_spread;
    return HIP_PROXY_OK;
}

int ca3dmm_sim_autotune_ozaki(ca3dmm_sim_p sim, int reps, int *tile_n, int *xcd_queue)
{
    if (!sim || !sim->ozaki) return HIP_PROXY_ERR_ARG;
    for (auto e : sim->eng)
        if (!e->workbuf_d && total_bytes(e)) return HIP_PROXY_ERR_ARG;
    int rc = 0;
    if (!sim->ozplan && (rc = build_ozaki(sim))) return rc;
    if ((rc = ca3::grouped::autotune_ozaki(sim->ozplan, std::max(1, reps)))) return rc;
    const ca3::grouped::OzTuning &t = ca3::grouped::tuning_ozaki(sim->ozplan);
    if (tile_n) *tile_n = t.tile_n;
    if (xcd_queue) *xcd_queue = t.xcd_queue;
    return hip_stream_sync(nullptr);
}

int ca3dmm_sim_set_tuning_ozaki(ca3dmm_sim_p sim, int tile_n, int xcd_queue)
{
    if (!sim || !sim->ozaki || (tile_n != 256 && tile_n != 128)) return HIP_PROXY_ERR_ARG;
    int rc = 0;
    if (!sim->ozplan && (rc = build_ozaki(sim))) return rc;
    ca3::grouped::OzTuning t;
    t.tile_n = tile_n;
    t.xcd_queue = xcd_queue ? 1 : 0;
    ca3::grouped::set_tuning_ozaki(sim->ozplan, t);
    return HIP_PROXY_OK;
}

int ca3dmm_sim_autotune(ca3dmm_sim_p sim, int reps, int *group_rows, int *wg_per_cu, int *tile)
{
    if (!sim) return HIP_PROXY_ERR_ARG;
    for (auto e : sim->eng)
        if (!e->workbuf_d && total_bytes(e)) return HIP_PROXY_ERR_ARG;
    int rc = 0;
    if (!sim->plan && (rc = build_grouped(sim))) return rc;
    if ((rc = ca3::grouped::autotune(sim->plan, std::max(1, reps)))) return rc;
    const ca3::grouped::Tuning &t = ca3::grouped::tuning(sim->plan);
    if (group_rows) *group_rows = t.group_rows;
    if (wg_per_cu) *wg_per_cu = t.wg_per_cu;
    if (tile) *tile = t.tile;
    return hip_stream_sync(nullptr);
}

int ca3dmm_sim_set_tuning(ca3dmm_sim_p sim, int group_rows, int wg_per_cu, int tile)
{
    if (!sim || group_rows < 1 || wg_per_cu < 1 || wg_per_cu > 2 || tile < 0 || tile > 3 ||
        ((tile == 1 || tile == 2) && wg_per_cu != 1))
        return HIP_PROXY_ERR_ARG;
    int rc = 0;
    if (!sim->plan && (rc = build_grouped(sim))) return rc;
    ca3::grouped::Tuning t;
    t.tile = tile;
    t.group_rows = group_rows;
    t.wg_per_cu = wg_per_cu;
    ca3::grouped::set_tuning(sim->plan, t);
    return HIP_PROXY_OK;
}

void ca3dmm_sim_free(ca3dmm_sim_p *sim)
{
    if (!sim || !*sim) return;
    drop_sequences(*sim);
    ca3::grouped::destroy_batch((*sim)->batch);
    if ((*sim)->plan) ca3::grouped::destroy((*sim)->plan);
    if ((*sim)->ozplan) {
        (void) hip_stream_sync(nullptr);
        ca3::grouped::destroy_ozaki((*sim)->ozplan);
    }
    for (auto &e : (*sim)->eng) ca3dmm_engine_free(&e);
    delete *sim;
    *sim = nullptr;
}

}  // extern "C"
