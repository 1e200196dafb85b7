// FP64 MFMA DGEMM for gfx950 -- the local DGEMM of CA3DMM's Cannon step.
//
// Replaces cublasDgemm behind cuda_cublas_dgemm (src/cuda_proxy.cu:211-230),
// called at src/cannon.c:418,586,616,665 and src/dev_type.c:212.
//
// Column-major C := alpha*op(A)*op(B) + beta*C, optionally followed by the
// k-layer partial-C fold of the reduce-scatter (src/ca3dmm.c:1311-1324):
// C := ((P0 + P1) + ...) + (alpha*op(A)*op(B) + beta*C).
//
// Design (MI355X / CDNA4):
//  * v_mfma_f64_16x16x4_f64.  A workgroup of 4 wave64s owns a 128x128 tile
//    of C; each wave owns 64x64 = 4x4 MFMA blocks (64 accumulators per lane).
//  * K advances in stages of 16.  A and B stage tiles are held in LDS as
//    [k][m] and [k][n] (m / n fastest), double-buffered: the next stage's
//    global loads are issued before the current stage's MFMAs and written to
//    the other buffer afterwards, one barrier per stage.
//  * Transposes are fused into the loads: an operand is either contiguous
//    along m/n (A as stored, B^T) or along k (A^T, B as stored); both map to
//    coalesced global reads, and the LDS image is identical, so the MFMA
//    loop is the same for all four cases.
//  * LDS addresses are XOR-swizzled per k row, so both the transposing
//    stores (16 k rows, one column) and the MFMA operand reads (16
//    consecutive m or n, two k rows per half wave) hit distinct banks.
//  * Tiles are visited XCD-contiguously (workgroup b runs on XCD b % 8) and
//    in groups of 16 tile-rows, so the tiles resident on one XCD share A and
//    B panels in that XCD's L2.
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>

#include "ca3dmm_ext.h"
#include "dgemm_tile.h"
#include "grouped.h"
#include "hip_common.h"

namespace {

using namespace ca3::gemm;

template <int TA, int TB, class TL>
__global__ __launch_bounds__(TL::THREADS, TL::MIN_BLOCKS) void dgemm_kernel(GemmArgs g)
{
    __shared__ double As[2][kBK * TL::BM];
    __shared__ double Bs[2][kBK * TL::BN];

    int tm, tn;
    tile_coords<TL>(blockIdx.x, gridDim.x, g.M, g.N, &tm, &tn, g.group);
    const int m0 = tm * TL::BM, n0 = tn * TL::BN;

    // A as stored (TA = 0) is contiguous along m; B as stored (TB = 0) is
    // contiguous along k, B^T along n.
    OperandLoader<!TA, TL::BM, TL::THREADS> la;
    OperandLoader<TB, TL::BN, TL::THREADS> lb;
    la.init(g.A, g.lda, m0, g.M);
    lb.init(g.B, g.ldb, n0, g.N);

    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    mainloop<TL>(la, lb, 0, g.K, m0 + TL::BM <= g.M && n0 + TL::BN <= g.N, As, Bs, acc);

    epilogue<TL>(g, m0, n0, acc);
}

// The 128 x 128 tile with direct-to-LDS stages for its whole k-stages (the
// k tail through the register-staged checked path).  WA / WB: 4 for 16-byte
// loads of an operand, 1 for 4-byte loads.  Edge
// tiles load copies of their last row / column in place of the missing
// ones; with 16-byte loads that needs an even count of existing rows
// (columns) where the operand is contiguous along m (n), otherwise the tile
// runs the checked path throughout.  Same products in the same order as
// dgemm_kernel<TA, TB, Tile128>.
template <int TA, int TB, int WA, int WB>
__global__ __launch_bounds__(kThreads, 2) void dgemm_dtl_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(1024))) double A0[kBK * kBM], A1[kBK * kBM];
    __shared__ __attribute__((aligned(1024))) double B0[kBK * kBN], B1[kBK * kBN];

    int tm, tn;
    tile_coords<Tile128>(blockIdx.x, gridDim.x, g.M, g.N, &tm, &tn, g.group);
    const int m0 = tm * kBM, n0 = tn * kBN;

    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    int kt0 = 0;
    const int mn = min(kBM, g.M - m0), nn = min(kBN, g.N - n0);
    if ((TA || WA == 1 || (mn & 1) == 0) && (!TB || WB == 1 || (nn & 1) == 0) && g.K >= kBK) {
        DtlLoader<!TA, WA> da;
        DtlLoader<TB, WB> db;
        da.init(g.A, g.lda, m0, mn);
        db.init(g.B, g.ldb, n0, nn);
        const int nfull = g.K / kBK;
        mainloop_dtl<!TA, TB>(da, db, g.lda, g.ldb, nfull, A0, A1, B0, B1, acc);
        kt0 = nfull * kBK;
    }
    if (kt0 < g.K) {
        OperandLoader<!TA> la;
        OperandLoader<TB> lb;
        la.init(g.A, g.lda, m0, g.M);
        lb.init(g.B, g.ldb, n0, g.N);
        mainloop_checked(la, lb, kt0, g.K, LdsPair{A0, A1}, LdsPair{B0, B1}, acc);
    }
    epilogue(g, m0, n0, acc);
}

// One problem of a batched launch (rank r's product C := sum over its
// Cannon rounds of A_s B_s, every operand as stored, NN): its tiles are
// linear tile indices [base, base + tiles_m tiles_n).
struct BatchProb {
    int M, N, nseg;
    int dtl;                                   // bit s: segment s may take the direct-to-LDS stages
    int64_t base;
    double *C;
    int64_t ldc;
    const double *A[ca3::grouped::kMaxSeg];
    const double *B[ca3::grouped::kMaxSeg];
    int64_t lda[ca3::grouped::kMaxSeg], ldb[ca3::grouped::kMaxSeg];
    int k[ca3::grouped::kMaxSeg];
};

// Every problem's 128 x 128 tiles in one launch, one workgroup per tile
// (gridDim.x = total tiles rounded up to a multiple of 8).  Linear tile
// indices are laid out XCD-contiguously over the whole batch, so each XCD
// works through its own problems' panels; inside a problem tiles are
// walked in groups of `group` tile-rows.  A tile's segments run in order.
// PER_SEG: each segment is its own product added by the epilogue (beta 0
// on the first, 1 after), the bits of one dgemm_dtl_kernel<0, 0> launch
// per segment.  Otherwise all segments accumulate in registers and the tile
// is stored once (C = the sum), the bits of the persistent grouped kernel.
template <bool PER_SEG>
__global__ __launch_bounds__(kThreads, 2) void dgemm_batch_kernel(const BatchProb *__restrict__ probs, int nprob,
                                                                    int64_t total, int group)
{
    __shared__ __attribute__((aligned(1024))) double A0[kBK * kBM], A1[kBK * kBM];
    __shared__ __attribute__((aligned(1024))) double B0[kBK * kBN], B1[kBK * kBN];

    const int64_t nb = gridDim.x;
    int64_t t = blockIdx.x;
    t = (t & 7) * (nb >> 3) + (t >> 3);
    if (t >= total) return;
    int lo = 0, hi = nprob - 1;                // last problem with base <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (probs[mid].base <= t) lo = mid;
        else hi = mid - 1;
    }
    const BatchProb &P = probs[lo];
    int tm, tn;
    tile_coords<Tile128>((int) (t - P.base), 1, P.M, P.N, &tm, &tn, group);   // nb = 1: no XCD remap here
    const int m0 = tm * kBM, n0 = tn * kBN;
    const int mn = min(kBM, P.M - m0), nn = min(kBN, P.N - n0);

    GemmArgs g;
    g.M = P.M; g.N = P.N;
    g.K = 0;                                  // not read by the epilogue
    g.alpha = 1.0;
    g.beta = 0.0;
    g.C = P.C; g.ldc = P.ldc;
    g.nsrc = 0;
    f64x4 acc[4][4];
    for (int s = 0; s < P.nseg; ++s) {
        if (PER_SEG || s == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
        const int K = P.k[s];
        int kt0 = 0;
        int sm0 = m0, sn0 = n0, smn = mn, snn = nn;
        asm volatile("" : "+s"(sm0), "+s"(sn0), "+s"(smn), "+s"(snn));
        if (((P.dtl >> s) & 1) && (smn & 1) == 0 && K >= kBK) {
            DtlLoader<1, 4> da;
            DtlLoader<0, 4> db;
            da.init(P.A[s], P.lda[s], sm0, smn);
            db.init(P.B[s], P.ldb[s], sn0, snn);
            const int nfull = K / kBK;
            mainloop_dtl<1, 0>(da, db, P.lda[s], P.ldb[s], nfull, A0, A1, B0, B1, acc);
            kt0 = nfull * kBK;
        }
        if (kt0 < K) {
            OperandLoader<1> la;
            OperandLoader<0> lb;
            la.init(P.A[s], P.lda[s], sm0, P.M);
            lb.init(P.B[s], P.ldb[s], sn0, P.N);
            mainloop_checked(la, lb, kt0, K, LdsPair{A0, A1}, LdsPair{B0, B1}, acc);
        }
        if (PER_SEG) {
            // the epilogue's addresses are the same on every segment: keep
            // the compiler from computing them once and holding them (in
            // vector registers) across the main loops
            double *C = P.C;
            int64_t ldc = P.ldc;
            int em0 = m0, en0 = n0;
            asm volatile("" : "+s"(C), "+s"(ldc), "+s"(em0), "+s"(en0));
            g.C = C;
            g.ldc = ldc;
            g.beta = s ? 1.0 : 0.0;
            epilogue(g, em0, en0, acc);
        }
    }
    if (!PER_SEG) {
        if (P.nseg == 0) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
        epilogue(g, m0, n0, acc);             // alpha = 1, beta = 0: C = the sum as accumulated
    }
}

template <int TA, int TB, int WA, int WB>
int launch_dtl(const GemmArgs &g)
{
    const int tiles = ((g.M + kBM - 1) / kBM) * ((g.N + kBN - 1) / kBN);
    dgemm_dtl_kernel<TA, TB, WA, WB><<<tiles, kThreads, 0, ca3::stream()>>>(g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "dgemm_dtl_kernel");
    return HIP_PROXY_OK;
}

template <int TA, int TB, class TL>
int launch_tile(const GemmArgs &g)
{
    const int tiles = ((g.M + TL::BM - 1) / TL::BM) * ((g.N + TL::BN - 1) / TL::BN);
    dgemm_kernel<TA, TB, TL><<<tiles, TL::THREADS, 0, ca3::stream()>>>(g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "dgemm_kernel");
    return HIP_PROXY_OK;
}

// Workgroup tile: 128 x 128 (4 waves, two workgroups per CU) unless
// CA3DMM_DGEMM_TILE selects 256x128 or 128x256 (8 waves, one per CU).
// The 128 x 128 tile loads its whole k-stages direct-to-LDS (6-16 % faster
// than register staging, same bits) with 16-byte loads, whatever the
// operand's leading dimension and its address modulo 16 bytes: gfx950
// serves 16-byte global loads from 8-byte aligned addresses (the compiler
// itself emits them for such pointers), so a leading dimension of 1333
// only splits some loads over two cache lines.
// CA3DMM_DGEMM_LOADS=reg keeps the register-staged stages throughout;
// CA3DMM_DGEMM_LOADS=dword takes 4-byte loads for both operands (a test
// form); CA3DMM_DGEMM_LOADS=aligned takes 16-byte loads only for 16-byte
// aligned operands with even leading dimensions, 4-byte loads for other
// operands contiguous along m or n, and register staging otherwise.
int tile_choice()
{
    const char *v = getenv("CA3DMM_DGEMM_TILE");
    return (v && !strcmp(v, "256x128")) ? 1 : (v && !strcmp(v, "128x256")) ? 2 : 0;
}

// The direct-to-LDS load width of an operand (4 or 1 words), 0 for none
// (the product is then register-staged).  xc: the operand is contiguous
// along m or n.
int dtl_words(const double *P, int64_t ld, bool xc, LoadMode mode)
{
    if (!dtl_ld_ok(ld) || ((uintptr_t) P % 8) != 0 || mode == kLoadsReg) return 0;
    if (mode == kLoadsDword) return 1;
    if (mode == kLoadsAligned && (((uintptr_t) P % 16) != 0 || (ld % 2) != 0)) return xc ? 1 : 0;
    return 4;
}

template <int TA, int TB>
int launch(const GemmArgs &g)
{
    const int tc = tile_choice();
    const LoadMode mode = load_mode();
    if (tc == 0 && mode != kLoadsReg) {
        const int wa = dtl_words(g.A, g.lda, !TA, mode), wb = dtl_words(g.B, g.ldb, TB, mode);
        if (wa == 4 && wb == 4) return launch_dtl<TA, TB, 4, 4>(g);
        if (wa == 4 && wb == 1) return launch_dtl<TA, TB, 4, 1>(g);
        if (wa == 1 && wb == 4) return launch_dtl<TA, TB, 1, 4>(g);
        if (wa == 1 && wb == 1) return launch_dtl<TA, TB, 1, 1>(g);
    }
    switch (tc) {
    case 1: return launch_tile<TA, TB, Tile256x128>(g);
    case 2: return launch_tile<TA, TB, Tile128x256>(g);
    default: return launch_tile<TA, TB, Tile128>(g);
    }
}

int run(hip_proxy_enum_t transA, hip_proxy_enum_t transB, int m, int n, int k, double alpha,
        const double *A, int ldA, const double *B, int ldB, double beta, double *C, int ldC,
        int nsrc, const double *const *P, const int *ldp)
{
    const int ta = (transA != HipNoTrans), tb = (transB != HipNoTrans);
    if (m < 0 || n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: negative size");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (!C || ldC < std::max(1, m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: bad C / ldC");
    if (k > 0) {
        if (!A || ldA < std::max(1, ta ? k : m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: bad A / ldA");
        if (!B || ldB < std::max(1, tb ? n : k)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: bad B / ldB");
    }
    if (nsrc < 0 || nsrc > kMaxFoldSrc) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: nsrc %d", nsrc);
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.M = m; g.N = n; g.K = k;
    g.alpha = alpha; g.beta = beta;
    g.A = A; g.lda = ldA; g.B = B; g.ldb = ldB; g.C = C; g.ldc = ldC;
    g.nsrc = nsrc;
    {
        const char *v = getenv("CA3DMM_DGEMM_GROUP");
        const int G = v ? atoi(v) : 16;
        g.group = (G >= 1 && G <= 256) ? G : 16;
    }
    for (int t = 0; t < nsrc; ++t) {
        if (!P || !ldp || !P[t] || ldp[t] < m) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm: partial %d", t);
        g.P[t] = P[t];
        g.ldp[t] = ldp[t];
    }
    int rc;
    if (!ta && !tb) rc = launch<0, 0>(g);
    else if (!ta && tb) rc = launch<0, 1>(g);
    else if (ta && !tb) rc = launch<1, 0>(g);
    else rc = launch<1, 1>(g);
    if (rc) return rc;
    return ca3::finish();
}

}  // namespace

// ---------------------------------------------------------------------
// Batched launch of the simulated grid's per-rank products (grouped.h)
// ---------------------------------------------------------------------
namespace ca3 {
namespace grouped {

struct BatchPlan {
    BatchProb *d = nullptr;
    int nprob = 0;
    int64_t total = 0;
    int group = 16;
    bool per_segment = true;
};

int build_batch(const std::vector<Problem> &probs, bool per_segment, BatchPlan **out)
{
    *out = nullptr;
    std::vector<BatchProb> h;
    int64_t base = 0;
    const gemm::LoadMode mode = gemm::load_mode();
    for (const Problem &q : probs) {
        if (q.M <= 0 || q.N <= 0) continue;
        if (q.nseg < 0 || q.nseg > kMaxSeg || !q.C || q.ldc < q.M)
            return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm batch: bad problem");
        BatchProb b;
        memset(&b, 0, sizeof(b));
        b.M = q.M; b.N = q.N; b.nseg = q.nseg;
        b.base = base;
        b.C = q.C; b.ldc = q.ldc;
        for (int s = 0; s < q.nseg; ++s) {
            const Seg &g = q.seg[s];
            if (g.k < 0 || (g.k > 0 && (!g.A || !g.B || g.lda < q.M || g.ldb < g.k)))
                return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm batch: bad segment %d", s);
            b.A[s] = g.A; b.B[s] = g.B;
            b.lda[s] = g.lda; b.ldb[s] = g.ldb;
            b.k[s] = g.k;
            if (dtl_words(g.A, g.lda, true, mode) == 4 && dtl_words(g.B, g.ldb, false, mode) == 4)
                b.dtl |= 1 << s;
        }
        h.push_back(b);
        base += (int64_t) ((q.M + gemm::kBM - 1) / gemm::kBM) * ((q.N + gemm::kBN - 1) / gemm::kBN);
    }
    if ((base + 7) / 8 * 8 > INT_MAX) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dgemm batch: %lld tiles", (long long) base);
    auto *p = new BatchPlan;
    p->per_segment = per_segment;
    p->nprob = (int) h.size();
    p->total = base;
    const char *v = getenv("CA3DMM_DGEMM_GROUP");
    const int G = v ? atoi(v) : 16;
    p->group = (G >= 1 && G <= 256) ? G : 16;
    if (!h.empty()) {
        hipError_t e = hipMalloc(&p->d, h.size() * sizeof(BatchProb));
        if (e == hipSuccess) e = hipMemcpy(p->d, h.data(), h.size() * sizeof(BatchProb), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            destroy_batch(p);
            return ca3::fail(e, "dgemm batch plan");
        }
    }
    *out = p;
    return HIP_PROXY_OK;
}

int run_batch(BatchPlan *p, int group)
{
    if (!p) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgemm batch: no plan");
    if (p->total == 0) return HIP_PROXY_OK;
    const int grid = (int) ((p->total + 7) / 8 * 8);
    const int G = (group >= 1 && group <= 256) ? group : p->group;
    if (p->per_segment)
        dgemm_batch_kernel<true><<<grid, gemm::kThreads, 0, ca3::stream()>>>(p->d, p->nprob, p->total, G);
    else
        dgemm_batch_kernel<false><<<grid, gemm::kThreads, 0, ca3::stream()>>>(p->d, p->nprob, p->total, G);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "dgemm_batch_kernel");
    return ca3::finish();
}

void destroy_batch(BatchPlan *p)
{
    if (!p) return;
    if (p->d) (void) hipFree(p->d);
    delete p;
}

}  // namespace grouped
}  // namespace ca3

ca3::gemm::LoadMode ca3::gemm::load_mode()
{
    const char *v = getenv("CA3DMM_DGEMM_LOADS");
    if (!v) return kLoads16;
    if (!strcmp(v, "reg")) return kLoadsReg;
    if (!strcmp(v, "dword")) return kLoadsDword;
    if (!strcmp(v, "aligned")) return kLoadsAligned;
    return kLoads16;
}

extern "C" {

int hip_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n,
              const int k, const double alpha, const double *A, const int ldA, const double *B,
              const int ldB, const double beta, double *C, const int ldC)
{
    return run(transA, transB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, 0, nullptr, nullptr);
}

int hip_dgemm_fold(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n,
                   const int k, const double alpha, const double *A, const int ldA,
                   const double *B, const int ldB, const double beta, double *C, const int ldC,
                   int nsrc, const double *const *P, const int *ldp)
{
    return run(transA, transB, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, nsrc, P, ldp);
}

}  // extern "C"
