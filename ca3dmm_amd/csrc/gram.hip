// Upper-triangle Gram product C = A^T A on FP64 MFMA (gfx950).
//
// The reference computes B^T B in full through the general Cannon DGEMM
// (src/ca3dmm.c:1172-1265) and the CholeskyQR driver uses it for
// S = B^T B (examples/example_CholeskyQR.c:109-117, :160-168).  Here only
// tiles (i <= j) of the upper triangle are computed: n(n+1)k flops instead
// of 2n^2 k.  Tall-skinny inputs (k >> n, e.g. 2e6 x 512) give few tiles,
// so k is split across workgroups; each split writes its partial tile and a
// second kernel adds the partials in split order (deterministic) and writes
// the upper triangle, optionally mirrored into the lower one.  Mirroring
// copies values, so the result is exactly symmetric.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "ca3dmm_ext.h"
#include "dgemm_tile.h"
#include "hip_common.h"

// Partial tiles of hip_dsyrk_gram_noalloc: a fixed array in device memory,
// 512 tiles of 128 x 128 doubles (64 MiB) per device, held until the process
// ends.  Host code must use the address hipGetSymbolAddress returns for the
// current device, never this name.
static __device__ __align__(256) double scratch[(size_t) 512 * 128 * 128];

namespace {

using namespace ca3::gemm;

struct GramArgs {
    int n, k, ntile, npairs, ksplit, kstep;
    const double *A;
    int64_t lda;
    double *P;      // partials: [ksplit][npairs][128 * 128]
};

// pair index -> (ti, tj), ti <= tj, row-major over the upper triangle
__device__ __forceinline__ void pair_of(int p, int nt, int *ti, int *tj)
{
    int i = 0, rem = p;
    while (rem >= nt - i) { rem -= nt - i; ++i; }
    *ti = i;
    *tj = i + rem;
}

// W: 4 for 16-byte direct-to-LDS loads (A 16-byte aligned, even lda), 0 for
// register-staged stages (both panels are k-contiguous, where 4-byte
// direct-to-LDS loads lose to register staging).
template <int W>
__global__ __launch_bounds__(kThreads, 2) void gram_partial_kernel(GramArgs g)
{
    __shared__ __attribute__((aligned(1024))) double A0[kBK * kBM], A1[kBK * kBM];
    __shared__ __attribute__((aligned(1024))) double B0[kBK * kBN], B1[kBK * kBN];
    // Workgroups are dealt round-robin over the 8 XCDs; the logical ids are
    // made XCD-contiguous (the grid is a multiple of 8) so that the pairs of
    // one k-range, which share their panels, run behind the same L2.
    const int nb = gridDim.x, total = g.npairs * g.ksplit;
    const int bid = (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
    if (bid >= total) return;
    const int pair = bid % g.npairs, ks = bid / g.npairs;
    int ti, tj;
    pair_of(pair, g.ntile, &ti, &tj);
    const int kbeg = ks * g.kstep, kend = min(g.k, kbeg + g.kstep);
    // op(A) = A^T and op(B) = A: both panels are contiguous along k
    OperandLoader<0> la, lb;
    la.init(g.A, g.lda, ti * kBM, g.n);
    lb.init(g.A, g.lda, tj * kBN, g.n);
    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    const bool interior = (ti + 1) * kBM <= g.n && (tj + 1) * kBN <= g.n;
    int kt0 = kbeg;
    if (W > 0 && kend - kbeg >= kBK) {
        // whole stages direct-to-LDS (kbeg is a multiple of 16, so the panels
        // keep A's alignment; edge tiles load copies of their last line in
        // place of the missing ones); the same products in the same order
        DtlLoader<0, 4> da, db;
        da.init(g.A + kbeg, g.lda, ti * kBM, min(kBM, g.n - ti * kBM));
        db.init(g.A + kbeg, g.lda, tj * kBN, min(kBN, g.n - tj * kBN));
        const int nfull = (kend - kbeg) / kBK;
        mainloop_dtl<0, 0>(da, db, g.lda, g.lda, nfull, A0, A1, B0, B1, acc);
        kt0 = kbeg + nfull * kBK;
    }
    if (kt0 < kend) mainloop(la, lb, kt0, kend, interior, LdsPair{A0, A1}, LdsPair{B0, B1}, acc);
    // store the full 128 x 128 partial (rows fastest)
    double *out = g.P + ((int64_t) ks * g.npairs + pair) * (kBM * kBN);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = wm * 64 + i * 16 + (lane & 15);
                const int col = wn * 64 + j * 16 + 4 * r + (lane >> 4);
                out[col * kBM + row] = acc[i][j][r];
            }
}

__global__ __launch_bounds__(256) void gram_reduce_kernel(const double *__restrict__ P, int n, int ntile, int npairs,
                                                          int ksplit, double alpha, double beta, double *C,
                                                          int64_t ldc, int mirror)
{
    const int pair = blockIdx.y;
    int ti, tj;
    pair_of(pair, ntile, &ti, &tj);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < kBM * kBN; e += gridDim.x * blockDim.x) {
        const int row = e % kBM, col = e / kBM;
        const int gi = ti * kBM + row, gj = tj * kBN + col;
        if (gi >= n || gj >= n || gi > gj) continue;
        double s = P[(int64_t) pair * (kBM * kBN) + e];
        for (int ks = 1; ks < ksplit; ++ks) s = ca3::add_rn(s, P[((int64_t) ks * npairs + pair) * (kBM * kBN) + e]);
        double *c = C + (int64_t) gj * ldc + gi;
        double v = ca3::mul_rn(alpha, s);
        if (beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(beta, *c));
        *c = v;
        if (mirror && gi != gj) C[(int64_t) gi * ldc + gj] = v;
    }
}

struct Ws {
    void *p = nullptr;
    size_t bytes = 0;
};
Ws g_gram_ws[16];
std::mutex g_gram_mu;

constexpr int kScratchTiles = 512;   // partial tiles the allocation-free entry can hold
double *g_scratch_ptr[16];          // device address of `scratch` per device (0: not looked up yet)

// Tiles, k-split and k-range per split for n x n from k.  Enough workgroups
// for one round of the chip at two per CU (at most kScratchTiles partial
// tiles on a 256-CU part), stages of >= 1024 k, whole 16-k stages per split.
void plan(int n, int k, GramArgs *g)
{
    memset(g, 0, sizeof(*g));
    const int ntile = (n + kBM - 1) / kBM, npairs = ntile * (ntile + 1) / 2;
    const int target = 2 * ca3::cu_count();
    int ksplit = std::max(1, std::min(target / npairs, (k + 1023) / 1024));
    if (k == 0) ksplit = 1;
    const int kstep = std::max(kBK, ((k + ksplit - 1) / ksplit + kBK - 1) / kBK * kBK);
    ksplit = std::max(1, (k + kstep - 1) / kstep);
    g->n = n; g->k = k; g->ntile = ntile; g->npairs = npairs; g->ksplit = ksplit; g->kstep = kstep;
}

int check_args(int n, int k, const double *A, int ldA, const double *C, int ldC)
{
    if (n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "gram: negative size");
    if (n > 0 && (!C || ldC < n)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "gram: bad C / ldC");
    if (n > 0 && k > 0 && (!A || ldA < std::max(1, k))) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "gram: bad A / ldA");
    return 0;
}

int launch(const GramArgs &g, double alpha, double beta, double *C, int ldC, int mirror)
{
    hipStream_t st = ca3::stream();
    const int grid = (g.npairs * g.ksplit + 7) / 8 * 8;
    // 16-byte direct-to-LDS loads from any 8-byte aligned A (as the DGEMM;
    // CA3DMM_DGEMM_LOADS=aligned keeps them to 16-byte aligned A with even
    // lda, "reg" and "dword" stage through registers)
    const LoadMode mode = load_mode();
    const bool aligned16 = ((uintptr_t) g.A % 16) == 0 && (g.lda % 2) == 0;
    if (dtl_ld_ok(g.lda) && ((uintptr_t) g.A % 8) == 0 && (mode == kLoads16 || (mode == kLoadsAligned && aligned16)))
        gram_partial_kernel<4><<<grid, kThreads, 0, st>>>(g);
    else
        gram_partial_kernel<0><<<grid, kThreads, 0, st>>>(g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "gram_partial_kernel");
    gram_reduce_kernel<<<dim3(16, g.npairs), 256, 0, st>>>(g.P, g.n, g.ntile, g.npairs, g.ksplit, alpha, beta, C,
                                                          ldC, mirror ? 1 : 0);
    return ca3::finish();
}

}  // namespace

extern "C" {

int hip_dsyrk_gram(const int n, const int k, const double alpha, const double *A, const int ldA,
                   const double beta, double *C, const int ldC, int mirror)
{
    int rc = check_args(n, k, A, ldA, C, ldC);
    if (rc || n == 0) return rc;
    GramArgs g;
    plan(n, k, &g);
    const size_t bytes = sizeof(double) * (size_t) g.ksplit * g.npairs * kBM * kBN;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    Ws &w = g_gram_ws[dev & 15];
    {
        std::lock_guard<std::mutex> lk(g_gram_mu);
        if (w.bytes < bytes) {
            if (w.p) {
                (void) hipStreamSynchronize(ca3::stream());
                (void) hipFree(w.p);
                w.p = nullptr;
                w.bytes = 0;
            }
            CA3_CHECK(hipMalloc(&w.p, bytes));
            w.bytes = bytes;
        }
    }
    g.A = A; g.lda = ldA; g.P = (double *) w.p;
    return launch(g, alpha, beta, C, ldC, mirror);
}

int hip_dsyrk_gram_noalloc(const int n, const int k, const double alpha, const double *A, const int ldA,
                           const double beta, double *C, const int ldC, int mirror)
{
    int rc = check_args(n, k, A, ldA, C, ldC);
    if (rc || n == 0) return rc;
    GramArgs g;
    plan(n, k, &g);
    static_assert(sizeof(scratch) == sizeof(double) * (size_t) kScratchTiles * kBM * kBN, "scratch size");
    if ((int64_t) g.ksplit * g.npairs > kScratchTiles)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG,
                             "gram_noalloc: n = %d, k = %d needs %lld partial 128 x 128 tiles; the fixed scratch "
                             "holds %d (n <= 3968)", n, k, (long long) g.ksplit * g.npairs, kScratchTiles);
    int dev = 0;
    CA3_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 16) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "gram_noalloc: device %d >= 16", dev);
    double *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_gram_mu);
        if (!g_scratch_ptr[dev]) {
            // once per device, with that device current: the host-side name
            // `scratch` is not a device address
            void *p = nullptr;
            CA3_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_ptr[dev] = (double *) p;
        }
        ptr = g_scratch_ptr[dev];
    }
    g.A = A; g.lda = ldA; g.P = ptr;
    return launch(g, alpha, beta, C, ldC, mirror);
}

}  // extern "C"
