// Persistent grouped FP64 GEMM for a simulated CA3DMM grid on one GPU.
//
// No reference counterpart as a kernel: it runs, in ONE launch, the work of
// every rank's Cannon step (src/cannon.c:319-433: q rounds of C += A_s B_s)
// and the k-layer reduction of the reduce-scatter (src/ca3dmm.c:1287-1335).
//
//  * Work items are 128 x 128 tiles of every rank's C block; a persistent
//    grid (wg_per_cu workgroups per CU) pulls items from a device queue
//    (one atomic counter), so uneven cuboids balance themselves.
//  * A tile accumulates all q rounds (k-segments whose A and B blocks live
//    in different ranks' buffers) in registers and is stored once.
//  * Ranks that share an mn-position across the kp k-layers form a fold
//    group.  After storing its tile, a workgroup bumps the tile's completion
//    counter (release fence first); the workgroup that completes the count
//    (acquire fence) adds the kp layer tiles in layer order and writes each
//    column to the C_out slice of the layer that owns it.  Nobody waits, so
//    every workgroup exits once the queue is empty.
//  * Counters and the queue head are reset by a fill kernel in the same
//    stream (no memset nodes), and both launches can be captured in a
//    hipGraph and replayed.
//  * The tuner times the queue order (tile rows walked together) and the
//    persistent grid size and keeps the fastest per problem signature.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "dgemm_tile.h"
#include "grouped.h"
#include "hip_common.h"

namespace ca3 {
namespace grouped {

using namespace ca3::gemm;

struct Plan {
    int nprob = 0, nfold = 0, total = 0, ncounters = 0;
    Problem *d_probs = nullptr;
    Fold *d_folds = nullptr;
    int *d_tile0 = nullptr;       // [nprob + 1] first item of each problem
    int *d_ftile0 = nullptr;      // [nfold] first counter of each fold group
    int *d_counters = nullptr;    // [ncounters + 1]; last slot is the queue head
    Tuning tune;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    Tuning graph_tune;
    std::string signature;
};

namespace {

__global__ void zero_ints(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

__device__ __forceinline__ void grouped_coords(int t, int tiles_m, int tiles_n, int G, int *tm, int *tn)
{
    const int per = G * tiles_n, grp = t / per, first = grp * G;
    const int gsz = min(tiles_m - first, G), r = t - grp * per;
    *tm = first + r % gsz;
    *tn = r / gsz;
}

__global__ __launch_bounds__(kThreads, 2) void grouped_kernel(const Problem *__restrict__ probs, int nprob,
                                                               const int *__restrict__ tile0,
                                                               const Fold *__restrict__ folds,
                                                               const int *__restrict__ ftile0, int *counters,
                                                               int *queue, int total, int G)
{
    __shared__ double As[2][kBK * kBM];
    __shared__ double Bs[2][kBK * kBN];
    __shared__ int s_item, s_last;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(queue, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= total) break;
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile0[mid] <= item) lo = mid; else hi = mid - 1;
        }
        const Problem &P = probs[lo];
        const int t = item - tile0[lo];
        const int tiles_m = (P.M + kBM - 1) / kBM, tiles_n = (P.N + kBN - 1) / kBN;
        int tm, tn;
        grouped_coords(t, tiles_m, tiles_n, G, &tm, &tn);
        const int m0 = tm * kBM, n0 = tn * kBN;

        f64x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < P.nseg; ++s) {
            OperandLoader<1> la;     // A block as stored: contiguous along m
            OperandLoader<0> lb;     // B block as stored: contiguous along k
            la.init(P.seg[s].A, P.seg[s].lda, m0, P.M);
            lb.init(P.seg[s].B, P.seg[s].ldb, n0, P.N);
            mainloop(la, lb, 0, P.seg[s].k, m0 + kBM <= P.M && n0 + kBN <= P.N, As, Bs, acc);
        }
        // store the tile (lane l, register r of block (i, j): row l % 16, column 4r + l / 16)
        const int lane = tid & 63, wave = tid >> 6, wm = wave & 1, wn = wave >> 1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = m0 + wm * 64 + i * 16 + (lane & 15);
            if (row >= P.M) continue;
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = n0 + wn * 64 + j * 16 + 4 * r + (lane >> 4);
                    if (col < P.N) P.C[(int64_t) col * P.ldc + row] = acc[i][j][r];
                }
        }
        if (P.fold < 0) continue;
        const Fold &F = folds[P.fold];
        __threadfence();                      // release this tile's stores
        __syncthreads();
        if (tid == 0) s_last = (atomicAdd(&counters[ftile0[P.fold] + t], 1) == F.kp - 1);
        __syncthreads();
        if (!s_last) continue;
        __threadfence();                      // acquire the other layers' tiles
        for (int e = tid; e < kBM * kBN; e += kThreads) {
            const int row = m0 + (e & (kBM - 1)), col = n0 + e / kBM;
            if (row >= F.M || col >= F.N) continue;
            const int64_t off = (int64_t) col * F.ldl + row;
            double s = F.Cl[0][off];
            for (int l = 1; l < F.kp; ++l) s = ca3::add_rn(s, F.Cl[l][off]);
            int sl = 0;
            while (col >= F.cs[sl + 1]) ++sl;
            F.Cout[sl][(int64_t) (col - F.cs[sl]) * F.ldo + row] = s;
        }
    }
}

std::map<std::string, Tuning> g_cache;
std::mutex g_cache_mu;

int launch(Plan *p, const Tuning &t, hipStream_t st)
{
    zero_ints<<<std::max(1, std::min(1024, (p->ncounters + 1 + 255) / 256)), 256, 0, st>>>(p->d_counters,
                                                                                       p->ncounters + 1);
    const int grid = std::max(1, std::min(p->total, ca3::cu_count() * t.wg_per_cu));
    grouped_kernel<<<grid, kThreads, 0, st>>>(p->d_probs, p->nprob, p->d_tile0, p->d_folds, p->d_ftile0,
                                              p->d_counters, p->d_counters + p->ncounters, p->total,
                                              t.group_rows);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ca3::fail(e, "grouped_kernel");
    return 0;
}

void drop_graph(Plan *p)
{
    if (p->exec) { (void) hipGraphExecDestroy(p->exec); p->exec = nullptr; }
    if (p->graph) { (void) hipGraphDestroy(p->graph); p->graph = nullptr; }
}

}  // namespace

int build(const std::vector<Problem> &probs, const std::vector<Fold> &folds, Plan **out)
{
    *out = nullptr;
    auto *p = new Plan;
    p->nprob = (int) probs.size();
    p->nfold = (int) folds.size();
    std::vector<int> tile0(p->nprob + 1, 0), ftile0(std::max(1, p->nfold), 0);
    std::vector<int> fold_tiles(std::max(1, p->nfold), -1);
    for (int i = 0; i < p->nprob; ++i) {
        const Problem &q = probs[i];
        if (q.nseg < 0 || q.nseg > kMaxSeg || q.M < 0 || q.N < 0) { delete p; return HIP_PROXY_ERR_ARG; }
        const int tiles = ((q.M + kBM - 1) / kBM) * ((q.N + kBN - 1) / kBN);
        tile0[i + 1] = tile0[i] + tiles;
        if (q.fold >= 0) {
            if (q.fold >= p->nfold) { delete p; return HIP_PROXY_ERR_ARG; }
            fold_tiles[q.fold] = tiles;
        }
        p->signature += std::to_string(q.M) + "x" + std::to_string(q.N) + "x" + std::to_string(q.nseg) + ";";
    }
    p->total = tile0[p->nprob];
    for (int f = 0; f < p->nfold; ++f) {
        if (folds[f].kp < 1 || folds[f].kp > kMaxKp || fold_tiles[f] < 0) { delete p; return HIP_PROXY_ERR_ARG; }
        ftile0[f] = p->ncounters;
        p->ncounters += fold_tiles[f];
    }
    const size_t nb = sizeof(Problem) * std::max(1, p->nprob), fb = sizeof(Fold) * std::max(1, p->nfold);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&p->d_probs, nb);
    if (e == hipSuccess) e = hipMalloc(&p->d_folds, fb);
    if (e == hipSuccess) e = hipMalloc(&p->d_tile0, sizeof(int) * (p->nprob + 1));
    if (e == hipSuccess) e = hipMalloc(&p->d_ftile0, sizeof(int) * ftile0.size());
    if (e == hipSuccess) e = hipMalloc(&p->d_counters, sizeof(int) * (p->ncounters + 1));
    hipStream_t st = ca3::stream();
    if (e == hipSuccess && p->nprob) e = hipMemcpyAsync(p->d_probs, probs.data(), sizeof(Problem) * p->nprob,
                                                        hipMemcpyHostToDevice, st);
    if (e == hipSuccess && p->nfold) e = hipMemcpyAsync(p->d_folds, folds.data(), sizeof(Fold) * p->nfold,
                                                        hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->d_tile0, tile0.data(), sizeof(int) * tile0.size(),
                                            hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(p->d_ftile0, ftile0.data(), sizeof(int) * ftile0.size(),
                                            hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        destroy(p);
        return ca3::fail(e, "grouped plan");
    }
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_cache.find(p->signature);
        if (it != g_cache.end()) p->tune = it->second;
    }
    *out = p;
    return 0;
}

int run(Plan *p, bool use_graph)
{
    hipStream_t st = ca3::stream();
    if (!use_graph) return launch(p, p->tune, st);
    if (p->exec && (p->graph_tune.group_rows != p->tune.group_rows || p->graph_tune.wg_per_cu != p->tune.wg_per_cu))
        drop_graph(p);
    if (!p->exec) {
        CA3_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = launch(p, p->tune, st);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return ca3::fail(e, "hipStreamEndCapture");
        p->graph = g;
        CA3_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        p->graph_tune = p->tune;
    }
    CA3_CHECK(hipGraphLaunch(p->exec, st));
    return 0;
}

int autotune(Plan *p, int reps)
{
    static const int rows[] = {1, 4, 8, 16};
    static const int wgs[] = {1, 2};
    hipStream_t st = ca3::stream();
    hipEvent_t e0, e1;
    CA3_CHECK(hipEventCreate(&e0));
    CA3_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    Tuning bt = p->tune;
    int rc = 0;
    for (int gr : rows) {
        for (int w : wgs) {
            Tuning t;
            t.group_rows = gr;
            t.wg_per_cu = w;
            if ((rc = launch(p, t, st))) break;                 // warm-up
            (void) hipEventRecord(e0, st);
            for (int r = 0; r < reps && !rc; ++r) rc = launch(p, t, st);
            (void) hipEventRecord(e1, st);
            (void) hipEventSynchronize(e1);
            float ms = 0.f;
            (void) hipEventElapsedTime(&ms, e0, e1);
            if (!rc && ms < best) { best = ms; bt = t; }
        }
        if (rc) break;
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    if (rc) return rc;
    p->tune = bt;
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_cache[p->signature] = bt;
    return 0;
}

const Tuning &tuning(const Plan *p) { return p->tune; }

void set_tuning(Plan *p, const Tuning &t) { p->tune = t; }

void destroy(Plan *p)
{
    if (!p) return;
    drop_graph(p);
    (void) hipFree(p->d_probs);
    (void) hipFree(p->d_folds);
    (void) hipFree(p->d_tile0);
    (void) hipFree(p->d_ftile0);
    (void) hipFree(p->d_counters);
    delete p;
}

struct SeqGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int capture_sequence(SeqGraph **out, const std::function<int()> &body)
{
    *out = nullptr;
    hipStream_t st = ca3::stream();
    CA3_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) { if (g) (void) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return ca3::fail(e, "hipStreamEndCapture");
    auto *sg = new SeqGraph;
    sg->graph = g;
    e = hipGraphInstantiate(&sg->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        destroy_sequence(sg);
        return ca3::fail(e, "hipGraphInstantiate");
    }
    *out = sg;
    return 0;
}

bool sync_mode() { return ca3::sync_mode(); }

int replay_sequence(SeqGraph *g)
{
    CA3_CHECK(hipGraphLaunch(g->exec, ca3::stream()));
    return 0;
}

void destroy_sequence(SeqGraph *g)
{
    if (!g) return;
    if (g->exec) (void) hipGraphExecDestroy(g->exec);
    if (g->graph) (void) hipGraphDestroy(g->graph);
    delete g;
}

}  // namespace grouped
}  // namespace ca3
