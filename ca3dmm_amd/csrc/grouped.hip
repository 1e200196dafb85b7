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
//  * The workgroup tile is 128 x 128 (4 waves, one or two workgroups per
//    CU), 256 x 128 or 128 x 256 (8 waves, one per CU).
//  * One more candidate is not persistent: every rank's 128 x 128 tiles
//    in one batched launch of the direct-to-LDS DGEMM (dgemm.hip), the
//    rounds accumulated in registers as here, then each fold group's
//    layers added in layer order by the fold kernel, column slice by column
//    slice; the same bits as the persistent kernel.
//  * The tuner times the tile shape, the queue order (tile rows walked
//    together) and the persistent grid size, and keeps the fastest per
//    problem signature, in memory and in a tuning table on disk
//    (CA3DMM_TUNE_CACHE, default ~/.cache/ca3dmm_amd/tuning.txt) that later
//    processes read back.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <unistd.h>

#include "ca3dmm_ext.h"
#include "dgemm_tile.h"
#include "grouped.h"
#include "hip_common.h"

namespace ca3 {
namespace grouped {

using namespace ca3::gemm;

constexpr int kTileShapes = 3;     // 128 x 128, 256 x 128, 128 x 256
constexpr int kTileBatched = 3;    // Tuning::tile of the batched direct-to-LDS launch

struct Plan {
    int nprob = 0, nfold = 0;
    Problem *d_probs = nullptr;
    Fold *d_folds = nullptr;
    // per tile shape: items (tiles) in all, fold counters, and device tables
    int total[kTileShapes] = {0, 0, 0}, ncounters[kTileShapes] = {0, 0, 0};
    int *d_tile0[kTileShapes] = {nullptr, nullptr, nullptr};     // [nprob + 1] first item of each problem
    int *d_ftile0[kTileShapes] = {nullptr, nullptr, nullptr};    // [nfold] first counter of each fold group
    int *d_counters = nullptr;    // [max ncounters + 1]; slot ncounters[tile] is the queue head
    BatchPlan *batch = nullptr;   // the batched candidate's tables
    std::vector<Fold> h_folds;    // its fold launches
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

template <class TL>
__global__ __launch_bounds__(TL::THREADS, TL::MIN_BLOCKS) void grouped_kernel(
    const Problem *__restrict__ probs, int nprob, const int *__restrict__ tile0, const Fold *__restrict__ folds,
    const int *__restrict__ ftile0, int *counters, int *queue, int total, int G)
{
    __shared__ double As[2][kBK * TL::BM];
    __shared__ double Bs[2][kBK * TL::BN];
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
        const int tiles_m = (P.M + TL::BM - 1) / TL::BM, tiles_n = (P.N + TL::BN - 1) / TL::BN;
        int tm, tn;
        grouped_coords(t, tiles_m, tiles_n, G, &tm, &tn);
        const int m0 = tm * TL::BM, n0 = tn * TL::BN;

        f64x4 acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
        for (int s = 0; s < P.nseg; ++s) {
            OperandLoader<1, TL::BM, TL::THREADS> la;     // A block as stored: contiguous along m
            OperandLoader<0, TL::BN, TL::THREADS> lb;     // B block as stored: contiguous along k
            la.init(P.seg[s].A, P.seg[s].lda, m0, P.M);
            lb.init(P.seg[s].B, P.seg[s].ldb, n0, P.N);
            mainloop<TL>(la, lb, 0, P.seg[s].k, m0 + TL::BM <= P.M && n0 + TL::BN <= P.N, As, Bs, acc);
        }
        // store the tile (lane l, register r of block (i, j): row l % 16, column 4r + l / 16)
        const int lane = tid & 63, wave = tid >> 6, wm = wave % TL::WM, wn = wave / TL::WM;
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
        for (int e = tid; e < TL::BM * TL::BN; e += TL::THREADS) {
            const int row = m0 + (e & (TL::BM - 1)), col = n0 + e / TL::BM;
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

int tile_bm(int tile) { return tile == 1 ? 256 : 128; }
int tile_bn(int tile) { return tile == 2 ? 256 : 128; }

// The batched candidate: the products, then every fold group's column
// slices through the fold kernel (layers in layer order).
int launch_batched(Plan *p, const Tuning &t)
{
    int rc = run_batch(p->batch, t.group_rows);
    for (const Fold &F : p->h_folds) {
        for (int sl = 0; sl < F.kp && !rc; ++sl) {
            const int ncol = F.cs[sl + 1] - F.cs[sl];
            if (F.M <= 0 || ncol <= 0) continue;
            const double *parts[kMaxKp];
            int lds[kMaxKp];
            for (int l = 0; l < F.kp; ++l) {
                parts[l] = F.Cl[l] + (int64_t) F.cs[sl] * F.ldl;
                lds[l] = (int) F.ldl;
            }
            rc = hip_fold_sum(F.kp, parts, lds, F.M, ncol, F.Cout[sl], (int) F.ldo);
        }
        if (rc) break;
    }
    return rc;
}

int launch(Plan *p, const Tuning &t, hipStream_t st)
{
    if (t.tile == kTileBatched) return launch_batched(p, t);
    const int ti = (t.tile >= 0 && t.tile < kTileShapes) ? t.tile : 0;
    const int nc = p->ncounters[ti], total = p->total[ti];
    zero_ints<<<std::max(1, std::min(1024, (nc + 1 + 255) / 256)), 256, 0, st>>>(p->d_counters, nc + 1);
    const int wg = ti == 0 ? t.wg_per_cu : 1;
    const int grid = std::max(1, std::min(total, ca3::cu_count() * wg));
    if (ti == 0)
        grouped_kernel<Tile128><<<grid, Tile128::THREADS, 0, st>>>(
            p->d_probs, p->nprob, p->d_tile0[0], p->d_folds, p->d_ftile0[0], p->d_counters, p->d_counters + nc,
            total, t.group_rows);
    else if (ti == 1)
        grouped_kernel<Tile256x128><<<grid, Tile256x128::THREADS, 0, st>>>(
            p->d_probs, p->nprob, p->d_tile0[1], p->d_folds, p->d_ftile0[1], p->d_counters, p->d_counters + nc,
            total, t.group_rows);
    else
        grouped_kernel<Tile128x256><<<grid, Tile128x256::THREADS, 0, st>>>(
            p->d_probs, p->nprob, p->d_tile0[2], p->d_folds, p->d_ftile0[2], p->d_counters, p->d_counters + nc,
            total, t.group_rows);
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

// ---------------------------------------------------------------------
// Tuning table: one line per (form, problem signature), read once per
// process from the file and rewritten whole after each autotune.
//   fp64 <signature> tile=<0|1|2> rows=<n> wg=<n>
//   crt  <signature> tile_n=<256|128> xcd=<0|1>
// ---------------------------------------------------------------------
namespace {

std::map<std::string, std::string> g_table;      // "form signature" -> settings
bool g_table_loaded = false;
std::mutex g_table_mu;

std::string table_path()
{
    const char *v = getenv("CA3DMM_TUNE_CACHE");
    if (v && *v) return v;
    const char *x = getenv("XDG_CACHE_HOME");
    const char *h = getenv("HOME");
    if (x && *x) return std::string(x) + "/ca3dmm_amd/tuning.txt";
    if (h && *h) return std::string(h) + "/.cache/ca3dmm_amd/tuning.txt";
    return "";
}

void load_table_locked()
{
    if (g_table_loaded) return;
    g_table_loaded = true;
    const std::string path = table_path();
    if (path.empty()) return;
    FILE *f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[8192];
    while (fgets(line, sizeof(line), f)) {
        char form[16], sig[4096];
        int off = 0;
        if (sscanf(line, "%15s %4095s %n", form, sig, &off) < 2) continue;
        std::string rest(line + off);
        while (!rest.empty() && (rest.back() == '\n' || rest.back() == '\r')) rest.pop_back();
        g_table[std::string(form) + " " + sig] = rest;
    }
    fclose(f);
}

void save_table_locked()
{
    const std::string path = table_path();
    if (path.empty()) return;
    // create the directory chain (best effort; a table that cannot be
    // written only costs a retune in the next process)
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1))
        (void) mkdir(path.substr(0, pos).c_str(), 0755);
    const std::string tmp = path + ".tmp" + std::to_string((long) getpid());
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return;
    for (const auto &kv : g_table) fprintf(f, "%s %s\n", kv.first.c_str(), kv.second.c_str());
    if (fclose(f) == 0) (void) rename(tmp.c_str(), path.c_str());
    else (void) remove(tmp.c_str());
}

}  // namespace

bool lookup_tuning(const char *form, const std::string &sig, Tuning *t)
{
    std::lock_guard<std::mutex> lk(g_table_mu);
    load_table_locked();
    auto it = g_table.find(std::string(form) + " " + sig);
    if (it == g_table.end()) return false;
    Tuning r;
    if (sscanf(it->second.c_str(), "tile=%d rows=%d wg=%d", &r.tile, &r.group_rows, &r.wg_per_cu) != 3) return false;
    if (r.tile < 0 || r.tile > kTileBatched || r.group_rows < 1 || r.wg_per_cu < 1 || r.wg_per_cu > 2) return false;
    *t = r;
    return true;
}

void store_tuning(const char *form, const std::string &sig, const Tuning &t)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "tile=%d rows=%d wg=%d", t.tile, t.group_rows, t.wg_per_cu);
    std::lock_guard<std::mutex> lk(g_table_mu);
    load_table_locked();
    g_table[std::string(form) + " " + sig] = buf;
    save_table_locked();
}

bool lookup_tuning_ozaki(const std::string &sig, OzTuning *t)
{
    std::lock_guard<std::mutex> lk(g_table_mu);
    load_table_locked();
    auto it = g_table.find("crt " + sig);
    if (it == g_table.end()) return false;
    OzTuning r;
    if (sscanf(it->second.c_str(), "tile_n=%d xcd=%d", &r.tile_n, &r.xcd_queue) != 2) return false;
    if ((r.tile_n != 256 && r.tile_n != 128) || r.xcd_queue < 0 || r.xcd_queue > 1) return false;
    *t = r;
    return true;
}

void store_tuning_ozaki(const std::string &sig, const OzTuning &t)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "tile_n=%d xcd=%d", t.tile_n, t.xcd_queue);
    std::lock_guard<std::mutex> lk(g_table_mu);
    load_table_locked();
    g_table["crt " + sig] = buf;
    save_table_locked();
}

int build(const std::vector<Problem> &probs, const std::vector<Fold> &folds, Plan **out)
{
    *out = nullptr;
    auto *p = new Plan;
    p->nprob = (int) probs.size();
    p->nfold = (int) folds.size();
    for (int i = 0; i < p->nprob; ++i) {
        const Problem &q = probs[i];
        if (q.nseg < 0 || q.nseg > kMaxSeg || q.M < 0 || q.N < 0 || (q.fold >= p->nfold)) {
            delete p;
            return HIP_PROXY_ERR_ARG;
        }
        p->signature += std::to_string(q.M) + "x" + std::to_string(q.N) + "x" + std::to_string(q.nseg) + ";";
    }
    for (int f = 0; f < p->nfold; ++f)
        if (folds[f].kp < 1 || folds[f].kp > kMaxKp) { delete p; return HIP_PROXY_ERR_ARG; }
    std::vector<int> tile0[kTileShapes], ftile0[kTileShapes];
    int maxc = 0;
    for (int ti = 0; ti < kTileShapes; ++ti) {
        const int bm = tile_bm(ti), bn = tile_bn(ti);
        tile0[ti].assign(p->nprob + 1, 0);
        ftile0[ti].assign(std::max(1, p->nfold), 0);
        std::vector<int> fold_tiles(std::max(1, p->nfold), -1);
        for (int i = 0; i < p->nprob; ++i) {
            const Problem &q = probs[i];
            const int tiles = ((q.M + bm - 1) / bm) * ((q.N + bn - 1) / bn);
            tile0[ti][i + 1] = tile0[ti][i] + tiles;
            if (q.fold >= 0) fold_tiles[q.fold] = tiles;
        }
        p->total[ti] = tile0[ti][p->nprob];
        for (int f = 0; f < p->nfold; ++f) {
            if (fold_tiles[f] < 0) { delete p; return HIP_PROXY_ERR_ARG; }
            ftile0[ti][f] = p->ncounters[ti];
            p->ncounters[ti] += fold_tiles[f];
        }
        maxc = std::max(maxc, p->ncounters[ti]);
    }
    const size_t nb = sizeof(Problem) * std::max(1, p->nprob), fb = sizeof(Fold) * std::max(1, p->nfold);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&p->d_probs, nb);
    if (e == hipSuccess) e = hipMalloc(&p->d_folds, fb);
    for (int ti = 0; ti < kTileShapes && e == hipSuccess; ++ti) {
        e = hipMalloc(&p->d_tile0[ti], sizeof(int) * (p->nprob + 1));
        if (e == hipSuccess) e = hipMalloc(&p->d_ftile0[ti], sizeof(int) * ftile0[ti].size());
    }
    if (e == hipSuccess) e = hipMalloc(&p->d_counters, sizeof(int) * (maxc + 1));
    hipStream_t st = ca3::stream();
    if (e == hipSuccess && p->nprob) e = hipMemcpyAsync(p->d_probs, probs.data(), sizeof(Problem) * p->nprob,
                                                        hipMemcpyHostToDevice, st);
    if (e == hipSuccess && p->nfold) e = hipMemcpyAsync(p->d_folds, folds.data(), sizeof(Fold) * p->nfold,
                                                        hipMemcpyHostToDevice, st);
    for (int ti = 0; ti < kTileShapes && e == hipSuccess; ++ti) {
        e = hipMemcpyAsync(p->d_tile0[ti], tile0[ti].data(), sizeof(int) * tile0[ti].size(), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(p->d_ftile0[ti], ftile0[ti].data(), sizeof(int) * ftile0[ti].size(),
                               hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        destroy(p);
        return ca3::fail(e, "grouped plan");
    }
    int rc = build_batch(probs, false, &p->batch);
    if (rc) {
        destroy(p);
        return rc;
    }
    for (const Fold &F : folds)
        if (F.ldl > INT_MAX || F.ldo > INT_MAX) {
            destroy(p);
            return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "grouped plan: fold leading dimension above INT_MAX");
        }
    p->h_folds = folds;
    Tuning t;
    if (lookup_tuning("fp64", p->signature, &t)) p->tune = t;
    *out = p;
    return 0;
}

int run(Plan *p, bool use_graph)
{
    hipStream_t st = ca3::stream();
    if (!use_graph) return launch(p, p->tune, st);
    if (p->exec && (p->graph_tune.group_rows != p->tune.group_rows || p->graph_tune.wg_per_cu != p->tune.wg_per_cu ||
                    p->graph_tune.tile != p->tune.tile))
        drop_graph(p);
    if (!p->exec) {
        CA3_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const bool sync_saved = ca3::sync_mode();
        hip_proxy_set_sync(0);                    // no stream synchronisation inside the capture
        int rc = launch(p, p->tune, st);
        hip_proxy_set_sync(sync_saved ? 1 : 0);
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
    hipStream_t st = ca3::stream();
    // the candidates are timed back to back: no synchronisation per launch
    const bool sync_saved = ca3::sync_mode();
    hip_proxy_set_sync(0);
    hipEvent_t e0, e1;
    CA3_CHECK(hipEventCreate(&e0));
    CA3_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    Tuning bt = p->tune;
    int rc = 0;
    for (int ti = 0; ti <= kTileBatched && !rc; ++ti) {
        for (int gr : rows) {
            for (int w = (ti == kTileBatched ? 2 : 1); w <= (ti == 0 || ti == kTileBatched ? 2 : 1); ++w) {
                Tuning t;
                t.tile = ti;
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
    }
    (void) hipEventDestroy(e0);
    (void) hipEventDestroy(e1);
    hip_proxy_set_sync(sync_saved ? 1 : 0);
    if (rc) return rc;
    p->tune = bt;
    store_tuning("fp64", p->signature, bt);
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
    for (int ti = 0; ti < kTileShapes; ++ti) {
        (void) hipFree(p->d_tile0[ti]);
        (void) hipFree(p->d_ftile0[ti]);
    }
    (void) hipFree(p->d_counters);
    destroy_batch(p->batch);
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
