// Internal interface of the persistent grouped GEMM (grouped.hip), used by
// the simulated-grid engine (engine.cpp).  No HIP types appear here.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace ca3 {
namespace grouped {

constexpr int kMaxSeg = 16;   // Cannon rounds per problem
constexpr int kMaxKp = 16;    // k layers per fold group

struct Seg {                  // one k-segment (Cannon round) of a problem
    const double *A;          // m x k, column-major, leading dimension lda
    const double *B;          // k x n, column-major, leading dimension ldb
    int64_t lda, ldb;
    int k;
};

struct Problem {              // C (M x N, ldc) := sum over segments of A_s B_s
    int M, N, nseg;
    Seg seg[kMaxSeg];
    double *C;
    int64_t ldc;
    int fold;                 // fold group index or -1
};

struct Fold {                 // k-layer reduction of the C blocks of kp problems
    int kp;
    double *Cl[kMaxKp];       // the layers' C blocks (same shape and ld)
    int64_t ldl;
    int M, N;
    int cs[kMaxKp + 1];       // column slice t = [cs[t], cs[t+1]) goes to Cout[t]
    double *Cout[kMaxKp];
    int64_t ldo;
};

struct Tuning {
    int tile = 0;             // workgroup tile: 0 = 128 x 128, 1 = 256 x 128, 2 = 128 x 256 (persistent,
                              // register-staged); 3 = one batched direct-to-LDS launch of 128 x 128
                              // tiles (rounds in registers), then the folds as separate launches
    int group_rows = 8;       // tile rows walked together in the queue order
    int wg_per_cu = 2;        // persistent workgroups per CU (1 for the 8-wave tiles)
};

struct Plan;

// Build device tables for the problems; 0 on success.
int build(const std::vector<Problem> &probs, const std::vector<Fold> &folds, Plan **out);
// Run once (counters reset inside the same stream order); use_graph replays
// a captured hipGraph.
int run(Plan *p, bool use_graph);
// Time the candidate tunings and keep the fastest; returns 0 on success.
int autotune(Plan *p, int reps);
const Tuning &tuning(const Plan *p);
void set_tuning(Plan *p, const Tuning &t);
void destroy(Plan *p);

// The same problems and fold groups through the CRT form of the Ozaki
// scheme with nmod moduli: each problem's product over all its k-segments
// (one exact INT8 product per modulus), reconstructed in FP64 and, for a
// fold group, added over the layers in layer order (ozaki.hip).
struct OzTuning {
    int tile_n = 256;         // workgroup tile 256 x tile_n: 256 (8 waves) or 128 (4 waves)
    int xcd_queue = 1;        // 1: items in 8 per-XCD ranges; 0: one counter over all items
};

// The tuning table shared by both forms: in memory, and on disk at
// CA3DMM_TUNE_CACHE (default ~/.cache/ca3dmm_amd/tuning.txt), one line per
// problem signature; lookups read the file once per process.
bool lookup_tuning(const char *form, const std::string &sig, Tuning *t);
void store_tuning(const char *form, const std::string &sig, const Tuning &t);

struct OzPlan;
int build_ozaki(const std::vector<Problem> &probs, const std::vector<Fold> &folds, int nmod, OzPlan **out);
// Time the candidate tunings of the grouped launch (tile 256 / 128 x queue
// order) and keep the fastest for this grid's shape signature.
int autotune_ozaki(OzPlan *p, int reps);
const OzTuning &tuning_ozaki(const OzPlan *p);
void set_tuning_ozaki(OzPlan *p, const OzTuning &t);
// grouped: the products of every problem in one persistent launch fed by a
// device queue (else one launch per problem); use_graph replays the whole
// sequence from a captured hipGraph.
int run_ozaki(OzPlan *p, bool grouped, bool use_graph);
void destroy_ozaki(OzPlan *p);
// Modulus count for the CRT products of these problems: the count rule of
// hip_ozaki_choose (ca3dmm_ext.h) with the largest problem k, on the spread
// statistics of the sampled entries samples[3 s .. 3 s + 2] = (problem,
// row, column), each over the whole k of its problem; *L (may be null) gets
// the spread statistic L the rule read.  Synchronous.
int choose_moduli(const std::vector<Problem> &probs, const std::vector<int> &samples, int k, int *nmod,
                  double *L);
bool lookup_tuning_ozaki(const std::string &sig, OzTuning *t);
void store_tuning_ozaki(const std::string &sig, const OzTuning &t);

// The same problems (every operand NN, as stored) as one non-persistent
// launch of the direct-to-LDS DGEMM (dgemm.hip): one workgroup per
// 128 x 128 tile of every problem, its k-segments in order.  per_segment:
// each segment is added to C as its own product (beta 0 on the first, 1
// after), so C gets the bits of one hip_dgemm call per segment; otherwise
// the segments accumulate in registers and C is stored once, the bits of
// the persistent grouped kernel.  Segments may have k = 0; folds are not
// part of it.  The load form (CA3DMM_DGEMM_LOADS) and the tile-row group
// (CA3DMM_DGEMM_GROUP; group > 0 overrides it per run) are read when the
// plan is built.
struct BatchPlan;
int build_batch(const std::vector<Problem> &probs, bool per_segment, BatchPlan **out);
int run_batch(BatchPlan *p, int group = 0);
void destroy_batch(BatchPlan *p);

// A whole enqueued sequence (every kernel `body` launches on the library
// stream) captured once into a hipGraph and replayed.  body must not
// synchronise, allocate or copy through the host while it is captured.
struct SeqGraph;
int capture_sequence(SeqGraph **out, const std::function<int()> &body);
// The library's sync mode (hip_proxy_set_sync), for code outside the HIP sources.
bool sync_mode();
int replay_sequence(SeqGraph *g);
void destroy_sequence(SeqGraph *g);

}  // namespace grouped
}  // namespace ca3
