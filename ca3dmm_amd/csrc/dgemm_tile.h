// Device building blocks of the FP64 MFMA GEMM tile, shared by the plain
// DGEMM, the grouped (simulated-grid) GEMM and the Gram kernel.
//
// MFMA operand / result layout used throughout (v_mfma_f64_16x16x4_f64,
// one wave64):  X(16x4): lane l holds X[l%16][l/16];  Y(4x16): lane l holds
// Y[l/16][l%16];  D = X*Y (16x16): lane l, register r holds D[4*r + l/16][l%16].
// The tile code feeds X = op(B)^T and Y = op(A)^T, so D is a block of C^T:
// lane l, register r holds C[m = l%16][n = 4*r + l/16].  For a fixed r the 16
// lanes of a row group then hold 16 consecutive rows of one column of C,
// which makes the column-major stores 128-byte contiguous.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "hip_common.h"

namespace ca3 {
namespace gemm {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kBK = 16, kThreads = 256;
constexpr int kMaxFoldSrc = 8;

// Workgroup tile BM x BN of C: waves of 64 x 64 (4 x 4 MFMA blocks each),
// WM along m and WN along n.  128 x 128 (4 waves, two workgroups per CU)
// is the default; 256 x 128 and 128 x 256 use 8 waves, one workgroup per CU.
template <int BM_, int BN_>
struct TileShape {
    static constexpr int BM = BM_, BN = BN_;
    static constexpr int WM = BM / 64, WN = BN / 64;
    static constexpr int NW = WM * WN, THREADS = 64 * NW;
    static constexpr int MIN_BLOCKS = THREADS == 256 ? 2 : 1;     // per CU, for __launch_bounds__
};
using Tile128 = TileShape<128, 128>;
using Tile256x128 = TileShape<256, 128>;
using Tile128x256 = TileShape<128, 256>;

struct GemmArgs {
    int M, N, K;
    double alpha, beta;
    const double *A;
    int64_t lda;
    const double *B;
    int64_t ldb;
    double *C;
    int64_t ldc;
    int group;                      // tile rows walked together (tile_coords)
    int nsrc;                       // partial-C fold in the epilogue
    const double *P[kMaxFoldSrc];
    int64_t ldp[kMaxFoldSrc];
};

// XOR swizzle of column x within LDS row kk (see dgemm.hip header).
__device__ __forceinline__ int swz(int kk, int x)
{
    return x ^ (((kk & 1) << 4) | ((kk >> 1) & 7));
}

// Linear workgroup id -> (tile row, tile col): XCD-contiguous ranges, then
// groups of G tile-rows (16 unless CA3DMM_DGEMM_GROUP says otherwise)
// walked column by column.
template <class TL = Tile128>
__device__ __forceinline__ void tile_coords(int bid, int nb, int M, int N, int *tm, int *tn, int G = 16)
{
    const int tiles_m = (M + TL::BM - 1) / TL::BM, tiles_n = (N + TL::BN - 1) / TL::BN;
    (void) tiles_n;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const int per = G * tiles_n;
    const int grp = bid / per;
    const int first = grp * G;
    const int gsz = min(tiles_m - first, G);
    const int r = bid - grp * per;
    *tm = first + r % gsz;
    *tn = r / gsz;
}

// Loads one 16 x W stage of an operand (W = the tile's extent along m or
// n) with T threads into registers and writes it to LDS as [k][x].  XC = 1:
// the operand is contiguous along x (m or n); XC = 0: contiguous along k.
// Out-of-range elements read as 0.  Each thread holds P = 16 W / T elements:
// XC = 1 at k rows kk0, kk0 + T/W, ...; XC = 0 at columns x, x + T/16, ...
template <int XC, int W = 128, int T = 256>
struct OperandLoader {
    static constexpr int P = W * kBK / T;
    static constexpr int KSTEP = T / W;         // XC: k rows per pass of the workgroup
    static constexpr int XSTEP = T / kBK;       // !XC: columns per pass
    static_assert(P >= 1 && P <= 16 && KSTEP * P == kBK && XSTEP * P == W, "loader shape");
    const double *base;   // element (x = this thread's first x, k = 0)
    int64_t ld;
    int kk0;              // XC: first k row of this thread; !XC: its k row
    unsigned xmask;       // bit p: x of element p is in range
    int xcol;             // this thread's (first) x
    double r[P];

    __device__ __forceinline__ void init(const double *M, int64_t ld_, int x0, int xdim)
    {
        const int t = threadIdx.x;
        ld = ld_;
        if (XC) {
            xcol = t % W;
            kk0 = t / W;
            xmask = (x0 + xcol < xdim) ? 0xffffu : 0u;
            base = M + (x0 + xcol);
        } else {
            kk0 = t & 15;
            const int xt = t >> 4;
            xcol = xt;
            xmask = 0;
#pragma unroll
            for (int p = 0; p < P; ++p)
                if (x0 + xt + XSTEP * p < xdim) xmask |= 1u << p;
            base = M + (int64_t) (x0 + xt) * ld;
        }
    }

    __device__ __forceinline__ void load(int k0, int kend)
    {
        if (XC) {
            const double *q = base + (int64_t) (k0 + kk0) * ld;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const bool ok = xmask && (k0 + kk0 + KSTEP * p < kend);
                r[p] = ok ? q[(int64_t) (KSTEP * p) * ld] : 0.0;
            }
        } else {
            const bool kok = (k0 + kk0 < kend);
            const double *q = base + (k0 + kk0);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const bool ok = kok && ((xmask >> p) & 1u);
                r[p] = ok ? q[(int64_t) (XSTEP * p) * ld] : 0.0;
            }
        }
    }

    // The same stage with every element in range (interior tile, whole
    // k-stage): no predicates, so the loads issue without branches.
    __device__ __forceinline__ void load_full(int k0)
    {
        if (XC) {
            const double *q = base + (int64_t) (k0 + kk0) * ld;
#pragma unroll
            for (int p = 0; p < P; ++p) r[p] = q[(int64_t) (KSTEP * p) * ld];
        } else {
            const double *q = base + (k0 + kk0);
#pragma unroll
            for (int p = 0; p < P; ++p) r[p] = q[(int64_t) (XSTEP * p) * ld];
        }
    }

    __device__ __forceinline__ void store(double *S) const
    {
        if (XC) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const int kk = kk0 + KSTEP * p;
                S[kk * W + swz(kk, xcol)] = r[p];
            }
        } else {
#pragma unroll
            for (int p = 0; p < P; ++p) S[kk0 * W + swz(kk0, xcol + XSTEP * p)] = r[p];
        }
    }
};

// k-slices [KS0, KS1) of the 16 k-steps of MFMAs on the LDS stage (4 x
// v_mfma_f64_16x16x4 k-slices in all).
template <class TL = Tile128, int KS0 = 0, int KS1 = kBK / 4>
__device__ __forceinline__ void stage_mfma(const double *As, const double *Bs, f64x4 (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % TL::WM, wn = wave / TL::WM;
    const int li = lane & 15;
#pragma unroll
    for (int ks = KS0; ks < KS1; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[kk * TL::BM + swz(kk, wm * 64 + i * 16 + li)];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[kk * TL::BN + swz(kk, wn * 64 + j * 16 + li)];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
}

// The two LDS stage buffers of an operand held as two separate arrays (the
// direct-to-LDS loop needs them so); As[i] as for a double[2][...] array.
struct LdsPair {
    double *b0, *b1;
    __device__ __forceinline__ double *operator[](int i) const { return i ? b1 : b0; }
};

// Accumulate op(A)[m0:+BM, kbeg:kend) * op(B)[kbeg:kend, n0:+BN) into acc
// with bounds checks on every element (edge tiles and partial k-stages).
// As / Bs: double[2][16 * BM] (or LdsPair), As[i] = stage buffer i.
template <class TL = Tile128, class LA, class LB, class BA, class BB>
__device__ __forceinline__ void mainloop_checked(LA &la, LB &lb, int kbeg, int kend, BA As, BB Bs,
                                                 f64x4 (&acc)[4][4])
{
    if (kend <= kbeg) return;
    const int nk = (kend - kbeg + kBK - 1) / kBK;
    la.load(kbeg, kend);
    lb.load(kbeg, kend);
    la.store(As[0]);
    lb.store(Bs[0]);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = (kt + 1 < nk);
        if (more) {
            la.load(kbeg + (kt + 1) * kBK, kend);
            lb.load(kbeg + (kt + 1) * kBK, kend);
        }
        stage_mfma<TL>(As[cur], Bs[cur], acc);
        if (more) {
            la.store(As[cur ^ 1]);
            lb.store(Bs[cur ^ 1]);
        }
        __syncthreads();
    }
}

// The same accumulation.  A tile that is interior in m and n (interior =
// m0 + BM <= M and n0 + BN <= N) runs its whole k-stages through a loop
// without predicates or data-dependent branches (one register stage in
// flight, the last stage's load repeated in bounds and never consumed);
// edge tiles and the k tail go through mainloop_checked.  Both add the
// k-steps in the same order, so results do not depend on the path.
template <class TL = Tile128, class LA, class LB, class BA, class BB>
__device__ __forceinline__ void mainloop(LA &la, LB &lb, int kbeg, int kend, bool interior, BA As, BB Bs,
                                         f64x4 (&acc)[4][4])
{
    const int nfull = (interior && kend > kbeg) ? (kend - kbeg) / kBK : 0;
    if (nfull > 0) {
        la.load_full(kbeg);
        lb.load_full(kbeg);
        la.store(As[0]);
        lb.store(Bs[0]);
        __syncthreads();
        for (int kt = 0; kt < nfull; ++kt) {
            const int cur = kt & 1;
            const int kn = kbeg + min(kt + 1, nfull - 1) * kBK;
            la.load_full(kn);
            lb.load_full(kn);
            // the next stage goes to the other buffer halfway through this
            // one's MFMAs (its loads have had half a stage to land), so the
            // LDS writes complete under the second half instead of before
            // the barrier
            stage_mfma<TL, 0, kBK / 8>(As[cur], Bs[cur], acc);
            la.store(As[cur ^ 1]);
            lb.store(Bs[cur ^ 1]);
            stage_mfma<TL, kBK / 8, kBK / 4>(As[cur], Bs[cur], acc);
            __syncthreads();
        }
    }
    const int kt0 = kbeg + nfull * kBK;
    if (kt0 < kend) mainloop_checked<TL>(la, lb, kt0, kend, As, Bs, acc);
}

// ---------------------------------------------------------------------
// Direct-to-LDS stages (whole k-stages of the 128 x 128 DGEMM, 16-byte
// aligned operands with even leading dimensions): global_load_lds_dwordx4
// writes lane l's 16 bytes (two doubles) at M0 + 16 l, so one wave
// instruction fills a 1 KB LDS block; a stage of an operand is 16 blocks
// (16 k x 128 x), four per wave.  No staging registers and no LDS stores:
// the loads of stage z + 1 are issued once stage z's operands are read
// from LDS, and have stage z's MFMAs to land.  Layouts (d = one double):
//   XC = 1 (contiguous along x): block = one k row of 128 x; 16-byte chunk
//     c of row k holds x = 2 (c ^ fx(k)), +1 with fx(k) = 8 (k & 1);
//   XC = 0 (contiguous along k): block = 8 x lines of 16 k; chunk c of line
//     x holds k = 2 (c ^ hx(x)), +1 with hx(x) = (x >> 1) & 7.
// Each lane fetches the global chunk its LDS slot holds, so the reads stay
// coalesced, and the MFMA operand reads (16 consecutive x at 4 k) hit 64
// distinct banks per 32 lanes.  The k-slices of the MFMAs are those of
// stage_mfma, so both paths add the same products in the same order.
// ---------------------------------------------------------------------
template <int XC>
__device__ __forceinline__ int dtl_index(int k, int x)
{
    if (XC) return k * 128 + 2 * ((x >> 1) ^ ((k & 1) << 3)) + (x & 1);
    return x * 16 + 2 * ((k >> 1) ^ ((x >> 1) & 7)) + (k & 1);
}

// Edge tiles: only xn <= 128 x of the tile exist.  A lane whose slot holds
// an x >= xn loads the last existing x instead, so every load stays inside
// the operand; those LDS rows hold copies whose products land in C rows or
// columns the epilogue never stores.  Addresses are a wave-uniform base per
// block (scalar registers) plus a 32-bit lane offset, which requires
// 7 * ld * 8 + 128 < 2^32 bytes for XC = 0 (dtl_ld_ok).
//
// WORDS = 4: one 16-byte load per lane and block (global_load_lds_dwordx4)
// of a double pair, from an 8-byte aligned address (an odd ld or an
// operand 8 bytes past a 16-byte boundary splits some loads over two
// cache lines); for XC = 1 it needs an even xn (a chunk holds an x pair).
// WORDS = 1: four 4-byte loads per lane and block (global_load_lds_dword,
// lane l of load q lands at M0 + 256 q + 4 l), for any xn; the LDS image
// is the same.
__host__ __device__ constexpr bool dtl_ld_ok(int64_t ld) { return ld >= 0 && ld <= (int64_t) 1 << 26; }

// Load forms of the direct-to-LDS stages (CA3DMM_DGEMM_LOADS): 16-byte
// loads from any 8-byte aligned operand (default); "aligned": 16-byte
// loads only for 16-byte aligned operands with even ld, 4-byte loads for
// other operands contiguous along m or n, register staging otherwise;
// "dword": 4-byte loads; "reg": register staging only.
enum LoadMode { kLoads16, kLoadsAligned, kLoadsDword, kLoadsReg };
LoadMode load_mode();

template <int XC, int WORDS = 4>
struct DtlLoader;

template <int XC>
struct DtlLoader<XC, 4> {
    const double *base[4];   // block i of this wave for k0 = 0 (wave-uniform)
    uint32_t off[4];         // this lane's byte offset from base[i]

    __device__ __forceinline__ void init(const double *M, int64_t ld, int x0, int xn = 128)
    {
        const int lane = threadIdx.x & 63;
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = 4 * w + i;
            if (XC) {
                // k row b of the stage; chunk c of the row holds x = 2c, 2c + 1
                base[i] = M + x0 + (int64_t) b * ld;
                off[i] = 16u * (uint32_t) min(lane ^ ((b & 1) << 3), (xn >> 1) - 1);
            } else {
                // lines x = 8b .. 8b + 7 of the tile; the lane's k pair is
                // fixed by its slot (line xl), the line read is x <= xn - 1
                const int xb = min(8 * b, xn - 1), xl = 8 * b + (lane >> 3), x = min(xl, xn - 1);
                base[i] = M + (int64_t) (x0 + xb) * ld;
                off[i] = (uint32_t) ((int64_t) (x - xb) * ld * 8) + 16u * (uint32_t) ((lane & 7) ^ ((xl >> 1) & 7));
            }
        }
    }

    // issue the stage starting at koff (elements) into S (16 x 128 doubles)
    __device__ __forceinline__ void issue(int64_t koff, double *S) const
    {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const char *g = reinterpret_cast<const char *>(base[i] + koff) + (uint64_t) off[i];
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *) g,
                                             (void __attribute__((address_space(3))) *) (S + (4 * w + i) * 128),
                                             16, 0, 0);
        }
    }
};

// Load q (0..3) of a block fills LDS bytes [256 q, 256 q + 256) of it: lane
// l writes the 4-byte half h = l & 1 of double j = (l >> 1) & 1 of 16-byte
// chunk c = 16 q + (l >> 2).  The offsets are rebuilt per stage from a few
// lane constants (the vector ALU is idle under the MFMAs), so the loader
// holds as many registers as the 16-byte one.
template <int XC>
struct DtlLoader<XC, 1> {
    const double *M;
    int64_t ld;
    int x0, xn;
    uint32_t t;     // l >> 2: the chunk within the load's 16
    uint32_t jh;    // 8 j + 4 h

    __device__ __forceinline__ void init(const double *M_, int64_t ld_, int x0_, int xn_ = 128)
    {
        const int lane = threadIdx.x & 63;
        M = M_; ld = ld_; x0 = x0_; xn = xn_;
        t = (uint32_t) (lane >> 2);
        jh = 8u * ((lane >> 1) & 1) + 4u * (lane & 1);
    }

    __device__ __forceinline__ void issue(int64_t koff, double *S) const
    {
        const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int b = 4 * w + i;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double *base;
                uint32_t off;
                if (XC) {
                    // k row b; chunk c = 16 q + t holds x pair p = c ^ 8 (b & 1)
                    base = M + x0 + (int64_t) b * ld + koff;
                    const uint32_t x = 2u * (uint32_t) ((16 * q + (int) t) ^ ((b & 1) << 3)) + (jh >> 3);
                    off = 8u * min(x, (uint32_t) xn - 1u) + (jh & 4u);
                } else {
                    // line xl = 8b + 2q + (t >> 3), k pair (t & 7) ^ ((xl >> 1) & 7)
                    const int xb = min(8 * b + 2 * q, xn - 1);
                    base = M + (int64_t) (x0 + xb) * ld + koff;
                    const int xl = 8 * b + 2 * q + (int) (t >> 3), x = min(xl, xn - 1);
                    off = (uint32_t) ((x - xb) * (int) ld) * 8u +
                          16u * (uint32_t) ((t & 7) ^ (uint32_t) ((xl >> 1) & 7)) + jh;
                }
                auto src = (const void __attribute__((address_space(1))) *) (reinterpret_cast<const char *>(base) +
                                                                             (uint64_t) off);
                auto dst = (void __attribute__((address_space(3))) *) (S + (4 * w + i) * 128 + 32 * q);
                __builtin_amdgcn_global_load_lds(src, dst, 4, 0, 0);
            }
        }
    }
};

template <int XCA, int XCB>
__device__ __forceinline__ void stage_mfma_dtl(const double *As, const double *Bs, f64x4 (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15;
#pragma unroll
    for (int ks = 0; ks < kBK / 4; ++ks) {
        const int kk = ks * 4 + (lane >> 4);
        double a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = As[dtl_index<XCA>(kk, wm * 64 + i * 16 + li)];
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = Bs[dtl_index<XCB>(kk, wn * 64 + j * 16 + li)];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
}

// The whole 16-deep k-stages [0, nfull * 16) of an interior 128 x 128 tile
// through the direct-to-LDS double buffer (koff(k) = k * ld for XC = 1, k
// for XC = 0); nothing is loaded for nfull = 0.  Ends with a barrier, after
// which the caller runs the k tail, if any, through mainloop_checked.
//
// The two buffers of each operand must be distinct __shared__ arrays and
// the loop is unrolled by two so that every access names its array at
// compile time: the compiler then knows the LDS reads of stage z do not
// touch the buffer the DMA of stage z + 1 writes.  With one array indexed
// by the stage's parity it cannot tell, and makes the first LDS read after
// each DMA wait for it, which exposes the whole load latency every stage.
template <int XCA, int XCB, int WA, int WB>
__device__ __forceinline__ void mainloop_dtl(const DtlLoader<XCA, WA> &la, const DtlLoader<XCB, WB> &lb, int64_t lda,
                                             int64_t ldb, int nfull, double *A0, double *A1, double *B0,
                                             double *B1, f64x4 (&acc)[4][4])
{
    auto koff = [](int xc, int64_t ld, int k0) -> int64_t { return xc ? (int64_t) k0 * ld : (int64_t) k0; };
    // own loads of the stage have landed (vmcnt counts the LDS DMA), then
    // every wave's; every wave has also finished reading the previous stage
    auto publish = []() __attribute__((always_inline)) {
        __builtin_amdgcn_s_waitcnt(0x0f70);     // vmcnt(0); expcnt (6:4) and lgkmcnt (11:8) not waited on
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };
    if (nfull > 0) {
        la.issue(koff(XCA, lda, 0), A0);
        lb.issue(koff(XCB, ldb, 0), B0);
    }
    for (int kt = 0; kt < nfull; kt += 2) {
        publish();                              // stage kt in buffer 0
        if (kt + 1 < nfull) {
            la.issue(koff(XCA, lda, (kt + 1) * kBK), A1);
            lb.issue(koff(XCB, ldb, (kt + 1) * kBK), B1);
        }
        stage_mfma_dtl<XCA, XCB>(A0, B0, acc);
        if (kt + 1 >= nfull) break;
        publish();                              // stage kt + 1 in buffer 1
        if (kt + 2 < nfull) {
            la.issue(koff(XCA, lda, (kt + 2) * kBK), A0);
            lb.issue(koff(XCB, ldb, (kt + 2) * kBK), B0);
        }
        stage_mfma_dtl<XCA, XCB>(A1, B1, acc);
    }
    __syncthreads();                            // the buffers are free for the caller's tail
}

// C := fold(P) + (alpha*acc + beta*C) for the tile at (m0, n0).
template <class TL = Tile128>
__device__ __forceinline__ void epilogue(const GemmArgs &g, int m0, int n0, const f64x4 (&acc)[4][4])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave % TL::WM, wn = wave / TL::WM;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = m0 + wm * 64 + i * 16 + (lane & 15);
        if (row >= g.M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = n0 + wn * 64 + j * 16 + 4 * r + (lane >> 4);
                if (col >= g.N) continue;
                double *c = g.C + (int64_t) col * g.ldc + row;
                double v = ca3::mul_rn(g.alpha, acc[i][j][r]);
                if (g.beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(g.beta, *c));
                if (g.nsrc > 0) {
                    double s = g.P[0][(int64_t) col * g.ldp[0] + row];
                    for (int t = 1; t < g.nsrc; ++t) s = ca3::add_rn(s, g.P[t][(int64_t) col * g.ldp[t] + row]);
                    v = ca3::add_rn(s, v);
                }
                *c = v;
            }
        }
    }
}

}  // namespace gemm
}  // namespace ca3
