// Ozaki-scheme split-precision DGEMM on INT8 MFMA (gfx950).
//
// No reference counterpart: an alternative to the FP64 MFMA DGEMM that
// keeps FP64 accuracy.  Specification (and the bit-exact oracle):
// oracle/ozaki.py, functions split8 / ozaki_dgemm8.
//
//   1. ozaki_max_kernel   per line (row of op(A), column of op(B)) the
//                          largest |x| (atomicMax on its bits), then
//      ozaki_scale_kernel  the line's scale exponent E: with f, e = frexp
//                          of that max, E = e + 1 if f <= 253/256, else e + 2;
//   2. ozaki_slice_kernel  x' = x * 2^-E (|x'| <= 253/512) and the s balanced
//                          base-256 digits of rint(x' * 2^(8s)) as int8
//                          planes, line-major with k contiguous, zero padded
//                          to whole tiles; LDS-tiled so both source layouts
//                          read coalesced;
//   3. ozaki_diag_kernel   diagonal d = s-1 .. 0: sum_{t+u=d} QA_t QB_u on
//                          v_mfma_i32_16x16x64_i8 with exact INT32
//                          accumulation, added in FP64 (x 2^-8(d+2)),
//                          smallest diagonal first; the last launch rescales
//                          by 2^EA_i 2^EB_j and applies alpha/beta.
// Each diagonal (and INT32-exact k-chunk) is one launch of an INT8 GEMM
// over the concatenated slice pairs: 256 x 128 tiles, 128 x 64 per wave,
// LDS double buffer fed from two register sets (prefetch distance two
// stages).  The FP64 accumulation of the diagonals runs through a workspace
// W that stays L2/HBM resident, which frees the registers the FP64
// accumulators would need.  INT32 exactness needs
// (d+1) * Kchunk * 128 * 128 < 2^31; for each diagonal the chunks are added
// in k order (the oracle follows the same order).
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ca3dmm_ext.h"
#include "grouped.h"
#include "hip_common.h"

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 256;     // padding of lines (a multiple of both tile sides)
constexpr int kKb = 64;        // k bytes per MFMA step (one LDS sub-block)
constexpr int kKp = 256;       // k padding and k-chunk granularity (whole pipeline rounds)
constexpr int kMaxSlices = 16;
constexpr int kExpEmpty = INT_MIN;
// A line holding an Inf or a NaN: it is sliced as a zero line (no undefined
// conversions), and every C entry of its row (op(A)) or column (op(B)) is
// written as NaN.  Bits of |x| at or above this mark are Inf or NaN.
constexpr int kExpNonFinite = INT_MIN + 1;
constexpr unsigned long long kNonFiniteBits = 0x7ff0000000000000ull;

// v * 2^(ea + eb) in one step (one rounding, no intermediate overflow): 0
// for an empty line, NaN for a line with a non-finite element.
__device__ __forceinline__ double oz_unscale(double v, int ea, int eb)
{
    if (ea == kExpNonFinite || eb == kExpNonFinite) return __builtin_nan("");
    if (ea == kExpEmpty || eb == kExpEmpty) return 0.0;
    return ldexp(v, ea + eb);
}

// ---------------------------------------------------------------------
// 1. line exponents
// ---------------------------------------------------------------------
// Line l of the operand: element (l, kk) at X[l*sl + kk*sk].  mx[l] holds
// the bits of the largest |x| seen (non-negative doubles order like their
// bit patterns, so an integer atomicMax is exact); 0 means an all-zero line.
// The maximum is taken on the bits, so an Inf (0x7ff0...) or a NaN (above
// it) is kept: bits >= kNonFiniteBits mark a non-finite line.
// Block (bx, by) of a gx x gy grid: 64 lines x (K / gy chunk); the thread
// layout follows the contiguous direction so loads coalesce.
__device__ __forceinline__ void oz_max_body(int bx, int by, int gy, const double *__restrict__ X, int lines, int K,
                                            int64_t sl, int64_t sk, unsigned long long *mx)
{
    const bool line_contig = (sl == 1);
    const int t = threadIdx.x;
    const int kchunk = (K + gy - 1) / gy;
    const int k0 = by * kchunk, k1 = min(K, k0 + kchunk);
    const int l0 = bx * 64;
    if (!line_contig) {
        // k-contiguous lines: each wave walks 16 of the block's lines, its
        // lanes reading consecutive elements (512-byte loads), then reduces
        // across the wave
        const int lane = t & 63, w = t >> 6;
        for (int r = 0; r < 16; ++r) {
            const int l = l0 + w * 16 + r;
            if (l >= lines) break;
            const double *xl = X + (int64_t) l * sl;
            unsigned long long m = 0;
            for (int kk = k0 + lane; kk < k1; kk += 64)
                m = max(m, (unsigned long long) __double_as_longlong(fabs(xl[kk])));
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
            if (lane == 0 && m != 0) atomicMax(&mx[l], m);
        }
        return;
    }
    __shared__ unsigned long long red[256];
    unsigned long long m = 0;
    const int myline = l0 + (t & 63);
    for (int kk = k0 + (t >> 6); kk < k1; kk += 4)
        if (myline < lines) m = max(m, (unsigned long long) __double_as_longlong(fabs(X[myline + (int64_t) kk * sk])));
    red[t] = m;
    __syncthreads();
    if (t < 64) {
        const int l = l0 + t;
        unsigned long long v = 0;
        for (int j = 0; j < 4; ++j) v = max(v, red[t + 64 * j]);
        if (l < lines && v != 0) atomicMax(&mx[l], v);
    }
}

__global__ __launch_bounds__(256) void ozaki_max_kernel(const double *__restrict__ X, int lines, int K,
                                                         int64_t sl, int64_t sk, unsigned long long *mx)
{
    oz_max_body(blockIdx.x, blockIdx.y, gridDim.y, X, lines, K, sl, sk, mx);
}

// E[l] = kExpEmpty for a zero line, kExpNonFinite for a line with an Inf or
// a NaN, else e + 1 or e + 2 as above.
__global__ void ozaki_scale_kernel(const unsigned long long *__restrict__ mx, int n, int *__restrict__ E)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    if (mx[l] >= kNonFiniteBits) {
        E[l] = kExpNonFinite;
        return;
    }
    const double v = __longlong_as_double((long long) mx[l]);
    if (v == 0.0) {
        E[l] = kExpEmpty;
        return;
    }
    int e;
    const double f = frexp(v, &e);
    E[l] = (f <= 253.0 / 256.0) ? e + 1 : e + 2;
}

__global__ void fill_int_kernel(int *p, int n, int v)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ---------------------------------------------------------------------
// 2. slices
// ---------------------------------------------------------------------
// Balanced base-256 digits of x * 2^-E, |.| <= 253/512, as digits of the integer
// N = rint(x * 2^(8s - E)) (digit s-1-t goes to slice t, t = 0 the most
// significant).  N = X * 256^off with X = M << (z mod 8) (M the 53-bit
// significand, |X| < 2^61) or, when the scale drops bits, X = rint(M 2^z),
// off = 0.  Adding the bias B = 0x8080...80 maps the balanced digits
// c_i in [-128, 127] of X to the plain bytes c_i + 128 of U = X + B (mod
// 2^64), so digit i is byte i of U xor 0x80 read as int8; eight bytes cover
// |X| < 2^61 and the digits outside the window are zero.
__device__ __forceinline__ void window_of(double x, int E, int s, uint64_t &U, int &off)
{
    const uint64_t kBias = 0x8080808080808080ull;
    const double a = ldexp(x, -E);
    off = 0;
    if (a == 0.0) {           // every byte 0x80: all digits zero
        U = kBias;
        return;
    }
    int ex;
    const double f = frexp(a, &ex);
    const int64_t M = (int64_t) ldexp(f, 53);        // exact signed 53-bit significand
    const int z = ex - 53 + 8 * s;
    int64_t X;
    if (z < 0) {
        X = (int64_t) rint(ldexp((double) M, z));     // < 2^53 in magnitude: exact
    } else {
        X = M << (z & 7);
        off = z >> 3;
    }
    U = (uint64_t) X + kBias;
}

// Tile of 64 lines x 64 k.  Output plane t: P[t * plane + line * Kp + kk].
__global__ __launch_bounds__(256) void ozaki_slice_kernel(const double *__restrict__ X, int lines, int K,
                                                           int64_t sl, int64_t sk, const int *__restrict__ ex,
                                                           int s, int8_t *__restrict__ P, int64_t plane, int Kp)
{
    __shared__ double tile[64][65];
    const int t = threadIdx.x;
    const int l0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
    if (sl == 1) {            // lines contiguous
        for (int p = 0; p < 16; ++p) {
            const int li = t & 63, ki = (t >> 6) + 4 * p;
            const int l = l0 + li, kk = k0 + ki;
            tile[li][ki] = (l < lines && kk < K) ? X[l + (int64_t) kk * sk] : 0.0;
        }
    } else {                  // k contiguous
        for (int p = 0; p < 16; ++p) {
            const int ki = t & 63, li = (t >> 6) + 4 * p;
            const int l = l0 + li, kk = k0 + ki;
            tile[li][ki] = (l < lines && kk < K) ? X[(int64_t) l * sl + kk] : 0.0;
        }
    }
    __syncthreads();
    const int li = t >> 2, kq = (t & 3) * 16;
    const int l = l0 + li;
    const int exl = ex[min(l, lines - 1)];
    const bool zero_line = (l >= lines) || exl == kExpEmpty || exl == kExpNonFinite;
    const int E = zero_line ? 0 : exl;
    uint64_t U[16];
    int off[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) window_of(zero_line ? 0.0 : tile[li][kq + j], E, s, U[j], off[j]);
    int8_t *base = P + (int64_t) l * Kp + k0 + kq;
#pragma unroll
    for (int t2 = 0; t2 < kMaxSlices; ++t2) {
        if (t2 < s) {
            u32x4 v;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t w = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int j = 4 * q + b;
                    const int i = s - 1 - t2 - off[j];       // byte of U holding digit s-1-t2
                    const uint32_t byte = (i >= 0 && i < 8) ? (((uint32_t) (U[j] >> (8 * i)) & 255u) ^ 0x80u) : 0u;
                    w |= byte << (8 * b);
                }
                v[q] = w;
            }
            *reinterpret_cast<u32x4 *>(base + (int64_t) t2 * plane) = v;
        }
    }
}

// ---------------------------------------------------------------------
// 3. one diagonal of slice products, FP64 recombination in the epilogue
// ---------------------------------------------------------------------
// A work item computes, for one diagonal d and one INT32-exact k-chunk c,
// the integer tile D = sum_{t=0..d} QA_t QB_{d-t} over that chunk (a plain
// INT8 GEMM over the concatenated slice pairs).  Two ways to add the items:
//  * W mode (one launch per (d, c), many output tiles): the epilogue folds D
//    into the FP64 accumulator W (m x n, ld Mp): W = D * 2^-8(d+2) on the
//    first launch, W + D * 2^-8(d+2) afterwards; the last launch writes
//    C = alpha * 2^eA 2^eB (W + ...) + beta * C instead;
//  * split mode (few output tiles, long k): one launch runs every (tile, d, c)
//    item and stores D as INT32 partials P32[d][c]; ozaki_fold_kernel then
//    adds them per element in the same order (d = s-1 .. 0, c ascending).
// Both give the same bits.
// Workgroup tiles: 2 x 2 waves of WTM x WTN.  The default (WTM = WTN = 128,
// 256 x 256 per workgroup, one workgroup per CU) keeps the 256 INT32
// accumulators of a wave in the AGPR half of the register file; 128 x 64
// (256 x 128 per workgroup, two per CU) is the smaller variant.

struct OzArgs {
    int M, N, Mp, Np, Kp;
    int d, kb0, nkb;        // W mode: diagonal, first k-block of the chunk, k-blocks in the chunk
    int first, last;
    double alpha, beta;
    const int8_t *PA;       // planes of Mp x Kp
    const int8_t *PB;       // planes of Np x Kp
    int64_t planeA, planeB;
    const int *eA, *eB;     // line scale exponents E (kExpEmpty for zero lines)
    double *W;              // FP64 accumulator, ld Mp
    double *C;
    int64_t ldc;
    // split mode
    int s, nchunks, kcb;    // slices, chunks per diagonal, k-blocks per chunk
    int ntiles, upper;      // output tiles per (d, c) item group; upper-triangle tiles only
    int ngroups;            // s * nchunks
    int *P32;               // INT32 partials, one Mp x Np plane per (d, c)
    int64_t p32plane;
    // CRT form: residue planes R_t (int8, Mp x Np, ld Mp), s = number of moduli
    int8_t *R8;
    int64_t rplane;
    // grouped CRT launch (in the launch's own arguments): first queue item
    // of each problem, the 8 per-XCD counters (the per-problem OzArgs come
    // as a separate kernel argument)
    const int *tile0;
    int nprob, total;
    int *queue;
    int flatq;              // 1: one counter over all items; 0: 8 per-XCD ranges
};

// Uniform values kept in scalar registers (readfirstlane of each 32-bit half).
__device__ __forceinline__ int64_t oz_rfl64(int64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t) v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t) ((uint64_t) v >> 32));
    return (int64_t) (((uint64_t) hi << 32) | lo);
}

template <class T>
__device__ __forceinline__ T *oz_rflp(T *p)
{
    return (T *) (uintptr_t) oz_rfl64((int64_t) (uintptr_t) p);
}

// 16-byte chunk swizzle within a 64-byte LDS row.  An MFMA operand read
// (ds_read_b128) has lanes 16 q .. 16 q + 15 on rows r..r+15 (r a multiple of
// 16) at logical chunk q.  Measured on gfx950 with SQ_LDS_BANK_CONFLICT over
// this exact read pattern: f(row) = (row >> 1) & 3 is conflict-free (4 LDS
// cycles per read, as a linear 1 KB read), while no swizzle, row & 3,
// (row >> 2) & 3 and the 3-bit mixes tried all take 8 cycles (two-way).
// Stores (4 lanes per row) are permutations within rows.
__device__ __forceinline__ int oz_swz(int row) { return (row >> 1) & 3; }

__device__ __forceinline__ int lds_off(int row, int chunk)
{
    return row * 64 + ((chunk ^ oz_swz(row)) << 4);
}

// One pipeline stage = KS sub-blocks of 64 bytes of k; per sub-block PA (PB)
// passes of 64 rows of A (B), 16 bytes per thread and pass (thread t: row
// t / 4 of the pass, 16-byte chunk t % 4).  Sub-block j of stage z is
// 64-byte block z * KS + j of the (slice pair, k-block) sequence of the
// chunk (nkb, the 64-byte blocks per chunk, is a multiple of KS).
template <int KS, int PA, int PB>
__device__ __forceinline__ void oz_gload(const OzArgs &g, const int8_t *gA, const int8_t *gB, int64_t row64,
                                         int d, int kb0, int nkb, int z, u32x4 (&ra)[KS * PA],
                                         u32x4 (&rb)[KS * PB])
{
#pragma unroll
    for (int j = 0; j < KS; ++j) {
        const int zz = z * KS + j;
        const int t = zz / nkb, b = zz - t * nkb;
        const int64_t koff = (int64_t) (kb0 + b) * kKb;
        const int8_t *pa = gA + (int64_t) t * g.planeA + koff;
        const int8_t *pb = gB + (int64_t) (d - t) * g.planeB + koff;
#pragma unroll
        for (int p = 0; p < PA; ++p) ra[j * PA + p] = *reinterpret_cast<const u32x4 *>(pa + p * row64);
#pragma unroll
        for (int p = 0; p < PB; ++p) rb[j * PB + p] = *reinterpret_cast<const u32x4 *>(pb + p * row64);
    }
}

// LDS sub-block j of a buffer: rows of 64 bytes, A at j * rowsA * 64.
template <int KS, int PA, int PB>
__device__ __forceinline__ void oz_sstore(int8_t *As, int8_t *Bs, int lrow, int lchunk, const u32x4 (&ra)[KS * PA],
                                          const u32x4 (&rb)[KS * PB])
{
#pragma unroll
    for (int j = 0; j < KS; ++j) {
#pragma unroll
        for (int p = 0; p < PA; ++p)
            *reinterpret_cast<u32x4 *>(&As[j * PA * 64 * 64 + lds_off(lrow + 64 * p, lchunk)]) = ra[j * PA + p];
#pragma unroll
        for (int p = 0; p < PB; ++p)
            *reinterpret_cast<u32x4 *>(&Bs[j * PB * 64 * 64 + lds_off(lrow + 64 * p, lchunk)]) = rb[j * PB + p];
    }
}

template <int FI, int FJ>
__device__ __forceinline__ void oz_compute(const int8_t *As, const int8_t *Bs, int wm, int wn, int li, int kq,
                                           i32x4 (&acc)[FI][FJ])
{
    i32x4 bf[FJ];
#pragma unroll
    for (int j = 0; j < FJ; ++j)
        bf[j] = *reinterpret_cast<const i32x4 *>(&Bs[lds_off(wn * FJ * 16 + j * 16 + li, kq)]);
#pragma unroll
    for (int i = 0; i < FI; ++i) {
        const i32x4 af = *reinterpret_cast<const i32x4 *>(&As[lds_off(wm * FI * 16 + i * 16 + li, kq)]);
#pragma unroll
        for (int j = 0; j < FJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[j], af, acc[i][j], 0, 0, 0);
    }
}

// The work item of this workgroup: output tile (tm, tn), diagonal d and the
// k-chunk [kb0, kb0 + nkb) in 64-byte blocks; false for an idle workgroup of
// the split form's last round.
template <bool SPLIT, int kBMo, int kBNo>
__device__ __forceinline__ bool oz_item(const OzArgs &g, int &tm, int &tn, int &d, int &kb0, int &nkb,
                                        int64_t &part_off)
{
    const int tiles_m = g.Mp / kBMo, tiles_n = g.Np / kBNo;
    part_off = 0;
    if (!SPLIT) {
        int bid = blockIdx.x;
        const int nb = gridDim.x;
        if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
        const int G = 4, per = G * tiles_n, grp = bid / per, first = grp * G;
        const int gsz = min(tiles_m - first, G), rr = bid - grp * per;
        tm = first + rr % gsz;
        tn = rr / gsz;
        d = g.d;
        kb0 = g.kb0;
        nkb = g.nkb;
    } else {
        // Workgroups are dealt to the 8 XCDs round robin: slot j of XCD x
        // takes item group (j / ntiles) * 8 + x and output tile j % ntiles,
        // so the tiles of one (d, c) group share that XCD's L2 and the
        // groups (heaviest diagonal first) spread evenly over the XCDs.
        const int x = blockIdx.x & 7, j = blockIdx.x >> 3;
        const int jq = j / g.ntiles;
        const int grp = jq * 8 + x;
        if (grp >= g.ngroups) return false;
        int tile = j - jq * g.ntiles;
        d = g.s - 1 - grp / g.nchunks;
        const int c = grp - (grp / g.nchunks) * g.nchunks;
        kb0 = c * g.kcb;
        nkb = min(g.kcb, g.Kp / kKb - kb0);
        part_off = (int64_t) (d * g.nchunks + c) * g.p32plane;
        if (g.upper) {        // row tm of tiles holds tn = tm * R .. tiles_n - 1 (tiles touching j >= i)
            constexpr int R = kBMo / kBNo;
            tm = 0;
            while (tm < tiles_m - 1 && tile >= tiles_n - R * tm) {
                tile -= tiles_n - R * tm;
                ++tm;
            }
            tn = R * tm + tile;
        } else {
            tm = tile / tiles_n;
            tn = tile - tm * tiles_n;
        }
    }
    return true;
}

// Opaque zeros: without the empty asm the compiler shares one zero register
// among the accumulators and then copies them inside the main loop.
template <int FI, int FJ>
__device__ __forceinline__ void oz_zero(i32x4 (&acc)[FI][FJ])
{
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
            acc[i][j] = i32x4{0, 0, 0, 0};
            asm volatile("" : "+a"(acc[i][j]));
        }
}

// Epilogue: lane l, register r of block (i, j) holds D[m][n] with m = l % 16,
// n = 4 * (l / 16) + r (operands fed as (B, A)).
template <bool SPLIT, int WTM, int WTN, int FI, int FJ>
__device__ __forceinline__ void oz_epilogue(const OzArgs &g, const i32x4 (&acc)[FI][FJ], int m0, int n0, int d,
                                            int64_t part_off, bool first, bool last)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, kq = lane >> 4;
    if (SPLIT) {              // the padded Mp x Np plane is written whole
        int *P = g.P32 + part_off;
#pragma unroll
        for (int i = 0; i < FI; ++i) {
            const int row = m0 + wm * WTM + i * 16 + li;
#pragma unroll
            for (int j = 0; j < FJ; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int col = n0 + wn * WTN + j * 16 + 4 * kq + r;
                    P[(int64_t) col * g.Mp + row] = acc[i][j][r];
                }
        }
        return;
    }
    const double w = ldexp(1.0, -8 * (d + 2));
#pragma unroll
    for (int i = 0; i < FI; ++i) {
        const int row = m0 + wm * WTM + i * 16 + li;
        if (row >= g.M) continue;
        const int ea = g.eA[row];
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = n0 + wn * WTN + j * 16 + 4 * kq + r;
                if (col >= g.N) continue;
                double *wp = g.W + (int64_t) col * g.Mp + row;
                double v = (double) acc[i][j][r] * w;
                if (!first) v = *wp + v;
                if (!last) {
                    *wp = v;
                    continue;
                }
                v = oz_unscale(v, ea, g.eB[col]);
                double *cp = g.C + (int64_t) col * g.ldc + row;
                v = ca3::mul_rn(g.alpha, v);
                if (g.beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(g.beta, *cp));
                *cp = v;
            }
        }
        // keep each row block's W traffic together (bounds the live registers)
        __builtin_amdgcn_sched_barrier(0);
    }
}


// W-only epilogue of the fused form: the whole padded tile (W is Mp x Np),
// W = D * 2^-8(d+2) for the first segment, W + D * 2^-8(d+2) afterwards.
template <int WTM, int WTN, int FI, int FJ>
__device__ __forceinline__ void oz_epilogue_w(const OzArgs &g, const i32x4 (&acc)[FI][FJ], int m0, int n0, int d,
                                              bool first)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, kq = lane >> 4;
    const double w = ldexp(1.0, -8 * (d + 2));
#pragma unroll
    for (int i = 0; i < FI; ++i) {
        const int row = m0 + wm * WTM + i * 16 + li;
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = n0 + wn * WTN + j * 16 + 4 * kq + r;
                double *wp = g.W + (int64_t) col * g.Mp + row;
                const double v = (double) acc[i][j][r] * w;
                *wp = first ? v : *wp + v;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// C := alpha * 2^EA_i 2^EB_j W + beta * C on the m x n result (the same
// operations as the last launch of the per-segment form).
__global__ __launch_bounds__(256) void ozaki_finalize_kernel(const double *__restrict__ W, int64_t ldw, int M, int N,
                                                              const int *__restrict__ eA, const int *__restrict__ eB,
                                                              double alpha, double beta, double *C, int64_t ldc)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    for (int j = blockIdx.y; j < N; j += gridDim.y) {
        if (i >= M) return;
        const double v0 = W[(int64_t) j * ldw + i];
        double v = oz_unscale(v0, eA[i], eB[j]);
        double *cp = C + (int64_t) j * ldc + i;
        v = ca3::mul_rn(alpha, v);
        if (beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(beta, *cp));
        *cp = v;
    }
}

// Register-staged form: global -> registers -> LDS, two LDS buffers.
template <bool SPLIT, int WTM, int WTN, int KS>
__global__ __launch_bounds__(256, (WTN == 128 ? 1 : 2)) void ozaki_diag_kernel(OzArgs g)
{
    constexpr int kBMo = 2 * WTM, kBNo = 2 * WTN;       // workgroup tile
    constexpr int FI = WTM / 16, FJ = WTN / 16;          // 16 x 16 blocks per wave
    constexpr int PA = kBMo / 64, PB = kBNo / 64;        // 64-row load passes per sub-block
    static_assert(kBMo % kBNo == 0 && kTile % kBMo == 0 && kTile % kBNo == 0, "tile shapes");
    static_assert((kKp / kKb) % KS == 0, "a stage divides the k padding");
    __shared__ __attribute__((aligned(16))) int8_t As[2][KS * kBMo * kKb];
    __shared__ __attribute__((aligned(16))) int8_t Bs[2][KS * kBNo * kKb];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn, d, kb0, nkb;
    int64_t part_off;
    if (!oz_item<SPLIT, kBMo, kBNo>(g, tm, tn, d, kb0, nkb, part_off)) return;
    const int m0 = tm * kBMo, n0 = tn * kBNo;

    const int lrow = tid >> 2, lchunk = tid & 3;
    const int8_t *gA = g.PA + (int64_t) (m0 + lrow) * g.Kp + lchunk * 16;
    const int8_t *gB = g.PB + (int64_t) (n0 + lrow) * g.Kp + lchunk * 16;
    const int64_t row64 = (int64_t) 64 * g.Kp;
    const int total = (d + 1) * nkb / KS;     // stages of KS 64-byte blocks over (slice pair, k-block)

    i32x4 acc[FI][FJ];
    oz_zero(acc);
    const int li = lane & 15, kq = lane >> 4;

    // Pipeline.  DEPTH 2 (small tiles): two register staging sets; at the
    // top of a half iteration the loads of the stage two ahead are issued,
    // then the current LDS buffer is consumed and the stage one ahead goes
    // from registers into the other buffer; the odd last stage is peeled
    // off the loop.  DEPTH 1 (large tiles, whose KS = 2 stages are long
    // enough to cover a load): one staging set, loads one stage ahead.
    // Either way one barrier per stage, no data-dependent branch inside the
    // loop (a branch there makes the compiler copy accumulators), and the
    // scheduling barrier keeps each stage's loads ahead of its MFMAs.
    // Loads past the last stage re-read the last stage (in bounds) and are
    // never consumed.
    constexpr int DEPTH = (KS == 1) ? 2 : 1;
    if (DEPTH == 2) {
        u32x4 ra0[KS * PA], rb0[KS * PB], ra1[KS * PA], rb1[KS * PB];
        oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, 0, ra0, rb0);
        oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, min(1, total - 1), ra1, rb1);
        oz_sstore<KS, PA, PB>(As[0], Bs[0], lrow, lchunk, ra0, rb0);
        __syncthreads();
        int z = 0;
        for (; z + 1 < total; z += 2) {
            // buffer 0 holds stage z, set 1 stage z+1
            oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, min(z + 2, total - 1), ra0, rb0);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KS; ++j)
                oz_compute(&As[0][j * kBMo * kKb], &Bs[0][j * kBNo * kKb], wm, wn, li, kq, acc);
            oz_sstore<KS, PA, PB>(As[1], Bs[1], lrow, lchunk, ra1, rb1);
            __syncthreads();
            oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, min(z + 3, total - 1), ra1, rb1);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KS; ++j)
                oz_compute(&As[1][j * kBMo * kKb], &Bs[1][j * kBNo * kKb], wm, wn, li, kq, acc);
            oz_sstore<KS, PA, PB>(As[0], Bs[0], lrow, lchunk, ra0, rb0);
            __syncthreads();
        }
        if (z < total) {      // odd last stage
#pragma unroll
            for (int j = 0; j < KS; ++j)
                oz_compute(&As[0][j * kBMo * kKb], &Bs[0][j * kBNo * kKb], wm, wn, li, kq, acc);
        }
    } else {
        u32x4 ra[KS * PA], rb[KS * PB];
        oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, 0, ra, rb);
        oz_sstore<KS, PA, PB>(As[0], Bs[0], lrow, lchunk, ra, rb);
        __syncthreads();
        for (int z = 0; z < total; ++z) {
            const int cur = z & 1;
            oz_gload<KS, PA, PB>(g, gA, gB, row64, d, kb0, nkb, min(z + 1, total - 1), ra, rb);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < KS; ++j)
                oz_compute(&As[cur][j * kBMo * kKb], &Bs[cur][j * kBNo * kKb], wm, wn, li, kq, acc);
            oz_sstore<KS, PA, PB>(As[cur ^ 1], Bs[cur ^ 1], lrow, lchunk, ra, rb);
            __syncthreads();
        }
    }

    oz_epilogue<SPLIT, WTM, WTN>(g, acc, m0, n0, d, part_off, g.first != 0, g.last != 0);
}

// Direct-to-LDS form: global_load_lds_dwordx4 writes lane l's 16 bytes at
// M0 + 16 l, so one instruction fills a 1 KB block of 16 LDS rows; lane l
// fetches row l / 4 of the block at the logical chunk (l % 4) ^ f(row), which
// lands it where lds_off puts that chunk.  No staging registers: NBUF LDS
// stages of 64 bytes of k, the loads of stage z + NBUF - 1 issued right
// after the barrier that opens stage z.  Each wave loads WA blocks of A and
// WB of B per stage; before the barrier it waits until its own loads of
// stage z are done (vmcnt <= (NBUF - 2) * (WA + WB)), the barrier then makes
// every wave's part visible.  Loads past the last stage re-read the last
// stage (in bounds) into a buffer that is never consumed.
// (One workgroup per CU: the 4-stage ring of a 256 x 128 tile takes 96 KB
// of the 160 KB LDS.)
template <bool SPLIT, int WTM, int WTN, int NBUF>
__global__ __launch_bounds__(256, 1) void ozaki_dtl_kernel(OzArgs g)
{
    constexpr int kBMo = 2 * WTM, kBNo = 2 * WTN;
    constexpr int FI = WTM / 16, FJ = WTN / 16;
    constexpr int WA = kBMo / 64, WB = kBNo / 64;       // 1 KB blocks per wave per stage
    constexpr int L = WA + WB;
    static_assert(kBMo % kBNo == 0 && kTile % kBMo == 0 && kTile % kBNo == 0, "tile shapes");
    static_assert(NBUF >= 3 && NBUF <= 4 && (NBUF - 2) * L <= 63, "ring size / vmcnt range");
    // one LDS variable per ring slot, so the compiler can tell a slot being
    // filled from the slot being read and need not drain the loads in flight
    __shared__ __attribute__((aligned(1024))) int8_t A0[kBMo * kKb], A1[kBMo * kKb], A2[kBMo * kKb];
    __shared__ __attribute__((aligned(1024))) int8_t B0[kBNo * kKb], B1[kBNo * kKb], B2[kBNo * kKb];
    __shared__ __attribute__((aligned(1024))) int8_t A3[NBUF > 3 ? kBMo * kKb : 16], B3[NBUF > 3 ? kBNo * kKb : 16];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn, d, kb0, nkb;
    int64_t part_off;
    if (!oz_item<SPLIT, kBMo, kBNo>(g, tm, tn, d, kb0, nkb, part_off)) return;
    const int m0 = tm * kBMo, n0 = tn * kBNo;

    // this lane's part of every 1 KB block: row r = 16 * blk + lane / 4
    const int lr = lane >> 2;
    const int lc = ((lane & 3) ^ oz_swz(lr)) << 4;
    const int8_t *gA = g.PA + (int64_t) (m0 + wave * WA * 16 + lr) * g.Kp + lc;
    const int8_t *gB = g.PB + (int64_t) (n0 + wave * WB * 16 + lr) * g.Kp + lc;
    const int64_t blk16 = (int64_t) 16 * g.Kp;
    const int total = (d + 1) * nkb;          // stages of 64 bytes of k over (slice pair, k-block)

    auto slotA = [&](auto J) -> int8_t * {
        if constexpr (J.value == 0) return A0;
        else if constexpr (J.value == 1) return A1;
        else if constexpr (J.value == 2) return A2;
        else return A3;
    };
    auto slotB = [&](auto J) -> int8_t * {
        if constexpr (J.value == 0) return B0;
        else if constexpr (J.value == 1) return B1;
        else if constexpr (J.value == 2) return B2;
        else return B3;
    };
    auto issue = [&](int z, auto J) __attribute__((always_inline)) {
        const int t = z / nkb, b = z - t * nkb;
        const int64_t koff = (int64_t) (kb0 + b) * kKb;
        const int8_t *pa = gA + (int64_t) t * g.planeA + koff;
        const int8_t *pb = gB + (int64_t) (d - t) * g.planeB + koff;
        int8_t *sa = slotA(J), *sb = slotB(J);
#pragma unroll
        for (int i = 0; i < WA; ++i)
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *) (pa + i * blk16),
                                             (void __attribute__((address_space(3))) *) (sa + (wave * WA + i) * 1024),
                                             16, 0, 0);
#pragma unroll
        for (int i = 0; i < WB; ++i)
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *) (pb + i * blk16),
                                             (void __attribute__((address_space(3))) *) (sb + (wave * WB + i) * 1024),
                                             16, 0, 0);
    };

    i32x4 acc[FI][FJ];
    oz_zero(acc);
    const int li = lane & 15, kq = lane >> 4;

    // s_waitcnt: vmcnt = (NBUF - 2) * L (bits 3:0 and 15:14), expcnt and lgkmcnt not waited on
    constexpr int kVm = (NBUF - 2) * L;
    constexpr int kWait = (kVm & 15) | (((kVm >> 4) & 3) << 14) | (7 << 4) | (15 << 8);
    // Stage z uses slot z % NBUF; after the barrier that opens it, the loads
    // of stage z + NBUF - 1 go into slot (z - 1) % NBUF.
    auto step = [&](int z, auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        // own loads of stage z landed; then every wave's (no fence: a fenced
        // barrier would also wait for the later stages still in flight).  A
        // wave's LDS reads of stage z - 1 have returned before it gets here,
        // since the MFMAs consumed them.
        __builtin_amdgcn_s_waitcnt(kWait);
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        issue(min(z + NBUF - 1, total - 1), std::integral_constant<int, (j + NBUF - 1) % NBUF>{});
        oz_compute(slotA(J), slotB(J), wm, wn, li, kq, acc);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    issue(0, I0{});
    issue(min(1, total - 1), I1{});
    if constexpr (NBUF > 3) issue(min(2, total - 1), I2{});
    // total = (d + 1) * nkb with nkb a multiple of kKp / kKb = 4, so the
    // ring turns a whole number of times (no remainder steps, which would
    // make the compiler copy accumulators)
    static_assert((kKp / kKb) % NBUF == 0, "ring size divides the k padding");
    for (int z = 0; z < total; z += NBUF) {
        step(z, I0{});
        step(z + 1, I1{});
        step(z + 2, I2{});
        if constexpr (NBUF > 3) step(z + 3, I3{});
    }
    // the repeated loads of the last stage may still be landing in LDS:
    // drain them before the workgroup can end
    __builtin_amdgcn_s_waitcnt(0);
    oz_epilogue<SPLIT, WTM, WTN>(g, acc, m0, n0, d, part_off, g.first != 0, g.last != 0);
}

// ---------------------------------------------------------------------
// CRT form (Ozaki scheme II): one INT8 product per modulus
// ---------------------------------------------------------------------
// Lines are scaled by 2^e to integers a' = rint(x 2^e) with ||a'||_2 <= 2^T,
// 2^(2T+1) < P = p_0 ... p_{N-1}, so |C'| < P/2 is fixed by its residues.
// Residue planes (balanced int8) go through the same pipelined INT8 kernel,
// one segment per (modulus, k-chunk), whose epilogue keeps Z mod p as int8
// planes R_t (ld Mp); ozaki2_crt_kernel turns the N residues of an element
// into balanced mixed-radix digits over the pair products Q_i = p_2i p_2i+1
// (Garner) and evaluates them in FP64 by Horner from the top digit (rounded
// multiply, then rounded add).
constexpr int kCrtMax = 25;
constexpr int kCrtMin = 4;
constexpr int kCrtKMax = 1 << 23;   // the integer norm sums stay below 2^63
// k-chunk: kchunk * 128 * 128 + 128 < 2^30 (the INT32 sum plus the previous
// residue stays within INT32), a multiple of kKp
constexpr int kCrtChunk = (int) ((((1u << 30) - 1 - 128) / 16384u) / kKp * kKp);

constexpr int kCrtModuli[kCrtMax] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211,
                                     199, 197, 193, 191, 181, 179, 173, 167, 163, 157, 151, 149};
// T(N) = (bitlen(P) - 2) / 2 for the first N moduli (N = 0 .. 25)
constexpr int kCrtT[kCrtMax + 1] = {0, 3, 7, 11, 15, 19, 23, 27, 31, 35, 39, 43, 46, 50,
                                    54, 58, 62, 65, 69, 73, 77, 80, 84, 88, 91, 95};

constexpr int crt_inv(int a, int p)                 // a^-1 mod p (extended Euclid)
{
    int t = 0, nt = 1, r = p, nr = a % p;
    while (nr != 0) {
        const int q = r / nr, t2 = t - q * nt, r2 = r - q * nr;
        t = nt; nt = t2; r = nr; nr = r2;
    }
    return t < 0 ? t + p : t;
}

constexpr int crt_pow2(int e, int p)                // 2^e mod p, balanced
{
    int r = 1;
    for (int i = 0; i < e; ++i) r = (2 * r) % p;
    return r > p / 2 ? r - p : r;
}

constexpr int64_t crt_bal64(int64_t x, int64_t q)  // balanced residue: [-(q-1)/2, (q-1)/2], [-q/2, q/2) for even q
{
    int64_t r = x % q;
    if (r < 0) r += q;
    return r > ((q & 1) ? (q - 1) / 2 : q / 2 - 1) ? r - q : r;
}

// Reconstruction over composite moduli: Q_i = p_2i p_2i+1 (i < N / 2) and,
// for odd N, the single p_N-1 as the last one; kCrtC composites at most.
constexpr int kCrtC = (kCrtMax + 1) / 2;

struct CrtTables {
    double pd[kCrtMax], invpd[kCrtMax];              // modulus, 1/p (rounded)
    double c42[kCrtMax], c32[kCrtMax], c64[kCrtMax]; // 2^42, 2^32, 2^64 mod p (balanced)
    // Garner over the pairs: with M_i = Q_0 ... Q_i-1, g_i = M_i^-1 mod Q_i
    // and the idempotents E_a = 1 (mod p_2i), 0 (mod p_2i+1) and E_b the
    // other way round, digit i is V_i = (r_2i ga_i + r_2i+1 gb_i +
    // sum_j<i V_j h_ij) mod Q_i with ga_i = E_a g_i, gb_i = E_b g_i and
    // h_ij = -M_j g_i, all balanced modulo Q_i.
    double q[kCrtC], invq[kCrtC];                    // Q_i, 1/Q_i (rounded)
    float ga[kCrtC], gb[kCrtC];
    double h[kCrtC][kCrtC];
    // the same for a single last modulus p_2i (odd N): g = M_i^-1 mod p_2i
    double qs[kCrtC], invqs[kCrtC];
    float gs[kCrtC];
    double hs[kCrtC][kCrtC];
    double bias0;                                    // 1/2 + 1/(2 Q_0): rounding offset of the even Q_0
};

constexpr CrtTables crt_tables()
{
    CrtTables c{};
    for (int t = 0; t < kCrtMax; ++t) {
        const int p = kCrtModuli[t];
        c.pd[t] = (double) p;
        c.invpd[t] = 1.0 / (double) p;
        c.c42[t] = crt_pow2(42, p);
        c.c32[t] = crt_pow2(32, p);
        c.c64[t] = crt_pow2(64, p);
    }
    int64_t Q[kCrtC] = {};
    for (int i = 0; i < kCrtC; ++i)
        Q[i] = 2 * i + 1 < kCrtMax ? (int64_t) kCrtModuli[2 * i] * kCrtModuli[2 * i + 1] : kCrtModuli[2 * i];
    for (int i = 0; i < kCrtC; ++i) {
        const int64_t pa = kCrtModuli[2 * i];
        for (int single = 0; single < 2; ++single) {
            if (!single && 2 * i + 1 >= kCrtMax) continue;
            const int64_t qi = single ? pa : Q[i];
            int64_t mi = 1;                          // M_i mod qi
            for (int j = 0; j < i; ++j) mi = mi * (Q[j] % qi) % qi;
            const int64_t gi = crt_inv((int) mi, (int) qi);
            int64_t mj = 1;                          // M_j mod qi
            for (int j = 0; j < kCrtC; ++j) {
                const double hv = j < i ? (double) crt_bal64(-mj * gi, qi) : 0.0;
                if (single) c.hs[i][j] = hv;
                else c.h[i][j] = hv;
                if (j < i) mj = mj * (Q[j] % qi) % qi;
            }
            if (single) {
                c.qs[i] = (double) qi;
                c.invqs[i] = 1.0 / (double) qi;
                c.gs[i] = (float) crt_bal64(gi, qi);
            } else {
                const int64_t pb = kCrtModuli[2 * i + 1];
                const int64_t ea = pb * crt_inv((int) (pb % pa), (int) pa) % qi;
                const int64_t eb = pa * crt_inv((int) (pa % pb), (int) pb) % qi;
                c.q[i] = (double) qi;
                c.invq[i] = 1.0 / (double) qi;
                c.ga[i] = (float) crt_bal64(ea * gi, qi);
                c.gb[i] = (float) crt_bal64(eb * gi, qi);
            }
        }
    }
    c.bias0 = 0.5 + 0.5 / (double) Q[0];
    return c;
}

__constant__ CrtTables c_crt = crt_tables();
// Q_0 = 256 * 255: E_a = -255, E_b = 256; Q_1 = 253 * 251: g_1 = 65280^-1
static_assert(crt_tables().ga[0] == -255.0f && crt_tables().gb[0] == 256.0f && crt_tables().ga[1] == 5271.0f &&
                  crt_tables().gb[1] == -17457.0f && crt_tables().h[1][0] == 12186.0,
              "composite Garner constants");

// Balanced residue modulo an odd p (149 <= p <= 255) of an INT32 x, exact in
// its low byte, in three vector instructions.  FP64 holds x exactly and
// |x / p| < 2^24, so x (1/p) errs by less than 2^-30, while x / p is at least
// 1/(2p) from a half-integer.  fma(x, 1/p, 1.5 2^52) therefore rounds once, to
// 1.5 2^52 + q with q = rint(x / p), and the low word of that double is
// q mod 2^32.  r = x - p q is the balanced residue, |r| <= 127.  The 24-bit
// multiply-add uses q mod 2^24 (exactly q when |q| < 2^23), which moves the
// result by a multiple of 2^24 and leaves its low byte.  negp = -p.
__device__ __forceinline__ int crt_mod_i32(int x, int negp, double invp)
{
    const double t = fma((double) x, invp, 0x1.8p52);
    return x + __mul24((int) __double2loint(t), negp);
}

// Low bytes of r0 .. r3 as one word (byte q = r_q), three byte permutes.
__device__ __forceinline__ uint32_t crt_pack4(int r0, int r1, int r2, int r3)
{
    const uint32_t lo = __builtin_amdgcn_perm((uint32_t) r1, (uint32_t) r0, 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm((uint32_t) r3, (uint32_t) r2, 0x0c0c0400u);
    return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Per line: S = sum over elements of ceil(m4^2 2^(2(ex - E) + 32)), with E
// the frexp exponent of the line's max |x| (from ozaki_max_kernel), ex that
// of the element and m4 = ceil(16 |frac|); x^2 <= 4^E 2^-40 * term.  Integer
// atomics: the sum does not depend on the order.
__device__ __forceinline__ unsigned long long crt_norm_term(double x, int E)
{
    if (x == 0.0) return 0ull;
    int ex;
    const double f = frexp(fabs(x), &ex);
    const unsigned long long m2 = (unsigned long long) ceil(f * 16.0) * (unsigned long long) ceil(f * 16.0);
    const int sh = 2 * (ex - E) + 32;                 // <= 32: ex <= E
    if (sh >= 0) return m2 << sh;
    return ((m2 - 1) >> min(-sh, 62)) + 1;
}

__device__ __forceinline__ void oz2_norm_body(int bx, int by, int gy, const double *__restrict__ X, int lines,
                                              int K, int64_t sl, int64_t sk,
                                              const unsigned long long *__restrict__ mx, unsigned long long *S)
{
    const bool line_contig = (sl == 1);
    const int t = threadIdx.x;
    const int kchunk = (K + gy - 1) / gy;
    const int k0 = by * kchunk, k1 = min(K, k0 + kchunk);
    const int l0 = bx * 64;
    if (!line_contig) {
        // k-contiguous lines: a wave per line at a time, as in oz_max_body
        const int lane = t & 63, w = t >> 6;
        for (int r = 0; r < 16; ++r) {
            const int l = l0 + w * 16 + r;
            if (l >= lines) break;
            if (mx[l] >= kNonFiniteBits) continue;          // flagged by its maximum; no norm needed
            int E;
            (void) frexp(__longlong_as_double((long long) mx[l]), &E);
            const double *xl = X + (int64_t) l * sl;
            unsigned long long acc = 0;
            for (int kk = k0 + lane; kk < k1; kk += 64) acc += crt_norm_term(xl[kk], E);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
            if (lane == 0 && acc != 0) atomicAdd(&S[l], acc);
        }
        return;
    }
    __shared__ unsigned long long red[256];
    unsigned long long acc = 0;
    const int myline = l0 + (t & 63);
    if (myline < lines && mx[myline] < kNonFiniteBits) {
        int E;
        (void) frexp(__longlong_as_double((long long) mx[myline]), &E);
        for (int kk = k0 + (t >> 6); kk < k1; kk += 4) acc += crt_norm_term(X[myline + (int64_t) kk * sk], E);
    }
    red[t] = acc;
    __syncthreads();
    if (t < 64) {
        const int l = l0 + t;
        unsigned long long v = 0;
        for (int j = 0; j < 4; ++j) v += red[t + 64 * j];
        if (l < lines && v != 0) atomicAdd(&S[l], v);
    }
}

__global__ __launch_bounds__(256) void ozaki2_norm_kernel(const double *__restrict__ X, int lines, int K,
                                                           int64_t sl, int64_t sk,
                                                           const unsigned long long *__restrict__ mx,
                                                           unsigned long long *S)
{
    oz2_norm_body(blockIdx.x, blockIdx.y, gridDim.y, X, lines, K, sl, sk, mx, S);
}

// e = T + 20 - E - h with h the smallest integer >= 0 such that
// S <= 4^h - ceil(2897 2^(2h - T)): then 4^(T-h) S <= (2^T - 2897/2)^2, so
// 2^e ||line||_2 plus the rounding slack sqrt(k)/2 <= 2897/2 (k <= 2^23)
// stays within 2^T.  The search starts at ceil(bitlen(S - 1) / 2) - 1,
// below which 4^h < S.  kExpEmpty for a zero line, kExpNonFinite for a line
// with an Inf or a NaN.
constexpr int kCrtSlack = 2897;     // ceil(sqrt(kCrtKMax))
static_assert((long long) kCrtSlack * kCrtSlack >= kCrtKMax &&
                  (long long) (kCrtSlack - 1) * (kCrtSlack - 1) < kCrtKMax,
              "rounding slack bound");
__global__ void ozaki2_scale_kernel(const unsigned long long *__restrict__ mx, const unsigned long long *__restrict__ S,
                                    int n, int T, int *__restrict__ e)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    if (mx[l] >= kNonFiniteBits) {
        e[l] = kExpNonFinite;
        return;
    }
    const double v = __longlong_as_double((long long) mx[l]);
    if (v == 0.0) {
        e[l] = kExpEmpty;
        return;
    }
    int E;
    (void) frexp(v, &E);
    typedef unsigned __int128 u128;
    const u128 s = S[l];
    const int b = s >= 2 ? 64 - __clzll((long long) (S[l] - 1)) : 0;
    int h = max(0, (b + 1) / 2 - 1);
    for (;; ++h) {
        const int sh = 2 * h - T;
        const u128 sub = sh >= 0 ? (u128) kCrtSlack << sh
                                 : ((u128) kCrtSlack + (((u128) 1 << -sh) - 1)) >> -sh;
        if (s + sub <= ((u128) 1 << (2 * h))) break;
    }
    e[l] = T + 20 - E - h;
}

// Residue planes: a' = rint(x 2^e) (an integer-valued double, |a'| <= 2^T)
// is cut exactly into FP64 pieces, NP = 2: a' = h 2^42 + l (T <= 84, |h|,
// |l| <= 2^42); NP = 3: a' = h2 2^64 + h1 2^32 + l (|h1|, |l| < 2^32).
// Then s = h c42 + l (or h2 c64 + h1 c32 + l) = a' (mod p) with c = 2^k mod
// p balanced: |s| < 2^50, exact in FP64, and its balanced remainder comes
// from one rounded quotient (|s / p| < 2^43, error < 2^-9 < 1/(2p)).  For
// p_0 = 256 the residue is a' mod 256 = l - 256 floor(l / 256).  Plane t:
// P[t * plane + line * Kp + kk].
template <int NP>
__device__ __forceinline__ void oz2_residue_tiles(int tau0, int step, int ntiles, int nbx,
                                                  const double *__restrict__ X, int lines, int K, int64_t sl,
                                                  int64_t sk, const int *__restrict__ ex, int nmod,
                                                  int8_t *__restrict__ P, int64_t plane, int Kp)
{
    __shared__ double tile[64][65];
    const int t = threadIdx.x;
    if (tau0 >= ntiles) return;
    // element p of this thread in a 64 x 64 tile: (line li, k ki), along the
    // operand's contiguous direction across the wave
    auto at = [&](int p, int &li, int &ki) __attribute__((always_inline)) {
        if (sl == 1) { li = t & 63; ki = (t >> 6) + 4 * p; }
        else { ki = t & 63; li = (t >> 6) + 4 * p; }
    };
    double v[16];
    int exn;                                            // exponent of this thread's line in the fetched tile
    auto fetch = [&](int tau) __attribute__((always_inline)) {
        const int l0 = (tau % nbx) * 64, k0 = (tau / nbx) * 64;
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            int li, ki;
            at(p, li, ki);
            const int l = l0 + li, kk = k0 + ki;
            v[p] = (l < lines && kk < K) ? X[sl == 1 ? l + (int64_t) kk * sk : (int64_t) l * sl + kk] : 0.0;
        }
        // loaded with the tile: a load issued after the prefetch and waited
        // for at once would also wait for the prefetch (vmcnt retires in
        // issue order)
        exn = ex[min(l0 + (t >> 2), lines - 1)];
    };
    fetch(tau0);
    for (int tau = tau0; tau < ntiles; tau += step) {
        __syncthreads();                                // the previous tile's LDS reads are done
#pragma unroll
        for (int p = 0; p < 16; ++p) {
            int li, ki;
            at(p, li, ki);
            tile[li][ki] = v[p];
        }
        __syncthreads();
        const int exc = exn;
        if (tau + step < ntiles) fetch(tau + step);     // lands while this tile is reduced
        const int l0 = (tau % nbx) * 64, k0 = (tau / nbx) * 64;
        const int li = t >> 2, kq = (t & 3) * 16;
        const int l = l0 + li;
        const bool zero_line = (l >= lines) || exc == kExpEmpty || exc == kExpNonFinite;
        const int e = zero_line ? 0 : exc;
        double ph[16], pm[16], pl[16];                  // NP = 2: (-, h, l); NP = 3: (h2, h1, l)
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            double a = zero_line ? 0.0 : rint(ldexp(tile[li][kq + j], e));
            if (NP == 3) {
                ph[j] = trunc(a * 0x1p-64);
                a = fma(-ph[j], 0x1p64, a);
                pm[j] = trunc(a * 0x1p-32);
                pl[j] = fma(-pm[j], 0x1p32, a);
            } else {
                pm[j] = trunc(a * 0x1p-42);
                pl[j] = fma(-pm[j], 0x1p42, a);
            }
        }
        int8_t *base = P + (int64_t) l * Kp + k0 + kq;
        {
            u32x4 w;                                    // p_0 = 256: the low byte of l
            int lo[16];                                 // a' mod 256 in [0, 256), exact
#pragma unroll
            for (int j = 0; j < 16; ++j) lo[j] = (int) fma(-floor(pl[j] * 0x1p-8), 256.0, pl[j]);
#pragma unroll
            for (int q = 0; q < 4; ++q) w[q] = crt_pack4(lo[4 * q], lo[4 * q + 1], lo[4 * q + 2], lo[4 * q + 3]);
            __builtin_nontemporal_store(w, reinterpret_cast<u32x4 *>(base));
        }
        for (int m = 1; m < nmod; ++m) {
            const double ch = c_crt.c64[m], cm = NP == 3 ? c_crt.c32[m] : c_crt.c42[m];
            const double p = c_crt.pd[m], ip = c_crt.invpd[m];
            u32x4 w;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int r[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int j = 4 * q + b;
                    double sv = fma(pm[j], cm, pl[j]);
                    if (NP == 3) sv = fma(ph[j], ch, sv);
                    r[b] = (int) fma(-rint(sv * ip), p, sv);
                }
                w[q] = crt_pack4(r[0], r[1], r[2], r[3]);
            }
            __builtin_nontemporal_store(w, reinterpret_cast<u32x4 *>(base + (int64_t) m * plane));
        }
    }
}

// One workgroup per (nbx * nby) / gridDim.x tiles of 64 lines x 64 k, tile
// tau = bx + by * nbx taken in grid-stride order.
template <int NP>
__global__ __launch_bounds__(256) void ozaki2_residue_kernel(const double *__restrict__ X, int lines, int K,
                                                              int64_t sl, int64_t sk, const int *__restrict__ ex,
                                                              int nmod, int8_t *__restrict__ P, int64_t plane, int Kp,
                                                              int nbx, int ntiles)
{
    oz2_residue_tiles<NP>(blockIdx.x, gridDim.x, ntiles, nbx, X, lines, K, sl, sk, ex, nmod, P, plane, Kp);
}

// Spread statistics for the slice / modulus count chooser: for sampled
// entry s of C = op(A) op(B), i_s = (s 2654435761 + 2135587861) mod 2^32 mod
// m and j_s = (s 2246822519 + 3266489917) mod 2^32 mod n, with a = |row i_s
// of op(A)| and b = |column j_s of op(B)|: sum a b, sum a^2, sum b^2,
// sum a, sum b, max a, max b (one workgroup per sample; an estimate, so the
// summation order is free).
constexpr int kSpreadSamples = 4096;

__device__ __forceinline__ void spread_accumulate(double x, double y, double v[7])
{
    v[0] = fma(x, y, v[0]);
    v[1] = fma(x, x, v[1]);
    v[2] = fma(y, y, v[2]);
    v[3] += x;
    v[4] += y;
    v[5] = fmax(v[5], x);
    v[6] = fmax(v[6], y);
}

// Workgroup tree over the 256 threads' partial statistics; thread q < 7
// writes statistic q.
__device__ __forceinline__ void spread_reduce_store(const double v[7], double *__restrict__ out)
{
    __shared__ double red[7][256];
#pragma unroll
    for (int q = 0; q < 7; ++q) red[q][threadIdx.x] = v[q];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) {
#pragma unroll
            for (int q = 0; q < 7; ++q) {
                const double o = red[q][threadIdx.x + w];
                red[q][threadIdx.x] = q >= 5 ? fmax(red[q][threadIdx.x], o) : red[q][threadIdx.x] + o;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) out[threadIdx.x] = red[threadIdx.x][0];
}
__global__ __launch_bounds__(256) void ozaki_spread_kernel(const double *__restrict__ A, int64_t a_sl, int64_t a_sk,
                                                            int m, const double *__restrict__ B, int64_t b_sl,
                                                            int64_t b_sk, int n, int k, double *__restrict__ out)
{
    const int s = blockIdx.x;
    const uint64_t i = (((uint64_t) s * 2654435761ull + 2135587861ull) & 0xffffffffull) % (uint64_t) m;
    const uint64_t j = (((uint64_t) s * 2246822519ull + 3266489917ull) & 0xffffffffull) % (uint64_t) n;
    const double *a = A + (int64_t) i * a_sl;
    const double *b = B + (int64_t) j * b_sl;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int l = threadIdx.x; l < k; l += 256)
        spread_accumulate(fabs(a[(int64_t) l * a_sk]), fabs(b[(int64_t) l * b_sk]), v);
    spread_reduce_store(v, out + (int64_t) s * 7);
}

// The statistics of sampled entries of the grouped launch's problems (the
// local products of a simulated grid): sample s is entry (i, j) of problem
// prob, and a, b run over row i of every k-segment's A and column j of its B,
// so the sums cover the problem's whole k (segments in order).
struct SpreadProb {
    const double *A[ca3::grouped::kMaxSeg];
    const double *B[ca3::grouped::kMaxSeg];
    int64_t lda[ca3::grouped::kMaxSeg], ldb[ca3::grouped::kMaxSeg];
    int k[ca3::grouped::kMaxSeg];
    int nseg;
};

__global__ __launch_bounds__(256) void ozaki_grid_spread_kernel(const SpreadProb *__restrict__ probs,
                                                                 const int *__restrict__ samples,
                                                                 double *__restrict__ out)
{
    const int s = blockIdx.x;
    const SpreadProb &P = probs[samples[3 * s]];
    const int i = samples[3 * s + 1], j = samples[3 * s + 2];
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int g = 0; g < P.nseg; ++g) {
        const double *a = P.A[g] + i, *b = P.B[g] + (int64_t) j * P.ldb[g];
        const int64_t lda = P.lda[g];
        for (int l = threadIdx.x; l < P.k[g]; l += 256)
            spread_accumulate(fabs(a[(int64_t) l * lda]), fabs(b[l]), v);
    }
    spread_reduce_store(v, out + (int64_t) s * 7);
}

// Exact FP64 pieces of a' = rint(x 2^e), |a'| <= 2^T (the residue kernel's
// bound): two pieces (a' = h 2^42 + l) for T <= 84, three 32-bit pieces
// for 84 < T <= 95.
constexpr int crt_pieces(int nmod) { return kCrtT[nmod] <= 84 ? 2 : 3; }

// Residue planes are stored in the MFMA's own order: a plane is Mp x Np
// int8 in 16 x 16 blocks of 256 contiguous bytes, block (rb, cb) at
// (cb * Mp / 16 + rb) * 256; inside a block, lane l's 4-byte word (at 4 l)
// holds rows rb * 16 + l % 16, columns cb * 16 + 4 (l / 16) + q for byte q,
// which is the accumulator layout of v_mfma_i32_16x16x64_i8 with operands
// fed (B, A).  A wave then stores and loads whole 256-byte runs.
__device__ __forceinline__ int64_t crt_block_word(int Mp, int rb, int cb)
{
    return ((int64_t) cb * (Mp >> 4) + rb) * 64;
}

// Epilogue of a CRT segment (modulus t, one k-chunk): R_t = (R_t + Z) mod p_t
// (R_t = Z mod p_t on the first chunk), over the whole padded tile.  For
// p_0 = 256 the balanced residue is the low byte.  The plane and block
// addresses are wave-uniform (scalar registers) and each lane adds only its
// 4-byte slot, so no per-block 64-bit address is kept live across the
// segment loop and the 32 stores of a lane stream without waits.
template <int WTM, int WTN, int FI, int FJ>
__device__ __forceinline__ void oz_epilogue_crt(const OzArgs &g, const i32x4 (&acc)[FI][FJ], int m0, int n0, int t,
                                                bool first)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave & 1, wn = wave >> 1;
    const int negp = -(int) c_crt.pd[t];
    const double ip = c_crt.invpd[t];
    typedef __attribute__((address_space(1))) uint32_t gu32;
    gu32 *Rt = (gu32 *) oz_rflp(reinterpret_cast<uint32_t *>(g.R8 + (int64_t) t * g.rplane));
    const int mp16 = __builtin_amdgcn_readfirstlane(g.Mp >> 4);
    const int rb0 = __builtin_amdgcn_readfirstlane((m0 + wm * WTM) >> 4);
    const int cb0 = __builtin_amdgcn_readfirstlane((n0 + wn * WTN) >> 4);
    // the first chunk loads nothing, so its stores never wait (vmcnt also
    // counts stores); later chunks add the stored residue back in
    auto run = [&](auto ODD, auto FIRST) __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < FI; ++i) {
#pragma unroll
            for (int j = 0; j < FJ; ++j) {
                gu32 *wp = Rt + (int64_t) ((cb0 + j) * mp16 + rb0 + i) * 64;
                const uint32_t prev = decltype(FIRST)::value ? 0u : wp[lane];
                int res[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int v = acc[i][j][r] + (int) (int8_t) (prev >> (8 * r));
                    res[r] = decltype(ODD)::value ? crt_mod_i32(v, negp, ip) : v;
                }
                wp[lane] = crt_pack4(res[0], res[1], res[2], res[3]);
            }
        }
    };
    if (first) {
        if (t == 0) run(std::false_type{}, std::true_type{});
        else run(std::true_type{}, std::true_type{});
    } else {
        if (t == 0) run(std::false_type{}, std::false_type{});
        else run(std::true_type{}, std::false_type{});
    }
}

// One lane's 4 elements of block (rb, cb): row rb * 16 + lane % 16,
// columns cb * 16 + 4 (lane / 16) + q.  Garner's balanced digits V_i over
// the composite moduli: the member residues enter in FP32 (|r ga + r' gb|
// < 2^24, exact), the digit sum and its reduction run in FP64 (|sum| <
// 2^35 and the rounded quotient errs by far less than 1/(2 Q_i), so rint
// finds the balanced remainder; floor with the offset 1/2 + 1/(2 Q_0) gives
// [-Q_0/2, Q_0/2) for the even Q_0).  Then x = V_0 + Q_0 (V_1 + Q_1 (...))
// in FP64 from the top digit, and 2^-(eA + eB) x (0 for an empty line, NaN
// for a line with a non-finite element).
template <int NM>
__device__ __forceinline__ void oz2_crt_block(const int8_t *__restrict__ R, int64_t rplane, int Mp, int rb, int cb,
                                              int lane, const int *__restrict__ eA, const int *__restrict__ eB,
                                              double (&out)[4])
{
    constexpr int NC = (NM + 1) / 2;
    // rb and cb are wave-uniform: the block's plane addresses stay in scalar
    // registers and a lane adds only its 4-byte slot
    typedef __attribute__((address_space(1))) const uint32_t gcu32;
    const uint32_t *W = reinterpret_cast<const uint32_t *>(R) + crt_block_word(Mp, rb, cb);
    uint32_t w[NM];
#pragma unroll
    for (int t = 0; t < NM; ++t) w[t] = __builtin_nontemporal_load((gcu32 *) oz_rflp(W + t * (rplane >> 2)) + lane);
    const int ea = eA[rb * 16 + (lane & 15)];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double V[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const bool pair = 2 * i + 1 < NM;
            const float ra = (float) (int8_t) (w[2 * i] >> (8 * q));
            const float s32 = pair ? fmaf((float) (int8_t) (w[2 * i + 1] >> (8 * q)), c_crt.gb[i], ra * c_crt.ga[i])
                                   : ra * c_crt.gs[i];
            double s = (double) s32;
#pragma unroll
            for (int j = 0; j < i; ++j) s = fma(V[j], pair ? c_crt.h[i][j] : c_crt.hs[i][j], s);
            if (i == 0) {
                V[i] = fma(-floor(fma(s, c_crt.invq[0], c_crt.bias0)), c_crt.q[0], s);
            } else {
                const double qi = pair ? c_crt.q[i] : c_crt.qs[i], iq = pair ? c_crt.invq[i] : c_crt.invqs[i];
                V[i] = fma(-rint(s * iq), qi, s);
            }
        }
        double x = V[NC - 1];
#pragma unroll
        for (int i = NC - 2; i >= 0; --i) x = ca3::add_rn(ca3::mul_rn(x, c_crt.q[i]), V[i]);
        const int eb = eB[cb * 16 + 4 * (lane >> 4) + q];
        out[q] = (ea == kExpNonFinite || eb == kExpNonFinite) ? __builtin_nan("")
                 : (ea != kExpEmpty && eb != kExpEmpty) ? ldexp(x, -(ea + eb)) : 0.0;
    }
}

// C = alpha * (the reconstructed product) (+ beta C).  A workgroup takes 4
// blocks (one per wave), blocks ordered column of blocks by column of blocks.
template <int NM>
__global__ __launch_bounds__(256) void ozaki2_crt_kernel(const int8_t *__restrict__ R, int64_t rplane, int Mp, int M,
                                                          int N, const int *__restrict__ eA,
                                                          const int *__restrict__ eB, double alpha, double beta,
                                                          double *C, int64_t ldc)
{
    const int nbm = (M + 15) >> 4, nbn = (N + 15) >> 4;
    const int64_t b = (int64_t) blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int) threadIdx.x >> 6);
    if (b >= (int64_t) nbm * nbn) return;
    const int lane = threadIdx.x & 63;
    const int cb = (int) (b / nbm), rb = (int) (b - (int64_t) cb * nbm);
    const int row = rb * 16 + (lane & 15);
    if (row >= M) return;
    double x[4];
    oz2_crt_block<NM>(R, rplane, Mp, rb, cb, lane, eA, eB, x);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int col = cb * 16 + 4 * (lane >> 4) + q;
        if (col >= N) break;
        double *cp = C + (int64_t) col * ldc + row;
        double v = ca3::mul_rn(alpha, x[q]);
        if (beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(beta, *cp));
        __builtin_nontemporal_store(v, cp);
    }
}

// Grouped reconstruction: one job per fold group of the simulated grid (kp
// layer products of the same C block, added in layer order, each column
// written to its owner's C_out slice) or per problem (kp = 1).  A job's
// workgroups take 4 blocks each, as ozaki2_crt_kernel.
constexpr int kGrpMaxKp = 16;
struct CrtJob {
    int kp, M, N, Mp, nbm, nwg;           // nbm: ceil(M / 16); nwg: workgroups of the job
    const int8_t *R[kGrpMaxKp];           // the layers' residue planes (rplane apart)
    int64_t rplane;
    const int *eA[kGrpMaxKp], *eB[kGrpMaxKp];
    int cs[kGrpMaxKp + 1];                // column slice t = [cs[t], cs[t+1]) goes to Cout[t]
    double *Cout[kGrpMaxKp];
    int64_t ldo;
};

template <int NM>
__global__ __launch_bounds__(256) void ozaki2_crt_grouped_kernel(const CrtJob *__restrict__ jobs,
                                                                  const int *__restrict__ blk0, int njobs)
{
    const int wg = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= wg) lo = mid; else hi = mid - 1;
    }
    const CrtJob &J = jobs[lo];
    const int64_t b = (int64_t) (wg - blk0[lo]) * 4 + __builtin_amdgcn_readfirstlane((int) threadIdx.x >> 6);
    const int nbn = (J.N + 15) >> 4;
    if (b >= (int64_t) J.nbm * nbn) return;
    const int lane = threadIdx.x & 63;
    const int cb = (int) (b / J.nbm), rb = (int) (b - (int64_t) cb * J.nbm);
    const int row = rb * 16 + (lane & 15);
    if (row >= J.M) return;
    double sum[4];
    for (int l = 0; l < J.kp; ++l) {
        double x[4];
        oz2_crt_block<NM>(J.R[l], J.rplane, J.Mp, rb, cb, lane, J.eA[l], J.eB[l], x);
#pragma unroll
        for (int q = 0; q < 4; ++q) sum[q] = (l == 0) ? x[q] : ca3::add_rn(sum[q], x[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int col = cb * 16 + 4 * (lane >> 4) + q;
        if (col >= J.N) break;
        int sl = 0;
        while (col >= J.cs[sl + 1]) ++sl;
        __builtin_nontemporal_store(sum[q], &J.Cout[sl][(int64_t) (col - J.cs[sl]) * J.ldo + row]);
    }
}

// Grouped preparation: one job per (problem, operand, k-segment); kind 0
// line maxima, 1 norm sums, 2 residue planes.  A block finds its job in the
// prefix of block counts.
struct PrepJob {
    const double *X;
    int64_t sl, sk;                        // element (line, kk) at X[line * sl + kk * sk]
    int lines, K;
    int rbx, rby;                          // grid of the max / norm passes
    int sbx, sby;                          // grid of the residue pass
    unsigned long long *mx, *S;            // this operand's lines
    int *ex;
    int8_t *P;                             // planes, already offset to the segment's k
    int64_t plane;
    int Kp;
};

template <int KIND, int NP = 3>
__global__ __launch_bounds__(256) void ozaki2_prep_grouped_kernel(const PrepJob *__restrict__ jobs,
                                                                   const int *__restrict__ blk0, int njobs,
                                                                   int nmod)
{
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk0[mid] <= b) lo = mid; else hi = mid - 1;
    }
    const PrepJob &J = jobs[lo];
    const int local = b - blk0[lo];
    if (KIND == 0) {
        oz_max_body(local % J.rbx, local / J.rbx, J.rby, J.X, J.lines, J.K, J.sl, J.sk, J.mx);
    } else if (KIND == 1) {
        oz2_norm_body(local % J.rbx, local / J.rbx, J.rby, J.X, J.lines, J.K, J.sl, J.sk, J.mx, J.S);
    } else {
        oz2_residue_tiles<NP>(local, 1 << 30, J.sbx * J.sby, J.sbx, J.X, J.lines, J.K, J.sl, J.sk, J.ex, nmod,
                              J.P, J.plane, J.Kp);
    }
}

// The fields a CRT segment loop reads, copied through readfirstlane so they
// live in scalar registers (as kernel arguments do) and the loop's address
// arithmetic stays scalar.
__device__ __forceinline__ OzArgs oz_uniform_crt_args(const OzArgs &a)
{
    OzArgs g;
    memset(&g, 0, sizeof(g));
    g.M = __builtin_amdgcn_readfirstlane(a.M);
    g.N = __builtin_amdgcn_readfirstlane(a.N);
    g.Mp = __builtin_amdgcn_readfirstlane(a.Mp);
    g.Np = __builtin_amdgcn_readfirstlane(a.Np);
    g.Kp = __builtin_amdgcn_readfirstlane(a.Kp);
    g.s = __builtin_amdgcn_readfirstlane(a.s);
    g.nchunks = __builtin_amdgcn_readfirstlane(a.nchunks);
    g.kcb = __builtin_amdgcn_readfirstlane(a.kcb);
    g.PA = oz_rflp(a.PA);
    g.PB = oz_rflp(a.PB);
    g.planeA = oz_rfl64(a.planeA);
    g.planeB = oz_rfl64(a.planeB);
    g.R8 = oz_rflp(a.R8);
    g.rplane = oz_rfl64(a.rplane);
    return g;
}

// Software-pipelined form (the default).  Direct-to-LDS loads fill a ring
// of four 64-byte stages; the fragments of stage z+1 are read from LDS into
// a second register set while the MFMAs of stage z run, so the MFMA pipe
// does not wait for LDS.  At the top of stage z a wave waits for its own
// loads of stage z+1 (vmcnt <= L: only stage z+2's loads may still be in
// flight), then one fence-free barrier makes every wave's part of stage z+1
// visible and frees the slot of stage z-1 (its fragments were read during
// stage z-2), into which the loads of stage z+3 then go.
template <int FI, int FJ>
__device__ __forceinline__ void oz_frags(const int8_t *As, const int8_t *Bs, int wm, int wn, int li, int kq,
                                         i32x4 (&af)[FI], i32x4 (&bf)[FJ])
{
#pragma unroll
    for (int j = 0; j < FJ; ++j)
        bf[j] = *reinterpret_cast<const i32x4 *>(&Bs[lds_off(wn * FJ * 16 + j * 16 + li, kq)]);
#pragma unroll
    for (int i = 0; i < FI; ++i)
        af[i] = *reinterpret_cast<const i32x4 *>(&As[lds_off(wm * FI * 16 + i * 16 + li, kq)]);
}

template <int FI, int FJ>
__device__ __forceinline__ void oz_mfmas(const i32x4 (&af)[FI], const i32x4 (&bf)[FJ], i32x4 (&acc)[FI][FJ])
{
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(bf[j], af[i], acc[i][j], 0, 0, 0);
}

// FUSED (per-diagonal form only): one launch for all diagonals; each
// workgroup walks its tile's (diagonal, k-chunk) segments in the order of
// the specification (d = s-1 .. 0, chunks ascending), folding each into W
// (the last into C), so every element sees the same FP64 additions as with
// one launch per segment, while one tile's W traffic overlaps the other
// tiles' MFMAs and only one launch tail remains.
// NW waves per workgroup in a 2 x NW/2 grid (wave w: rows (w & 1) * WTM,
// columns (w >> 1) * WTN); with NW = 8 two waves share each SIMD and cover
// each other's waits.
template <bool SPLIT, int WTM, int WTN, bool FUSED, int NW, bool CRT = false, bool GROUPED = false>
__global__ __launch_bounds__(64 * NW, 1) void ozaki_pipe_kernel(OzArgs g0, const OzArgs *table = nullptr)
{
    static_assert(!(SPLIT && FUSED), "the split form has its own fold");
    static_assert(!CRT || (FUSED && !SPLIT), "the CRT form runs fused");
    static_assert(!GROUPED || CRT, "the grouped launch runs the CRT form");
    constexpr int NBUF = 4;
    constexpr int kBMo = 2 * WTM, kBNo = (NW / 2) * WTN;
    constexpr int FI = WTM / 16, FJ = WTN / 16;
    constexpr int WA = kBMo / (16 * NW), WB = kBNo / (16 * NW);  // 1 KB blocks per wave per stage
    static_assert(WA * 16 * NW == kBMo && WB * 16 * NW == kBNo && WA > 0 && WB > 0, "load split");
    // 4 waves: fragments of stage z+1 prefetched into a second register set;
    // 8 waves: no room for it (128 VGPRs per wave), the other wave of the
    // SIMD covers the fragment reads
    constexpr bool PREFETCH = (NW == 4);
    // CRT form with 8 waves: the load stream runs on across segments (the
    // next modulus' first stages are loaded during the last stages of the
    // current one and land while its epilogue runs).  Not in the grouped
    // kernel, where the extra state makes the compiler shuffle accumulators.
    constexpr bool XSEG = CRT && !PREFETCH && !GROUPED;
    constexpr int L = WA + WB;
    static_assert(kBMo % kBNo == 0 && kTile % kBMo == 0 && kTile % kBNo == 0, "tile shapes");
    static_assert((kKp / kKb) % NBUF == 0 && L <= 63, "ring turns per chunk / vmcnt range");
    __shared__ __attribute__((aligned(1024))) int8_t A0[kBMo * kKb], A1[kBMo * kKb], A2[kBMo * kKb], A3[kBMo * kKb];
    __shared__ __attribute__((aligned(1024))) int8_t B0[kBNo * kKb], B1[kBNo * kKb], B2[kBNo * kKb], B3[kBNo * kKb];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);     // uniform: LDS targets stay scalar
    const int wm = wave & 1, wn = wave >> 1;
    // one output tile (tm, tn) of the product described by g
    auto body = [&](const OzArgs &g, const int tm, const int tn, int d, int kb0, int nkb, const int64_t part_off)
        __attribute__((always_inline)) {
    const int m0 = tm * kBMo, n0 = tn * kBNo;

    const int lr = lane >> 2;
    const int lc = ((lane & 3) ^ oz_swz(lr)) << 4;
    const int8_t *gA = g.PA + (int64_t) (m0 + wave * WA * 16 + lr) * g.Kp + lc;
    const int8_t *gB = g.PB + (int64_t) (n0 + wave * WB * 16 + lr) * g.Kp + lc;
    const int64_t blk16 = (int64_t) 16 * g.Kp;
    const int nseg = FUSED ? g.s * g.nchunks : 1;
    int it = 0, ib = 0, inext = 0;
    int idd = 0, icc = 0;                     // XSEG: modulus and chunk of the next stage to load
    for (int sg = 0; sg < nseg; ++sg) {
    bool first = g.first != 0, last = g.last != 0;
    if (FUSED) {
        // slices: diagonal d = s-1 .. 0; CRT: modulus d = 0 .. s-1 (pair (d, d))
        d = CRT ? sg / g.nchunks : g.s - 1 - sg / g.nchunks;
        const int c = sg - (sg / g.nchunks) * g.nchunks;
        kb0 = c * g.kcb;
        nkb = min(g.kcb, g.Kp / kKb - kb0);
        first = CRT ? (c == 0) : (sg == 0);
    }
    const int total = (CRT ? 1 : d + 1) * nkb;  // 64-byte stages; a multiple of NBUF
    if (!XSEG) {
        it = 0;
        ib = 0;
        inext = 0;
    }

    auto slotA = [&](auto J) -> int8_t * {
        if constexpr (J.value == 0) return A0;
        else if constexpr (J.value == 1) return A1;
        else if constexpr (J.value == 2) return A2;
        else return A3;
    };
    auto slotB = [&](auto J) -> int8_t * {
        if constexpr (J.value == 0) return B0;
        else if constexpr (J.value == 1) return B1;
        else if constexpr (J.value == 2) return B2;
        else return B3;
    };
    // issue pointer: (pair it, block ib) of the next stage to load, advanced
    // by one stage per issue and held at the last stage past the end
    auto issue = [&](int /*z*/, auto J) __attribute__((always_inline)) {
        const int8_t *pa, *pb;
        if constexpr (XSEG) {
            // (modulus idd, chunk icc, block ib): on to the next segment at
            // the end of one, held at the last stage past the end of the last
            const int kb0i = icc * g.kcb;
            const int nkbi = min(g.kcb, g.Kp / kKb - kb0i);
            const int b = ib, nb = ib + 1;
            const bool wrap = (nb == nkbi);
            const bool last_chunk = (icc + 1 == g.nchunks);
            const bool more_seg = !last_chunk || (idd + 1 < g.s);
            const int64_t koff = (int64_t) (kb0i + b) * kKb;
            pa = gA + (int64_t) idd * g.planeA + koff;
            pb = gB + (int64_t) idd * g.planeB + koff;
            ib = wrap ? (more_seg ? 0 : ib) : nb;
            const bool step_seg = wrap && more_seg;
            idd = (step_seg && last_chunk) ? idd + 1 : idd;
            icc = step_seg ? (last_chunk ? 0 : icc + 1) : icc;
        } else {
            const int t = it, b = ib;
            const bool more = (inext + 1 < total);
            inext = more ? inext + 1 : inext;
            const int nb = ib + 1;
            ib = more ? (nb == nkb ? 0 : nb) : ib;
            it = more ? (nb == nkb ? it + 1 : it) : it;
            const int64_t koff = (int64_t) (kb0 + b) * kKb;
            pa = gA + (int64_t) (CRT ? d : t) * g.planeA + koff;
            pb = gB + (int64_t) (d - t) * g.planeB + koff;
        }
        int8_t *sa = slotA(J), *sb = slotB(J);
#pragma unroll
        for (int i = 0; i < WA; ++i)
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *) (pa + i * blk16),
                                             (void __attribute__((address_space(3))) *) (sa + (wave * WA + i) * 1024),
                                             16, 0, 0);
#pragma unroll
        for (int i = 0; i < WB; ++i)
            __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *) (pb + i * blk16),
                                             (void __attribute__((address_space(3))) *) (sb + (wave * WB + i) * 1024),
                                             16, 0, 0);
    };
    constexpr int kWaitL = (L & 15) | (((L >> 4) & 3) << 14) | (7 << 4) | (15 << 8);      // vmcnt <= L
    constexpr int kWaitL2 = ((2 * L) & 15) | ((((2 * L) >> 4) & 3) << 14) | (7 << 4) | (15 << 8);
    auto sync = [&](int wait) __attribute__((always_inline)) {
        (void) wait;
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
    };

    i32x4 acc[FI][FJ];
    oz_zero(acc);
    const int li = lane & 15, kq = lane >> 4;
    i32x4 af0[FI], bf0[FJ], af1[FI], bf1[FJ];

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    // prologue: stages 0, 1, 2 in flight; fragments of stage 0 (XSEG: only
    // before the first segment, later ones were loaded by the previous one)
    if (!XSEG || sg == 0) {
        issue(0, I0{});
        issue(min(1, total - 1), I1{});
        issue(min(2, total - 1), I2{});
    }
    if constexpr (PREFETCH) {
        __builtin_amdgcn_s_waitcnt(kWaitL2);  // own loads of stage 0 done
        sync(0);
        oz_frags(A0, B0, wm, wn, li, kq, af0, bf0);
    }

    // stage z (slot j = z % 4).  PREFETCH: publish stage z+1, refill slot
    // (z-1) % 4 with stage z+3, read the fragments of stage z+1 into set
    // (z+1) % 2, run the MFMAs of stage z from set z % 2.  Otherwise: publish
    // stage z, refill slot (z-1) % 4, read and run stage z.
    auto step = [&](int z, auto J, i32x4 (&afc)[FI], i32x4 (&bfc)[FJ], i32x4 (&afn)[FI], i32x4 (&bfn)[FJ])
        __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        // scheduling barriers keep everything of stage z after the LDS
        // barrier (MFMAs touch only registers, so the compiler would
        // otherwise move them across it)
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(PREFETCH ? kWaitL : kWaitL2);
        sync(0);
        __builtin_amdgcn_sched_barrier(0);
        issue(z + 3, std::integral_constant<int, (j + 3) % 4>{});
        if constexpr (PREFETCH) {
            using JN = std::integral_constant<int, (j + 1) % 4>;
            oz_frags(slotA(JN{}), slotB(JN{}), wm, wn, li, kq, afn, bfn);
            oz_mfmas(afc, bfc, acc);
        } else {
            oz_frags(slotA(J), slotB(J), wm, wn, li, kq, afc, bfc);
            oz_mfmas(afc, bfc, acc);
        }
        // One wave per SIMD: nothing else fills the MFMA pipe, so the reads
        // of stage z+1 and the loads of stage z+3 are threaded between the
        // MFMAs of stage z (one after each MFMA), then the rest of the MFMAs
        // run back to back.  Masks: 0x008 MFMA, 0x100 LDS read, 0x020 VMEM read.
        if constexpr (PREFETCH) {
#pragma unroll
            for (int q = 0; q < FI + FJ; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
            }
#pragma unroll
            for (int q = 0; q < L; ++q) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
            }
            __builtin_amdgcn_sched_group_barrier(0x008, FI * FJ - FI - FJ - L, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    for (int z = 0; z < total; z += 4) {
        step(z, I0{}, af0, bf0, af1, bf1);
        step(z + 1, I1{}, af1, bf1, af0, bf0);
        step(z + 2, I2{}, af0, bf0, af1, bf1);
        step(z + 3, I3{}, af1, bf1, af0, bf0);
    }
    // the loads and fragment reads past the last stage were repeats of it;
    // drain the loads before the workgroup can end (or reuse the ring).
    // XSEG: the loads in flight are the next segment's first stages; they
    // are drained after the epilogue (vmcnt also counts its stores)
    if (!XSEG) __builtin_amdgcn_s_waitcnt(0);
    if (CRT) {
        oz_epilogue_crt<WTM, WTN>(g, acc, m0, n0, d, first);
        if (XSEG) __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    } else if (FUSED) {
        oz_epilogue_w<WTM, WTN>(g, acc, m0, n0, d, first);
        __syncthreads();                      // every wave is done with the ring
    } else {
        oz_epilogue<SPLIT, WTM, WTN>(g, acc, m0, n0, d, part_off, first, last);
    }
    }
    };

    if constexpr (GROUPED) {
        // persistent: items (problem, tile), problem p owning items
        // [tile0[p], tile0[p + 1]).  The items are cut into 8 contiguous
        // ranges, one per XCD (workgroup b runs on XCD b % 8), each with its
        // own counter in queue[0..7], so neighbouring tiles share an L2; a
        // workgroup whose range is used up takes items from the next ranges.
        // Every workgroup leaves once all 8 ranges are used up.
        __shared__ int s_item;
        const int xcd = blockIdx.x & 7;
        for (;;) {
            if (tid == 0) {
                int got = -1;
                if (g0.flatq) {
                    const int my = atomicAdd(&g0.queue[0], 1);
                    got = my < g0.total ? my : -1;
                } else {
                    for (int v = 0; v < 8 && got < 0; ++v) {
                        const int x = (xcd + v) & 7;
                        const int lo = (int) ((int64_t) g0.total * x / 8);
                        const int hi = (int) ((int64_t) g0.total * (x + 1) / 8);
                        if (lo >= hi) continue;
                        const int my = lo + atomicAdd(&g0.queue[x], 1);
                        if (my < hi) got = my;
                    }
                }
                s_item = got;
            }
            __syncthreads();
            // uniform values in scalar registers: the problem's fields then
            // come by scalar loads and its address arithmetic stays scalar
            const int item = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();
            if (item < 0) break;
            int lo = 0, hi = g0.nprob - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (g0.tile0[mid] <= item) lo = mid; else hi = mid - 1;
            }
            const OzArgs g = oz_uniform_crt_args(table[__builtin_amdgcn_readfirstlane(lo)]);
            const int t = item - g0.tile0[lo];
            const int tiles_m = g.Mp / kBMo;
            const int G = 4, per = G * (g.Np / kBNo), grp = t / per, first = grp * G;
            const int gsz = min(tiles_m - first, G), rr = t - grp * per;
            body(g, first + rr % gsz, rr / gsz, 0, 0, 0, 0);
        }
    } else {
        int tm, tn, d, kb0, nkb;
        int64_t part_off;
        if (!oz_item<SPLIT, kBMo, kBNo>(g0, tm, tn, d, kb0, nkb, part_off)) return;
        body(g0, tm, tn, d, kb0, nkb, part_off);
    }
}

// Split mode, step 2: per element, the INT32 partials added in FP64 in the
// order d = s-1 .. 0, chunks ascending, then scaled and written like the W
// mode's last launch.  upper: only j >= i; mirror: C(j, i) := C(i, j).
__global__ __launch_bounds__(256) void ozaki_fold_kernel(const int *__restrict__ P32, int64_t plane, int s,
                                                          int nchunks, int Mp, int M, int N,
                                                          const int *__restrict__ eA, const int *__restrict__ eB,
                                                          double alpha, double beta, double *C, int64_t ldc,
                                                          int upper, int mirror)
{
    const int i = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (i >= M || j >= N || (upper && i > j)) return;
    const int *p = P32 + (int64_t) j * Mp + i;
    double acc = 0.0;
    for (int d = s - 1; d >= 0; --d) {
        const double w = ldexp(1.0, -8 * (d + 2));
        for (int c = 0; c < nchunks; ++c) acc = ca3::add_rn(acc, (double) p[(int64_t) (d * nchunks + c) * plane] * w);
    }
    double v = oz_unscale(acc, eA[i], eB[j]);
    double *cp = C + (int64_t) j * ldc + i;
    v = ca3::mul_rn(alpha, v);
    if (beta != 0.0) v = ca3::add_rn(v, ca3::mul_rn(beta, *cp));
    *cp = v;
    if (mirror && i != j) C[(int64_t) i * ldc + j] = v;
}

// Library-owned, grow-only workspace for the planes and exponents.
struct Workspace {
    void *p = nullptr;
    size_t bytes = 0;
    int dev = -1;
};
Workspace g_ws[16];
std::mutex g_ws_mu;

void *workspace(size_t bytes, int *rc)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    Workspace &w = g_ws[dev & 15];
    std::lock_guard<std::mutex> lk(g_ws_mu);
    if (w.bytes < bytes) {
        if (w.p) {
            (void) hipStreamSynchronize(ca3::stream());
            (void) hipFree(w.p);
            w.p = nullptr;
            w.bytes = 0;
        }
        hipError_t e = hipMalloc(&w.p, bytes);
        if (e != hipSuccess) { *rc = ca3::fail(e, "ozaki workspace"); return nullptr; }
        w.bytes = bytes;
    }
    *rc = 0;
    return w.p;
}

// L = max log2(||a||_2 ||b||_2 / a.b) and L_s = max log2((max(a) ||b||_1 +
// max(b) ||a||_1) / a.b) over S samples of the seven spread statistics
// (sum ab, sum a^2, sum b^2, sum a, sum b, max a, max b); samples with an
// empty line are skipped, and both are 0 when every sample is.
void spread_exponents(const double *h, int S, double *L, double *Ls)
{
    double l = -INFINITY, ls = -INFINITY;
    for (int q = 0; q < S; ++q) {
        const double *v = h + (size_t) q * 7;
        if (!(v[1] > 0.0 && v[2] > 0.0)) continue;         // an empty line: nothing to lose
        const double lg = std::log2(v[0]);                  // -inf for disjoint supports
        l = std::max(l, 0.5 * std::log2(v[1]) + 0.5 * std::log2(v[2]) - lg);
        ls = std::max(ls, std::log2(v[5] * v[4] + v[6] * v[3]) - lg);
    }
    if (l == -INFINITY) l = ls = 0.0;
    *L = l;
    *Ls = ls;
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Sizes and mode of one product.  gram: op(B) = op(A)^T shares A's planes.
struct OzPlan {
    int Mp, Np, Kp, kchunk, nchunks, ntiles;
    int bmo, bno;             // workgroup tile
    int loads;                // 0 pipelined, 1 direct-to-LDS ring, 2 register-staged
    int waves;                // pipelined form: waves per workgroup (4 or 8)
    bool split, upper;
    size_t planesA, planesB, exps, maxes, acc, total;
};

OzPlan oz_plan(int m, int n, int k, int s, bool gram)
{
    OzPlan p;
    // CA3DMM_OZAKI_TILE=256x128 selects the smaller workgroup tile;
    // CA3DMM_OZAKI_LOADS selects the main loop: pipe (default: direct-to-LDS
    // ring with fragment prefetch), dtl (direct-to-LDS ring) or reg
    // (register-staged loads)
    const char *lcfg = getenv("CA3DMM_OZAKI_LOADS");
    p.loads = 0;
    if (lcfg && strcmp(lcfg, "dtl") == 0) p.loads = 1;
    if (lcfg && strcmp(lcfg, "reg") == 0) p.loads = 2;
    // CA3DMM_OZAKI_WAVES=4 runs the pipelined form with 4 waves of 128 x 128
    // instead of 8 waves of 128 x 64 per 256 x 256 tile
    const char *wcfg = getenv("CA3DMM_OZAKI_WAVES");
    p.waves = (wcfg && strcmp(wcfg, "4") == 0) ? 4 : 8;
    const char *tcfg = getenv("CA3DMM_OZAKI_TILE");
    p.bmo = 256;
    p.bno = (tcfg && strcmp(tcfg, "256x128") == 0) ? 128 : 256;
    p.Mp = (int) round_up(m, kTile);
    p.Np = gram ? p.Mp : (int) round_up(n, kTile);
    p.Kp = (int) round_up(std::max(k, 1), kKp);
    // INT32 exactness of one diagonal: s * kchunk * 128 * 128 < 2^31
    p.kchunk = std::min(p.Kp, (int) (((1u << 31) - 1) / (16384u * (unsigned) s)) / kKp * kKp);
    p.nchunks = (p.Kp + p.kchunk - 1) / p.kchunk;
    p.upper = gram;
    const int tiles_m = p.Mp / p.bmo, tiles_n = p.Np / p.bno;
    p.ntiles = tiles_m * tiles_n;
    if (gram) {
        p.ntiles = 0;
        for (int tm = 0; tm < tiles_m; ++tm) p.ntiles += std::max(0, tiles_n - tm * (p.bmo / p.bno));
    }
    // split-K when the output tiles alone cannot fill the chip;
    // CA3DMM_OZAKI_SPLIT=0/1 forces the mode of a non-Gram product
    p.split = gram || (p.nchunks > 1 && p.ntiles < ca3::cu_count());
    if (!gram) {
        const char *f = getenv("CA3DMM_OZAKI_SPLIT");
        if (f && (f[0] == '0' || f[0] == '1')) p.split = (f[0] == '1');
    }
    p.planesA = (size_t) s * p.Mp * p.Kp;
    p.planesB = gram ? 0 : (size_t) s * p.Np * p.Kp;
    p.exps = sizeof(int) * (size_t) (p.Mp + p.Np);
    p.maxes = sizeof(unsigned long long) * (size_t) (p.Mp + p.Np);
    p.acc = p.split ? sizeof(int) * (size_t) s * p.nchunks * p.Mp * p.Np : sizeof(double) * (size_t) p.Mp * p.Np;
    p.total = p.planesA + p.planesB + p.exps + p.maxes + 256 + p.acc;
    return p;
}

// C := alpha op(A) op(B) + beta C; gram: op(A) = A^T (A is k x m), op(B) = A,
// only j >= i computed (mirror copies it to j < i).
int ozaki_core(int ta, int tb, int m, int n, int k, double alpha, const double *A, int ldA, const double *B,
               int ldB, double beta, double *C, int ldC, int s, bool gram, bool mirror)
{
    const OzPlan pl = oz_plan(m, n, k, s, gram);
    const int Mp = pl.Mp, Np = pl.Np, Kp = pl.Kp;
    int rc = 0;
    char *ws = (char *) workspace(pl.total, &rc);
    if (rc) return rc;
    int8_t *PA = (int8_t *) ws;
    int8_t *PB = gram ? PA : PA + pl.planesA;
    int *eA = (int *) (ws + pl.planesA + pl.planesB);
    int *eB = gram ? eA : eA + Mp;
    unsigned long long *mxA = (unsigned long long *) (ws + pl.planesA + pl.planesB + pl.exps);
    unsigned long long *mxB = gram ? mxA : mxA + Mp;
    void *accp = (void *) (((uintptr_t) (ws + pl.planesA + pl.planesB + pl.exps + pl.maxes) + 255) &
                           ~(uintptr_t) 255);
    hipStream_t st = ca3::stream();
    const int nl = gram ? Mp : Mp + Np;       // lines with a scale (padding lines stay empty)
    fill_int_kernel<<<(2 * nl + 255) / 256, 256, 0, st>>>((int *) mxA, 2 * nl, 0);
    // op(A): line i, element (i, l): A[i + l*ldA] (no trans) or A[l + i*ldA]
    const int64_t a_sl = ta ? ldA : 1, a_sk = ta ? 1 : ldA;
    // op(B): line j, element (l, j): B[l + j*ldB] (no trans) or B[j + l*ldB]
    const int64_t b_sl = tb ? 1 : ldB, b_sk = tb ? ldB : 1;
    const int ky = std::min(64, std::max(1, k / 512));
    if (k > 0) {
        ozaki_max_kernel<<<dim3((m + 63) / 64, ky), 256, 0, st>>>(A, m, k, a_sl, a_sk, mxA);
        if (!gram) ozaki_max_kernel<<<dim3((n + 63) / 64, ky), 256, 0, st>>>(B, n, k, b_sl, b_sk, mxB);
    }
    ozaki_scale_kernel<<<(nl + 255) / 256, 256, 0, st>>>(mxA, nl, eA);
    ozaki_slice_kernel<<<dim3(Mp / 64, Kp / 64), 256, 0, st>>>(A, m, k, a_sl, a_sk, eA, s, PA,
                                                               (int64_t) Mp * Kp, Kp);
    if (!gram)
        ozaki_slice_kernel<<<dim3(Np / 64, Kp / 64), 256, 0, st>>>(B, n, k, b_sl, b_sk, eB, s, PB,
                                                                   (int64_t) Np * Kp, Kp);
    OzArgs g;
    memset(&g, 0, sizeof(g));
    g.M = m; g.N = n; g.Mp = Mp; g.Np = Np; g.Kp = Kp;
    g.alpha = alpha; g.beta = beta;
    g.PA = PA; g.PB = PB;
    g.planeA = (int64_t) Mp * Kp; g.planeB = (int64_t) Np * Kp;
    g.eA = eA; g.eB = eB;
    g.C = C; g.ldc = ldC;
    if (pl.split) {
        g.s = s;
        g.nchunks = pl.nchunks;
        g.kcb = pl.kchunk / kKb;
        g.ntiles = pl.ntiles;
        g.upper = pl.upper ? 1 : 0;
        g.ngroups = s * pl.nchunks;
        g.P32 = (int *) accp;
        g.p32plane = (int64_t) Mp * Np;
        const int64_t slots = (int64_t) ((g.ngroups + 7) / 8) * g.ntiles;
        const unsigned nwg = (unsigned) (slots * 8);
        if (pl.loads == 0) {
            if (pl.bno == 256 && pl.waves == 8) ozaki_pipe_kernel<true, 128, 64, false, 8><<<nwg, 512, 0, st>>>(g);
            else if (pl.bno == 256) ozaki_pipe_kernel<true, 128, 128, false, 4><<<nwg, 256, 0, st>>>(g);
            else ozaki_pipe_kernel<true, 128, 64, false, 4><<<nwg, 256, 0, st>>>(g);
        } else if (pl.loads == 1) {
            if (pl.bno == 256) ozaki_dtl_kernel<true, 128, 128, 4><<<nwg, 256, 0, st>>>(g);
            else ozaki_dtl_kernel<true, 128, 64, 4><<<nwg, 256, 0, st>>>(g);
        } else {
            if (pl.bno == 256) ozaki_diag_kernel<true, 128, 128, 2><<<nwg, 256, 0, st>>>(g);
            else ozaki_diag_kernel<true, 128, 64, 1><<<nwg, 256, 0, st>>>(g);
        }
        ozaki_fold_kernel<<<dim3((m + 255) / 256, n), 256, 0, st>>>(g.P32, g.p32plane, s, pl.nchunks, Mp, m, n,
                                                                     eA, eB, alpha, beta, C, ldC,
                                                                     pl.upper ? 1 : 0, mirror ? 1 : 0);
        return ca3::finish();
    }
    g.W = (double *) accp;
    const int grid = (Mp / pl.bmo) * (Np / pl.bno);
    if (pl.loads == 0) {      // all segments in one launch
        g.s = s;
        g.nchunks = pl.nchunks;
        g.kcb = pl.kchunk / kKb;
        if (pl.bno == 256 && pl.waves == 8) ozaki_pipe_kernel<false, 128, 64, true, 8><<<grid, 512, 0, st>>>(g);
        else if (pl.bno == 256) ozaki_pipe_kernel<false, 128, 128, true, 4><<<grid, 256, 0, st>>>(g);
        else ozaki_pipe_kernel<false, 128, 64, true, 4><<<grid, 256, 0, st>>>(g);
        ozaki_finalize_kernel<<<dim3((m + 255) / 256, std::min(n, 65535)), 256, 0, st>>>(g.W, Mp, m, n, eA, eB, alpha,
                                                                                       beta, C, ldC);
        return ca3::finish();
    }
    for (int d = s - 1; d >= 0; --d) {
        for (int c = 0; c < pl.nchunks; ++c) {
            g.d = d;
            g.kb0 = c * (pl.kchunk / kKb);
            g.nkb = std::min(pl.kchunk, Kp - c * pl.kchunk) / kKb;
            g.first = (d == s - 1 && c == 0);
            g.last = (d == 0 && c == pl.nchunks - 1);
            if (pl.loads == 1) {
                if (pl.bno == 256) ozaki_dtl_kernel<false, 128, 128, 4><<<grid, 256, 0, st>>>(g);
                else ozaki_dtl_kernel<false, 128, 64, 4><<<grid, 256, 0, st>>>(g);
            } else {
                if (pl.bno == 256) ozaki_diag_kernel<false, 128, 128, 2><<<grid, 256, 0, st>>>(g);
                else ozaki_diag_kernel<false, 128, 64, 1><<<grid, 256, 0, st>>>(g);
            }
        }
    }
    return ca3::finish();
}

// Sizes of one CRT-form product.
// Workgroups of ozaki2_residue_kernel<NP> resident at once on the current
// device (per-CU occupancy times CUs), queried once per device: the
// grid-stride launch uses exactly that many, so no partial last round.
template <int NP>
int residue_slots()
{
    static std::atomic<int> cache[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    int v = cache[dev & 15].load();
    if (v > 0) return v;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&ozaki2_residue_kernel<NP>),
                                                      256, 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    v = per_cu * cus;
    cache[dev & 15].store(v);
    return v;
}

struct Oz2Plan {
    int Mp, Np, Kp, nchunks;
    size_t planesA, planesB, exps, words, res, total;
};

Oz2Plan oz2_plan(int m, int n, int k, int nmod)
{
    Oz2Plan p;
    p.Mp = (int) round_up(m, kTile);
    p.Np = (int) round_up(n, kTile);
    p.Kp = (int) round_up(std::max(k, 1), kKp);
    p.nchunks = (p.Kp + kCrtChunk - 1) / kCrtChunk;
    p.planesA = (size_t) nmod * p.Mp * p.Kp;
    p.planesB = (size_t) nmod * p.Np * p.Kp;
    p.exps = sizeof(int) * (size_t) (p.Mp + p.Np);
    p.words = 2 * sizeof(unsigned long long) * (size_t) (p.Mp + p.Np);   // line maxima, then norm sums
    p.res = (size_t) nmod * p.Mp * p.Np;
    p.total = p.planesA + p.planesB + p.exps + p.words + 256 + p.res;
    return p;
}

template <int NM>
void launch_crt(dim3 grid, hipStream_t st, const int8_t *R, int64_t rplane, int Mp, int m, int n, const int *eA,
                const int *eB, double alpha, double beta, double *C, int64_t ldc)
{
    ozaki2_crt_kernel<NM><<<grid, 256, 0, st>>>(R, rplane, Mp, m, n, eA, eB, alpha, beta, C, ldc);
}

template <int... Ns>
void launch_crt_n(int nmod, std::integer_sequence<int, Ns...>, dim3 grid, hipStream_t st, const int8_t *R,
                  int64_t rplane, int Mp, int m, int n, const int *eA, const int *eB, double alpha, double beta,
                  double *C, int64_t ldc)
{
    ((nmod == Ns + kCrtMin ? launch_crt<Ns + kCrtMin>(grid, st, R, rplane, Mp, m, n, eA, eB, alpha, beta, C, ldc)
                           : void()), ...);
}

// C := alpha op(A) op(B) + beta C through the CRT form with nmod moduli.
int ozaki2_core(int ta, int tb, int m, int n, int k, double alpha, const double *A, int ldA, const double *B,
                int ldB, double beta, double *C, int ldC, int nmod)
{
    const Oz2Plan pl = oz2_plan(m, n, k, nmod);
    const int Mp = pl.Mp, Np = pl.Np, Kp = pl.Kp;
    int rc = 0;
    char *ws = (char *) workspace(pl.total, &rc);
    if (rc) return rc;
    int8_t *PA = (int8_t *) ws;
    int8_t *PB = PA + pl.planesA;
    int *eA = (int *) (ws + pl.planesA + pl.planesB);
    int *eB = eA + Mp;
    unsigned long long *mx = (unsigned long long *) (ws + pl.planesA + pl.planesB + pl.exps);
    unsigned long long *S = mx + (Mp + Np);
    int8_t *R = (int8_t *) (((uintptr_t) (ws + pl.planesA + pl.planesB + pl.exps + pl.words) + 255) &
                            ~(uintptr_t) 255);
    hipStream_t st = ca3::stream();
    const int nl = Mp + Np;
    fill_int_kernel<<<(4 * nl + 255) / 256, 256, 0, st>>>((int *) mx, 4 * nl, 0);
    const int64_t a_sl = ta ? ldA : 1, a_sk = ta ? 1 : ldA;
    const int64_t b_sl = tb ? 1 : ldB, b_sk = tb ? ldB : 1;
    const int ky = std::min(64, std::max(1, k / 512));
    if (k > 0) {
        ozaki_max_kernel<<<dim3((m + 63) / 64, ky), 256, 0, st>>>(A, m, k, a_sl, a_sk, mx);
        ozaki_max_kernel<<<dim3((n + 63) / 64, ky), 256, 0, st>>>(B, n, k, b_sl, b_sk, mx + Mp);
        ozaki2_norm_kernel<<<dim3((m + 63) / 64, ky), 256, 0, st>>>(A, m, k, a_sl, a_sk, mx, S);
        ozaki2_norm_kernel<<<dim3((n + 63) / 64, ky), 256, 0, st>>>(B, n, k, b_sl, b_sk, mx + Mp, S + Mp);
    }
    ozaki2_scale_kernel<<<(nl + 255) / 256, 256, 0, st>>>(mx, S, nl, kCrtT[nmod], eA);
    auto residues = [&](auto NP) {
        const int slots = residue_slots<decltype(NP)::value>();
        const int ta = (Mp / 64) * (Kp / 64), tb = (Np / 64) * (Kp / 64);
        ozaki2_residue_kernel<decltype(NP)::value><<<std::min(ta, slots), 256, 0, st>>>(
            A, m, k, a_sl, a_sk, eA, nmod, PA, (int64_t) Mp * Kp, Kp, Mp / 64, ta);
        ozaki2_residue_kernel<decltype(NP)::value><<<std::min(tb, slots), 256, 0, st>>>(
            B, n, k, b_sl, b_sk, eB, nmod, PB, (int64_t) Np * Kp, Kp, Np / 64, tb);
    };
    if (crt_pieces(nmod) == 2) residues(std::integral_constant<int, 2>{});
    else residues(std::integral_constant<int, 3>{});
    OzArgs g;
    memset(&g, 0, sizeof(g));
    g.M = m; g.N = n; g.Mp = Mp; g.Np = Np; g.Kp = Kp;
    g.alpha = alpha; g.beta = beta;
    g.PA = PA; g.PB = PB;
    g.planeA = (int64_t) Mp * Kp; g.planeB = (int64_t) Np * Kp;
    g.eA = eA; g.eB = eB;
    g.C = C; g.ldc = ldC;
    g.s = nmod;
    g.nchunks = pl.nchunks;
    g.kcb = kCrtChunk / kKb;
    g.R8 = R;
    g.rplane = (int64_t) Mp * Np;
    // CA3DMM_OZAKI_WAVES=4: 256 x 128 tiles of 4 waves instead of 256 x 256 of 8
    static const bool waves4 = [] {
        const char *w = getenv("CA3DMM_OZAKI_WAVES");
        return w && strcmp(w, "4") == 0;
    }();
    if (waves4) ozaki_pipe_kernel<false, 128, 64, true, 4, true><<<(Mp / 256) * (Np / 128), 256, 0, st>>>(g);
    else ozaki_pipe_kernel<false, 128, 64, true, 8, true><<<(Mp / 256) * (Np / 256), 512, 0, st>>>(g);
    const dim3 grid((unsigned) ((((int64_t) (m + 15) / 16) * ((n + 15) / 16) + 3) / 4));
    launch_crt_n(nmod, std::make_integer_sequence<int, kCrtMax - kCrtMin + 1>{}, grid, st, R, g.rplane, Mp, m, n,
                 eA, eB, alpha, beta, C, ldC);
    return ca3::finish();
}

}  // namespace

// ---------------------------------------------------------------------
// Grouped CRT form for the simulated grid (declared in grouped.h)
// ---------------------------------------------------------------------
namespace ca3 {
namespace grouped {

struct OzPlan {
    int nmod = 0, nprob = 0, nlines = 0;
    int total[2] = {0, 0};                   // queue items for 256 x 256 / 256 x 128 tiles
    int njobs = 0, rblocks = 0, sblocks = 0, ncrt = 0, crtblocks = 0;
    void *planes = nullptr;                  // operand residue planes and product residue planes
    unsigned long long *words = nullptr;     // [2 * nlines]: line maxima, then norm sums
    int *e = nullptr;                        // [nlines] line exponents
    int *queue = nullptr;
    PrepJob *d_jobs = nullptr;
    int *d_blk_r = nullptr, *d_blk_s = nullptr;
    OzArgs *d_args = nullptr;
    int *d_tile0[2] = {nullptr, nullptr};
    CrtJob *d_crt = nullptr;
    int *d_crt_blk = nullptr;
    std::vector<OzArgs> h_args;              // the per-problem launches
    std::vector<int> h_tiles[2];
    OzTuning tune;
    std::string signature;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_grouped = false;
    OzTuning graph_tune;
};

namespace {

void drop_oz_graph(OzPlan *p)
{
    if (p->exec) { (void) hipGraphExecDestroy(p->exec); p->exec = nullptr; }
    if (p->graph) { (void) hipGraphDestroy(p->graph); p->graph = nullptr; }
}

template <int... Ns>
void launch_crt_grouped_n(int nmod, std::integer_sequence<int, Ns...>, int blocks, hipStream_t st,
                          const CrtJob *jobs, const int *blk0, int njobs)
{
    ((nmod == Ns + kCrtMin ? (void) (ozaki2_crt_grouped_kernel<Ns + kCrtMin><<<blocks, 256, 0, st>>>(jobs, blk0,
                                                                                                   njobs))
                           : void()), ...);
}


int oz_sequence(OzPlan *p, const OzTuning &tn, bool grouped, hipStream_t st)
{
    const int v = (tn.tile_n == 128) ? 1 : 0;
    fill_int_kernel<<<std::max(1, std::min(4096, (4 * p->nlines + 255) / 256)), 256, 0, st>>>((int *) p->words,
                                                                                               4 * p->nlines, 0);
    fill_int_kernel<<<1, 64, 0, st>>>(p->queue, 8, 0);
    if (p->rblocks) {
        ozaki2_prep_grouped_kernel<0><<<p->rblocks, 256, 0, st>>>(p->d_jobs, p->d_blk_r, p->njobs, p->nmod);
        ozaki2_prep_grouped_kernel<1><<<p->rblocks, 256, 0, st>>>(p->d_jobs, p->d_blk_r, p->njobs, p->nmod);
    }
    ozaki2_scale_kernel<<<(p->nlines + 255) / 256, 256, 0, st>>>(p->words, p->words + p->nlines, p->nlines,
                                                                  kCrtT[p->nmod], p->e);
    if (p->sblocks) {
        if (crt_pieces(p->nmod) == 2)
            ozaki2_prep_grouped_kernel<2, 2><<<p->sblocks, 256, 0, st>>>(p->d_jobs, p->d_blk_s, p->njobs, p->nmod);
        else
            ozaki2_prep_grouped_kernel<2, 3><<<p->sblocks, 256, 0, st>>>(p->d_jobs, p->d_blk_s, p->njobs, p->nmod);
    }
    if (grouped) {
        OzArgs g0;
        memset(&g0, 0, sizeof(g0));
        g0.tile0 = p->d_tile0[v];
        g0.nprob = p->nprob;
        g0.total = p->total[v];
        g0.queue = p->queue;
        g0.flatq = tn.xcd_queue ? 0 : 1;
        // one workgroup per CU (128 KB of LDS for 256 x 256, 64 KB for
        // 256 x 128 at one per CU too: the queue keeps every CU busy); every
        // one leaves when the queue passes the last item
        const int grid = std::max(1, std::min(g0.total, ca3::cu_count()));
        if (g0.total) {
            if (v == 0) ozaki_pipe_kernel<false, 128, 64, true, 8, true, true><<<grid, 512, 0, st>>>(g0, p->d_args);
            else ozaki_pipe_kernel<false, 128, 64, true, 4, true, true><<<grid, 256, 0, st>>>(g0, p->d_args);
        }
    } else {
        for (int r = 0; r < p->nprob; ++r) {
            if (!p->h_tiles[v][r]) continue;
            if (v == 0)
                ozaki_pipe_kernel<false, 128, 64, true, 8, true, false><<<p->h_tiles[0][r], 512, 0, st>>>(p->h_args[r]);
            else
                ozaki_pipe_kernel<false, 128, 64, true, 4, true, false><<<p->h_tiles[1][r], 256, 0, st>>>(p->h_args[r]);
        }
    }
    if (p->crtblocks)
        launch_crt_grouped_n(p->nmod, std::make_integer_sequence<int, kCrtMax - kCrtMin + 1>{}, p->crtblocks, st,
                             p->d_crt, p->d_crt_blk, p->ncrt);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : ca3::fail(e, "grouped ozaki launch");
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &v, hipStream_t st)
{
    hipError_t e = hipMalloc((void **) dst, sizeof(T) * std::max<size_t>(1, v.size()));
    if (e == hipSuccess && !v.empty()) e = hipMemcpyAsync(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, st);
    return e;
}

}  // namespace

int build_ozaki(const std::vector<Problem> &probs, const std::vector<Fold> &folds, int nmod, OzPlan **out)
{
    *out = nullptr;
    if (nmod < kCrtMin || nmod > kCrtMax)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "grouped ozaki: nmoduli %d outside %d..%d", nmod, kCrtMin, kCrtMax);
    auto *p = new OzPlan;
    p->nmod = nmod;
    p->nprob = (int) probs.size();
    struct Lay {
        int Mp, Np, Kp, lineA, lineB;
        size_t pa, pb, r;
    };
    std::vector<Lay> lay(p->nprob);
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t at = off; off += (bytes + 255) / 256 * 256; return at; };
    std::vector<std::vector<int>> kofs(p->nprob);
    for (int i = 0; i < p->nprob; ++i) {
        const Problem &q = probs[i];
        if (q.nseg < 0 || q.nseg > kMaxSeg || q.M < 0 || q.N < 0) { delete p; return HIP_PROXY_ERR_ARG; }
        Lay &L = lay[i];
        L.Mp = (int) round_up(std::max(q.M, 1), kTile);
        L.Np = (int) round_up(std::max(q.N, 1), kTile);
        int64_t kk = 0;
        for (int s2 = 0; s2 < q.nseg; ++s2) {
            kofs[i].push_back((int) kk);
            kk += round_up(q.seg[s2].k, 64);
        }
        if (kk > kCrtKMax) { delete p; return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "grouped ozaki: k above 2^23"); }
        L.Kp = (int) round_up(std::max<int64_t>(kk, 1), kKp);
        L.pa = take((size_t) nmod * L.Mp * L.Kp);
        L.pb = take((size_t) nmod * L.Np * L.Kp);
        L.r = take((size_t) nmod * L.Mp * L.Np);
        L.lineA = p->nlines;
        p->nlines += L.Mp;
        L.lineB = p->nlines;
        p->nlines += L.Np;
    }
    hipStream_t st = ca3::stream();
    hipError_t e = hipMalloc(&p->planes, std::max<size_t>(off, 256));
    if (e == hipSuccess) e = hipMalloc((void **) &p->words, sizeof(unsigned long long) * 2 * std::max(1, p->nlines));
    if (e == hipSuccess) e = hipMalloc((void **) &p->e, sizeof(int) * std::max(1, p->nlines));
    if (e == hipSuccess) e = hipMalloc((void **) &p->queue, 8 * sizeof(int));
    if (e != hipSuccess) { destroy_ozaki(p); return ca3::fail(e, "grouped ozaki workspace"); }
    int8_t *base = (int8_t *) p->planes;
    unsigned long long *mx = p->words, *S = p->words + p->nlines;

    std::vector<PrepJob> jobs;
    std::vector<int> blk_r(1, 0), blk_s(1, 0);
    std::vector<int> tile0[2] = {std::vector<int>(1, 0), std::vector<int>(1, 0)};
    std::vector<OzArgs> args;
    std::map<const double *, int> by_c;
    for (int i = 0; i < p->nprob; ++i) {
        const Problem &q = probs[i];
        const Lay &L = lay[i];
        by_c[q.C] = i;
        const int nseg = std::max(q.nseg, 1);              // no segment: one zero-filling job per operand
        for (int s2 = 0; s2 < nseg; ++s2) {
            const bool has = s2 < q.nseg;
            const int k = has ? q.seg[s2].k : 0;
            const int ko = has ? kofs[i][s2] : 0;
            const bool lastseg = (s2 == nseg - 1);
            const int kext = lastseg ? L.Kp - ko : (int) round_up(k, 64);
            for (int op = 0; op < 2; ++op) {
                PrepJob J;
                memset(&J, 0, sizeof(J));
                J.K = k;
                J.lines = op ? q.N : q.M;
                if (op == 0) { J.X = has ? q.seg[s2].A : nullptr; J.sl = 1; J.sk = has ? q.seg[s2].lda : 1; }
                else { J.X = has ? q.seg[s2].B : nullptr; J.sl = has ? q.seg[s2].ldb : 1; J.sk = 1; }
                const int line0 = op ? L.lineB : L.lineA, Lp = op ? L.Np : L.Mp;
                J.rbx = (J.lines + 63) / 64;
                J.rby = (k > 0 && J.lines > 0) ? std::min(64, std::max(1, k / 512)) : 0;
                J.sbx = Lp / 64;
                J.sby = kext / 64;
                J.mx = mx + line0;
                J.S = S + line0;
                J.ex = p->e + line0;
                J.P = base + (op ? L.pb : L.pa) + ko;
                J.plane = (int64_t) Lp * L.Kp;
                J.Kp = L.Kp;
                jobs.push_back(J);
                blk_r.push_back(blk_r.back() + J.rbx * J.rby);
                blk_s.push_back(blk_s.back() + J.sbx * J.sby);
            }
        }
        OzArgs g;
        memset(&g, 0, sizeof(g));
        g.M = q.M; g.N = q.N; g.Mp = L.Mp; g.Np = L.Np; g.Kp = L.Kp;
        g.PA = base + L.pa; g.PB = base + L.pb;
        g.planeA = (int64_t) L.Mp * L.Kp; g.planeB = (int64_t) L.Np * L.Kp;
        g.eA = p->e + L.lineA; g.eB = p->e + L.lineB;
        g.s = nmod;
        g.nchunks = (L.Kp + kCrtChunk - 1) / kCrtChunk;
        g.kcb = kCrtChunk / kKb;
        g.R8 = base + L.r;
        g.rplane = (int64_t) L.Mp * L.Np;
        args.push_back(g);
        for (int v = 0; v < 2; ++v) {
            const int tiles = (q.M > 0 && q.N > 0) ? (L.Mp / 256) * (L.Np / (v ? 128 : 256)) : 0;
            p->h_tiles[v].push_back(tiles);
            tile0[v].push_back(tile0[v].back() + tiles);
        }
        p->signature += std::to_string(q.M) + "x" + std::to_string(q.N) + "x" + std::to_string(L.Kp) + ";";
    }
    p->signature += "m" + std::to_string(nmod);
    p->h_args = args;
    p->total[0] = tile0[0].back();
    p->total[1] = tile0[1].back();
    p->njobs = (int) jobs.size();
    p->rblocks = blk_r.back();
    p->sblocks = blk_s.back();
    blk_r.pop_back();
    blk_s.pop_back();
    tile0[0].pop_back();
    tile0[1].pop_back();

    // reconstruction jobs: fold groups, then problems outside any fold group
    std::vector<CrtJob> crt;
    std::vector<int> cblk(1, 0);
    std::vector<char> folded(p->nprob, 0);
    auto add_job = [&](CrtJob &J) {
        J.nbm = (J.M + 15) / 16;
        J.nwg = (int) (((int64_t) J.nbm * ((J.N + 15) / 16) + 3) / 4);
        crt.push_back(J);
        cblk.push_back(cblk.back() + J.nwg);
    };
    for (const Fold &f : folds) {
        if (f.kp < 1 || f.kp > kGrpMaxKp || f.kp > kMaxKp) { destroy_ozaki(p); return HIP_PROXY_ERR_ARG; }
        CrtJob J;
        memset(&J, 0, sizeof(J));
        J.kp = f.kp; J.M = f.M; J.N = f.N;
        for (int l = 0; l < f.kp; ++l) {
            auto it = by_c.find(f.Cl[l]);
            if (it == by_c.end() || probs[it->second].M != f.M || probs[it->second].N != f.N) {
                destroy_ozaki(p);
                return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "grouped ozaki: fold layer without its product");
            }
            const int i = it->second;
            folded[i] = 1;
            J.Mp = lay[i].Mp;
            J.rplane = (int64_t) lay[i].Mp * lay[i].Np;
            J.R[l] = base + lay[i].r;
            J.eA[l] = p->e + lay[i].lineA;
            J.eB[l] = p->e + lay[i].lineB;
            J.Cout[l] = f.Cout[l];
            J.cs[l] = f.cs[l];
        }
        J.cs[f.kp] = f.cs[f.kp];
        J.ldo = f.ldo;
        if (J.M > 0 && J.N > 0) add_job(J);
    }
    for (int i = 0; i < p->nprob; ++i) {
        const Problem &q = probs[i];
        if (folded[i] || q.M == 0 || q.N == 0) continue;
        CrtJob J;
        memset(&J, 0, sizeof(J));
        J.kp = 1; J.M = q.M; J.N = q.N; J.Mp = lay[i].Mp;
        J.rplane = (int64_t) lay[i].Mp * lay[i].Np;
        J.R[0] = base + lay[i].r;
        J.eA[0] = p->e + lay[i].lineA;
        J.eB[0] = p->e + lay[i].lineB;
        J.Cout[0] = q.C;
        J.cs[0] = 0;
        J.cs[1] = q.N;
        J.ldo = q.ldc;
        add_job(J);
    }
    p->ncrt = (int) crt.size();
    p->crtblocks = cblk.back();
    cblk.pop_back();

    e = upload(&p->d_jobs, jobs, st);
    if (e == hipSuccess) e = upload(&p->d_blk_r, blk_r, st);
    if (e == hipSuccess) e = upload(&p->d_blk_s, blk_s, st);
    if (e == hipSuccess) e = upload(&p->d_args, args, st);
    if (e == hipSuccess) e = upload(&p->d_tile0[0], tile0[0], st);
    if (e == hipSuccess) e = upload(&p->d_tile0[1], tile0[1], st);
    if (e == hipSuccess) e = upload(&p->d_crt, crt, st);
    if (e == hipSuccess) e = upload(&p->d_crt_blk, cblk, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { destroy_ozaki(p); return ca3::fail(e, "grouped ozaki plan"); }
    {
        OzTuning t;
        if (lookup_tuning_ozaki(p->signature, &t)) p->tune = t;
    }
    *out = p;
    return 0;
}

int run_ozaki(OzPlan *p, bool grouped, bool use_graph)
{
    hipStream_t st = ca3::stream();
    if (!use_graph) return oz_sequence(p, p->tune, grouped, st);
    if (p->exec && (p->graph_grouped != grouped || p->graph_tune.tile_n != p->tune.tile_n ||
                    p->graph_tune.xcd_queue != p->tune.xcd_queue))
        drop_oz_graph(p);
    if (!p->exec) {
        CA3_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int rc = oz_sequence(p, p->tune, grouped, st);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (rc) { if (g) (void) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return ca3::fail(e, "hipStreamEndCapture");
        p->graph = g;
        CA3_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
        p->graph_grouped = grouped;
        p->graph_tune = p->tune;
    }
    CA3_CHECK(hipGraphLaunch(p->exec, st));
    return 0;
}

int autotune_ozaki(OzPlan *p, int reps)
{
    static const int tiles[] = {256, 128};
    static const int queues[] = {1, 0};
    hipStream_t st = ca3::stream();
    hipEvent_t e0, e1;
    CA3_CHECK(hipEventCreate(&e0));
    CA3_CHECK(hipEventCreate(&e1));
    float best = 1e30f;
    OzTuning bt = p->tune;
    int rc = 0;
    for (int tn : tiles) {
        for (int q : queues) {
            OzTuning t;
            t.tile_n = tn;
            t.xcd_queue = q;
            if ((rc = oz_sequence(p, t, true, st))) break;               // warm-up
            (void) hipEventRecord(e0, st);
            for (int r = 0; r < reps && !rc; ++r) rc = oz_sequence(p, t, true, st);
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
    store_tuning_ozaki(p->signature, bt);
    return 0;
}

const OzTuning &tuning_ozaki(const OzPlan *p) { return p->tune; }

void set_tuning_ozaki(OzPlan *p, const OzTuning &t) { p->tune = t; }

void destroy_ozaki(OzPlan *p)
{
    if (!p) return;
    drop_oz_graph(p);
    (void) hipFree(p->planes);
    (void) hipFree(p->words);
    (void) hipFree(p->e);
    (void) hipFree(p->queue);
    (void) hipFree(p->d_jobs);
    (void) hipFree(p->d_blk_r);
    (void) hipFree(p->d_blk_s);
    (void) hipFree(p->d_args);
    (void) hipFree(p->d_tile0[0]);
    (void) hipFree(p->d_tile0[1]);
    (void) hipFree(p->d_crt);
    (void) hipFree(p->d_crt_blk);
    delete p;
}

int choose_moduli(const std::vector<Problem> &probs, const std::vector<int> &samples, int k, int *nmod,
                  double *Lout)
{
    const int ns = (int) samples.size() / 3;
    for (int q = 0; q < ns; ++q) {
        const int pi = samples[3 * q], i = samples[3 * q + 1], j = samples[3 * q + 2];
        if (pi < 0 || pi >= (int) probs.size() || i < 0 || i >= probs[pi].M || j < 0 || j >= probs[pi].N ||
            probs[pi].nseg < 0 || probs[pi].nseg > kMaxSeg)
            return ca3::fail_msg(HIP_PROXY_ERR_ARG, "grouped ozaki: spread sample %d outside its problem", q);
    }
    double L = 0.0, Ls = 0.0;
    if (ns > 0) {
        std::vector<SpreadProb> tab(probs.size());
        for (size_t t = 0; t < probs.size(); ++t) {
            SpreadProb &d = tab[t];
            memset(&d, 0, sizeof(d));
            d.nseg = probs[t].nseg;
            for (int g = 0; g < d.nseg; ++g) {
                d.A[g] = probs[t].seg[g].A;
                d.B[g] = probs[t].seg[g].B;
                d.lda[g] = probs[t].seg[g].lda;
                d.ldb[g] = probs[t].seg[g].ldb;
                d.k[g] = probs[t].seg[g].k;
            }
        }
        const size_t tb = (tab.size() * sizeof(SpreadProb) + 255) / 256 * 256;
        const size_t sb = ((size_t) ns * 3 * sizeof(int) + 255) / 256 * 256;
        int rc = 0;
        char *w = (char *) workspace(tb + sb + (size_t) ns * 7 * sizeof(double), &rc);
        if (rc) return rc;
        hipStream_t st = ca3::stream();
        std::vector<double> h((size_t) ns * 7);
        hipError_t e = hipMemcpyAsync(w, tab.data(), tab.size() * sizeof(SpreadProb), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(w + tb, samples.data(), (size_t) ns * 3 * sizeof(int), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            ozaki_grid_spread_kernel<<<ns, 256, 0, st>>>((const SpreadProb *) w, (const int *) (w + tb),
                                                         (double *) (w + tb + sb));
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(h.data(), w + tb + sb, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return ca3::fail(e, "grouped ozaki: spread");
        spread_exponents(h.data(), ns, &L, &Ls);
    }
    if (Lout) *Lout = L;
    return hip_ozaki_counts(L, Ls, k, nullptr, nmod);
}

}  // namespace grouped
}  // namespace ca3

extern "C" {

size_t hip_ozaki2_workspace_bytes(int m, int n, int k, int nmoduli)
{
    if (m <= 0 || n <= 0 || k <= 0 || nmoduli < kCrtMin || nmoduli > kCrtMax || k > kCrtKMax) return 0;
    return oz2_plan(m, n, k, nmoduli).total;
}

int hip_ozaki2_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n,
                     const int k, const double alpha, const double *A, const int ldA, const double *B,
                     const int ldB, const double beta, double *C, const int ldC, int nmoduli)
{
    const int ta = (transA != HipNoTrans), tb = (transB != HipNoTrans);
    if (m < 0 || n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: negative size");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (nmoduli < kCrtMin || nmoduli > kCrtMax)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: nmoduli %d outside %d..%d", nmoduli, kCrtMin, kCrtMax);
    if (k > kCrtKMax) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: k %d above %d", k, kCrtKMax);
    if (!C || ldC < std::max(1, m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: bad C / ldC");
    if (k > 0) {
        if (!A || ldA < std::max(1, ta ? k : m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: bad A / ldA");
        if (!B || ldB < std::max(1, tb ? n : k)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki2: bad B / ldB");
    }
    return ozaki2_core(ta, tb, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, nmoduli);
}

int hip_ozaki_counts(double L, double Ls, int k, int *nslices, int *nmoduli)
{
    if (std::isnan(L)) L = INFINITY;
    if (std::isnan(Ls)) Ls = INFINITY;
    int N = kCrtMax;
    const double cm = 53.5 + (k > 1024 ? 0.5 * std::log2(k / 1024.0) : 0.0);
    for (int c = kCrtMin; c <= kCrtMax; ++c)
        if (kCrtT[c] >= L + cm) {
            N = c;
            break;
        }
    const double off = 54.0 + (k > 1024 ? 2.33 * std::log2(k / 1024.0) : 0.0);
    int sl = kMaxSlices;
    for (int c = 1; c <= kMaxSlices; ++c)
        if (8.0 * c >= Ls + off) {
            sl = c;
            break;
        }
    if (nslices) *nslices = sl;
    if (nmoduli) *nmoduli = N;
    return HIP_PROXY_OK;
}

int hip_ozaki_choose(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n, const int k,
                     const double *A, const int ldA, const double *B, const int ldB, int *nslices, int *nmoduli)
{
    const int ta = (transA != HipNoTrans), tb = (transB != HipNoTrans);
    if (m < 0 || n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki_choose: negative size");
    if (m == 0 || n == 0 || k == 0) return hip_ozaki_counts(0.0, 0.0, k, nslices, nmoduli);
    if (!A || ldA < std::max(1, ta ? k : m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki_choose: bad A / ldA");
    if (!B || ldB < std::max(1, tb ? n : k)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki_choose: bad B / ldB");
    int rc = 0;
    double *d = (double *) workspace((size_t) kSpreadSamples * 7 * sizeof(double), &rc);
    if (rc) return rc;
    const int64_t a_sl = ta ? ldA : 1, a_sk = ta ? 1 : ldA;
    const int64_t b_sl = tb ? 1 : ldB, b_sk = tb ? ldB : 1;
    hipStream_t st = ca3::stream();
    ozaki_spread_kernel<<<kSpreadSamples, 256, 0, st>>>(A, a_sl, a_sk, m, B, b_sl, b_sk, n, k, d);
    std::vector<double> h((size_t) kSpreadSamples * 7);
    hipError_t e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return ca3::fail(e, "ozaki_choose");
    double L = 0.0, Ls = 0.0;
    spread_exponents(h.data(), kSpreadSamples, &L, &Ls);
    return hip_ozaki_counts(L, Ls, k, nslices, nmoduli);
}

size_t hip_ozaki_workspace_bytes(int m, int n, int k, int nslices)
{
    if (m <= 0 || n <= 0 || k <= 0 || nslices < 1 || nslices > kMaxSlices) return 0;
    return oz_plan(m, n, k, nslices, false).total;
}

int hip_ozaki_dgemm(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n,
                    const int k, const double alpha, const double *A, const int ldA, const double *B,
                    const int ldB, const double beta, double *C, const int ldC, int nslices)
{
    const int ta = (transA != HipNoTrans), tb = (transB != HipNoTrans);
    if (m < 0 || n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki: negative size");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (nslices < 1 || nslices > kMaxSlices)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki: nslices %d outside 1..%d", nslices, kMaxSlices);
    if (!C || ldC < std::max(1, m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki: bad C / ldC");
    if (k > 0) {
        if (!A || ldA < std::max(1, ta ? k : m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki: bad A / ldA");
        if (!B || ldB < std::max(1, tb ? n : k)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki: bad B / ldB");
    }
    return ozaki_core(ta, tb, m, n, k, alpha, A, ldA, B, ldB, beta, C, ldC, nslices, false, false);
}

int hip_ozaki_gram(const int n, const int k, const double alpha, const double *A, const int ldA,
                   const double beta, double *C, const int ldC, int mirror, int nslices)
{
    if (n < 0 || k < 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki gram: negative size");
    if (n == 0) return HIP_PROXY_OK;
    if (nslices < 1 || nslices > kMaxSlices)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki gram: nslices %d outside 1..%d", nslices, kMaxSlices);
    if (!C || ldC < std::max(1, n)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki gram: bad C / ldC");
    if (k > 0 && (!A || ldA < std::max(1, k))) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "ozaki gram: bad A / ldA");
    return ozaki_core(1, 0, n, n, k, alpha, A, ldA, A, ldA, beta, C, ldC, nslices, true, mirror != 0);
}

}  // extern "C"
