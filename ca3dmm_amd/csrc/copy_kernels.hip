// HBM-bound data-movement kernels of the CA3DMM local path.
//
//   hip_copy_matrix_block / hip_copy_matrix_blocks
//       mat_redist pack & unpack (src/mat_redist.c:317-348, :389-416) and the
//       single-block copy of src/cuda_proxy.cu:130-149.  Every block of a
//       pack or unpack goes into ONE launch: block descriptors travel as
//       kernel arguments, each workgroup copies one tile of whole rows of one
//       block, rows are flattened so a short row does not idle lanes, and
//       16-byte words are used whenever every row start allows it.
//   hip_fold_sum / hip_op_sum
//       the reduce-scatter's element-wise sum (src/mpi_op_omp.c:6-23) applied
//       to all kp partials in one pass in a fixed order.
//   hip_transpose_cm / hip_dgeam
//       utils.c:181-196 and the cublasDgeam transpose idiom
//       (src/ca3dmm.c:1036-1059), LDS-tiled.
//   hip_dcopy
//       cuda_proxy.cu:176-185.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <vector>

#include "ca3dmm_ext.h"
#include "hip_common.h"

namespace {

constexpr int kCopyThreads = 256;
constexpr int kCopyUnroll = 4;
constexpr int kMaxDescs = 48;     // per launch; longer lists are split
constexpr int kTileWords = kCopyThreads * kCopyUnroll * 4;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

struct CopyDesc {
    const char *src;
    char *dst;
    int64_t lds, ldd;   // in words
    int32_t nrow, ncol; // ncol in words
    int32_t rows_per_tile;
    int32_t tile_start;
    int32_t chunks;     // > 1: rows longer than a tile, cut into this many tiles each
    double inv_ncol;    // 1 / ncol, for the row of a flattened index
};

struct CopyBatch {
    int32_t n;
    int32_t total_tiles;
    CopyDesc d[kMaxDescs];
};

// Row and column of flattened index e (< 2^31) in rows of ncol words,
// without an integer division: the product e * (1/ncol) is within 2^-21 of
// e / ncol for ncol < 2^20, so its truncation is the row or one less, and
// one correction step fixes it.
__device__ __forceinline__ void row_col(unsigned e, unsigned ncol, double inv, unsigned &r, unsigned &c)
{
    r = (unsigned) ((double) e * inv);
    c = e - r * ncol;
    if ((int) c < 0) { --r; c += ncol; }
    if (c >= ncol) { ++r; c -= ncol; }
}

// Streaming copies: the data is touched once, so the loads and stores are
// non-temporal (measured ~10 % more bandwidth than cached accesses here).
template <typename W>
__global__ __launch_bounds__(kCopyThreads) void copy_blocks_kernel(const CopyBatch b)
{
    const int t = blockIdx.x;
    int lo = 0, hi = b.n - 1;       // last descriptor whose tile_start <= t
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (b.d[mid].tile_start <= t) lo = mid; else hi = mid - 1;
    }
    const CopyDesc &d = b.d[lo];
    int r0, nr;
    int64_t c0 = 0;
    unsigned ncol = (unsigned) d.ncol;
    double inv = d.inv_ncol;
    if (d.chunks > 1) {             // one tile-long piece of one long row
        const int q = t - d.tile_start;
        r0 = q / d.chunks;
        c0 = (int64_t) (q - r0 * d.chunks) * kTileWords;
        ncol = (unsigned) min((int64_t) kTileWords, (int64_t) d.ncol - c0);
        nr = 1;
        inv = 1.0 / (double) ncol;
    } else {
        r0 = (t - d.tile_start) * d.rows_per_tile;
        nr = min(d.rows_per_tile, d.nrow - r0);
    }
    const unsigned total = (unsigned) nr * ncol;
    const W *__restrict__ s = reinterpret_cast<const W *>(d.src) + (int64_t) r0 * d.lds + c0;
    W *__restrict__ o = reinterpret_cast<W *>(d.dst) + (int64_t) r0 * d.ldd + c0;
    for (unsigned e0 = threadIdx.x; e0 < total; e0 += kCopyThreads * kCopyUnroll) {
        W v[kCopyUnroll];
        unsigned rr[kCopyUnroll], cc[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const unsigned e = e0 + u * kCopyThreads;
            row_col(e, ncol, inv, rr[u], cc[u]);
            if (e < total) v[u] = __builtin_nontemporal_load(s + (int64_t) rr[u] * d.lds + cc[u]);
        }
#pragma unroll
        for (int u = 0; u < kCopyUnroll; ++u) {
            const unsigned e = e0 + u * kCopyThreads;
            if (e < total) __builtin_nontemporal_store(v[u], o + (int64_t) rr[u] * d.ldd + cc[u]);
        }
    }
}

// Copy descriptors expressed in words of `w` bytes.
int launch_copy(const hip_blk_copy_t *descs, int nblk, size_t dt, size_t w)
{
    CopyBatch b;
    int i = 0;
    while (i < nblk) {
        memset(&b, 0, sizeof(b));
        int tiles = 0;
        for (; i < nblk && b.n < kMaxDescs; ++i) {
            const hip_blk_copy_t &s = descs[i];
            if (s.nrow <= 0 || s.ncol <= 0) continue;
            CopyDesc &d = b.d[b.n];
            d.src = (const char *) s.src;
            d.dst = (char *) s.dst;
            d.lds = (int64_t) (s.lds * dt / w);
            d.ldd = (int64_t) (s.ldd * dt / w);
            d.nrow = s.nrow;
            d.ncol = (int32_t) ((size_t) s.ncol * dt / w);
            d.inv_ncol = 1.0 / (double) d.ncol;
            d.tile_start = tiles;
            if (d.ncol > kTileWords) {
                // a row longer than a tile is cut into tile-long pieces, so a
                // single long row (the allgather's blocks) still spreads over
                // the chip
                d.chunks = (d.ncol + kTileWords - 1) / kTileWords;
                d.rows_per_tile = 1;
                tiles += d.nrow * d.chunks;
            } else {
                d.chunks = 1;
                d.rows_per_tile = std::max(1, kTileWords / d.ncol);
                tiles += (d.nrow + d.rows_per_tile - 1) / d.rows_per_tile;
            }
            ++b.n;
        }
        if (b.n == 0) continue;
        b.total_tiles = tiles;
        hipStream_t st = ca3::stream();
        if (w == 16) copy_blocks_kernel<u32x4><<<tiles, kCopyThreads, 0, st>>>(b);
        else if (w == 8) copy_blocks_kernel<u32x2><<<tiles, kCopyThreads, 0, st>>>(b);
        else copy_blocks_kernel<unsigned int><<<tiles, kCopyThreads, 0, st>>>(b);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return ca3::fail(e, "copy_blocks_kernel");
    }
    return HIP_PROXY_OK;
}

bool aligned_to(const hip_blk_copy_t &s, size_t dt, size_t w)
{
    return ((uintptr_t) s.src % w) == 0 && ((uintptr_t) s.dst % w) == 0 &&
           ((size_t) s.lds * dt) % w == 0 && ((size_t) s.ldd * dt) % w == 0 &&
           ((size_t) s.ncol * dt) % w == 0;
}

// ---------------------------------------------------------------------
// Fold sum: dst = ((P0 + P1) + P2) + ...   (column-major blocks)
// ---------------------------------------------------------------------
constexpr int kMaxFold = 16;
struct FoldArgs {
    int32_t nsrc;
    const double *P[kMaxFold];
    int64_t ldp[kMaxFold];
};

// Block (x, y): column y (and every gridDim.y-th after it), words
// [x * 256 * U, (x + 1) * 256 * U) of it, U independent words per thread, all
// sources loaded before the sums; V = 2 uses 16-byte words (nrow even and
// every start aligned).  Non-temporal: every byte is used once.
constexpr int kFoldU = 4;
template <int V>
__global__ __launch_bounds__(256) void fold_sum_kernel(FoldArgs f, int nrow, int ncol, double *dst,
                                                       int64_t ldd)
{
    typedef typename std::conditional<V == 2, f64x2, double>::type T;
    const int rowsv = nrow / V;
    const int r0 = blockIdx.x * 256 * kFoldU + threadIdx.x;
    for (int64_t c = blockIdx.y; c < ncol; c += gridDim.y) {
        T acc[kFoldU];
#pragma unroll
        for (int u = 0; u < kFoldU; ++u) {
            const int r = r0 + u * 256;
            if (r < rowsv) acc[u] = __builtin_nontemporal_load(reinterpret_cast<const T *>(f.P[0] + c * f.ldp[0]) + r);
        }
        for (int t = 1; t < f.nsrc; ++t) {
            const T *src = reinterpret_cast<const T *>(f.P[t] + c * f.ldp[t]);
            T x[kFoldU];
#pragma unroll
            for (int u = 0; u < kFoldU; ++u) {
                const int r = r0 + u * 256;
                if (r < rowsv) x[u] = __builtin_nontemporal_load(src + r);
            }
#pragma unroll
            for (int u = 0; u < kFoldU; ++u) {
                if constexpr (V == 2) {
                    acc[u].x = ca3::add_rn(acc[u].x, x[u].x);
                    acc[u].y = ca3::add_rn(acc[u].y, x[u].y);
                } else {
                    acc[u] = ca3::add_rn(acc[u], x[u]);
                }
            }
        }
        T *out = reinterpret_cast<T *>(dst + c * ldd);
#pragma unroll
        for (int u = 0; u < kFoldU; ++u) {
            const int r = r0 + u * 256;
            if (r < rowsv) __builtin_nontemporal_store(acc[u], out + r);
        }
    }
}

template <typename T>
__global__ void op_sum_kernel(int len, const T *__restrict__ in, T *__restrict__ inout)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < len; i += gridDim.x * blockDim.x)
        inout[i] = inout[i] + in[i];
}

// ---------------------------------------------------------------------
// C(m x n) := alpha*op(A) + beta*op(B), column-major, 64x64 LDS tiles.
// ---------------------------------------------------------------------
constexpr int kTT = 64;

__global__ __launch_bounds__(256) void geam_kernel(int m, int n, double alpha, const double *A,
                                                   int64_t lda, int ta, double beta,
                                                   const double *B, int64_t ldb, int tb,
                                                   double *Cm, int64_t ldc)
{
    __shared__ double ta_s[kTT][kTT + 1];
    __shared__ double tb_s[kTT][kTT + 1];
    const int i0 = blockIdx.x * kTT, j0 = blockIdx.y * kTT;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    // Transposed operands are read along their own columns (coalesced) into
    // LDS and consumed along rows.
    if (ta) {
        for (int jj = ty; jj < kTT; jj += 4) {   // jj indexes op(A) row i0+jj
            int gi = i0 + jj, gj = j0 + tx;
            ta_s[jj][tx] = (gi < m && gj < n) ? A[(int64_t) gi * lda + gj] : 0.0;
        }
    }
    if (beta != 0.0 && tb) {
        for (int jj = ty; jj < kTT; jj += 4) {
            int gi = i0 + jj, gj = j0 + tx;
            tb_s[jj][tx] = (gi < m && gj < n) ? B[(int64_t) gi * ldb + gj] : 0.0;
        }
    }
    __syncthreads();
    for (int jj = ty; jj < kTT; jj += 4) {
        int gi = i0 + tx, gj = j0 + jj;
        if (gi >= m || gj >= n) continue;
        double a = ta ? ta_s[tx][jj] : A[(int64_t) gj * lda + gi];
        double v = ca3::mul_rn(alpha, a);
        if (beta != 0.0) {
            double b = tb ? tb_s[tx][jj] : B[(int64_t) gj * ldb + gi];
            v = ca3::add_rn(v, ca3::mul_rn(beta, b));
        }
        Cm[(int64_t) gj * ldc + gi] = v;
    }
}

__global__ void dcopy_kernel(int n, const double *x, int incx, double *y, int incy)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        y[(int64_t) i * incy] = x[(int64_t) i * incx];
}

unsigned grid_for(int64_t work, int per_block)
{
    int64_t b = (work + per_block - 1) / per_block;
    int64_t cap = (int64_t) ca3::cu_count() * 32;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (unsigned) b;
}

}  // namespace

extern "C" {

int hip_copy_matrix_blocks(size_t dt_size, int nblk, const hip_blk_copy_t *descs)
{
    if (dt_size != 4 && dt_size != 8)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "copy: dt_size must be 4 or 8 (got %zu)", dt_size);
    if (nblk < 0 || (nblk > 0 && !descs)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "copy: bad list");
    // each block goes with the widest word its addresses and leading
    // dimensions allow (16 bytes, else the element size): one launch list per
    // word size
    std::vector<hip_blk_copy_t> wide, narrow;
    for (int i = 0; i < nblk; ++i) {
        const hip_blk_copy_t &s = descs[i];
        if (s.nrow < 0 || s.ncol < 0 || (s.nrow > 0 && s.ncol > 0 && (!s.src || !s.dst)))
            return ca3::fail_msg(HIP_PROXY_ERR_ARG, "copy: block %d invalid", i);
        if (s.nrow > 0 && s.ncol > 0 && (s.lds < s.ncol || s.ldd < s.ncol))
            return ca3::fail_msg(HIP_PROXY_ERR_ARG, "copy: block %d leading dim < ncol", i);
        if (s.nrow <= 0 || s.ncol <= 0) continue;
        if ((int64_t) s.ncol * (int64_t) dt_size / 4 > 0x7fffffffLL)
            return ca3::fail_msg(HIP_PROXY_ERR_ARG, "copy: block %d row too long", i);
        (aligned_to(s, dt_size, 16) ? wide : narrow).push_back(s);
    }
    int rc = wide.empty() ? 0 : launch_copy(wide.data(), (int) wide.size(), dt_size, 16);
    if (!rc && !narrow.empty()) rc = launch_copy(narrow.data(), (int) narrow.size(), dt_size, dt_size);
    if (rc) return rc;
    return ca3::finish();
}

int hip_copy_matrix_block(size_t dt_size, const int nrow, const int ncol, const void *src,
                          const int lds, void *dst, const int ldd)
{
    hip_blk_copy_t d;
    d.src = src; d.dst = dst; d.lds = lds; d.ldd = ldd; d.nrow = nrow; d.ncol = ncol;
    return hip_copy_matrix_blocks(dt_size, 1, &d);
}

int hip_fold_sum(int nsrc, const double *const *P, const int *ldp, int nrow, int ncol,
                 double *dst, int ldd)
{
    if (nsrc < 1 || !P || !ldp || !dst || nrow < 0 || ncol < 0 || ldd < nrow)
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "fold_sum: bad arguments (nsrc=%d)", nsrc);
    for (int t = 0; t < nsrc; ++t)
        if (!P[t] || ldp[t] < nrow) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "fold_sum: source %d", t);
    if (nrow == 0 || ncol == 0) return HIP_PROXY_OK;
    // More sources than one launch takes: the first kMaxFold into dst, then
    // dst and the next kMaxFold - 1, and so on.  Each launch adds left to
    // right onto the running value, so the order (and the bits) are those
    // of one left-to-right fold.
    for (int t0 = 0; t0 < nsrc;) {
        FoldArgs f;
        memset(&f, 0, sizeof(f));
        int n = 0;
        if (t0 > 0) {
            f.P[0] = dst;
            f.ldp[0] = ldd;
            n = 1;
        }
        const int take = std::min(nsrc - t0, kMaxFold - n);
        for (int t = 0; t < take; ++t, ++n) {
            f.P[n] = P[t0 + t];
            f.ldp[n] = ldp[t0 + t];
        }
        f.nsrc = n;
        t0 += take;
        bool vec = (nrow % 2 == 0) && ((uintptr_t) dst % 16 == 0) && (ldd % 2 == 0);
        for (int t = 0; t < n; ++t) vec = vec && ((uintptr_t) f.P[t] % 16 == 0) && (f.ldp[t] % 2 == 0);
        const int rowsv = nrow / (vec ? 2 : 1);
        const dim3 grid((rowsv + 256 * kFoldU - 1) / (256 * kFoldU), std::min(ncol, 65535));
        if (vec) fold_sum_kernel<2><<<grid, 256, 0, ca3::stream()>>>(f, nrow, ncol, dst, ldd);
        else fold_sum_kernel<1><<<grid, 256, 0, ca3::stream()>>>(f, nrow, ncol, dst, ldd);
        const int rc = ca3::finish();
        if (rc) return rc;
    }
    return HIP_PROXY_OK;
}

int hip_op_sum(int dtype, int len, const void *invec, void *inoutvec)
{
    if (len < 0 || (len > 0 && (!invec || !inoutvec)))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "op_sum: bad arguments");
    if (len == 0) return HIP_PROXY_OK;
    unsigned g = grid_for(len, 256 * 4);
    hipStream_t st = ca3::stream();
    if (dtype == 0) op_sum_kernel<double><<<g, 256, 0, st>>>(len, (const double *) invec, (double *) inoutvec);
    else if (dtype == 1) op_sum_kernel<float><<<g, 256, 0, st>>>(len, (const float *) invec, (float *) inoutvec);
    else if (dtype == 2) op_sum_kernel<int><<<g, 256, 0, st>>>(len, (const int *) invec, (int *) inoutvec);
    else return ca3::fail_msg(HIP_PROXY_ERR_ARG, "op_sum: dtype %d", dtype);
    return ca3::finish();
}

int hip_dgeam(hip_proxy_enum_t transA, hip_proxy_enum_t transB, const int m, const int n,
              const double alpha, const double *A, const int ldA, const double beta,
              const double *B, const int ldB, double *C, const int ldC)
{
    int ta = (transA != HipNoTrans), tb = (transB != HipNoTrans);
    if (m < 0 || n < 0 || !C || ldC < std::max(1, m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgeam: bad C");
    if (m == 0 || n == 0) return HIP_PROXY_OK;
    if (!A || ldA < std::max(1, ta ? n : m)) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgeam: bad A");
    if (beta != 0.0 && (!B || ldB < std::max(1, tb ? n : m)))
        return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dgeam: bad B");
    dim3 grid((m + kTT - 1) / kTT, (n + kTT - 1) / kTT);
    if (grid.y > 65535) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dgeam: n too large");
    geam_kernel<<<grid, 256, 0, ca3::stream()>>>(m, n, alpha, A, ldA, ta, beta, B, ldB, tb, C, ldC);
    return ca3::finish();
}

int hip_transpose_cm(int nrow, int ncol, const double *A, int ldA, double *AT, int ldAT)
{
    return hip_dgeam(HipTrans, HipNoTrans, nrow, ncol, 1.0, A, ldA, 0.0, nullptr, 1, AT, ldAT);
}

int hip_dcopy(const int n, const double *x, const int incx, double *y, const int incy)
{
    if (n <= 0) return HIP_PROXY_OK;
    if (!x || !y || incx == 0 || incy == 0) return ca3::fail_msg(HIP_PROXY_ERR_ARG, "dcopy: bad arguments");
    if (incx < 0 || incy < 0) return ca3::fail_msg(HIP_PROXY_ERR_UNSUPP, "dcopy: negative increments");
    if (incx == 1 && incy == 1) return hip_copy_matrix_block(8, 1, n, x, n, y, n);
    dcopy_kernel<<<grid_for(n, 1024), 256, 0, ca3::stream()>>>(n, x, incx, y, incy);
    return ca3::finish();
}

}  // extern "C"
