// Internal plumbing shared by the HIP sources of libca3dmm_amd.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "hip_proxy.h"

namespace ca3 {

// Library-owned stream of the current device (created on first use, default
// priority, non-blocking with respect to the null stream).
hipStream_t stream();

// 1 when entry points synchronise the library stream before returning.
bool sync_mode();

// Record a HIP failure (text kept for hip_proxy_last_error) and map it to a
// status code.
int fail(hipError_t e, const char *what);

// Record an argument / support failure with a message.
int fail_msg(int code, const char *fmt, ...);

// Called at the end of every entry: check the launch, then synchronise the
// stream when sync mode is on.
int finish();

// Number of CUs of the current device (cached).
int cu_count();

// X := B R^{-1} (B, X m x n, R n x n upper triangular, non-unit) with X not
// overlapping B: R^{-1} and one FP64 MFMA DGEMM per 128-column block of X
// straight into X (the in-place hip_dtrsm needs a copy per block).
int trsm_right_upper_out(int m, int n, const double *R, int ldR, const double *B, int ldB, double *X, int ldX);

// FP64 multiply and add rounded separately.  hipcc contracts a * b + c into
// one fused multiply-add by default (also through the __dmul_rn / __dadd_rn
// wrappers, which are plain operators); the epilogues whose rounding order
// is part of the result use these instead.
__device__ __forceinline__ double mul_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a * b;
}

__device__ __forceinline__ double add_rn(double a, double b)
{
#pragma clang fp contract(off)
    return a + b;
}

}  // namespace ca3

#define CA3_CHECK(expr)                                   \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return ca3::fail(_e, #expr); \
    } while (0)

// Packed FP32 instructions are disabled for the whole library at build time
// (-Xclang -target-feature -Xclang -packed-fp32-ops); no kernel here uses
// them by hand either.
