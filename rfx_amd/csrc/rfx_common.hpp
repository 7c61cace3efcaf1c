// Shared helpers of the HIP translation units: status mapping and the
// packed upper-triangle index.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "rfx_prox.h"

#define RFX_HIP_TRY(expr)                     \
  do {                                        \
    hipError_t rfx_e_ = (expr);               \
    if (rfx_e_ != hipSuccess) return RFX_ERR_HIP; \
  } while (0)

// Launch status: a failed launch is reported, never ignored.
inline int rfx_launch_status() { return hipGetLastError() == hipSuccess ? RFX_OK : RFX_ERR_HIP; }

// Offset of row i in the packed row-major upper triangle with diagonal.
__host__ __device__ inline int64_t rfx_upper_row(int64_t i, int64_t n) { return i * n - (i * (i - 1)) / 2; }

// A count written to the output: added to what is there when accumulating,
// else stored nontemporally -- the kernels write each output word once and
// never read it back, so it need not displace the member runs that other
// rows still read from L2 (the long-row counting kernel at covtype-like
// n = 200,000, T = 1000: 62.4 -> 55.0 ms).
__device__ __forceinline__ void rfx_emit_count(int32_t* p, int32_t v, int32_t accumulate) {
  if (accumulate)
    *p += v;
  else
    __builtin_nontemporal_store(v, p);
}

__host__ __device__ inline int64_t rfx_upper_index(int64_t i, int64_t j, int64_t n) {
  return rfx_upper_row(i, n) + (j - i);
}

inline unsigned rfx_ceil_div(int64_t a, int64_t b) { return static_cast<unsigned>((a + b - 1) / b); }

// LDS of one workgroup on gfx950 (160 KiB).  Every launch with dynamic LDS
// asserts at compile time that its largest dynamic request plus the kernel's
// static LDS stays within it: a larger dispatch is rejected by the runtime
// and aborts the queue.
constexpr int64_t kLdsLimitBytes = 163840;
