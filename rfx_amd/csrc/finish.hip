// A5: finishprox on the GPU, bit-identical to rf::cpu_finishprox
// (src/rf_finishprox.cpp:12-74).
//
// The reference copies the accumulated FP64 matrix, sets the diagonal to
// nout, replaces each off-diagonal pair by its mean a = (P_ij + P_ji) / 2,
// multiplies row i by 1/nout_i (rows with nout_i = 0 are left alone), then
// averages the pair again and sets the diagonal to 1.  Per element that is
//     S_ij = 0.5 * (a * r_i + a * r_j),   r_k = 1.0 / nout_k  (or 1 if nout_k = 0)
// evaluated in this order with contraction off, so every double matches the
// CPU result bit for bit.  Counts are exact integers in double.
//
// FULL inputs need P_ij and P_ji: a 64 x 64 output tile stages the
// transposed input tile through LDS so both reads and the column-major
// writes stay coalesced.  UPPER inputs already hold P_ij + P_ji.

#include "rfx_common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kT = 64;

__device__ __forceinline__ double recip_or_one(int32_t c) { return c > 0 ? 1.0 / static_cast<double>(c) : 1.0; }

__device__ __forceinline__ double finish_value(int64_t pair_sum, double ri, double rj) {
  const double a = static_cast<double>(pair_sum) / 2.0;
  const double ai = a * ri;
  const double aj = a * rj;
  return 0.5 * (ai + aj);
}

// FULL (column-major) counts -> FULL or UPPER float64.
__global__ __launch_bounds__(256) void finish_full_kernel(const int32_t* __restrict__ P,
                                                          const int32_t* __restrict__ nout, int64_t n,
                                                          double* __restrict__ S, int32_t out_layout) {
  __shared__ int32_t tr[kT][kT + 1];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kT, j0 = static_cast<int64_t>(blockIdx.y) * kT;
  if (out_layout == RFX_LAYOUT_UPPER && i0 > j0 + kT - 1) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // tr[a][b] = P[j0 + a, i0 + b] (element (j, i) at j + i*n): contiguous in a.
  for (int b = ty; b < kT; b += 4) {
    const int64_t i = i0 + b, j = j0 + tx;
    tr[tx][b] = (i < n && j < n) ? P[j + i * n] : 0;
  }
  __syncthreads();
  for (int c = ty; c < kT; c += 4) {
    const int64_t j = j0 + c, i = i0 + tx;
    if (i >= n || j >= n) continue;
    double v;
    if (i == j) {
      v = 1.0;
    } else {
      const int64_t pij = P[i + j * n];
      const int64_t pji = tr[c][tx];
      v = finish_value(pij + pji, recip_or_one(nout[i]), recip_or_one(nout[j]));
    }
    if (out_layout == RFX_LAYOUT_FULL) {
      S[i + j * n] = v;
    } else if (j >= i) {
      S[rfx_upper_index(i, j, n)] = v;
    }
  }
}

// UPPER counts (P_ij + P_ji) -> UPPER float64, one row segment per block.
__global__ __launch_bounds__(256) void finish_upper_kernel(const int32_t* __restrict__ U,
                                                           const int32_t* __restrict__ nout, int64_t n,
                                                           double* __restrict__ S) {
  const int64_t i = blockIdx.x;
  const int64_t j = i + static_cast<int64_t>(blockIdx.y) * 256 + threadIdx.x;
  if (j >= n) return;
  const int64_t o = rfx_upper_index(i, j, n);
  S[o] = (i == j) ? 1.0 : finish_value(U[o], recip_or_one(nout[i]), recip_or_one(nout[j]));
}

// UPPER counts -> FULL float64 (column-major, symmetric).
__global__ __launch_bounds__(256) void finish_upper_to_full_kernel(const int32_t* __restrict__ U,
                                                                   const int32_t* __restrict__ nout, int64_t n,
                                                                   double* __restrict__ S) {
  const int64_t i = static_cast<int64_t>(blockIdx.y) * 256 + threadIdx.x;
  const int64_t j = blockIdx.x;
  if (i >= n) return;
  double v;
  if (i == j) {
    v = 1.0;
  } else {
    const int64_t a = i < j ? i : j, b = i < j ? j : i;
    v = finish_value(U[rfx_upper_index(a, b, n)], recip_or_one(nout[i]), recip_or_one(nout[j]));
  }
  S[i + j * n] = v;
}

}  // namespace

extern "C" int rfx_finishprox(hipStream_t st, const int32_t* counts, int32_t in_layout, const int32_t* nout, int64_t n,
                              double* proxsym, int32_t out_layout) {
  if (!counts || !nout || !proxsym || n <= 0 || n > RFX_MAX_SAMPLES) return RFX_ERR_INVALID;
  if ((in_layout != RFX_LAYOUT_FULL && in_layout != RFX_LAYOUT_UPPER) ||
      (out_layout != RFX_LAYOUT_FULL && out_layout != RFX_LAYOUT_UPPER))
    return RFX_ERR_INVALID;
  if (in_layout == RFX_LAYOUT_FULL) {
    dim3 grid(rfx_ceil_div(n, kT), rfx_ceil_div(n, kT));
    hipLaunchKernelGGL(finish_full_kernel, grid, dim3(256), 0, st, counts, nout, n, proxsym, out_layout);
  } else if (out_layout == RFX_LAYOUT_UPPER) {
    dim3 grid(static_cast<unsigned>(n), rfx_ceil_div(n, 256));
    hipLaunchKernelGGL(finish_upper_kernel, grid, dim3(256), 0, st, counts, nout, n, proxsym);
  } else {
    dim3 grid(static_cast<unsigned>(n), rfx_ceil_div(n, 256));
    hipLaunchKernelGGL(finish_upper_to_full_kernel, grid, dim3(256), 0, st, counts, nout, n, proxsym);
  }
  return rfx_launch_status();
}
