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

__device__ __forceinline__ double finish_value(double pair_sum, double ri, double rj) {
  const double a = pair_sum / 2.0;
  const double ai = a * ri;
  const double aj = a * rj;
  return 0.5 * (ai + aj);
}

// Pair sum P_ij + P_ji as the reference forms it in double (exact for
// int32 counts, whose sum is formed in 64-bit integers first).
__device__ __forceinline__ double pair_sum(int32_t a, int32_t b) {
  return static_cast<double>(static_cast<int64_t>(a) + static_cast<int64_t>(b));
}
__device__ __forceinline__ double pair_sum(double a, double b) { return a + b; }

// FULL (column-major) counts, int32 or float64 -> FULL or UPPER float64.
template <typename In>
__global__ __launch_bounds__(256) void finish_full_kernel(const In* __restrict__ P, const int32_t* __restrict__ nout,
                                                          int64_t n, double* __restrict__ S, int32_t out_layout) {
  __shared__ In tr[kT][kT + 1];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kT, j0 = static_cast<int64_t>(blockIdx.y) * kT;
  if (out_layout == RFX_LAYOUT_UPPER && i0 > j0 + kT - 1) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  // tr[a][b] = P[j0 + a, i0 + b] (element (j, i) at j + i*n): contiguous in a.
  for (int b = ty; b < kT; b += 4) {
    const int64_t i = i0 + b, j = j0 + tx;
    tr[tx][b] = (i < n && j < n) ? P[j + i * n] : In(0);
  }
  __syncthreads();
  for (int c = ty; c < kT; c += 4) {
    const int64_t j = j0 + c, i = i0 + tx;
    if (i >= n || j >= n) continue;
    double v;
    if (i == j) {
      v = 1.0;
    } else {
      v = finish_value(pair_sum(P[i + j * n], tr[c][tx]), recip_or_one(nout[i]), recip_or_one(nout[j]));
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
    v = finish_value(static_cast<double>(U[rfx_upper_index(a, b, n)]), recip_or_one(nout[i]), recip_or_one(nout[j]));
  }
  S[i + j * n] = v;
}

template <typename Out>
__global__ void convert_kernel(const int32_t* __restrict__ in, int64_t len, Out* __restrict__ out) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < len;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[e] = static_cast<Out>(in[e]);
}

// int32 -> float (exact below 2^24) -> half, round to nearest even
__global__ void convert_f16_kernel(const int32_t* __restrict__ in, int64_t len, _Float16* __restrict__ out) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < len;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x)
    out[e] = static_cast<_Float16>(static_cast<float>(in[e]));
}

}  // namespace

extern "C" int rfx_counts_convert(hipStream_t st, const int32_t* counts, int64_t len, int32_t format, void* out) {
  if (!counts || !out || len < 0) return RFX_ERR_INVALID;
  if (len == 0) return RFX_OK;
  int64_t blocks = (len + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  const dim3 g(static_cast<unsigned>(blocks)), b(256);
  switch (format) {
    case RFX_PREC_FP64:
      hipLaunchKernelGGL(convert_kernel<double>, g, b, 0, st, counts, len, static_cast<double*>(out));
      break;
    case RFX_PREC_FP32:
      hipLaunchKernelGGL(convert_kernel<float>, g, b, 0, st, counts, len, static_cast<float*>(out));
      break;
    case RFX_PREC_FP16:
      hipLaunchKernelGGL(convert_f16_kernel, g, b, 0, st, counts, len, static_cast<_Float16*>(out));
      break;
    default:
      return RFX_ERR_INVALID;
  }
  return rfx_launch_status();
}

extern "C" int rfx_finishprox(hipStream_t st, const int32_t* counts, int32_t in_layout, const int32_t* nout, int64_t n,
                              double* proxsym, int32_t out_layout) {
  if (!counts || !nout || !proxsym || n <= 0 || n > RFX_MAX_SAMPLES) return RFX_ERR_INVALID;
  if ((in_layout != RFX_LAYOUT_FULL && in_layout != RFX_LAYOUT_UPPER) ||
      (out_layout != RFX_LAYOUT_FULL && out_layout != RFX_LAYOUT_UPPER))
    return RFX_ERR_INVALID;
  if (in_layout == RFX_LAYOUT_FULL) {
    dim3 grid(rfx_ceil_div(n, kT), rfx_ceil_div(n, kT));
    hipLaunchKernelGGL(finish_full_kernel<int32_t>, grid, dim3(256), 0, st, counts, nout, n, proxsym, out_layout);
  } else if (out_layout == RFX_LAYOUT_UPPER) {
    dim3 grid(static_cast<unsigned>(n), rfx_ceil_div(n, 256));
    hipLaunchKernelGGL(finish_upper_kernel, grid, dim3(256), 0, st, counts, nout, n, proxsym);
  } else {
    dim3 grid(static_cast<unsigned>(n), rfx_ceil_div(n, 256));
    hipLaunchKernelGGL(finish_upper_to_full_kernel, grid, dim3(256), 0, st, counts, nout, n, proxsym);
  }
  return rfx_launch_status();
}

extern "C" int rfx_finishprox_f64(hipStream_t st, const double* prox, const int32_t* nout, int64_t n, double* proxsym,
                                  int32_t out_layout) {
  if (!prox || !nout || !proxsym || n <= 0 || n > RFX_MAX_SAMPLES) return RFX_ERR_INVALID;
  if (out_layout != RFX_LAYOUT_FULL && out_layout != RFX_LAYOUT_UPPER) return RFX_ERR_INVALID;
  dim3 grid(rfx_ceil_div(n, kT), rfx_ceil_div(n, kT));
  hipLaunchKernelGGL(finish_full_kernel<double>, grid, dim3(256), 0, st, prox, nout, n, proxsym, out_layout);
  return rfx_launch_status();
}
