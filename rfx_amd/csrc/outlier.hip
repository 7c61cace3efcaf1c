// Outlier scores from a finished proximity matrix, RandomForest::
// compute_outlier_scores (src/rf_random_forest.cpp:2489-2520, read through
// get_outlier_scores, include/rf_random_forest.hpp:210):
//     s_i = 1 / sum_j P[i n + j]^2   (0 when the sum is not positive)
//     score_i = s_i / median(s)       (median = sorted s at index n / 2; no
//                                      division when it is not positive)
// The reference accumulates in float, each step rounding the double
// fused multiply-add P^2 + (double)sum back to float (its GNU C++ -O3 build
// contracts the product into the sum), in j order; every row here runs that
// same chain, so the scores match bit for bit.
//
// One wavefront owns 64 rows.  Tiles of 64 rows x 64 columns are read row by
// row (512 contiguous bytes per load) into LDS, and each lane then runs its
// row's chain over the tile.  The median of the n floats is selected on the
// host (nth_element), then one scaling kernel divides.
#include <algorithm>
#include <vector>

#include "rfx_common.hpp"

namespace {

constexpr int kOt = 64;

__global__ __launch_bounds__(64) void outlier_rows_kernel(const double* __restrict__ prox, int64_t n,
                                                          float* __restrict__ raw) {
  __shared__ double tile[kOt][kOt + 1];
  const int lane = threadIdx.x;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kOt;
  const int64_t i = i0 + lane;
  float s = 0.0f;
  for (int64_t j0 = 0; j0 < n; j0 += kOt) {
    const int64_t jl = min<int64_t>(kOt, n - j0);
#pragma unroll 8
    for (int r = 0; r < kOt; ++r) {
      const int64_t row = i0 + r;
      tile[r][lane] = (row < n && lane < jl) ? prox[row * n + j0 + lane] : 0.0;
    }
    __syncthreads();
    if (i < n)
      for (int c = 0; c < jl; ++c) {
        const double v = tile[lane][c];
        s = static_cast<float>(__fma_rn(v, v, static_cast<double>(s)));
      }
    __syncthreads();
  }
  if (i < n) raw[i] = s > 0.0f ? 1.0f / s : 0.0f;
}

__global__ void outlier_scale_kernel(float* __restrict__ scores, int64_t n, float median) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) scores[i] = scores[i] / median;
}

}  // namespace

extern "C" int rfx_outlier_scores(hipStream_t st, const double* prox, int64_t n, float* scores, float* median_host) {
  if (!prox || !scores || n <= 0 || n > RFX_MAX_SAMPLES) return RFX_ERR_INVALID;
  hipLaunchKernelGGL(outlier_rows_kernel, dim3(static_cast<unsigned>(rfx_ceil_div(n, kOt))), dim3(kOt), 0, st, prox,
                     n, scores);
  int rc = rfx_launch_status();
  if (rc) return rc;
  std::vector<float> h(static_cast<size_t>(n));
  RFX_HIP_TRY(hipMemcpyAsync(h.data(), scores, sizeof(float) * h.size(), hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  std::nth_element(h.begin(), h.begin() + n / 2, h.end());
  const float median = h[static_cast<size_t>(n / 2)];
  if (median_host) *median_host = median;
  if (median > 0.0f) {
    hipLaunchKernelGGL(outlier_scale_kernel, dim3(static_cast<unsigned>(rfx_ceil_div(n, 256))), dim3(256), 0, st,
                       scores, n, median);
    rc = rfx_launch_status();
  }
  return rc;
}
