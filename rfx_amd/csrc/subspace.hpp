// Block subspace iteration with Rayleigh-Ritz for the top eigenpairs of a
// double-centred squared-distance operator B = -1/2 J E J, shared by the
// dense (count-matrix) and low-rank (factor) MDS paths.
//
// The caller supplies E V (the un-centred product) through an operator
// callback, plus the FP64 row means r of E and their mean g.  Then
//   B V = -1/2 (E V - r s^T - 1 w^T + g 1 s^T),  s = 1^T V,  w = r^T V
// (E symmetric).  Orthonormalisation is CholeskyQR2 and the k x k
// Rayleigh-Ritz problem is solved in FP64 on the host, whatever the storage
// precision of V (float or double).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <functional>
#include <vector>

#include "linalg_host.hpp"
#include "rfx_common.hpp"

namespace rfx {

template <typename Real>
__global__ void ss_init_kernel(Real* V, int64_t n, int k, uint64_t seed) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  uint64_t z = seed * 0x9e3779b97f4a7c15ull + static_cast<uint64_t>(e) + 0x632be59bd9b4e019ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  // uniform in (-1, 1): a dense random start block is all the iteration needs
  V[e] = static_cast<Real>((static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0);
}

// partial[b][a*kb + c] = sum over rows of block b of A[i,a] * B[i,c]
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void ss_gram_kernel(const TA* __restrict__ A, int ka, const TB* __restrict__ B,
                                                      int kb, int64_t n, int64_t rows_per_block,
                                                      double* __restrict__ partial) {
  extern __shared__ double red[];
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  const int pairs = ka * kb;
  for (int pc = threadIdx.x; pc < pairs; pc += blockDim.x) red[pc] = 0.0;
  __syncthreads();
  // each thread owns (a, c) pairs and walks the rows of the block
  for (int pc = threadIdx.x; pc < pairs; pc += blockDim.x) {
    const int a = pc / kb, c = pc % kb;
    double s = 0.0;
    for (int64_t i = r0; i < r1; ++i) s += static_cast<double>(A[i * ka + a]) * static_cast<double>(B[i * kb + c]);
    red[pc] = s;
  }
  __syncthreads();
  for (int pc = threadIdx.x; pc < pairs; pc += blockDim.x) partial[static_cast<int64_t>(blockIdx.x) * pairs + pc] = red[pc];
}

static __global__ void ss_reduce_partials_kernel(const double* __restrict__ partial, int nblocks, int pairs,
                                          double* __restrict__ out) {
  const int pc = blockIdx.x * blockDim.x + threadIdx.x;
  if (pc >= pairs) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += partial[static_cast<int64_t>(b) * pairs + pc];
  out[pc] = s;
}

// out[i, :] = V[i, :] * M   (M k x kout, FP64, row-major)
template <typename Real>
__global__ void ss_rotate_kernel(const Real* __restrict__ V, int k, const double* __restrict__ M, int kout,
                                 int64_t n, Real* __restrict__ out) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kout) return;
  const int64_t i = e / kout;
  const int c = static_cast<int>(e % kout);
  double s = 0.0;
  for (int a = 0; a < k; ++a) s += static_cast<double>(V[i * k + a]) * M[a * kout + c];
  out[e] = static_cast<Real>(s);
}

// W <- -1/2 (W - r s^T - 1 w^T + g 1 s^T) + shift V; sw = [s; w] (2 x k, FP64)
template <typename Real>
__global__ void ss_center_kernel(Real* __restrict__ W, const Real* __restrict__ V, const double* __restrict__ r,
                                 const double* __restrict__ sw, double g, double shift, int64_t n, int k) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int c = static_cast<int>(e % k);
  const double s = sw[c], w = sw[k + c];
  const double v = static_cast<double>(W[e]) - r[i] * s - w + g * s;
  W[e] = static_cast<Real>(-0.5 * v + shift * static_cast<double>(V[e]));
}

// [1, r_i] rows for the s and w column sums
static __global__ void ss_ones_r_kernel(const double* __restrict__ r, int64_t n, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = 1.0;
  out[2 * i + 1] = r[i];
}

// X[i, c] = Vtop[i, c] * sqrt(lambda_c) in FP64, row-major n x kout
template <typename Real>
__global__ void ss_coords_kernel(const Real* __restrict__ V, int k, const double* __restrict__ scale, int kout,
                                 int64_t n, double* __restrict__ X) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kout) return;
  const int64_t i = e / kout;
  const int c = static_cast<int>(e % kout);
  X[e] = static_cast<double>(V[i * k + c]) * scale[c];
}

struct SubspaceResult {
  std::vector<double> eigvals;  // Ritz values of B, descending
  int iterations = 0;
  double residual = 0.0;  // max relative Ritz-value change of the wanted ones at exit
};

// Device scratch for the solver (all allocated by the caller's workspace).
template <typename Real>
class SubspaceSolver {
 public:
  using Apply = std::function<int(const Real* V, Real* W, int k)>;

  SubspaceSolver(hipStream_t st, int64_t n, int k) : st_(st), n_(n), k_(k) {}

  static size_t workspace_bytes(int64_t n, int k) {
    const int64_t nb = gram_blocks(n);
    return al(sizeof(Real) * n * k) * 3 + al(8 * 2 * n) + al(8 * nb * k * k) + al(8 * k * k) * 2 + al(8 * 2 * k) +
           al(8 * k);
  }

  void carve(void* ws) {
    char* p = static_cast<char*>(ws);
    V_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    W_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    T_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    onesr_ = reinterpret_cast<double*>(p);
    p += al(8 * 2 * n_);
    partial_ = reinterpret_cast<double*>(p);
    p += al(8 * gram_blocks(n_) * k_ * k_);
    kk_ = reinterpret_cast<double*>(p);
    p += al(8 * k_ * k_);
    M_ = reinterpret_cast<double*>(p);
    p += al(8 * k_ * k_);
    sw_ = reinterpret_cast<double*>(p);
    p += al(8 * 2 * k_);
    scale_ = reinterpret_cast<double*>(p);
  }

  // Top `want` eigenpairs of B; coords (device, FP64 n x want) = V sqrt(lambda).
  // The iteration runs on B + shift I (a positive shift moves large negative
  // eigenvalues of a non-Euclidean B out of the way of the positive ones);
  // reported eigenvalues are those of B.
  int run(const Apply& apply, const double* r_dev, double g, int want, int max_iter, double tol, uint64_t seed,
          double* coords, SubspaceResult& res, double shift = 0.0) {
    shift_ = shift;
    const unsigned gb = rfx_ceil_div(n_ * k_, 256);
    hipLaunchKernelGGL(ss_init_kernel<Real>, dim3(gb), dim3(256), 0, st_, V_, n_, k_, seed);
    hipLaunchKernelGGL(ss_ones_r_kernel, dim3(rfx_ceil_div(n_, 256)), dim3(256), 0, st_, r_dev, n_, onesr_);
    int rc = orthonormalize(V_);
    if (rc) return rc;
    std::vector<double> prev(want, 0.0), H, w, Z;
    res.residual = 1.0;
    int it = 0;
    for (; it < max_iter; ++it) {
      rc = apply_b(apply, r_dev, g, V_, W_);
      if (rc) return rc;
      rc = gram(V_, k_, W_, k_, H);  // H = V^T B V
      if (rc) return rc;
      for (int a = 0; a < k_; ++a)
        for (int c = a + 1; c < k_; ++c) {
          const double m = 0.5 * (H[a * k_ + c] + H[c * k_ + a]);
          H[a * k_ + c] = H[c * k_ + a] = m;
        }
      jacobi_eigh(H, k_, w, Z);
      // V <- orth(W Z): the rotated block B V Z spans the next subspace
      rc = upload(M_, Z.data(), static_cast<size_t>(k_) * k_);
      if (rc) return rc;
      hipLaunchKernelGGL(ss_rotate_kernel<Real>, dim3(gb), dim3(256), 0, st_, W_, k_, M_, k_, n_, V_);
      rc = orthonormalize(V_);
      if (rc) return rc;
      double change = 0.0;
      for (int c = 0; c < want; ++c) change = std::max(change, std::fabs(w[c] - prev[c]) / std::max(std::fabs(w[c]), 1e-300));
      prev.assign(w.begin(), w.begin() + want);
      res.residual = change;
      if (it >= 2 && change < tol) {
        ++it;
        break;
      }
    }
    // final Rayleigh-Ritz on the converged block
    rc = apply_b(apply, r_dev, g, V_, W_);
    if (rc) return rc;
    rc = gram(V_, k_, W_, k_, H);
    if (rc) return rc;
    for (int a = 0; a < k_; ++a)
      for (int c = a + 1; c < k_; ++c) {
        const double m = 0.5 * (H[a * k_ + c] + H[c * k_ + a]);
        H[a * k_ + c] = H[c * k_ + a] = m;
      }
    jacobi_eigh(H, k_, w, Z);
    for (double& v : w) v -= shift_;
    res.eigvals = w;
    res.iterations = it;
    std::vector<double> Zw(static_cast<size_t>(k_) * want), sc(want);
    for (int a = 0; a < k_; ++a)
      for (int c = 0; c < want; ++c) Zw[a * want + c] = Z[a * k_ + c];
    for (int c = 0; c < want; ++c) sc[c] = w[c] > 0 ? std::sqrt(w[c]) : 0.0;
    rc = upload(M_, Zw.data(), static_cast<size_t>(k_) * want);
    if (rc) return rc;
    rc = upload(scale_, sc.data(), want);
    if (rc) return rc;
    hipLaunchKernelGGL(ss_rotate_kernel<Real>, dim3(rfx_ceil_div(n_ * want, 256)), dim3(256), 0, st_, V_, k_, M_,
                       want, n_, T_);
    hipLaunchKernelGGL(ss_coords_kernel<Real>, dim3(rfx_ceil_div(n_ * want, 256)), dim3(256), 0, st_, T_, want,
                       scale_, want, n_, coords);
    RFX_HIP_TRY(hipStreamSynchronize(st_));
    return rfx_launch_status();
  }

  static size_t al(size_t b) { return (b + 255) & ~size_t(255); }

  // Small host -> device upload; waits so the host buffer may be reused.
  int upload(double* dst, const double* src, size_t count) {
    RFX_HIP_TRY(hipMemcpyAsync(dst, src, 8 * count, hipMemcpyHostToDevice, st_));
    RFX_HIP_TRY(hipStreamSynchronize(st_));
    return RFX_OK;
  }

  static int64_t gram_blocks(int64_t n) {
    int64_t b = (n + 511) / 512;
    return b < 1 ? 1 : (b > 2048 ? 2048 : b);
  }

  template <typename TA, typename TB>
  int gram(const TA* A, int ka, const TB* B, int kb, std::vector<double>& out) {
    const int64_t nb = gram_blocks(n_);
    const int64_t rpb = (n_ + nb - 1) / nb;
    const int pairs = ka * kb;
    hipLaunchKernelGGL((ss_gram_kernel<TA, TB>), dim3(static_cast<unsigned>(nb)), dim3(256), 8 * pairs, st_, A, ka,
                       B, kb, n_, rpb, partial_);
    hipLaunchKernelGGL(ss_reduce_partials_kernel, dim3(rfx_ceil_div(pairs, 256)), dim3(256), 0, st_, partial_,
                       static_cast<int>(nb), pairs, kk_);
    out.resize(pairs);
    RFX_HIP_TRY(hipMemcpyAsync(out.data(), kk_, 8 * pairs, hipMemcpyDeviceToHost, st_));
    RFX_HIP_TRY(hipStreamSynchronize(st_));
    return rfx_launch_status();
  }

  Real* block_v() { return V_; }
  Real* block_w() { return W_; }
  Real* block_t() { return T_; }
  double* small_matrix() { return M_; }

  // CholeskyQR2 in FP64: twice V <- V R^{-1} with V^T V = R^T R.
  int orthonormalize(Real* V) {
    std::vector<double> G;
    for (int pass = 0; pass < 2; ++pass) {
      int rc = gram(V, k_, V, k_, G);
      if (rc) return rc;
      if (!cholesky_lower(G, k_)) {
        // rank-deficient block: shift the Gram diagonal and retry once
        double tr = 0;
        for (int a = 0; a < k_; ++a) tr += G[a * k_ + a];
        rc = gram(V, k_, V, k_, G);
        if (rc) return rc;
        for (int a = 0; a < k_; ++a) G[a * k_ + a] += 1e-12 * (tr > 0 ? tr : 1.0);
        if (!cholesky_lower(G, k_)) return RFX_ERR_INVALID;
      }
      std::vector<double> Rinv = inv_upper_from_lower(G, k_);
      rc = upload(M_, Rinv.data(), static_cast<size_t>(k_) * k_);
      if (rc) return rc;
      hipLaunchKernelGGL(ss_rotate_kernel<Real>, dim3(rfx_ceil_div(n_ * k_, 256)), dim3(256), 0, st_, V, k_, M_, k_,
                         n_, T_);
      RFX_HIP_TRY(hipMemcpyAsync(V, T_, sizeof(Real) * n_ * k_, hipMemcpyDeviceToDevice, st_));
    }
    return rfx_launch_status();
  }

 private:
  int apply_b(const Apply& apply, const double* r_dev, double g, const Real* V, Real* W) {
    int rc = apply(V, W, k_);
    if (rc) return rc;
    std::vector<double> sw;
    rc = gram(onesr_, 2, V, k_, sw);  // rows: s (ones^T V), w (r^T V)
    if (rc) return rc;
    rc = upload(sw_, sw.data(), 2 * static_cast<size_t>(k_));
    if (rc) return rc;
    hipLaunchKernelGGL(ss_center_kernel<Real>, dim3(rfx_ceil_div(n_ * k_, 256)), dim3(256), 0, st_, W, V, r_dev, sw_,
                       g, shift_, n_, k_);
    return rfx_launch_status();
  }

  hipStream_t st_;
  int64_t n_;
  int k_;
  double shift_ = 0.0;
  Real *V_ = nullptr, *W_ = nullptr, *T_ = nullptr;
  double *onesr_ = nullptr, *partial_ = nullptr, *kk_ = nullptr, *M_ = nullptr, *sw_ = nullptr, *scale_ = nullptr;
};

}  // namespace rfx
