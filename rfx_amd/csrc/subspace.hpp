// Block subspace iteration with Rayleigh-Ritz for the top eigenpairs of a
// double-centred squared-distance operator B = -1/2 J E J, shared by the
// dense (count-matrix) and low-rank (factor) MDS paths.
//
// The caller supplies W = E V (the un-centred product) through an operator
// callback, plus the FP64 row means r of E and their mean g.  With s = 1^T V,
// w = r^T V and q = w - g s the centred product is
//   B V = Wc = -1/2 (W - r s^T - 1 q^T)          (E symmetric)
// and is never formed: one fused pass over (V, W) per iteration gathers in
// FP64 the Grams V^T V, V^T W, W^T W and the sums 1^T V, r^T V, 1^T W, r^T W,
// from which the host builds H = V^T Wc and Wc^T Wc exactly.  The k x k
// work is then FP64 on the host: generalised Rayleigh-Ritz H z = theta
// (V^T V) z (so a slightly non-orthonormal V costs nothing), the Ritz
// residual norms ||Wc z - theta V z||^2 = z^T (Wc^T Wc) z - theta^2 as the
// stopping test, and the next block V <- Wc Z R^{-1} (Cholesky of
// Z^T Wc^T Wc Z) written by one kernel.  The iteration runs on B + shift I.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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

constexpr int kGramRows = 32;    // rows per LDS tile
constexpr int kGramBlocks = 240; // row-range partials (about one per CU)
constexpr int kGramTile = 64;    // output columns of A and of B per workgroup

// partial[b][a*kb + c] = sum over the rows of row block b (blockIdx.x) of
// A[i,a] * B[i,c], for the 64 x 64 output tile blockIdx.y (A tiles outer):
// rows staged through LDS kGramRows at a time, each thread accumulating a
// 4 x 4 FP64 sub-tile whose columns are 16 apart (consecutive lanes read
// consecutive LDS words).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void ss_gram_kernel(const TA* __restrict__ A, int ka, const TB* __restrict__ B,
                                                      int kb, int64_t n, int64_t rows_per_block,
                                                      double* __restrict__ partial) {
  __shared__ double sa[kGramRows][kGramTile];
  __shared__ double sb[kGramRows][kGramTile];
  const int ntc = (kb + kGramTile - 1) / kGramTile;
  const int a0 = (blockIdx.y / ntc) * kGramTile, c0 = (blockIdx.y % ntc) * kGramTile;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  const int ta = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0;
  for (int64_t i0 = r0; i0 < r1; i0 += kGramRows) {
    const int rows = static_cast<int>(min<int64_t>(kGramRows, r1 - i0));
    __syncthreads();
    for (int e = threadIdx.x; e < kGramRows * kGramTile; e += 256) {
      const int q = e / kGramTile, c = e % kGramTile;
      const int64_t i = i0 + q;
      sa[q][c] = q < rows && a0 + c < ka ? static_cast<double>(A[i * ka + a0 + c]) : 0.0;
      sb[q][c] = q < rows && c0 + c < kb ? static_cast<double>(B[i * kb + c0 + c]) : 0.0;
    }
    __syncthreads();
    for (int q = 0; q < rows; ++q) {
      double va[4], vb[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        va[x] = sa[q][ta + 16 * x];
        vb[x] = sb[q][tc + 16 * x];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] += va[x] * vb[y];
    }
  }
  const int pairs = ka * kb;
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int a = a0 + ta + 16 * x, c = c0 + tc + 16 * y;
      if (a < ka && c < kb) partial[static_cast<int64_t>(blockIdx.x) * pairs + a * kb + c] = acc[x][y];
    }
}

// out[pc] = sum over blocks of partial[b][pc]: one wavefront per entry,
// lanes stride the blocks, fixed-order tree reduction (deterministic).
static __global__ void ss_reduce_partials_kernel(const double* __restrict__ partial, int nblocks, int pairs,
                                                 double* __restrict__ out) {
  const int pc = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pc >= pairs) return;
  double s = 0.0;
  for (int b = lane; b < nblocks; b += 64) s += partial[static_cast<int64_t>(b) * pairs + pc];
  for (int d = 32; d >= 1; d >>= 1) s += __shfl_down(s, d, 64);
  if (lane == 0) out[pc] = s;
}

// Fused per-iteration sums over rows of [1, r, V, W]^T [V, W]: entry
// (a, c), a < 2 + 2k (0 = ones, 1 = r, then V, then W), c < 2k (V then W).
constexpr int kFusedPerThread = 20;

template <typename Real>
__global__ __launch_bounds__(256) void ss_fused_gram_kernel(const Real* __restrict__ V, const Real* __restrict__ W,
                                                            const double* __restrict__ r, int k, int64_t n,
                                                            int64_t rows_per_block, double* __restrict__ partial) {
  extern __shared__ double tile[];  // [kGramRows][2 + 2k]
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  const int wdt = 2 + 2 * k, nc = 2 * k, pairs = wdt * nc;
  double acc[kFusedPerThread];
#pragma unroll
  for (int q = 0; q < kFusedPerThread; ++q) acc[q] = 0.0;
  for (int64_t i0 = r0; i0 < r1; i0 += kGramRows) {
    const int rows = static_cast<int>(min<int64_t>(kGramRows, r1 - i0));
    __syncthreads();
    for (int e = threadIdx.x; e < rows * wdt; e += 256) {
      const int q = e / wdt, c = e % wdt;
      const int64_t i = i0 + q;
      double v;
      if (c == 0)
        v = 1.0;
      else if (c == 1)
        v = r[i];
      else if (c < 2 + k)
        v = static_cast<double>(V[i * k + (c - 2)]);
      else
        v = static_cast<double>(W[i * k + (c - 2 - k)]);
      tile[q * wdt + c] = v;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kFusedPerThread; ++s) {
      const int pc = threadIdx.x + s * 256;
      if (pc < pairs) {
        const int a = pc / nc, c = 2 + pc % nc;
        double t = 0.0;
        for (int q = 0; q < rows; ++q) t += tile[q * wdt + a] * tile[q * wdt + c];
        acc[s] += t;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kFusedPerThread; ++s) {
    const int pc = threadIdx.x + s * 256;
    if (pc < pairs) partial[static_cast<int64_t>(blockIdx.x) * pairs + pc] = acc[s];
  }
}

// The same sums for a block of 16 columns on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): X = [1, r, V, W] (34 columns, padded to 48) and
// Y = [V, W] (32 columns) as 3 x 2 output tiles of 16 x 16, K = 4 rows per
// instruction; the four wavefronts of a block combine in LDS and write the
// block's partial in the layout of ss_fused_gram_kernel.
// Register layout (wave64): A[i][k] at lane i + 16k, B[k][j] at lane j + 16k,
// D[i][j] at lane j + 16 * (i % 4), element i / 4 (the f64 form; checked with
// a lane probe on gfx950).
typedef double ss_f64x4 __attribute__((ext_vector_type(4)));

template <typename Real>
__device__ __forceinline__ double ss_xcol(const Real* __restrict__ V, const Real* __restrict__ W,
                                          const double* __restrict__ r, int64_t i, int c) {
  if (c == 0) return 1.0;
  if (c == 1) return r[i];
  if (c < 18) return static_cast<double>(V[i * 16 + (c - 2)]);
  if (c < 34) return static_cast<double>(W[i * 16 + (c - 18)]);
  return 0.0;
}

template <typename Real>
__global__ __launch_bounds__(256) void ss_fused_gram16_mfma_kernel(const Real* __restrict__ V,
                                                                   const Real* __restrict__ W,
                                                                   const double* __restrict__ r, int64_t n,
                                                                   int64_t rows_per_block,
                                                                   double* __restrict__ partial) {
  __shared__ double red[4][6][256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int64_t b0 = static_cast<int64_t>(blockIdx.x) * rows_per_block;
  const int64_t b1 = min(n, b0 + rows_per_block);
  const int64_t per = (b1 - b0 + 3) / 4;
  const int64_t r0 = b0 + wave * per, r1 = min(b1, r0 + per);
  ss_f64x4 acc[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) acc[t] = ss_f64x4{0.0, 0.0, 0.0, 0.0};
  for (int64_t j0 = r0; j0 < r1; j0 += 4) {
    const int64_t i = j0 + kq;
    const bool ok = i < r1;
    double xa[3], yb[2];
#pragma unroll
    for (int tx = 0; tx < 3; ++tx) xa[tx] = ok ? ss_xcol(V, W, r, i, tx * 16 + l16) : 0.0;
    yb[0] = ok ? static_cast<double>(V[i * 16 + l16]) : 0.0;
    yb[1] = ok ? static_cast<double>(W[i * 16 + l16]) : 0.0;
#pragma unroll
    for (int tx = 0; tx < 3; ++tx)
#pragma unroll
      for (int ty = 0; ty < 2; ++ty)
        acc[tx * 2 + ty] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[tx], yb[ty], acc[tx * 2 + ty], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 6; ++t)
#pragma unroll
    for (int e = 0; e < 4; ++e) red[wave][t][lane * 4 + e] = acc[t][e];
  __syncthreads();
  constexpr int nc = 32, pairs = 34 * 32;
  for (int q = tid; q < 6 * 256; q += 256) {
    const int t = q / 256, le = q % 256, ln = le / 4, e = le % 4;
    const int a = (t / 2) * 16 + (ln >> 4) + 4 * e, c = (t % 2) * 16 + (ln & 15);
    if (a < 34) {
      const double v = red[0][t][le] + red[1][t][le] + red[2][t][le] + red[3][t][le];
      partial[static_cast<int64_t>(blockIdx.x) * pairs + a * nc + c] = v;
    }
  }
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

// The same product for wide blocks (k and kout of 32 or more, kout up to
// 160): a 64-row tile of V and a 16-row slab of M staged in LDS, 4 rows x 10
// columns per thread.  Each output is the same FP64 fma chain over a = 0 ..
// k-1 as ss_rotate_kernel's, so the results are identical; the simple
// kernel re-reads M (166 KB at k = kout = 144, more than L1 holds) from L2
// for every multiply-add and takes 13 ms at n = 1,000,000.
constexpr int kRotRows = 64;
constexpr int kRotK = 16;
constexpr int kRotJ = 10;
template <typename Real>
__global__ __launch_bounds__(256) void ss_rotate_tiled_kernel(const Real* __restrict__ V, int k,
                                                              const double* __restrict__ M, int kout, int64_t n,
                                                              Real* __restrict__ out) {
  __shared__ double vs[kRotK][kRotRows + 1];
  __shared__ double ms[kRotK][16 * kRotJ];
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kRotRows;
  double acc[4][kRotJ];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int j = 0; j < kRotJ; ++j) acc[x][j] = 0.0;
  for (int a0 = 0; a0 < k; a0 += kRotK) {
    const int ka = min(kRotK, k - a0);
    __syncthreads();
    for (int e = threadIdx.x; e < kRotRows * kRotK; e += 256) {
      const int q = e / kRotK, a = e % kRotK;
      const int64_t i = r0 + q;
      vs[a][q] = (i < n && a < ka) ? static_cast<double>(V[i * k + a0 + a]) : 0.0;
    }
    for (int e = threadIdx.x; e < kRotK * 16 * kRotJ; e += 256) {
      const int a = e / (16 * kRotJ), c = e % (16 * kRotJ);
      ms[a][c] = (a < ka && c < kout) ? M[static_cast<int64_t>(a0 + a) * kout + c] : 0.0;
    }
    __syncthreads();
    for (int a = 0; a < ka; ++a) {
      double v[4], m[kRotJ];
#pragma unroll
      for (int x = 0; x < 4; ++x) v[x] = vs[a][tr * 4 + x];
#pragma unroll
      for (int j = 0; j < kRotJ; ++j) m[j] = ms[a][tc + 16 * j];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int j = 0; j < kRotJ; ++j) acc[x][j] = fma(v[x], m[j], acc[x][j]);
    }
  }
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int64_t i = r0 + tr * 4 + x;
    if (i >= n) continue;
#pragma unroll
    for (int j = 0; j < kRotJ; ++j) {
      const int c = tc + 16 * j;
      if (c < kout) out[i * kout + c] = static_cast<Real>(acc[x][j]);
    }
  }
}

// out (n x kout) = V (n x k) M (k x kout), M in FP64 on the device.
template <typename Real>
inline void ss_rotate(hipStream_t st, const Real* V, int k, const double* M, int kout, int64_t n, Real* out) {
  if (k >= 32 && kout >= 32 && kout <= 16 * kRotJ)
    hipLaunchKernelGGL(ss_rotate_tiled_kernel<Real>, dim3(static_cast<unsigned>(rfx_ceil_div(n, kRotRows))),
                       dim3(256), 0, st, V, k, M, kout, n, out);
  else
    hipLaunchKernelGGL(ss_rotate_kernel<Real>, dim3(static_cast<unsigned>(rfx_ceil_div(n * kout, 256))), dim3(256),
                       0, st, V, k, M, kout, n, out);
}

// Next block from the centred product without forming it:
//   V[i, c] = sum_a W[i,a] M[a,c] + r_i ar[c] + ao[c] + sum_a Vold[i,a] Ms[a,c]
// with prm = [M (k*k), ar (k), ao (k), Ms (k*k)]; one thread per (i, c).
template <typename Real>
__global__ void ss_update_kernel(const Real* __restrict__ W, const Real* __restrict__ Vold, Real* __restrict__ Vnew,
                                 const double* __restrict__ r, const double* __restrict__ prm, int k, int64_t n,
                                 int with_shift) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int c = static_cast<int>(e % k);
  const double* M = prm;
  const double* ar = prm + k * k;
  const double* ao = ar + k;
  const double* Ms = ao + k;
  double s = r[i] * ar[c] + ao[c];
  for (int a = 0; a < k; ++a) s += static_cast<double>(W[i * k + a]) * M[a * k + c];
  if (with_shift)
    for (int a = 0; a < k; ++a) s += static_cast<double>(Vold[i * k + a]) * Ms[a * k + c];
  Vnew[e] = static_cast<Real>(s);
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
  double residual = 0.0;  // max relative Ritz residual of the wanted pairs at exit
};

template <typename Real>
class SubspaceSolver {
 public:
  using Apply = std::function<int(const Real* V, Real* W, int k)>;

  SubspaceSolver(hipStream_t st, int64_t n, int k) : st_(st), n_(n), k_(k) {}

  static size_t al(size_t b) { return (b + 255) & ~size_t(255); }
  // Entries of the largest Gram: the fused per-iteration sums of run() (k <= 32
  // only) or a k x k block Gram.
  static int pairs_max(int k) {
    const int fused = k <= 32 ? (2 + 2 * k) * 2 * k : 0;
    return fused > k * k ? fused : k * k;
  }

  static size_t workspace_bytes(int64_t n, int k) {
    return al(sizeof(Real) * n * k) * 3 + al(8 * static_cast<size_t>(kGramBlocks) * pairs_max(k)) +
           al(8 * static_cast<size_t>(pairs_max(k))) + al(8 * static_cast<size_t>(2 * k * k + 2 * k)) + al(8 * k);
  }

  void carve(void* ws) {
    char* p = static_cast<char*>(ws);
    V_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    W_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    T_ = reinterpret_cast<Real*>(p);
    p += al(sizeof(Real) * n_ * k_);
    partial_ = reinterpret_cast<double*>(p);
    p += al(8 * static_cast<size_t>(kGramBlocks) * pairs_max(k_));
    kk_ = reinterpret_cast<double*>(p);
    p += al(8 * static_cast<size_t>(pairs_max(k_)));
    M_ = reinterpret_cast<double*>(p);
    p += al(8 * static_cast<size_t>(2 * k_ * k_ + 2 * k_));
    scale_ = reinterpret_cast<double*>(p);
  }

  Real* block_v() { return V_; }
  Real* block_w() { return W_; }
  Real* block_t() { return T_; }
  double* small_matrix() { return M_; }

  // Small host -> device upload; waits so the host buffer may be reused.
  int upload(double* dst, const double* src, size_t count) {
    RFX_HIP_TRY(hipMemcpyAsync(dst, src, 8 * count, hipMemcpyHostToDevice, st_));
    RFX_HIP_TRY(hipStreamSynchronize(st_));
    return RFX_OK;
  }

  template <typename TA, typename TB>
  int gram(const TA* A, int ka, const TB* B, int kb, std::vector<double>& out) {
    const int pairs = ka * kb;
    if (ka < 1 || kb < 1 || pairs > pairs_max(k_)) return RFX_ERR_UNSUPPORTED;
    const int64_t nb = std::min<int64_t>(kGramBlocks, (n_ + kGramRows - 1) / kGramRows);
    const int64_t rpb = (n_ + nb - 1) / nb;
    const unsigned tiles = static_cast<unsigned>(rfx_ceil_div(ka, kGramTile) * rfx_ceil_div(kb, kGramTile));
    hipLaunchKernelGGL((ss_gram_kernel<TA, TB>), dim3(static_cast<unsigned>(nb), tiles), dim3(256), 0, st_, A, ka,
                       B, kb, n_, rpb, partial_);
    hipLaunchKernelGGL(ss_reduce_partials_kernel, dim3(rfx_ceil_div(pairs, 4)), dim3(256), 0, st_, partial_,
                       static_cast<int>(nb), pairs, kk_);
    out.resize(pairs);
    RFX_HIP_TRY(hipMemcpyAsync(out.data(), kk_, 8 * pairs, hipMemcpyDeviceToHost, st_));
    RFX_HIP_TRY(hipStreamSynchronize(st_));
    return rfx_launch_status();
  }

  // CholeskyQR2 in FP64: twice V <- V R^{-1} with V^T V = R^T R.
  int orthonormalize(Real* V) {
    std::vector<double> G;
    for (int pass = 0; pass < 2; ++pass) {
      int rc = gram(V, k_, V, k_, G);
      if (rc) return rc;
      double tr = 0;
      for (int a = 0; a < k_; ++a) tr += G[a * k_ + a];
      if (!cholesky_lower(G, k_)) {
        // rank-deficient block: shift the Gram diagonal and retry once
        rc = gram(V, k_, V, k_, G);
        if (rc) return rc;
        for (int a = 0; a < k_; ++a) G[a * k_ + a] += 1e-12 * (tr > 0 ? tr : 1.0);
        if (!cholesky_lower(G, k_)) return RFX_ERR_INVALID;
      }
      std::vector<double> Rinv = inv_upper_from_lower(G, k_);
      rc = upload(M_, Rinv.data(), static_cast<size_t>(k_) * k_);
      if (rc) return rc;
      ss_rotate<Real>(st_, V, k_, M_, k_, n_, T_);
      RFX_HIP_TRY(hipMemcpyAsync(V, T_, sizeof(Real) * n_ * k_, hipMemcpyDeviceToDevice, st_));
    }
    return rfx_launch_status();
  }

  // Top `want` eigenpairs of B; coords (device, FP64 n x want) = V sqrt(lambda).
  // The iteration runs on B + shift I (a positive shift moves large negative
  // eigenvalues of a non-Euclidean B out of the way of the positive ones);
  // reported eigenvalues are those of B.  With start_from_v the block already
  // in block_v() is the starting subspace (a refinement sweep).
  int run(const Apply& apply, const double* r_dev, double g, int want, int max_iter, double tol, uint64_t seed,
          double* coords, SubspaceResult& res, double shift = 0.0, bool start_from_v = false) {
    const int k = k_;
    if (k > 32 || want > k) return RFX_ERR_UNSUPPORTED;
    if (!start_from_v)
      hipLaunchKernelGGL(ss_init_kernel<Real>, dim3(rfx_ceil_div(n_ * k, 256)), dim3(256), 0, st_, V_, n_, k, seed);
    int rc = orthonormalize(V_);
    if (rc) return rc;
    // constants of the centring algebra: r^T r and 1^T r
    double rr0 = 0.0, rr1 = 0.0;
    {
      std::vector<double> rh(n_);
      RFX_HIP_TRY(hipMemcpyAsync(rh.data(), r_dev, 8 * n_, hipMemcpyDeviceToHost, st_));
      RFX_HIP_TRY(hipStreamSynchronize(st_));
      for (int64_t i = 0; i < n_; ++i) {
        rr0 += rh[i] * rh[i];
        rr1 += rh[i];
      }
    }
    const int wdt = 2 + 2 * k, nc = 2 * k, pairs = wdt * nc;
    const int64_t nb = std::min<int64_t>(kGramBlocks, (n_ + kGramRows - 1) / kGramRows);
    const int64_t rpb = (n_ + nb - 1) / nb;
    const size_t kk = static_cast<size_t>(k) * k;
    std::vector<double> S(pairs), Gvv(kk), Gvw(kk), Gww(kk), sV(k), wV(k), sW(k), wW(k), q(k);
    std::vector<double> H(kk), Gc(kk), theta, Y, Z(kk), prm(2 * kk + 2 * k);
    const double sh = shift;
    res.residual = 1.0;
    for (int it = 0;; ++it) {
      rc = apply(V_, W_, k);
      if (rc) return rc;
      if (k == 16)
        hipLaunchKernelGGL((ss_fused_gram16_mfma_kernel<Real>), dim3(static_cast<unsigned>(nb)), dim3(256), 0, st_,
                           V_, W_, r_dev, n_, rpb, partial_);
      else
        hipLaunchKernelGGL((ss_fused_gram_kernel<Real>), dim3(static_cast<unsigned>(nb)), dim3(256),
                           8 * kGramRows * wdt, st_, V_, W_, r_dev, k, n_, rpb, partial_);
      hipLaunchKernelGGL(ss_reduce_partials_kernel, dim3(rfx_ceil_div(pairs, 4)), dim3(256), 0, st_, partial_,
                         static_cast<int>(nb), pairs, kk_);
      RFX_HIP_TRY(hipMemcpyAsync(S.data(), kk_, 8 * pairs, hipMemcpyDeviceToHost, st_));
      RFX_HIP_TRY(hipStreamSynchronize(st_));
      rc = rfx_launch_status();
      if (rc) return rc;
      for (int c = 0; c < k; ++c) {
        sV[c] = S[c];
        sW[c] = S[k + c];
        wV[c] = S[nc + c];
        wW[c] = S[nc + k + c];
      }
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          Gvv[a * k + c] = S[(2 + a) * nc + c];
          Gvw[a * k + c] = S[(2 + a) * nc + k + c];
          Gww[a * k + c] = S[(2 + k + a) * nc + k + c];
        }
      for (int c = 0; c < k; ++c) q[c] = wV[c] - g * sV[c];
      // H = V^T (Wc + sh V), Wc = -1/2 (W - r s^T - 1 q^T)
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c)
          H[a * k + c] = -0.5 * (Gvw[a * k + c] - wV[a] * sV[c] - sV[a] * q[c]) + sh * Gvv[a * k + c];
      for (int a = 0; a < k; ++a)
        for (int c = a + 1; c < k; ++c) H[a * k + c] = H[c * k + a] = 0.5 * (H[a * k + c] + H[c * k + a]);
      // Gc = (Wc + sh V)^T (Wc + sh V)
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double v = Gww[a * k + c] - wW[a] * sV[c] - sW[a] * q[c] - sV[a] * wW[c] + sV[a] * rr0 * sV[c] +
                     sV[a] * rr1 * q[c] - q[a] * sW[c] + q[a] * rr1 * sV[c] +
                     q[a] * static_cast<double>(n_) * q[c];
          v *= 0.25;
          if (sh != 0.0) {
            const double vwc_ac = -0.5 * (Gvw[a * k + c] - wV[a] * sV[c] - sV[a] * q[c]);
            const double vwc_ca = -0.5 * (Gvw[c * k + a] - wV[c] * sV[a] - sV[c] * q[a]);
            v += sh * (vwc_ac + vwc_ca) + sh * sh * Gvv[a * k + c];
          }
          Gc[a * k + c] = v;
        }
      // generalised Rayleigh-Ritz: L L^T = V^T V, eig((L^T)^-T H (L^T)^-1)
      std::vector<double> L = Gvv;
      if (!cholesky_lower(L, k)) return RFX_ERR_INVALID;
      const std::vector<double> Li = inv_upper_from_lower(L, k);
      std::vector<double> tmp(kk, 0.0), Ht(kk, 0.0);
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += H[a * k + b] * Li[b * k + c];
          tmp[a * k + c] = s;
        }
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += Li[b * k + a] * tmp[b * k + c];
          Ht[a * k + c] = s;
        }
      jacobi_eigh(Ht, k, theta, Y);
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += Li[a * k + b] * Y[b * k + c];
          Z[a * k + c] = s;
        }
      // Ritz residuals ||(B + sh) x - theta x||^2 = z^T Gc z - theta^2
      double worst = 0.0;
      for (int c = 0; c < want; ++c) {
        double zgz = 0;
        for (int a = 0; a < k; ++a)
          for (int b = 0; b < k; ++b) zgz += Z[a * k + c] * Gc[a * k + b] * Z[b * k + c];
        const double r2 = std::max(zgz - theta[c] * theta[c], 0.0);
        worst = std::max(worst, std::sqrt(r2) / std::max(std::fabs(theta[c] - sh), 1e-300));
      }
      res.residual = worst;
      if ((it >= (start_from_v ? 0 : 1) && worst < tol) || it + 1 >= max_iter) {
        // Ritz vectors X = V Z (orthonormal: Z^T Gvv Z = I), scaled by sqrt(lambda)
        std::vector<double> Zw(static_cast<size_t>(k) * want), sc(want);
        for (int a = 0; a < k; ++a)
          for (int c = 0; c < want; ++c) Zw[a * want + c] = Z[a * k + c];
        res.eigvals.resize(k);
        for (int c = 0; c < k; ++c) res.eigvals[c] = theta[c] - sh;
        for (int c = 0; c < want; ++c) sc[c] = res.eigvals[c] > 0 ? std::sqrt(res.eigvals[c]) : 0.0;
        rc = upload(M_, Zw.data(), static_cast<size_t>(k) * want);
        if (rc) return rc;
        rc = upload(scale_, sc.data(), want);
        if (rc) return rc;
        ss_rotate<Real>(st_, V_, k, M_, want, n_, T_);
        hipLaunchKernelGGL(ss_coords_kernel<Real>, dim3(rfx_ceil_div(n_ * want, 256)), dim3(256), 0, st_, T_, want,
                           scale_, want, n_, coords);
        RFX_HIP_TRY(hipStreamSynchronize(st_));
        res.iterations = it + 1;
        return rfx_launch_status();
      }
      // next block: (Wc + sh V) Z R^{-1}, R^T R = Z^T Gc Z
      std::vector<double> t2(kk, 0.0), Gz(kk, 0.0);
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += Gc[a * k + b] * Z[b * k + c];
          t2[a * k + c] = s;
        }
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += Z[b * k + a] * t2[b * k + c];
          Gz[a * k + c] = s;
        }
      for (int a = 0; a < k; ++a)
        for (int c = a + 1; c < k; ++c) Gz[a * k + c] = Gz[c * k + a] = 0.5 * (Gz[a * k + c] + Gz[c * k + a]);
      if (!cholesky_lower(Gz, k)) return RFX_ERR_INVALID;
      const std::vector<double> Ri = inv_upper_from_lower(Gz, k);
      std::vector<double> Mz(kk, 0.0);
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          double s = 0;
          for (int b = 0; b < k; ++b) s += Z[a * k + b] * Ri[b * k + c];
          Mz[a * k + c] = s;
        }
      // V' = -1/2 W Mz + 1/2 r (s^T Mz) + 1/2 1 (q^T Mz) + sh V Mz
      for (int a = 0; a < k; ++a)
        for (int c = 0; c < k; ++c) {
          prm[a * k + c] = -0.5 * Mz[a * k + c];
          prm[kk + 2 * k + a * k + c] = sh * Mz[a * k + c];
        }
      for (int c = 0; c < k; ++c) {
        double sm = 0, qm = 0;
        for (int a = 0; a < k; ++a) {
          sm += sV[a] * Mz[a * k + c];
          qm += q[a] * Mz[a * k + c];
        }
        prm[kk + c] = 0.5 * sm;
        prm[kk + k + c] = 0.5 * qm;
      }
      rc = upload(M_, prm.data(), prm.size());
      if (rc) return rc;
      hipLaunchKernelGGL(ss_update_kernel<Real>, dim3(rfx_ceil_div(n_ * k, 256)), dim3(256), 0, st_, W_, V_, T_,
                         r_dev, M_, k, n_, sh != 0.0 ? 1 : 0);
      RFX_HIP_TRY(hipMemcpyAsync(V_, T_, sizeof(Real) * n_ * k, hipMemcpyDeviceToDevice, st_));
    }
  }

 private:
  hipStream_t st_;
  int64_t n_;
  int k_;
  Real *V_ = nullptr, *W_ = nullptr, *T_ = nullptr;
  double *partial_ = nullptr, *kk_ = nullptr, *M_ = nullptr, *scale_ = nullptr;
};

}  // namespace rfx
