// A6 + A7: classical (Torgerson) 3-D MDS of the forest proximity on the GPU.
//
// The reference (src/rf_mds_cpu.cpp:88-341) builds D = max(M) - M with a zero
// diagonal, squares it, double-centres it in FP64 and hands the n x n matrix
// to LAPACK dsyevd for all n eigenpairs, keeping the top three with
// lambda > 1e-10.  Here nothing n x n is materialised besides the int32
// counts: the squared distances E = D o D are formed on the fly from the
// counts, tile by tile, inside the operator kernel, and the top eigenpairs
// come from block subspace iteration (subspace.hpp) with FP64 row means,
// FP64 CholeskyQR2 and FP64 Rayleigh-Ritz.  The operator runs in FP32 or FP64.
//
// Proximity forms (the MDS input M):
//   COUNTS     M_ij = c_ij * scale (e.g. ALL counts / T), any symmetric count
//              matrix, UPPER or FULL layout
//   REFERENCE  the pybind path's input (python/randomforest_py.cpp:1138-1146,
//              1221-1252; src/rf_mds_cpu.cpp:154-186): OOB_UPPER counts ->
//              finishprox S (FP64) -> float32 -> M_ij = 1/2 (S_ij/nout_i +
//              S_ij/nout_j), M_ii = 1 (rows with nout_i = 0 are not divided)
//   MATRIX     a given FP64 n x n column-major proximity S, as
//              rf::compute_mds_3d_cpu receives it: with nout, M_ij = 1/2
//              (S_ij/nout_i + S_ji/nout_j) and M_ii = 1
//              (src/rf_mds_cpu.cpp:154-186); without, M = S as it is
//   RFGAP      a given FP64 RF-GAP matrix P with the reference's use_rfgap
//              renormalisation (src/rf_mds_cpu.cpp:125-153): off the
//              diagonal as MATRIX with nout, M_ii = P_ii/nout_i (rows with
//              nout_i = 0 undivided)

#include <cstring>
#include <type_traits>
#include <vector>

#include "rfx_common.hpp"
#include "subspace.hpp"

namespace {

constexpr int kRows = 256;  // rows per workgroup = threads
constexpr int kCols = 32;   // columns per LDS tile
constexpr int kBlockK = 16; // subspace block size (VALU operator; Ritz values reported)
// Block of the count forms on the matrix cores: 32 columns from n =
// kMfWideMinN.  At n = 200,000 (covtype-like, T = 1000) the top eigenvalues
// sit close together (lambda3 / lambda4 = 1.007), so the iteration count
// follows lambda(k+1) / lambda3; each count, formed into E once, feeds one
// product per 16 columns, so a wider block costs 1.3-1.5 x per application
// for 0.56 x the iterations (dense MDS 10.1 -> 7.5 s FP64, 4.7 -> 3.8 s
// FP32).  Below, the per-iteration work outside the operator (the 32-wide
// Rayleigh-Ritz on the host, the Gram) outweighs it: the 64 instances of
// 1,000 to 50,000 samples took 10.2 s at 32 columns against 5.6 s at 16.
constexpr int kMfWideK = 32;
constexpr int64_t kMfWideMinN = 100000;
constexpr int kDenseMaxK = kMfWideK > kBlockK ? kMfWideK : kBlockK;

struct ProxForm {
  int32_t mode;    // RFX_MDS_COUNTS / _REFERENCE / _MATRIX / _RFGAP
  int32_t layout;  // RFX_LAYOUT_UPPER / RFX_LAYOUT_FULL
  double scale;
  double mx;       // max of M
  const int32_t* nout;
  const double* mat;  // MATRIX, RFGAP: the given matrix, column-major
  int64_t n;
};

__device__ __forceinline__ int64_t src_index(int32_t layout, int64_t i, int64_t j, int64_t n) {
  if (layout == RFX_LAYOUT_FULL) return i + j * n;
  const int64_t a = i < j ? i : j, b = i < j ? j : i;
  return rfx_upper_index(a, b, n);
}

__device__ __forceinline__ double recip_or_one(int32_t c) { return c > 0 ? 1.0 / static_cast<double>(c) : 1.0; }

// M_ij (i != j) from the stored count c_ij (MATRIX: from S itself)
__device__ __forceinline__ double prox_value(const ProxForm& f, int32_t c, int64_t i, int64_t j) {
  if (f.mode == RFX_MDS_COUNTS) return static_cast<double>(c) * f.scale;
  if (f.mode == RFX_MDS_MATRIX || f.mode == RFX_MDS_RFGAP) {
    const double sij = f.mat[i + j * f.n];
    if (!f.nout) return sij;
    const double sji = f.mat[j + i * f.n];
    const int32_t ni = f.nout[i], nj = f.nout[j];
    const double ti = ni > 0 ? sij / static_cast<double>(ni) : sij;
    const double tj = nj > 0 ? sji / static_cast<double>(nj) : sji;
    return 0.5 * (ti + tj);
  }
  const int32_t ni = f.nout[i], nj = f.nout[j];
  const double a = static_cast<double>(c) / 2.0;
  const double S = 0.5 * (a * recip_or_one(ni) + a * recip_or_one(nj));
  const double Sf = static_cast<double>(static_cast<float>(S));
  const double ti = ni > 0 ? Sf / static_cast<double>(ni) : Sf;
  const double tj = nj > 0 ? Sf / static_cast<double>(nj) : Sf;
  return 0.5 * (ti + tj);
}

__device__ __forceinline__ double diag_value(const ProxForm& f, int32_t c, int64_t i) {
  if (f.mode == RFX_MDS_MATRIX) return f.nout ? 1.0 : f.mat[i + i * f.n];
  if (f.mode == RFX_MDS_RFGAP) {
    const double sii = f.mat[i + i * f.n];
    const int32_t ni = f.nout[i];
    return ni > 0 ? sii / static_cast<double>(ni) : sii;
  }
  return f.mode == RFX_MDS_COUNTS ? static_cast<double>(c) * f.scale : 1.0;
}

__device__ __forceinline__ int32_t count_at(const ProxForm& f, const int32_t* counts, int64_t idx) {
  return (f.mode == RFX_MDS_MATRIX || f.mode == RFX_MDS_RFGAP) ? 0 : counts[idx];
}

// Max of M over all entries (diagonal included), FP64 bits via atomicMax.
__global__ void mds_max_kernel(const int32_t* __restrict__ counts, int64_t n, ProxForm f,
                               unsigned long long* __restrict__ out) {
  const int64_t i = blockIdx.x;
  double m = 0.0;
  const int64_t j0 = f.layout == RFX_LAYOUT_UPPER ? i : 0;
  for (int64_t j = j0 + threadIdx.x; j < n; j += blockDim.x) {
    // row i of the symmetric matrix, read along the contiguous direction
    const int32_t c = count_at(f, counts, f.layout == RFX_LAYOUT_UPPER ? src_index(f.layout, i, j, n) : j + i * n);
    const double v = (i == j) ? diag_value(f, c, i) : prox_value(f, c, i, j);
    m = v > m ? v : m;
  }
  for (int d = 32; d > 0; d >>= 1) {
    const double o = __shfl_down(m, d, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

// Y_part[s][i][c] = sum over j in split s of E_ij * V[j][c]  (ROWSUM: V = 1, K = 1)
template <typename Real, int K, bool ROWSUM>
__global__ __launch_bounds__(kRows) void mds_apply_kernel(const int32_t* __restrict__ counts, int64_t n, ProxForm f,
                                                          const Real* __restrict__ V, int64_t jlen,
                                                          Real* __restrict__ Ypart) {
  __shared__ Real Et[kCols][kRows];
  __shared__ Real Vt[kCols][K];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kRows;
  const int64_t jb = static_cast<int64_t>(blockIdx.y) * jlen;
  const int64_t je = min(n, jb + jlen);
  const int tid = threadIdx.x;
  const int64_t i = i0 + tid;
  Real acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = Real(0);

  for (int64_t j0 = jb; j0 < je; j0 += kCols) {
    __syncthreads();
    // Stage E for rows [i0, i0+256) x columns [j0, j0+32); choose the thread
    // mapping that reads the packed source contiguously.
    const bool right = f.layout == RFX_LAYOUT_UPPER && j0 >= i0 + kRows;
    for (int e = tid; e < kRows * kCols; e += kRows) {
      int r, q;
      if (right) {
        r = e / kCols;
        q = e % kCols;
      } else {
        q = e / kRows;
        r = e % kRows;
      }
      const int64_t ii = i0 + r, jj = j0 + q;
      Real v = Real(0);
      if (ii < n && jj < je && ii != jj) {
        const int32_t c = count_at(f, counts, src_index(f.layout, ii, jj, n));
        const double d = f.mx - prox_value(f, c, ii, jj);
        v = static_cast<Real>(d * d);
      }
      Et[q][r] = v;
    }
    if (!ROWSUM) {
      for (int e = tid; e < kCols * K; e += kRows) {
        const int q = e / K, c = e % K;
        const int64_t jj = j0 + q;
        Vt[q][c] = jj < je ? V[jj * K + c] : Real(0);
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int q = 0; q < kCols; ++q) {
      const Real e = Et[q][tid];
      if (ROWSUM) {
        acc[0] += e;
      } else {
#pragma unroll
        for (int c = 0; c < K; ++c) acc[c] += e * Vt[q][c];
      }
    }
  }
  if (i < n) {
    Real* out = Ypart + (static_cast<int64_t>(blockIdx.y) * n + i) * K;
#pragma unroll
    for (int c = 0; c < K; ++c) out[c] = acc[c];
  }
}

// The operator on the matrix cores (v_mfma_f32_16x16x4f32 for the FP32
// operator, v_mfma_f64_16x16x4f64 for FP64), for the count forms (COUNTS,
// REFERENCE).  A workgroup owns 128 rows (4 waves x 2
// strips of 16) and a split of the columns, 64 columns per step: the step's
// 128 x 64 counts are staged in LDS, read along the packed triangle's
// contiguous direction (along a row right of the block rows, along a column
// left of them and in the FULL column-major layout), V's 64 rows beside them.
// Each wave then forms E_ij = (mx - M_ij)^2 straight into A operands (one
// element per lane and instruction, A[row l16][k g]) and multiplies by V on
// the matrix cores, so each E element feeds 16 multiply-adds in one
// instruction slot.  M is formed in FP64 as the VALU operator forms it, with
// per-row reciprocals of nout for REFERENCE (precomputed: 1 / nout or 1).
// The output row of a (lane, element) slot differs between the two products
// (MdsMfma::row; both checked with lane probes on gfx950).  Out-of-range elements load element 0 (always valid) and
// are zeroed by selects, so the loads issue without branches.
typedef float md_f32x4 __attribute__((ext_vector_type(4)));
constexpr int kMfRowsWg = 128;  // rows per workgroup
constexpr int kMfStrips = kMfRowsWg / 64;  // 16-row strips per wave
constexpr int kMfStep = 64;     // columns per step
// padded LDS stride of the staged counts (int32): 68 = 4 mod 64, so the 64
// lanes' reads of one product (rows l16, columns g) hit 64 different banks
constexpr int kMfCStride = kMfStep + 4;
constexpr int kMfBatch = 8;     // loads in flight per thread before their LDS stores

template <int MODE, typename Real>
__device__ __forceinline__ Real mf_e_value(const ProxForm& f, int32_t c, double ri, double rj) {
  double m;
  if (MODE == RFX_MDS_COUNTS) {
    m = static_cast<double>(c) * f.scale;
  } else {
    const double a = static_cast<double>(c) / 2.0;
    const double S = 0.5 * (a * ri + a * rj);
    const double Sf = static_cast<double>(static_cast<float>(S));
    m = 0.5 * (Sf * ri + Sf * rj);
  }
  const double d = f.mx - m;
  return static_cast<Real>(d * d);
}

typedef double md_f64x4 __attribute__((ext_vector_type(4)));
// One 16 x 16 x 4 product: FP32 (v_mfma_f32_16x16x4f32) or FP64
// (v_mfma_f64_16x16x4f64); A[i][k] and B[k][j] sit in lanes i + 16k and
// j + 16k for both.
template <typename Real>
struct MdsMfma;
template <>
struct MdsMfma<float> {
  typedef md_f32x4 Acc;
  static __device__ __forceinline__ Acc mma(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D[i][j] in lane j + 16 (i / 4), element i % 4
  static __device__ __forceinline__ int row(int g, int e) { return 4 * g + e; }
};
template <>
struct MdsMfma<double> {
  typedef md_f64x4 Acc;
  static __device__ __forceinline__ Acc mma(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D[i][j] in lane j + 16 (i % 4), element i / 4
  static __device__ __forceinline__ int row(int g, int e) { return 4 * e + g; }
};

// KG groups of 16 block columns: each E element, formed once, feeds KG
// products.
template <typename Real, int MODE, int KG>
__global__ __launch_bounds__(256) void mds_apply_mfma_kernel(const int32_t* __restrict__ counts, int64_t n, ProxForm f,
                                                             const double* __restrict__ rinv,
                                                             const Real* __restrict__ V, int64_t jlen,
                                                             Real* __restrict__ Ypart) {
  using M = MdsMfma<Real>;
  constexpr int K = 16 * KG;
  __shared__ int32_t Cs[kMfRowsWg][kMfCStride];  // counts of the step, [row][column]
  __shared__ Real Vs[KG][kMfStep][16];            // V rows of the step in 16-column planes: the 64 lanes
                                                  // (c, g) of a product read 64 consecutive words
  __shared__ double Rs[kMfStep];                  // their reciprocals (REFERENCE)
  __shared__ int64_t rowbase[kMfRowsWg];          // packed-triangle offset of each block row
  __shared__ int64_t colbase[kMfStep];            // ... and of each step column (left of the rows)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, g = lane >> 4;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kMfRowsWg;
  const int64_t jb = static_cast<int64_t>(blockIdx.y) * jlen;
  const int64_t je = min(n, jb + jlen);
  const bool full = f.layout == RFX_LAYOUT_FULL;
  if (tid < kMfRowsWg) rowbase[tid] = rfx_upper_row(i0 + tid, n);
  typename M::Acc acc[kMfStrips][KG];
  double ri[kMfStrips];
#pragma unroll
  for (int s = 0; s < kMfStrips; ++s) {
#pragma unroll
    for (int kg = 0; kg < KG; ++kg) acc[s][kg] = typename M::Acc{Real(0), Real(0), Real(0), Real(0)};
    const int64_t i = i0 + wave * (16 * kMfStrips) + s * 16 + l16;
    ri[s] = (MODE == RFX_MDS_REFERENCE && i < n) ? rinv[i] : 1.0;
  }
  for (int64_t j0 = jb; j0 < je; j0 += kMfStep) {
    const bool leftside = full || j0 + kMfStep <= i0;
    __syncthreads();  // the previous step's operands are consumed (and rowbase written)
    if (leftside && !full) {
      if (tid < kMfStep) colbase[tid] = rfx_upper_row(j0 + tid, n);
      __syncthreads();
    }
    if (leftside) {
      // left of the block rows (or column-major): thread t reads row t % 128 of
      // columns t / 128 + 2k, consecutive threads one column's consecutive rows
      const int r = tid & (kMfRowsWg - 1);
      const int64_t ii = i0 + r;
#pragma unroll 1
      for (int k0 = 0; k0 < kMfStep / 2; k0 += kMfBatch) {
        int32_t v[kMfBatch];
#pragma unroll
        for (int k = 0; k < kMfBatch; ++k) {
          const int q = (tid >> 7) + 2 * (k0 + k);
          const int64_t jj = j0 + q;
          const bool ok = ii < n && jj < je;
          const int64_t idx = full ? ii + jj * n : colbase[q] + (ii - jj);
          const int32_t c = counts[ok ? idx : 0];
          v[k] = ok ? c : 0;
        }
#pragma unroll
        for (int k = 0; k < kMfBatch; ++k) Cs[r][(tid >> 7) + 2 * (k0 + k)] = v[k];
      }
    } else {
      // right of (or across) the block rows: thread t reads column t % 64 of
      // rows t / 64 + 4k, consecutive threads one row's consecutive columns
      const bool diag = j0 < i0 + kMfRowsWg;
      const int q = tid & 63;
      const int64_t jj = j0 + q;
#pragma unroll 1
      for (int k0 = 0; k0 < kMfRowsWg / 4; k0 += kMfBatch) {
        int32_t v[kMfBatch];
#pragma unroll
        for (int k = 0; k < kMfBatch; ++k) {
          const int r = (tid >> 6) + 4 * (k0 + k);
          const int64_t ii = i0 + r;
          bool ok = ii < n && jj < je;
          int64_t idx = rowbase[r] + (jj - ii);
          if (diag) {
            ok = ok && ii != jj;
            idx = ii < jj ? idx : rfx_upper_row(jj, n) + (ii - jj);
          }
          const int32_t c = counts[ok ? idx : 0];
          v[k] = ok ? c : 0;
        }
#pragma unroll
        for (int k = 0; k < kMfBatch; ++k) Cs[(tid >> 6) + 4 * (k0 + k)][q] = v[k];
      }
    }
    for (int e = tid; e < kMfStep * K; e += 256) {
      const int q = e / K, c = e % K;
      const int64_t jj = j0 + q;
      Vs[c >> 4][q][c & 15] = jj < je ? V[jj * K + c] : Real(0);
    }
    if (MODE == RFX_MDS_REFERENCE && tid < kMfStep) Rs[tid] = j0 + tid < je ? rinv[j0 + tid] : 1.0;
    __syncthreads();
    // 16 products of K = 4: k slot g of product m is column 4m + g of the step
#pragma unroll 2
    for (int m = 0; m < kMfStep / 4; ++m) {
      const int q = 4 * m + g;
      const int64_t jj = j0 + q;
      Real b[KG];
#pragma unroll
      for (int kg = 0; kg < KG; ++kg) b[kg] = Vs[kg][q][l16];
      const double rj = MODE == RFX_MDS_REFERENCE ? Rs[q] : 1.0;
#pragma unroll
      for (int s = 0; s < kMfStrips; ++s) {
        const int r = wave * (16 * kMfStrips) + s * 16 + l16;
        const int64_t ii = i0 + r;
        const Real ev = mf_e_value<MODE, Real>(f, Cs[r][q], ri[s], rj);
        const Real a = (ii < n && jj < je && ii != jj) ? ev : Real(0);
#pragma unroll
        for (int kg = 0; kg < KG; ++kg) acc[s][kg] = M::mma(a, b[kg], acc[s][kg]);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kMfStrips; ++s) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t i = i0 + wave * (16 * kMfStrips) + s * 16 + M::row(g, e);
#pragma unroll
      for (int kg = 0; kg < KG; ++kg)
        if (i < n) Ypart[(static_cast<int64_t>(blockIdx.y) * n + i) * K + 16 * kg + l16] = acc[s][kg][e];
    }
  }
}

// Per-row reciprocals of nout for the REFERENCE form: 1 / nout_i, or 1 when
// nout_i = 0 (such rows are not divided).
__global__ void mds_rinv_kernel(const int32_t* __restrict__ nout, int64_t n, double* __restrict__ rinv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) rinv[i] = nout && nout[i] > 0 ? 1.0 / static_cast<double>(nout[i]) : 1.0;
}

template <typename Real, typename Out>
__global__ void mds_sum_splits_kernel(const Real* __restrict__ Ypart, int splits, int64_t nk, Out* __restrict__ Y) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nk) return;
  double s = 0.0;
  for (int q = 0; q < splits; ++q) s += static_cast<double>(Ypart[static_cast<int64_t>(q) * nk + e]);
  Y[e] = static_cast<Out>(s);
}

__global__ void mds_zero_kernel(unsigned long long* p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0ull;
}

__global__ void mds_scale_kernel(double* __restrict__ r, int64_t n, double inv) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) r[i] *= inv;
}

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

int splits_for(int64_t n) {
  // enough workgroups to cover 256 CUs a few times over
  const int64_t row_blocks = (n + kRows - 1) / kRows;
  int64_t s = (2048 + row_blocks - 1) / row_blocks;
  const int64_t max_s = (n + kCols - 1) / kCols;
  if (s > max_s) s = max_s;
  if (s > 64) s = 64;
  return static_cast<int>(s < 1 ? 1 : s);
}

template <typename Real>
size_t dense_ws(int64_t n) {
  const int sp = splits_for(n);
  return al(8 * n) + al(8 * 4) + al(sizeof(Real) * sp * n * kDenseMaxK) + al(8 * sp * n) + al(8 * n) +
         rfx::SubspaceSolver<Real>::workspace_bytes(n, kDenseMaxK);
}

template <typename Real>
int run_dense(hipStream_t st, const int32_t* counts, const double* mat, int32_t layout, int64_t n, int32_t mode,
              double scale, const int32_t* nout, int32_t max_iter, double tol, void* ws, double* coords, double* eigvals_host,
              int32_t* iters_host) {
  char* p = static_cast<char*>(ws);
  double* r = reinterpret_cast<double*>(p);
  p += al(8 * n);
  unsigned long long* mx_bits = reinterpret_cast<unsigned long long*>(p);
  p += al(8 * 4);
  const int sp = splits_for(n);
  Real* Ypart = reinterpret_cast<Real*>(p);
  p += al(sizeof(Real) * sp * n * kDenseMaxK);
  double* rpart = reinterpret_cast<double*>(p);
  p += al(8 * sp * n);
  double* rinv = reinterpret_cast<double*>(p);
  p += al(8 * n);
  ProxForm f{mode, layout, scale, 0.0, nout, mat, n};

  hipLaunchKernelGGL(mds_zero_kernel, dim3(1), dim3(64), 0, st, mx_bits);
  hipLaunchKernelGGL(mds_max_kernel, dim3(static_cast<unsigned>(n)), dim3(256), 0, st, counts, n, f, mx_bits);
  unsigned long long bits = 0;
  RFX_HIP_TRY(hipMemcpyAsync(&bits, mx_bits, 8, hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  std::memcpy(&f.mx, &bits, 8);

  const int64_t jlen = ((n + sp - 1) / sp + kCols - 1) / kCols * kCols;
  dim3 grid(rfx_ceil_div(n, kRows), static_cast<unsigned>(sp));
  // FP64 row means of E (double-centring)
  hipLaunchKernelGGL((mds_apply_kernel<double, 1, true>), grid, dim3(kRows), 0, st, counts, n, f,
                     static_cast<const double*>(nullptr), jlen, rpart);
  hipLaunchKernelGGL((mds_sum_splits_kernel<double, double>), dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, rpart,
                     sp, n, r);
  hipLaunchKernelGGL(mds_scale_kernel, dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, r, n,
                     1.0 / static_cast<double>(n));
  std::vector<double> rh(n);
  RFX_HIP_TRY(hipMemcpyAsync(rh.data(), r, 8 * n, hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  double g = 0.0;
  for (int64_t i = 0; i < n; ++i) g += rh[i];
  g /= static_cast<double>(n);

  // The operator of the count forms runs on the matrix cores.
  const bool mfma = mode == RFX_MDS_COUNTS || mode == RFX_MDS_REFERENCE;
  const int K = mfma && n >= kMfWideMinN ? kMfWideK : kBlockK;
  rfx::SubspaceSolver<Real> solver(st, n, K);
  solver.carve(p);
  if (mode == RFX_MDS_REFERENCE)
    hipLaunchKernelGGL(mds_rinv_kernel, dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, nout, n, rinv);
  const dim3 mgrid(rfx_ceil_div(n, kMfRowsWg), static_cast<unsigned>(sp));
  auto apply = [&](const Real* V, Real* W, int k) -> int {
    (void)k;
    if (mfma) {
      if (mode == RFX_MDS_COUNTS && K == kMfWideK)
        hipLaunchKernelGGL((mds_apply_mfma_kernel<Real, RFX_MDS_COUNTS, kMfWideK / 16>), mgrid, dim3(256), 0, st,
                           counts, n, f, rinv, V, jlen, Ypart);
      else if (mode == RFX_MDS_COUNTS)
        hipLaunchKernelGGL((mds_apply_mfma_kernel<Real, RFX_MDS_COUNTS, 1>), mgrid, dim3(256), 0, st, counts, n, f,
                           rinv, V, jlen, Ypart);
      else if (K == kMfWideK)
        hipLaunchKernelGGL((mds_apply_mfma_kernel<Real, RFX_MDS_REFERENCE, kMfWideK / 16>), mgrid, dim3(256), 0, st,
                           counts, n, f, rinv, V, jlen, Ypart);
      else
        hipLaunchKernelGGL((mds_apply_mfma_kernel<Real, RFX_MDS_REFERENCE, 1>), mgrid, dim3(256), 0, st, counts, n,
                           f, rinv, V, jlen, Ypart);
      hipLaunchKernelGGL((mds_sum_splits_kernel<Real, Real>), dim3(rfx_ceil_div(n * K, 256)), dim3(256), 0, st,
                         Ypart, sp, n * K, W);
      return rfx_launch_status();
    }
    hipLaunchKernelGGL((mds_apply_kernel<Real, kBlockK, false>), grid, dim3(kRows), 0, st, counts, n, f, V, jlen,
                       Ypart);
    hipLaunchKernelGGL((mds_sum_splits_kernel<Real, Real>), dim3(rfx_ceil_div(n * kBlockK, 256)), dim3(256), 0, st,
                       Ypart, sp, n * kBlockK, W);
    return rfx_launch_status();
  };
  rfx::SubspaceResult res;
  int rc = solver.run(apply, r, g, 3, max_iter, tol, 0x5eedull, coords, res);
  if (rc) return rc;
  // Fewer than three positive Ritz values while large negative ones fill the
  // block: rerun on B + sigma I with sigma = |most negative Ritz value|.
  if (res.eigvals.size() >= 3 && !(res.eigvals[2] > 1e-10) && res.eigvals.back() < 0.0) {
    const double sigma = -res.eigvals.back();
    rc = solver.run(apply, r, g, 3, max_iter, tol, 0x5eedull, coords, res, sigma);
    if (rc) return rc;
  }
  if (eigvals_host)
    for (int c = 0; c < kBlockK; ++c) eigvals_host[c] = res.eigvals[c];
  if (iters_host) *iters_host = res.iterations;
  if (res.eigvals.size() < 3 || !(res.eigvals[2] > 1e-10)) return RFX_ERR_INVALID;
  return RFX_OK;
}

}  // namespace

extern "C" size_t rfx_mds_dense_workspace_bytes(int64_t n, int32_t precision) {
  if (n <= 0) return 0;
  return precision == RFX_PREC_FP64 ? dense_ws<double>(n) : dense_ws<float>(n);
}

extern "C" int rfx_mds_dense(hipStream_t st, const int32_t* counts, int32_t layout, int64_t n, int32_t mode,
                             double scale, const int32_t* nout, int32_t precision, int32_t max_iter, double tol,
                             void* workspace, size_t workspace_bytes, double* coords, double* eigvals_host,
                             int32_t* iters_host) {
  if (!counts || !coords || !workspace || n < 4 || n > RFX_MAX_SAMPLES || max_iter < 1) return RFX_ERR_INVALID;
  if (layout != RFX_LAYOUT_UPPER && layout != RFX_LAYOUT_FULL) return RFX_ERR_INVALID;
  if (mode != RFX_MDS_COUNTS && mode != RFX_MDS_REFERENCE) return RFX_ERR_INVALID;
  if (mode == RFX_MDS_REFERENCE && (!nout || layout != RFX_LAYOUT_UPPER)) return RFX_ERR_INVALID;
  if (precision != RFX_PREC_FP32 && precision != RFX_PREC_FP64) return RFX_ERR_INVALID;
  if (workspace_bytes < rfx_mds_dense_workspace_bytes(n, precision)) return RFX_ERR_WORKSPACE;
  if (precision == RFX_PREC_FP64)
    return run_dense<double>(st, counts, nullptr, layout, n, mode, scale, nout, max_iter, tol, workspace, coords,
                             eigvals_host, iters_host);
  return run_dense<float>(st, counts, nullptr, layout, n, mode, scale, nout, max_iter, tol, workspace, coords,
                          eigvals_host, iters_host);
}

extern "C" int rfx_mds_dense_matrix(hipStream_t st, const double* prox, int64_t n, const int32_t* nout,
                                    int32_t precision, int32_t max_iter, double tol, void* workspace,
                                    size_t workspace_bytes, double* coords, double* eigvals_host, int32_t* iters_host) {
  if (!prox || !coords || !workspace || n < 4 || n > RFX_MAX_SAMPLES || max_iter < 1) return RFX_ERR_INVALID;
  if (precision != RFX_PREC_FP32 && precision != RFX_PREC_FP64) return RFX_ERR_INVALID;
  if (workspace_bytes < rfx_mds_dense_workspace_bytes(n, precision)) return RFX_ERR_WORKSPACE;
  if (precision == RFX_PREC_FP64)
    return run_dense<double>(st, nullptr, prox, RFX_LAYOUT_FULL, n, RFX_MDS_MATRIX, 1.0, nout, max_iter, tol,
                             workspace, coords, eigvals_host, iters_host);
  return run_dense<float>(st, nullptr, prox, RFX_LAYOUT_FULL, n, RFX_MDS_MATRIX, 1.0, nout, max_iter, tol, workspace,
                          coords, eigvals_host, iters_host);
}

extern "C" int rfx_mds_dense_rfgap(hipStream_t st, const double* prox, int64_t n, const int32_t* nout,
                                   int32_t precision, int32_t max_iter, double tol, void* workspace,
                                   size_t workspace_bytes, double* coords, double* eigvals_host, int32_t* iters_host) {
  if (!prox || !nout || !coords || !workspace || n < 4 || n > RFX_MAX_SAMPLES || max_iter < 1)
    return RFX_ERR_INVALID;
  if (precision != RFX_PREC_FP32 && precision != RFX_PREC_FP64) return RFX_ERR_INVALID;
  if (workspace_bytes < rfx_mds_dense_workspace_bytes(n, precision)) return RFX_ERR_WORKSPACE;
  if (precision == RFX_PREC_FP64)
    return run_dense<double>(st, nullptr, prox, RFX_LAYOUT_FULL, n, RFX_MDS_RFGAP, 1.0, nout, max_iter, tol,
                             workspace, coords, eigvals_host, iters_host);
  return run_dense<float>(st, nullptr, prox, RFX_LAYOUT_FULL, n, RFX_MDS_RFGAP, 1.0, nout, max_iter, tol, workspace,
                          coords, eigvals_host, iters_host);
}
