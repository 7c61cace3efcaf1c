// A8 + A9: low-rank proximity factors and 3-D MDS from the factors.
//
// Factors.  P = C / T with C = sum_t Z_t Z_t^T (Z_t the sample-to-leaf
// indicator of tree t) is positive semidefinite, so P ~ U U^T with U the top
// eigenvectors scaled by sqrt(eigenvalue).  They come from randomized
// subspace iteration (Halko, Martinsson & Tropp 2011, Alg. 4.4 with a final
// Rayleigh-Ritz step) where every product C V is matrix-free through the
// leaf buckets: per tree chunk, leaf sums S_L = sum_{j in L} V_j
// (bucket-major, members in ascending order so the float sums are
// deterministic), then Y_i += sum_t S_{leaf_t(i)} (sample-major gather).
// No n x n array exists at any point, so the same path serves n = 10^6.
// The reference's low-rank object (cuda/rf_proximity_lowrank.cu:414-1779)
// builds rank-1 power-iteration columns per tree; this path approximates
// the accumulated proximity itself and is judged against FP64.
//
// Storage.  U is kept in FP32, BF16, INT8 or FP8-e4m3 (OCP), the last two
// with per-column FP32 scales (u = q * s_c).
//
// Factor MDS.  With P~ = U U^T and the exact unit diagonal, the squared
// distances are E_ij = (1 - P~_ij)^2 (i != j), E_ii = 0, and
//   E v = (1^T v) 1 - 2 U (U^T v) + q(v) - (1 - |u_i|^2)^2 v_i,
//   q_i(v) = sum_j (u_i . u_j)^2 v_j = x_i^T diag(w) Z,  Z = X^T v,
// where x_i collects u_ia u_ib (a <= b), u_ia, and 1, and w holds the
// symmetric multiplicities (1 or 2), -2 and 1.  X is never stored: both
// products regenerate it from U row by row.  The double-centred operator,
// FP64 row means, CholeskyQR2 and Rayleigh-Ritz are shared with the dense
// path (subspace.hpp).  Replaces LowRankProximityMatrix::compute_mds_from_factors
// (cuda/rf_proximity_lowrank.cu:1935-2531).

#include <cmath>
#include <cstring>
#include <vector>

#include "bucket.hpp"
#include "rfx_common.hpp"
#include "subspace.hpp"

namespace {

// ---------------------------------------------------------------------------
// storage formats
// ---------------------------------------------------------------------------
__host__ __device__ inline float bf16_to_f32(uint16_t h) {
  uint32_t b = static_cast<uint32_t>(h) << 16;
  float f;
  memcpy(&f, &b, 4);
  return f;
}
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
  b += 0x7fffu + ((b >> 16) & 1u);                      // round to nearest even
  return static_cast<uint16_t>(b >> 16);
}
// FP8 e4m3 (OCP): bias 7, no infinities, max 448, subnormal step 2^-9.
__host__ __device__ inline float e4m3_to_f32(uint8_t c) {
  const int e = (c >> 3) & 15, m = c & 7;
  float v = e == 0 ? ldexpf(static_cast<float>(m), -9) : ldexpf(1.0f + static_cast<float>(m) * 0.125f, e - 7);
  return (c & 0x80) ? -v : v;
}
__host__ __device__ inline uint8_t f32_to_e4m3(float f) {
  const uint8_t sign = f < 0.0f ? 0x80 : 0x00;
  float a = fabsf(f);
  if (!(a == a)) return 0x7f;
  if (a >= 448.0f) return sign | 0x7e;
  if (a < 0.015625f) {  // subnormal range: multiples of 2^-9, ties to even
    const int m = static_cast<int>(rintf(a * 512.0f));
    return sign | static_cast<uint8_t>(m);  // m == 8 is the smallest normal
  }
  int e;
  frexpf(a, &e);
  e -= 1;  // a = 1.f * 2^e
  int q = static_cast<int>(rintf((ldexpf(a, -e) - 1.0f) * 8.0f));
  if (q == 8) {
    q = 0;
    e += 1;
  }
  int code = ((e + 7) << 3) | q;
  if (code >= 0x7f) code = 0x7e;
  return sign | static_cast<uint8_t>(code);
}

template <int FMT>
struct Store;
template <>
struct Store<RFX_PREC_FP32> {
  using T = float;
  __device__ static float load(const float* U, int64_t idx, const float*, int) { return U[idx]; }
};
template <>
struct Store<RFX_PREC_BF16> {
  using T = uint16_t;
  __device__ static float load(const uint16_t* U, int64_t idx, const float* s, int c) {
    return bf16_to_f32(U[idx]) * s[c];
  }
};
template <>
struct Store<RFX_PREC_INT8> {
  using T = int8_t;
  __device__ static float load(const int8_t* U, int64_t idx, const float* s, int c) {
    return static_cast<float>(U[idx]) * s[c];
  }
};
template <>
struct Store<RFX_PREC_FP8> {
  using T = uint8_t;
  __device__ static float load(const uint8_t* U, int64_t idx, const float* s, int c) {
    return e4m3_to_f32(U[idx]) * s[c];
  }
};

size_t fmt_bytes(int fmt) {
  switch (fmt) {
    case RFX_PREC_FP32: return 4;
    case RFX_PREC_BF16: return 2;
    case RFX_PREC_INT8:
    case RFX_PREC_FP8: return 1;
    default: return 0;
  }
}

// stride is a multiple of r, so each thread stays on one column
__global__ void colabsmax_kernel(const float* __restrict__ U, int64_t n, int r, int64_t stride,
                                 unsigned* __restrict__ amax) {
  const int64_t e0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e0 >= stride) return;
  float m = 0.0f;
  for (int64_t e = e0; e < n * r; e += stride) m = fmaxf(m, fabsf(U[e]));
  if (e0 < n * r) atomicMax(&amax[e0 % r], __float_as_uint(m));
}

__global__ void scales_kernel(const unsigned* __restrict__ amax, int r, int fmt, float* __restrict__ scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= r) return;
  const float a = __uint_as_float(amax[c]);
  float s = 1.0f;
  if (fmt == RFX_PREC_INT8) s = a > 0.0f ? a / 127.0f : 1.0f;
  if (fmt == RFX_PREC_FP8) s = a > 0.0f ? a / 448.0f : 1.0f;
  scale[c] = s;
}

__global__ void quantize_kernel(const float* __restrict__ U, int64_t n, int r, int fmt,
                                const float* __restrict__ scale, void* __restrict__ out) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * r) return;
  const int c = static_cast<int>(e % r);
  const float v = U[e] / scale[c];
  if (fmt == RFX_PREC_FP32) {
    static_cast<float*>(out)[e] = U[e];
  } else if (fmt == RFX_PREC_BF16) {
    static_cast<uint16_t*>(out)[e] = f32_to_bf16(v);
  } else if (fmt == RFX_PREC_INT8) {
    const float q = fminf(127.0f, fmaxf(-127.0f, rintf(v)));
    static_cast<int8_t*>(out)[e] = static_cast<int8_t>(q);
  } else {
    static_cast<uint8_t*>(out)[e] = f32_to_e4m3(v);
  }
}

}  // namespace

#include "factor_mfma.hpp"

namespace {

// ---------------------------------------------------------------------------
// matrix-free C V through the leaf buckets
// ---------------------------------------------------------------------------
// S[b - lo][c] = sum over members of bucket b (ascending) of V[m][c]
// One wavefront per node (four per workgroup): the member ids are
// wavefront-uniform scalar loads, taken kLeafSumUnroll at a time so that their
// rows of V (lanes over columns, up to three 64-column groups per pass) are
// in flight together, the next batch's ids loading meanwhile; each column is
// then summed in member order, as a sequential loop would.  A thread per
// (node, column) walking the members one dependent load after another took
// 92 ms per application at n = 1,000,000, T = 32, L = 144; batches of 8 ids
// 18.9 ms, with the prefetch 15.1 ms; batches of 16 with the prefetch bring
// the whole rank-128 factorization from 0.166 s to 0.139 s.
constexpr int kLeafSumUnroll = 16;
constexpr int kLeafSumTail = 8;
// G column groups of 64 per pass (ceil(L / 64), at most 3).  The members left
// after the full batches (fewer than 16) go in at most two batches of 8 whose
// ids past the leaf repeat its last id and whose rows add +0.0f, which leaves
// every sum exactly as the sequential loop makes it.
template <int G>
__global__ __launch_bounds__(256) void leafsum_kernel(const uint32_t* __restrict__ bstart,
                                                      const uint32_t* __restrict__ bend,
                                                      const uint32_t* __restrict__ members, int64_t lo, int64_t hi,
                                                      const float* __restrict__ V, int L, float* __restrict__ S) {
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int64_t b = lo + static_cast<int64_t>(blockIdx.x) * 4 + wave;
  if (b >= hi) return;
  const int lane = static_cast<int>(threadIdx.x & 63);
  const uint32_t q0 = bstart[b], q1 = bend[b];
  float* out = S + (b - lo) * L;
  for (int c0 = 0; c0 < L; c0 += 64 * G) {
    // lanes past the row read its last column (an in-row address, so the
    // loads need no branch between them) and never store
    int col[G];
    float s[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      col[g] = min(c0 + 64 * g + lane, L - 1);
      s[g] = 0.0f;
    }
    uint32_t q = q0;
    uint32_t mem[kLeafSumUnroll];
    if (q + kLeafSumUnroll <= q1) {
#pragma unroll
      for (int u = 0; u < kLeafSumUnroll; ++u) mem[u] = members[q + u];
    }
    for (; q + kLeafSumUnroll <= q1; q += kLeafSumUnroll) {
      float v[kLeafSumUnroll][G];
#pragma unroll
      for (int u = 0; u < kLeafSumUnroll; ++u) {
        const float* row = V + static_cast<int64_t>(mem[u] >> 1) * L;
#pragma unroll
        for (int g = 0; g < G; ++g) v[u][g] = row[col[g]];
      }
      if (q + 2 * kLeafSumUnroll <= q1) {  // the next batch's ids, while the rows load
#pragma unroll
        for (int u = 0; u < kLeafSumUnroll; ++u) mem[u] = members[q + kLeafSumUnroll + u];
      }
#pragma unroll
      for (int u = 0; u < kLeafSumUnroll; ++u)
#pragma unroll
        for (int g = 0; g < G; ++g) s[g] += v[u][g];
    }
    for (; q < q1; q += kLeafSumTail) {
      const uint32_t cnt = min(static_cast<uint32_t>(kLeafSumTail), q1 - q);
      uint32_t mt[kLeafSumTail];
#pragma unroll
      for (int u = 0; u < kLeafSumTail; ++u) mt[u] = members[q + min(static_cast<uint32_t>(u), cnt - 1)];
      float v[kLeafSumTail][G];
#pragma unroll
      for (int u = 0; u < kLeafSumTail; ++u) {
        const float* row = V + static_cast<int64_t>(mt[u] >> 1) * L;
#pragma unroll
        for (int g = 0; g < G; ++g) v[u][g] = row[col[g]];
      }
#pragma unroll
      for (int u = 0; u < kLeafSumTail; ++u)
#pragma unroll
        for (int g = 0; g < G; ++g) s[g] += static_cast<uint32_t>(u) < cnt ? v[u][g] : 0.0f;
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (c0 + 64 * g + lane < L) out[c0 + 64 * g + lane] = s[g];
  }
}

// Y[i][c] (+)= sum over trees t in [t0, t1) of S[leaf_t(i) - lo][c].  Keys
// are taken kGatherBatch trees at a time and their leaf rows loaded together
// (an off-leaf key reads row 0 and is selected away), then added in tree
// order: a loop of one key and one dependent row load per tree waited twice
// per tree.
constexpr int kGatherBatch = 8;
__global__ void leafgather_kernel(const int32_t* __restrict__ keys, int64_t n, int32_t t0, int32_t t1, int64_t lo,
                                  const float* __restrict__ S, int L, int32_t accumulate, float* __restrict__ Y) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * L) return;
  const int64_t i = e / L;
  const int c = static_cast<int>(e % L);
  float s = accumulate ? Y[e] : 0.0f;
  int32_t t = t0;
  for (; t + kGatherBatch <= t1; t += kGatherBatch) {
    int32_t k[kGatherBatch];
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) k[u] = keys[static_cast<int64_t>(t + u) * n + i];
    float v[kGatherBatch];
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) {
      const int64_t r = k[u] >= 0 ? (static_cast<int64_t>(k[u]) >> 1) - lo : 0;
      v[u] = S[r * L + c];
    }
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) s = k[u] >= 0 ? s + v[u] : s;
  }
  for (; t < t1; ++t) {
    const int32_t k = keys[static_cast<int64_t>(t) * n + i];
    if (k >= 0) s += S[((static_cast<int64_t>(k) >> 1) - lo) * L + c];
  }
  Y[e] = s;
}

__global__ void scale_cols_kernel(float* __restrict__ Y, int64_t count, float a) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < count) Y[e] *= a;
}

struct CvPlan {
  std::vector<int32_t> t0;  // tree chunk boundaries
  int64_t max_nodes = 0;    // largest node range of a chunk
};

CvPlan plan_chunks(const int64_t* node_off, int32_t ntree, int L, size_t s_budget_bytes) {
  CvPlan p;
  const int64_t max_nodes = static_cast<int64_t>(s_budget_bytes / (4 * static_cast<size_t>(L)));
  int32_t t = 0;
  p.t0.push_back(0);
  while (t < ntree) {
    int32_t u = t + 1;
    while (u < ntree && node_off[u + 1] - node_off[t] <= max_nodes) ++u;
    p.max_nodes = std::max(p.max_nodes, node_off[u] - node_off[t]);
    p.t0.push_back(u);
    t = u;
  }
  return p;
}

constexpr size_t kLeafSumBudget = size_t(2) << 30;  // bytes of leaf sums per tree chunk

// ---------------------------------------------------------------------------
// factor MDS operator
// ---------------------------------------------------------------------------
constexpr int kPairChunk = 256;
constexpr int kFacMaxRank = 128;  // factor MDS ranks (the VALU kernels are built for RMAX = 64 and 128)
// Pair chunk of the apply kernel: 128 for FP64 at RMAX = 128 so that its
// LDS (RMAX + 1 factor columns of 256 rows + a chunk of Z) stays within
// the 160 KiB of a workgroup.
template <typename Real, int RMAX>
constexpr int fac_pair_chunk() {
  return (RMAX > 64 && sizeof(Real) == 8) ? 128 : kPairChunk;
}

struct PairTable {
  const uint8_t* a;  // first index (r = "one")
  const uint8_t* b;  // second index
  const float* w;    // multiplicity
  int count;
};

// partial[rb][p][c] = sum over rows of row block rb of x_ip V[i][c]
template <int FMT, typename Real, int K>
__global__ __launch_bounds__(256) void fac_reduce_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                         const float* __restrict__ scale, int64_t n, int r,
                                                         PairTable pt, const Real* __restrict__ V, int64_t rows_per_blk,
                                                         Real* __restrict__ partial) {
  __shared__ float Us[32][kFacMaxRank + 1];
  __shared__ Real Vs[32][K];
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int64_t r0 = static_cast<int64_t>(blockIdx.y) * rows_per_blk;
  const int64_t r1 = min(n, r0 + rows_per_blk);
  const int a = p < pt.count ? pt.a[p] : r, b = p < pt.count ? pt.b[p] : r;
  Real acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = Real(0);
  for (int64_t i0 = r0; i0 < r1; i0 += 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * (r + 1); e += 256) {
      const int q = e / (r + 1), col = e % (r + 1);
      const int64_t i = i0 + q;
      float v = 0.0f;
      if (i < r1) v = col < r ? Store<FMT>::load(U, i * r + col, scale, col) : 1.0f;
      Us[q][col] = v;
    }
    for (int e = threadIdx.x; e < 32 * K; e += 256) {
      const int q = e / K, c = e % K;
      const int64_t i = i0 + q;
      Vs[q][c] = i < r1 ? V[i * K + c] : Real(0);
    }
    __syncthreads();
    const int rows = static_cast<int>(min<int64_t>(32, r1 - i0));
    for (int q = 0; q < rows; ++q) {
      const Real x = static_cast<Real>(Us[q][a]) * static_cast<Real>(Us[q][b]);
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] += x * Vs[q][c];
    }
  }
  if (p < pt.count) {
    Real* out = partial + (static_cast<int64_t>(blockIdx.y) * pt.count + p) * K;
#pragma unroll
    for (int c = 0; c < K; ++c) out[c] = acc[c];
  }
}

// Z[p][c] = w_p * sum over row blocks of partial (FP64 sum, fixed order)
template <typename Real, int K>
__global__ void fac_reduce_partials_kernel(const Real* __restrict__ partial, int nrb, PairTable pt,
                                           Real* __restrict__ Z) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pt.count * K) return;
  double s = 0.0;
  for (int q = 0; q < nrb; ++q) s += static_cast<double>(partial[static_cast<int64_t>(q) * pt.count * K + e]);
  Z[e] = static_cast<Real>(s * static_cast<double>(pt.w[e / K]));
}

// Y[i][c] = x_i^T Z[:, c] - (1 - |u_i|^2)^2 V[i][c]
template <int FMT, typename Real, int K, int RMAX>
__global__ __launch_bounds__(256) void fac_apply_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                        const float* __restrict__ scale, int64_t n, int r,
                                                        PairTable pt, const Real* __restrict__ Z,
                                                        const Real* __restrict__ V, Real* __restrict__ Y) {
  constexpr int PC = fac_pair_chunk<Real, RMAX>();
  __shared__ float Us[RMAX + 1][256];
  __shared__ Real Zs[PC][K];
  __shared__ uint8_t As[PC], Bs[PC];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * 256;
  const int tid = threadIdx.x;
  const int64_t i = i0 + tid;
  for (int e = tid; e < (r + 1) * 256; e += 256) {
    const int col = e / 256, q = e % 256;
    const int64_t ii = i0 + q;
    float v = 0.0f;
    if (ii < n) v = col < r ? Store<FMT>::load(U, ii * r + col, scale, col) : 1.0f;
    Us[col][q] = v;
  }
  Real acc[K];
#pragma unroll
  for (int c = 0; c < K; ++c) acc[c] = Real(0);
  for (int p0 = 0; p0 < pt.count; p0 += PC) {
    const int pc = min(PC, pt.count - p0);
    __syncthreads();
    for (int e = tid; e < pc * K; e += 256) Zs[e / K][e % K] = Z[static_cast<int64_t>(p0) * K + e];
    for (int e = tid; e < pc; e += 256) {
      As[e] = pt.a[p0 + e];
      Bs[e] = pt.b[p0 + e];
    }
    __syncthreads();
    for (int q = 0; q < pc; ++q) {
      const Real x = static_cast<Real>(Us[As[q]][tid]) * static_cast<Real>(Us[Bs[q]][tid]);
#pragma unroll
      for (int c = 0; c < K; ++c) acc[c] += x * Zs[q][c];
    }
  }
  if (i < n) {
    double nrm = 0.0;
    for (int col = 0; col < r; ++col) nrm += static_cast<double>(Us[col][tid]) * Us[col][tid];
    const double d = (1.0 - nrm) * (1.0 - nrm);
#pragma unroll
    for (int c = 0; c < K; ++c)
      Y[i * K + c] = static_cast<Real>(static_cast<double>(acc[c]) - d * static_cast<double>(V[i * K + c]));
  }
}

size_t al(size_t b) { return (b + 255) & ~size_t(255); }

constexpr int kMdsK = 16;
constexpr int kStartDoubles = kFacMaxRank + kFacMaxRank * kMdsK;  // column means + r x 16 projection

int64_t fac_rows_per_blk(int64_t n) {
  int64_t rb = (n + 127) / 128;  // about 128 row blocks
  rb = (rb + 31) / 32 * 32;
  return rb < 32 ? 32 : rb;
}

struct PairHost {
  std::vector<uint8_t> a, b;
  std::vector<float> w;
};

PairHost make_pairs(int r) {
  PairHost h;
  for (int a = 0; a < r; ++a)
    for (int b = a; b < r; ++b) {
      h.a.push_back(static_cast<uint8_t>(a));
      h.b.push_back(static_cast<uint8_t>(b));
      h.w.push_back(a == b ? 1.0f : 2.0f);
    }
  for (int a = 0; a < r; ++a) {
    h.a.push_back(static_cast<uint8_t>(a));
    h.b.push_back(static_cast<uint8_t>(r));
    h.w.push_back(-2.0f);
  }
  h.a.push_back(static_cast<uint8_t>(r));
  h.b.push_back(static_cast<uint8_t>(r));
  h.w.push_back(1.0f);
  return h;
}

template <typename Real>
size_t fac_mds_ws(int64_t n, int r) {
  const int P = r * (r + 1) / 2 + r + 1;
  const int64_t nrb = (n + fac_rows_per_blk(n) - 1) / fac_rows_per_blk(n);
  return al(8 * n) + al(P * 6 + 16) + al(sizeof(Real) * nrb * P * kMdsK) + al(8 * nrb * P) +
         al(sizeof(Real) * P * kMdsK) + al(8 * P) + al(8 * n) + al(8 * kStartDoubles) +
         rfx::SubspaceSolver<Real>::workspace_bytes(n, kMdsK);
}

// Workspace of the VALU operator path (also the FP32 refinement of the mixed path).
template <typename Real>
struct FacWs {
  double* rdev;
  PairTable pt;
  int64_t nrb;
  Real* partial;
  double* partial64;
  Real* Z;
  double* Z64;
  double* ones;
  double* start;  // column means and projection of the start block (kStartDoubles)
  char* solver;
};

template <typename Real>
int fac_ws_setup(hipStream_t st, void* ws, int64_t n, int r, FacWs<Real>& f) {
  const int P = r * (r + 1) / 2 + r + 1;
  const int64_t rpb = fac_rows_per_blk(n);
  f.nrb = (n + rpb - 1) / rpb;
  char* p = static_cast<char*>(ws);
  f.rdev = reinterpret_cast<double*>(p);
  p += al(8 * n);
  uint8_t* pa = reinterpret_cast<uint8_t*>(p);
  uint8_t* pb = pa + P;
  float* pw = reinterpret_cast<float*>(p + ((2 * P + 15) / 16) * 16);
  p += al(P * 6 + 16);
  f.partial = reinterpret_cast<Real*>(p);
  p += al(sizeof(Real) * f.nrb * P * kMdsK);
  f.partial64 = reinterpret_cast<double*>(p);
  p += al(8 * f.nrb * P);
  f.Z = reinterpret_cast<Real*>(p);
  p += al(sizeof(Real) * P * kMdsK);
  f.Z64 = reinterpret_cast<double*>(p);
  p += al(8 * P);
  f.ones = reinterpret_cast<double*>(p);
  p += al(8 * n);
  f.start = reinterpret_cast<double*>(p);
  p += al(8 * kStartDoubles);
  f.solver = p;
  PairHost h = make_pairs(r);
  RFX_HIP_TRY(hipMemcpyAsync(pa, h.a.data(), P, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(pb, h.b.data(), P, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(pw, h.w.data(), 4 * P, hipMemcpyHostToDevice, st));
  std::vector<double> onesh(n, 1.0);
  RFX_HIP_TRY(hipMemcpyAsync(f.ones, onesh.data(), 8 * n, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  f.pt = PairTable{pa, pb, pw, P};
  return RFX_OK;
}

// FP64 row means of E (r_i = (E 1)_i / n) into f.rdev and their mean g.
template <int FMT, typename Real>
int fac_row_means(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, const FacWs<Real>& f,
                  double* g) {
  using ST = typename Store<FMT>::T;
  const ST* U = static_cast<const ST*>(Uq);
  dim3 gr(rfx_ceil_div(f.pt.count, 256), static_cast<unsigned>(f.nrb));
  hipLaunchKernelGGL((fac_reduce_kernel<FMT, double, 1>), gr, dim3(256), 0, st, U, scale, n, r, f.pt, f.ones,
                     fac_rows_per_blk(n), f.partial64);
  hipLaunchKernelGGL((fac_reduce_partials_kernel<double, 1>), dim3(rfx_ceil_div(f.pt.count, 256)), dim3(256), 0, st,
                     f.partial64, static_cast<int>(f.nrb), f.pt, f.Z64);
  if (r <= 64)
    hipLaunchKernelGGL((fac_apply_kernel<FMT, double, 1, 64>), dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, U, scale,
                       n, r, f.pt, f.Z64, f.ones, f.rdev);
  else
    hipLaunchKernelGGL((fac_apply_kernel<FMT, double, 1, kFacMaxRank>), dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st,
                       U, scale, n, r, f.pt, f.Z64, f.ones, f.rdev);
  std::vector<double> rh(n);
  RFX_HIP_TRY(hipMemcpyAsync(rh.data(), f.rdev, 8 * n, hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    rh[i] /= static_cast<double>(n);
    s += rh[i];
  }
  *g = s / static_cast<double>(n);
  RFX_HIP_TRY(hipMemcpyAsync(f.rdev, rh.data(), 8 * n, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  return RFX_OK;
}

// ---------------------------------------------------------------------------
// Start block of the factor MDS.  E_ij = (1 - u_i.u_j)^2 = 1 - 2 u_i.u_j +
// (u_i.u_j)^2, so B = -1/2 J E J = J U U^T J - 1/2 J (U o U)(U o U)^T J:
// where the proximities u_i.u_j are small the first term leads, and the
// leading eigenvectors of B lie close to the leading left singular vectors
// of the column-centred factors J U.  The iteration starts from
// V0 = J U Y, Y the leading 16-dimensional subspace of the r x r Gram
// (J U)^T (J U) (a few block power steps on the host) instead of a random
// block.  Needs r >= 16.
// ---------------------------------------------------------------------------
constexpr int kGramTileRows = 32;

// part[blk][m]: sum over the block's rows of u_a u_b (a <= b, packed upper
// triangle in row order) and of u_a (last r entries), FP64.
template <int FMT, int RMAX>
__global__ __launch_bounds__(256) void fac_gram_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                       const float* __restrict__ scale, int64_t n, int r,
                                                       int64_t rows_per_blk, double* __restrict__ part) {
  __shared__ float us[kGramTileRows][RMAX];
  const int m = r * (r + 1) / 2 + r;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * rows_per_blk;
  const int64_t i1 = min(n, i0 + rows_per_blk);
  constexpr int kPer = (RMAX * (RMAX + 1) / 2 + RMAX + 255) / 256;
  double acc[kPer];
  int pa[kPer], pb[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    acc[q] = 0.0;
    const int e = threadIdx.x + q * 256;
    pa[q] = -1;
    pb[q] = -1;
    if (e < r * (r + 1) / 2) {  // (a, b), a <= b, row order
      int a = 0, rem = e;
      while (rem >= r - a) {
        rem -= r - a;
        ++a;
      }
      pa[q] = a;
      pb[q] = a + rem;
    } else if (e < m) {
      pa[q] = e - r * (r + 1) / 2;
    }
  }
  for (int64_t t0 = i0; t0 < i1; t0 += kGramTileRows) {
    __syncthreads();
    for (int e = threadIdx.x; e < kGramTileRows * r; e += 256) {
      const int row = e / r, col = e % r;
      const int64_t i = t0 + row;
      us[row][col] = i < i1 ? Store<FMT>::load(U, i * r + col, scale, col) : 0.f;
    }
    __syncthreads();
    const int rows = static_cast<int>(min<int64_t>(kGramTileRows, i1 - t0));
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      if (pa[q] < 0) continue;
      double s = 0.0;
      if (pb[q] >= 0)
        for (int row = 0; row < rows; ++row)
          s += static_cast<double>(us[row][pa[q]]) * static_cast<double>(us[row][pb[q]]);
      else
        for (int row = 0; row < rows; ++row) s += static_cast<double>(us[row][pa[q]]);
      acc[q] += s;
    }
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int e = threadIdx.x + q * 256;
    if (e < m) part[static_cast<int64_t>(blockIdx.x) * m + e] = acc[q];
  }
}

__global__ void fac_gram_reduce_kernel(const double* __restrict__ part, int nblk, int m, double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[static_cast<int64_t>(b) * m + e];
  out[e] = s;
}

// V[i][c] = sum_a (u_ia - mu_a) Y[a][c]; st = [mu (kFacMaxRank) | Y (r x 16)]
template <int FMT, typename Real>
__global__ void fac_start_kernel(const typename Store<FMT>::T* __restrict__ U, const float* __restrict__ scale,
                                 int64_t n, int r, const double* __restrict__ st, Real* __restrict__ V) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n * kMdsK) return;
  const int64_t i = e / kMdsK;
  const int c = static_cast<int>(e % kMdsK);
  double s = 0.0;
  for (int a = 0; a < r; ++a)
    s += (static_cast<double>(Store<FMT>::load(U, i * r + a, scale, a)) - st[a]) * st[kFacMaxRank + a * kMdsK + c];
  V[e] = static_cast<Real>(s);
}

// Orthonormal columns (modified Gram-Schmidt) of the r x k row-major Y.
void host_orthonormalize(std::vector<double>& Y, int r, int k) {
  for (int c = 0; c < k; ++c) {
    for (int d = 0; d < c; ++d) {
      double dot = 0.0;
      for (int a = 0; a < r; ++a) dot += Y[a * k + c] * Y[a * k + d];
      for (int a = 0; a < r; ++a) Y[a * k + c] -= dot * Y[a * k + d];
    }
    double nn = 0.0;
    for (int a = 0; a < r; ++a) nn += Y[a * k + c] * Y[a * k + c];
    nn = std::sqrt(nn);
    for (int a = 0; a < r; ++a) Y[a * k + c] = nn > 0 ? Y[a * k + c] / nn : 0.0;
  }
}

// used = false (and V untouched) when the Gram's 16-dimensional leading
// subspace is numerically rank-deficient: the caller then starts at random.
template <int FMT, typename Real>
int fac_start_block(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, const FacWs<Real>& f,
                    Real* V, bool& used) {
  used = false;
  using ST = typename Store<FMT>::T;
  const ST* U = static_cast<const ST*>(Uq);
  const int m = r * (r + 1) / 2 + r;
  const int64_t rpb = fac_rows_per_blk(n);
  if (r <= 64)
    hipLaunchKernelGGL((fac_gram_kernel<FMT, 64>), dim3(static_cast<unsigned>(f.nrb)), dim3(256), 0, st, U, scale, n,
                       r, rpb, f.partial64);
  else
    hipLaunchKernelGGL((fac_gram_kernel<FMT, kFacMaxRank>), dim3(static_cast<unsigned>(f.nrb)), dim3(256), 0, st, U,
                       scale, n, r, rpb, f.partial64);
  hipLaunchKernelGGL(fac_gram_reduce_kernel, dim3(rfx_ceil_div(m, 256)), dim3(256), 0, st, f.partial64,
                     static_cast<int>(f.nrb), m, f.Z64);
  std::vector<double> G(m);
  RFX_HIP_TRY(hipMemcpyAsync(G.data(), f.Z64, 8 * m, hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  // centred Gram C = U^T U - n mu mu^T
  const int tri = r * (r + 1) / 2;
  std::vector<double> mu(r), C(static_cast<size_t>(r) * r);
  for (int a = 0; a < r; ++a) mu[a] = G[tri + a] / static_cast<double>(n);
  for (int a = 0, e = 0; a < r; ++a)
    for (int b = a; b < r; ++b, ++e) C[a * r + b] = C[b * r + a] = G[e] - static_cast<double>(n) * mu[a] * mu[b];
  // leading 16-dimensional subspace of C: block power steps from a fixed start
  std::vector<double> Y(static_cast<size_t>(r) * kMdsK), T(Y.size());
  uint64_t z = 0x243f6a8885a308d3ull;
  for (auto& y : Y) {
    z = z * 6364136223846793005ull + 1442695040888963407ull;
    y = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  }
  host_orthonormalize(Y, r, kMdsK);
  for (int it = 0; it < 8; ++it) {
    for (int a = 0; a < r; ++a)
      for (int c = 0; c < kMdsK; ++c) {
        double s = 0.0;
        for (int b = 0; b < r; ++b) s += C[a * r + b] * Y[b * kMdsK + c];
        T[a * kMdsK + c] = s;
      }
    Y = T;
    host_orthonormalize(Y, r, kMdsK);
  }
  double qmax = 0.0, qmin = 1e300;
  for (int c = 0; c < kMdsK; ++c) {
    double q = 0.0;
    for (int a = 0; a < r; ++a)
      for (int b = 0; b < r; ++b) q += Y[a * kMdsK + c] * C[a * r + b] * Y[b * kMdsK + c];
    qmax = std::max(qmax, q);
    qmin = std::min(qmin, q);
  }
  if (!(qmax > 0.0) || qmin <= 1e-10 * qmax) return RFX_OK;
  std::vector<double> up(kStartDoubles, 0.0);
  for (int a = 0; a < r; ++a) up[a] = mu[a];
  for (size_t e = 0; e < Y.size(); ++e) up[kFacMaxRank + e] = Y[e];
  RFX_HIP_TRY(hipMemcpyAsync(f.start, up.data(), 8 * kStartDoubles, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL((fac_start_kernel<FMT, Real>), dim3(rfx_ceil_div(n * kMdsK, 256)), dim3(256), 0, st, U, scale, n,
                     r, f.start, V);
  RFX_HIP_TRY(hipStreamSynchronize(st));  // the host buffer `up` is local
  used = true;
  return rfx_launch_status();
}

// E V on the VALU (FP32 or FP64 arithmetic) through the Khatri-Rao pairs.
template <int FMT, typename Real>
int fac_valu_apply(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, const FacWs<Real>& f,
                   const Real* V, Real* W) {
  using ST = typename Store<FMT>::T;
  const ST* U = static_cast<const ST*>(Uq);
  dim3 gr(rfx_ceil_div(f.pt.count, 256), static_cast<unsigned>(f.nrb));
  hipLaunchKernelGGL((fac_reduce_kernel<FMT, Real, kMdsK>), gr, dim3(256), 0, st, U, scale, n, r, f.pt, V,
                     fac_rows_per_blk(n), f.partial);
  hipLaunchKernelGGL((fac_reduce_partials_kernel<Real, kMdsK>), dim3(rfx_ceil_div(f.pt.count * kMdsK, 256)),
                     dim3(256), 0, st, f.partial, static_cast<int>(f.nrb), f.pt, f.Z);
  if (r <= 64)
    hipLaunchKernelGGL((fac_apply_kernel<FMT, Real, kMdsK, 64>), dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, U,
                       scale, n, r, f.pt, f.Z, V, W);
  else
    hipLaunchKernelGGL((fac_apply_kernel<FMT, Real, kMdsK, kFacMaxRank>), dim3(rfx_ceil_div(n, 256)), dim3(256), 0,
                       st, U, scale, n, r, f.pt, f.Z, V, W);
  return rfx_launch_status();
}

template <int FMT, typename Real>
int fac_mds_run(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, int max_iter, double tol,
                const FacWs<Real>& f, double* coords, rfx::SubspaceResult& res) {
  double g = 0.0;
  int rc = fac_row_means<FMT, Real>(st, Uq, scale, n, r, f, &g);
  if (rc) return rc;
  rfx::SubspaceSolver<Real> solver(st, n, kMdsK);
  solver.carve(f.solver);
  auto apply = [&](const Real* V, Real* W, int) -> int {
    return fac_valu_apply<FMT, Real>(st, Uq, scale, n, r, f, V, W);
  };
  bool informed = false;
  if (r >= kMdsK && n >= 2 * kMdsK) {
    rc = fac_start_block<FMT, Real>(st, Uq, scale, n, r, f, solver.block_v(), informed);
    if (rc) return rc;
  }
  return solver.run(apply, f.rdev, g, 3, max_iter, tol, 0x5eedull, coords, res, 0.0, informed);
}

template <typename Real>
int fac_mds_dispatch(hipStream_t st, const void* Uq, const float* scale, int fmt, int64_t n, int r, int max_iter,
                     double tol, void* ws, double* coords, rfx::SubspaceResult& res) {
  FacWs<Real> f;
  int rc = fac_ws_setup<Real>(st, ws, n, r, f);
  if (rc) return rc;
  switch (fmt) {
    case RFX_PREC_FP32:
      return fac_mds_run<RFX_PREC_FP32, Real>(st, Uq, scale, n, r, max_iter, tol, f, coords, res);
    case RFX_PREC_BF16:
      return fac_mds_run<RFX_PREC_BF16, Real>(st, Uq, scale, n, r, max_iter, tol, f, coords, res);
    case RFX_PREC_INT8:
      return fac_mds_run<RFX_PREC_INT8, Real>(st, Uq, scale, n, r, max_iter, tol, f, coords, res);
    case RFX_PREC_FP8:
      return fac_mds_run<RFX_PREC_FP8, Real>(st, Uq, scale, n, r, max_iter, tol, f, coords, res);
    default:
      return RFX_ERR_INVALID;
  }
}

// ---------------------------------------------------------------------------
// mixed path: bf16 MFMA operator on the stored factors, then FP32 refinement
// sweeps of the VALU operator on the FP32 factors from the converged block
// ---------------------------------------------------------------------------
int64_t mf_rows_per_chunk(int64_t n) {
  int64_t rpc = (n + 255) / 256;  // about 256 chunks per tile subset
  rpc = (rpc + 31) / 32 * 32;
  return rpc < kMfStage ? kMfStage : rpc;
}

int64_t mf_nchunks(int64_t n) { return (n + mf_rows_per_chunk(n) - 1) / mf_rows_per_chunk(n); }

// workspace of the matrix-core operator: partials P and Pm, S planes, m and s
size_t mf_ws_bytes(int64_t n, int r) {
  const int WT = mf_wt(r), W = 32 * WT;
  const int64_t nch = mf_nchunks(n);
  return al(4 * static_cast<size_t>(nch) * kMfK * mf_ntiles(WT) * 256) +
         al(4 * static_cast<size_t>(nch) * kMfK * (W + 1)) + al(2 * 2 * static_cast<size_t>(kMfK) * 2 * WT * WT * 512) +
         al(4 * static_cast<size_t>(kMfK) * (W + 1));
}

struct MfWs {
  float* P;
  float* Pm;
  uint16_t* Sb;
  float* mv;
};

MfWs mf_carve(char* p, int64_t n, int r) {
  const int WT = mf_wt(r), W = 32 * WT;
  const int64_t nch = mf_nchunks(n);
  MfWs m;
  m.P = reinterpret_cast<float*>(p);
  p += al(4 * static_cast<size_t>(nch) * kMfK * mf_ntiles(WT) * 256);
  m.Pm = reinterpret_cast<float*>(p);
  p += al(4 * static_cast<size_t>(nch) * kMfK * (W + 1));
  m.Sb = reinterpret_cast<uint16_t*>(p);
  p += al(2 * 2 * static_cast<size_t>(kMfK) * 2 * WT * WT * 512);
  m.mv = reinterpret_cast<float*>(p);
  return m;
}

template <int FMT, int WT, bool SPLIT>
int mf_apply_wt(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, const float* V, float* W,
                const MfWs& m) {
  using ST = typename Store<FMT>::T;
  const ST* U = static_cast<const ST*>(Uq);
  const int64_t rpc = mf_rows_per_chunk(n);
  const int64_t nch = mf_nchunks(n);
  constexpr int ntiles = mf_ntiles(WT);
  constexpr int RT = WT == 1 ? 4 : 2;  // row tiles per wave: registers for the B operands
  hipLaunchKernelGGL((fac_mfma_reduce_kernel<FMT, WT, SPLIT>), dim3(static_cast<unsigned>(nch), mf_nsub(WT)),
                     dim3(1024), 0, st, U, scale, n, r, V, rpc, m.P, m.Pm);
  const int64_t fin = static_cast<int64_t>(kMfK) * ntiles * 256 + static_cast<int64_t>(kMfK) * (32 * WT + 1);
  hipLaunchKernelGGL((fac_mfma_finalize_kernel<WT>), dim3(rfx_ceil_div(fin, 256)), dim3(256), 0, st, m.P, m.Pm,
                     static_cast<int>(nch), m.Sb, m.mv);
  hipLaunchKernelGGL((fac_mfma_apply_kernel<FMT, WT, SPLIT, RT>), dim3(rfx_ceil_div(n, 64 * RT)), dim3(256), 0, st, U,
                     scale, n, r, m.Sb, m.mv, V, W);
  return rfx_launch_status();
}

template <int FMT, bool SPLIT = false>
int mf_apply(hipStream_t st, const void* Uq, const float* scale, int64_t n, int r, const float* V, float* W,
             const MfWs& m) {
  switch (mf_wt(r)) {
    case 1: return mf_apply_wt<FMT, 1, SPLIT>(st, Uq, scale, n, r, V, W, m);
    case 2: return mf_apply_wt<FMT, 2, SPLIT>(st, Uq, scale, n, r, V, W, m);
    case 3: return mf_apply_wt<FMT, 3, SPLIT>(st, Uq, scale, n, r, V, W, m);
    case 4: return mf_apply_wt<FMT, 4, SPLIT>(st, Uq, scale, n, r, V, W, m);
    default: return RFX_ERR_UNSUPPORTED;
  }
}

template <int FMT>
int mixed_run(hipStream_t st, const void* Uq, const float* scale, const float* U32, int64_t n, int r, int max_iter,
              double tol, int refine_max, double refine_tol, void* ws, double* coords, rfx::SubspaceResult& res,
              int32_t* iters) {
  FacWs<float> f;
  int rc = fac_ws_setup<float>(st, ws, n, r, f);
  if (rc) return rc;
  const MfWs mws = mf_carve(static_cast<char*>(ws) + fac_mds_ws<float>(n, r), n, r);

  double g = 0.0;
  rc = fac_row_means<FMT, float>(st, Uq, scale, n, r, f, &g);
  if (rc) return rc;
  rfx::SubspaceSolver<float> solver(st, n, kMdsK);
  solver.carve(f.solver);
  auto apply_mf = [&](const float* V, float* W, int) -> int { return mf_apply<FMT>(st, Uq, scale, n, r, V, W, mws); };
  bool informed = false;
  if (r >= kMdsK && n >= 2 * kMdsK) {
    rc = fac_start_block<FMT, float>(st, Uq, scale, n, r, f, solver.block_v(), informed);
    if (rc) return rc;
  }
  rc = solver.run(apply_mf, f.rdev, g, 3, max_iter, tol, 0x5eedull, coords, res, 0.0, informed);
  if (rc) return rc;
  iters[0] = res.iterations;
  iters[1] = 0;
  if (!U32 || refine_max <= 0) return RFX_OK;
  rc = fac_row_means<RFX_PREC_FP32, float>(st, U32, nullptr, n, r, f, &g);
  if (rc) return rc;
  // FP32-accurate matrix-core operator on the FP32 factors (u as bf16 hi + lo)
  auto apply_32 = [&](const float* V, float* W, int) -> int {
    return mf_apply<RFX_PREC_FP32, true>(st, U32, nullptr, n, r, V, W, mws);
  };
  rc = solver.run(apply_32, f.rdev, g, 3, refine_max, refine_tol, 0x5eedull, coords, res, 0.0, true);
  if (rc) return rc;
  iters[1] = res.iterations;
  return RFX_OK;
}

// ---------------------------------------------------------------------------
// P = U U^T as a dense FP64 n x n column-major matrix (the factors' own
// proximity): 64 x 64 output tiles, the two row blocks of U dequantised to
// FP32 in LDS, each thread a 4 x 4 register tile accumulated in FP64 (the
// FP32 products are exact in FP64, so the result is the exact FP64 sum of
// the stored factors' products, in column order).
// ---------------------------------------------------------------------------
constexpr int kRcTile = 64;
constexpr int kRcMaxRank = 128;

template <int FMT>
__global__ __launch_bounds__(256) void reconstruct_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                          const float* __restrict__ scale, int64_t n, int r,
                                                          double* __restrict__ out) {
  __shared__ float Ui[kRcMaxRank][kRcTile + 1];
  __shared__ float Uj[kRcMaxRank][kRcTile + 1];
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kRcTile, j0 = static_cast<int64_t>(blockIdx.y) * kRcTile;
  const int tid = threadIdx.x;
  for (int e = tid; e < kRcTile * r; e += 256) {
    const int row = e / r, col = e % r;  // consecutive threads read one factor row
    const int64_t i = i0 + row, j = j0 + row;
    Ui[col][row] = i < n ? Store<FMT>::load(U, i * r + col, scale, col) : 0.f;
    Uj[col][row] = j < n ? Store<FMT>::load(U, j * r + col, scale, col) : 0.f;
  }
  __syncthreads();
  const int ti = tid & 15, tj = tid >> 4;  // rows ti + 16a, columns tj + 16b
  double acc[4][4] = {};
  for (int c = 0; c < r; ++c) {
    double a[4], b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = Ui[c][ti + 16 * q];
      b[q] = Uj[c][tj + 16 * q];
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * b[y];
  }
#pragma unroll
  for (int y = 0; y < 4; ++y) {
    const int64_t j = j0 + tj + 16 * y;
    if (j >= n) continue;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int64_t i = i0 + ti + 16 * x;
      if (i < n) out[i + j * n] = acc[x][y];
    }
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" int rfx_reconstruct_factors(hipStream_t st, const void* Uq, const float* scales, int32_t format, int64_t n,
                                       int32_t rank, double* out) {
  if (!Uq || !out || n <= 0 || rank < 1 || rank > kRcMaxRank || fmt_bytes(format) == 0) return RFX_ERR_INVALID;
  if (format != RFX_PREC_FP32 && !scales) return RFX_ERR_INVALID;
  const int64_t nb = (n + kRcTile - 1) / kRcTile;
  if (nb > 65535) return RFX_ERR_UNSUPPORTED;
  const dim3 grid(static_cast<unsigned>(nb), static_cast<unsigned>(nb));
  switch (format) {
    case RFX_PREC_FP32:
      hipLaunchKernelGGL(reconstruct_kernel<RFX_PREC_FP32>, grid, dim3(256), 0, st, static_cast<const float*>(Uq),
                         scales, n, rank, out);
      break;
    case RFX_PREC_BF16:
      hipLaunchKernelGGL(reconstruct_kernel<RFX_PREC_BF16>, grid, dim3(256), 0, st, static_cast<const uint16_t*>(Uq),
                         scales, n, rank, out);
      break;
    case RFX_PREC_INT8:
      hipLaunchKernelGGL(reconstruct_kernel<RFX_PREC_INT8>, grid, dim3(256), 0, st, static_cast<const int8_t*>(Uq),
                         scales, n, rank, out);
      break;
    default:
      hipLaunchKernelGGL(reconstruct_kernel<RFX_PREC_FP8>, grid, dim3(256), 0, st, static_cast<const uint8_t*>(Uq),
                         scales, n, rank, out);
      break;
  }
  return rfx_launch_status();
}

extern "C" size_t rfx_lowrank_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes, int32_t rank,
                                              int32_t oversample) {
  if (n <= 0 || ntree <= 0 || total_nodes <= 0 || rank <= 0 || oversample < 0) return 0;
  const int L = rank + oversample;
  const size_t sbytes = std::min(kLeafSumBudget, static_cast<size_t>(4) * L * static_cast<size_t>(total_nodes));
  return bucket_ws_bytes(n, ntree, total_nodes) + al(sbytes + 4 * static_cast<size_t>(L) * 1024 * 1024) +
         rfx::SubspaceSolver<float>::workspace_bytes(n, L);
}

extern "C" int rfx_lowrank_factors(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                   int64_t total_nodes, const int64_t* node_offset_host, int32_t rank,
                                   int32_t oversample, int32_t power_iters, uint64_t seed, void* ws, size_t ws_bytes,
                                   float* U, double* eigvals_host) {
  if (!keys || !node_offset_host || !ws || !U || n < 2 || ntree <= 0 || total_nodes <= 0 ||
      total_nodes > RFX_MAX_TOTAL_NODES || rank < 1 || rank > 128 || oversample < 0 || power_iters < 0)
    return RFX_ERR_INVALID;
  const int L = rank + oversample;
  if (L > n) return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  if (ws_bytes < rfx_lowrank_workspace_bytes(n, ntree, total_nodes, rank, oversample)) return RFX_ERR_WORKSPACE;
  char* p = static_cast<char*>(ws);
  const BucketWs w = bucket_ws_carve(p, n, ntree, total_nodes);
  p += bucket_ws_bytes(n, ntree, total_nodes);
  const size_t sbytes = std::min(kLeafSumBudget, static_cast<size_t>(4) * L * static_cast<size_t>(total_nodes));
  const CvPlan plan = plan_chunks(node_offset_host, ntree, L, sbytes);
  if (static_cast<size_t>(plan.max_nodes) * 4 * L > sbytes + 4 * static_cast<size_t>(L) * 1024 * 1024)
    return RFX_ERR_UNSUPPORTED;  // a single tree larger than the leaf-sum budget
  float* S = reinterpret_cast<float*>(p);
  p += al(sbytes + 4 * static_cast<size_t>(L) * 1024 * 1024);
  rfx::SubspaceSolver<float> solver(st, n, L);
  solver.carve(p);
  float* V = solver.block_v();
  float* Y = solver.block_w();

  int rc = bucket_build(st, keys, n, ntree, total_nodes, w);
  if (rc) return rc;

  auto cv = [&](const float* Vin, float* Yout) -> int {
    for (size_t q = 0; q + 1 < plan.t0.size(); ++q) {
      const int32_t t0 = plan.t0[q], t1 = plan.t0[q + 1];
      const int64_t lo = node_offset_host[t0], hi = node_offset_host[t1];
      const dim3 lsgrid(static_cast<unsigned>(rfx_ceil_div(hi - lo, 4)));
      if (L <= 64)
        hipLaunchKernelGGL(leafsum_kernel<1>, lsgrid, dim3(256), 0, st, w.bstart, w.bend, w.members, lo, hi, Vin, L, S);
      else if (L <= 128)
        hipLaunchKernelGGL(leafsum_kernel<2>, lsgrid, dim3(256), 0, st, w.bstart, w.bend, w.members, lo, hi, Vin, L, S);
      else
        hipLaunchKernelGGL(leafsum_kernel<3>, lsgrid, dim3(256), 0, st, w.bstart, w.bend, w.members, lo, hi, Vin, L, S);
      hipLaunchKernelGGL(leafgather_kernel, dim3(rfx_ceil_div(n * L, 256)), dim3(256), 0, st, keys, n, t0, t1, lo, S,
                         L, static_cast<int32_t>(q > 0), Yout);
    }
    return rfx_launch_status();
  };

  hipLaunchKernelGGL(rfx::ss_init_kernel<float>, dim3(rfx_ceil_div(n * L, 256)), dim3(256), 0, st, V, n, L, seed);
  rc = cv(V, Y);
  if (rc) return rc;
  // scale by 1/T so the block stays O(1) (C/T has unit diagonal)
  hipLaunchKernelGGL(scale_cols_kernel, dim3(rfx_ceil_div(n * L, 256)), dim3(256), 0, st, Y, n * L,
                     1.0f / static_cast<float>(ntree));
  RFX_HIP_TRY(hipMemcpyAsync(V, Y, 4 * n * L, hipMemcpyDeviceToDevice, st));
  rc = solver.orthonormalize(V);
  if (rc) return rc;
  for (int it = 0; it < power_iters; ++it) {
    rc = cv(V, Y);
    if (rc) return rc;
    hipLaunchKernelGGL(scale_cols_kernel, dim3(rfx_ceil_div(n * L, 256)), dim3(256), 0, st, Y, n * L,
                       1.0f / static_cast<float>(ntree));
    RFX_HIP_TRY(hipMemcpyAsync(V, Y, 4 * n * L, hipMemcpyDeviceToDevice, st));
    rc = solver.orthonormalize(V);
    if (rc) return rc;
  }
  // Rayleigh-Ritz: H = Q^T (C/T) Q, U = Q Z_r sqrt(theta_r)
  rc = cv(V, Y);
  if (rc) return rc;
  hipLaunchKernelGGL(scale_cols_kernel, dim3(rfx_ceil_div(n * L, 256)), dim3(256), 0, st, Y, n * L,
                     1.0f / static_cast<float>(ntree));
  std::vector<double> H, wv, Z;
  rc = solver.gram(V, L, Y, L, H);
  if (rc) return rc;
  for (int a = 0; a < L; ++a)
    for (int c = a + 1; c < L; ++c) {
      const double m = 0.5 * (H[a * L + c] + H[c * L + a]);
      H[a * L + c] = H[c * L + a] = m;
    }
  rfx::jacobi_eigh(H, L, wv, Z);
  std::vector<double> M(static_cast<size_t>(L) * rank);
  for (int a = 0; a < L; ++a)
    for (int c = 0; c < rank; ++c) M[a * rank + c] = Z[a * L + c] * std::sqrt(std::max(wv[c], 0.0));
  rc = solver.upload(solver.small_matrix(), M.data(), M.size());
  if (rc) return rc;
  rfx::ss_rotate<float>(st, V, L, solver.small_matrix(), rank, n, U);
  if (eigvals_host)
    for (int c = 0; c < rank; ++c) eigvals_host[c] = wv[c];
  RFX_HIP_TRY(hipStreamSynchronize(st));
  return rfx_launch_status();
}

extern "C" int rfx_quantize_factors(hipStream_t st, const float* U, int64_t n, int32_t rank, int32_t format,
                                    void* Uq, float* scales) {
  if (!U || !Uq || !scales || n <= 0 || rank <= 0 || rank > 128 || fmt_bytes(format) == 0) return RFX_ERR_INVALID;
  unsigned* amax = reinterpret_cast<unsigned*>(scales);  // reuse the scale array as the max buffer
  RFX_HIP_TRY(hipMemsetAsync(amax, 0, 4 * rank, st));
  const int64_t threads = 256LL * 1024;
  const int64_t stride = (threads / rank) * rank;
  hipLaunchKernelGGL(colabsmax_kernel, dim3(rfx_ceil_div(stride, 256)), dim3(256), 0, st, U, n, rank, stride, amax);
  hipLaunchKernelGGL(scales_kernel, dim3(rfx_ceil_div(rank, 64)), dim3(64), 0, st, amax, rank, format, scales);
  hipLaunchKernelGGL(quantize_kernel, dim3(rfx_ceil_div(n * rank, 256)), dim3(256), 0, st, U, n, rank, format,
                     scales, Uq);
  return rfx_launch_status();
}

extern "C" size_t rfx_mds_factors_workspace_bytes(int64_t n, int32_t rank, int32_t precision) {
  if (n <= 0 || rank <= 0) return 0;
  return precision == RFX_PREC_FP64 ? fac_mds_ws<double>(n, rank) : fac_mds_ws<float>(n, rank);
}

extern "C" int rfx_mds_factors(hipStream_t st, const void* Uq, const float* scales, int32_t format, int64_t n,
                               int32_t rank, int32_t precision, int32_t max_iter, double tol, void* ws,
                               size_t ws_bytes, double* coords, double* eigvals_host, int32_t* iters_host) {
  if (!Uq || !scales || !ws || !coords || n < 4 || rank < 1 || rank > kFacMaxRank || fmt_bytes(format) == 0 ||
      max_iter < 1)
    return RFX_ERR_INVALID;
  if (precision != RFX_PREC_FP32 && precision != RFX_PREC_FP64) return RFX_ERR_INVALID;
  if (ws_bytes < rfx_mds_factors_workspace_bytes(n, rank, precision)) return RFX_ERR_WORKSPACE;
  rfx::SubspaceResult res;
  int rc = precision == RFX_PREC_FP64
               ? fac_mds_dispatch<double>(st, Uq, scales, format, n, rank, max_iter, tol, ws, coords, res)
               : fac_mds_dispatch<float>(st, Uq, scales, format, n, rank, max_iter, tol, ws, coords, res);
  if (rc) return rc;
  if (eigvals_host)
    for (int c = 0; c < kMdsK; ++c) eigvals_host[c] = res.eigvals[c];
  if (iters_host) *iters_host = res.iterations;
  if (!(res.eigvals[2] > 1e-10)) return RFX_ERR_INVALID;
  return RFX_OK;
}

extern "C" size_t rfx_mds_factors_mixed_workspace_bytes(int64_t n, int32_t rank) {
  if (n <= 0 || rank <= 0 || rank > kFacMaxRank) return 0;
  return fac_mds_ws<float>(n, rank) + mf_ws_bytes(n, rank);
}

extern "C" int rfx_mds_factors_mixed(hipStream_t st, const void* Uq, const float* scales, int32_t format,
                                     const float* U32, int64_t n, int32_t rank, int32_t max_iter, double tol,
                                     int32_t refine_max, double refine_tol, void* ws, size_t ws_bytes,
                                     double* coords, double* eigvals_host, int32_t* iters_host) {
  if (!Uq || !scales || !ws || !coords || n < 4 || rank < 1 || rank > kFacMaxRank || fmt_bytes(format) == 0 ||
      max_iter < 1 || refine_max < 0)
    return RFX_ERR_INVALID;
  if (ws_bytes < rfx_mds_factors_mixed_workspace_bytes(n, rank)) return RFX_ERR_WORKSPACE;
  rfx::SubspaceResult res;
  int32_t iters[2] = {0, 0};
  int rc;
  switch (format) {
    case RFX_PREC_FP32:
      rc = mixed_run<RFX_PREC_FP32>(st, Uq, scales, U32, n, rank, max_iter, tol, refine_max, refine_tol, ws, coords,
                                    res, iters);
      break;
    case RFX_PREC_BF16:
      rc = mixed_run<RFX_PREC_BF16>(st, Uq, scales, U32, n, rank, max_iter, tol, refine_max, refine_tol, ws, coords,
                                    res, iters);
      break;
    case RFX_PREC_INT8:
      rc = mixed_run<RFX_PREC_INT8>(st, Uq, scales, U32, n, rank, max_iter, tol, refine_max, refine_tol, ws, coords,
                                    res, iters);
      break;
    default:
      rc = mixed_run<RFX_PREC_FP8>(st, Uq, scales, U32, n, rank, max_iter, tol, refine_max, refine_tol, ws, coords,
                                   res, iters);
      break;
  }
  if (rc) return rc;
  if (eigvals_host)
    for (int c = 0; c < kMdsK; ++c) eigvals_host[c] = res.eigvals[c];
  if (iters_host) {
    iters_host[0] = iters[0];
    iters_host[1] = iters[1];
  }
  if (!(res.eigvals[2] > 1e-10)) return RFX_ERR_INVALID;
  return RFX_OK;
}
