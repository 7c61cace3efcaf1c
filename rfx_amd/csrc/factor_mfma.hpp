// Factor-MDS operator on the matrix cores (gfx950 bf16 MFMA, 16 x 16 x 32,
// FP32 accumulation), for ranks up to 128.
//
// For a block V of 16 columns and E_ij = (1 - u_i.u_j)^2 (i != j), E_ii = 0:
//   (E V)_ic = sum_j v_jc (1 - u_i.u_j)^2 - (1 - |u_i|^2)^2 v_ic
//            = u_i^T S_c u_i - 2 u_i . m_c + s_c - (1 - |u_i|^2)^2 v_ic,
//   S_c = sum_j v_jc u_j u_j^T (W x W, symmetric), m_c = sum_j v_jc u_j,
//   s_c = sum_j v_jc,
// with the factors padded with zero columns to W = 32 * WT >= rank.
//   reduce    S_c = (U o v_c)^T U over row chunks (K = rows): upper 16 x 16
//             tiles, split over blockIdx.y in subsets of kMfSubTiles so a
//             wave's accumulators fit its registers; rows ta*16.. of m_c come
//             from the same A operands against a B operand of ones, in the
//             subset that holds diagonal tile (ta, ta); s_c from the staged V
//             (subset 0).  FP32 partials per chunk.
//   finalize  FP64 sums of the partials over the chunks in a fixed order;
//             S_c written as the apply kernel's A operands, split into bf16
//             hi + lo planes; m_c and s_c kept in FP32.
//   apply     T_c = U S_c per 16-row tile (K = W), then y_ic = sum_b
//             T_c[i][b] u_ib in registers, minus 2 u_i . m_c, plus s_c.
// Operands are bf16: factors stored in BF16 are exact, INT8/FP8 factors are
// dequantised (q * s_c) and rounded to bf16.  The products u * v and the
// entries of S are carried as unevaluated bf16 pairs hi + lo (two MFMAs), so
// the operator error is set by the stored factors, not by the arithmetic.
// SPLIT (FP32 factors) also carries u as a pair hi + lo (three MFMAs per
// product: hi*hi, lo*hi, hi*lo), an FP32-accurate operator on the matrix
// cores that the refinement sweeps use.
//
// MFMA 16x16x32 bf16 (v_mfma_f32_16x16x32_bf16), wave64 registers: lane
// l16 + 16g holds 8 elements of A row l16 (B column l16) and
// D[4g + e][l16] in element e of its accumulator.  The product does not
// depend on which K index a (lane group, element) slot holds as long as A and
// B agree, so each kernel picks its own K order:
//   reduce   K slot (g, e) = row 8g + e of the 32 staged rows (16 contiguous
//            bytes of the column-major stage);
//   apply    K slot (g, e) = column 16 (e / 4) + 4g + e % 4 of the 32-column
//            block, so the element of u_i that multiplies D[4g + e'][i] of
//            output tile tn sits in the same lane, element (tn % 2) * 4 + e'.
// (Both orders were checked against a host product of the raw lane data.)
#pragma once

#include "rfx_common.hpp"

namespace {

typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfK = 16;         // columns of the MDS block
constexpr int kMfStage = 256;    // rows staged in LDS per step
constexpr int kMfStride = 264;   // padded LDS row stride (bf16 elements) of the column-major stage
constexpr int kMfMaxWT = 4;      // W = 32 * WT >= rank, rank <= 128

__host__ __device__ constexpr int mf_ntiles(int WT) { return (2 * WT) * (2 * WT + 1) / 2; }
// 16 x 16 tiles of S_c accumulated by one wave: a divisor of the tile count
// (3, 10, 21, 36 for WT = 1..4), so every slot of every subset is a tile,
// small enough that the accumulators fit the 128 registers of a wave at
// 1024 threads
__host__ __device__ constexpr int mf_sub_tiles(int WT) { return WT == 1 ? 3 : WT == 2 ? 5 : WT == 3 ? 7 : 9; }
inline int mf_wt(int r) { return (r + 31) / 32; }
inline int mf_nsub(int WT) { return mf_ntiles(WT) / mf_sub_tiles(WT); }

__device__ __forceinline__ uint16_t mf_bf16(float f) {
  uint32_t b = __float_as_uint(f);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  b += 0x7fffu + ((b >> 16) & 1u);
  return static_cast<uint16_t>(b >> 16);
}
__device__ __forceinline__ float mf_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ mf_bf16x8 mf_from_u4(uint4 w) { return __builtin_bit_cast(mf_bf16x8, w); }
__device__ __forceinline__ uint16_t mf_elem(uint4 w, int e) {
  const uint32_t d = e < 2 ? w.x : e < 4 ? w.y : e < 6 ? w.z : w.w;
  return static_cast<uint16_t>((e & 1) ? d >> 16 : d & 0xffffu);
}
__device__ __forceinline__ uint4 mf_pack8(const uint16_t (&h)[8]) {
  return make_uint4(h[0] | (static_cast<uint32_t>(h[1]) << 16), h[2] | (static_cast<uint32_t>(h[3]) << 16),
                    h[4] | (static_cast<uint32_t>(h[5]) << 16), h[6] | (static_cast<uint32_t>(h[7]) << 16));
}

__device__ __forceinline__ mf_f32x4 mf_mma(uint4 a, uint4 b, mf_f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(mf_from_u4(a), mf_from_u4(b), c, 0, 0, 0);
}

// diagonal tile (q, q) in the row-major upper order of NT x NT tiles
__host__ __device__ constexpr int mf_diag_tile(int q, int NT) { return q * NT - q * (q - 1) / 2; }
// the first row whose diagonal tile is at or after tile t0, and how many
// diagonal tiles lie in [t0, t0 + ST)
__host__ __device__ constexpr int mf_first_diag(int NT, int t0) {
  int q = 0;
  while (q < NT && mf_diag_tile(q, NT) < t0) ++q;
  return q;
}
__host__ __device__ constexpr int mf_ndiag(int NT, int t0, int ST) {
  int k = 0;
  for (int q = 0; q < NT; ++q) k += mf_diag_tile(q, NT) >= t0 && mf_diag_tile(q, NT) < t0 + ST;
  return k;
}

// upper tile index t -> (ta, tb), ta <= tb, row-major over the NT x NT tiles;
// a fixed-count loop, so a constant t folds to constants after unrolling
template <int NT>
__host__ __device__ __forceinline__ void mf_tile(int t, int& ta, int& tb) {
  int a = 0;
#pragma unroll
  for (int q = 1; q < NT; ++q) a += mf_diag_tile(q, NT) <= t;
  ta = a;
  tb = a + (t - mf_diag_tile(a, NT));
}

// Partials of one row chunk: P[chunk][c][tile][lane * 4 + e] (tile = upper
// 16 x 16 tile of S_c, accumulator layout) and Pm[chunk][c][0..W) = m_c,
// Pm[chunk][c][W] = s_c.  Workgroup (chunk, subset): 16 waves, wave c
// accumulates the subset's tiles of S_c and the m_c rows of the subset's
// diagonal tiles; subset 0 also s_c.
template <int FMT, int WT, bool SPLIT, int SUB>
__device__ __forceinline__ void fac_mfma_reduce_subset(uint16_t (&Us)[SPLIT ? 2 : 1][WT * 32][kMfStride],
                                                       float (&Vs)[kMfK][kMfStride],
                                                       const typename Store<FMT>::T* __restrict__ U,
                                                       const float* __restrict__ scale, int64_t n, int r,
                                                       const float* __restrict__ V, int64_t rows_per_chunk,
                                                       float* __restrict__ P, float* __restrict__ Pm) {
  constexpr int W = WT * 32;
  constexpr int NT = 2 * WT;
  constexpr int ntiles = mf_ntiles(WT);
  constexpr int ST = mf_sub_tiles(WT);
  static_assert(mf_ntiles(WT) % ST == 0, "subsets must tile the upper triangle");
  constexpr int sub = SUB;
  constexpr int t0 = SUB * ST;  // compile-time tiles: the (ta, tb) walk folds away
  const int tid = threadIdx.x, lane = tid & 63, c = tid >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_chunk;
  const int64_t r1 = min(n, r0 + rows_per_chunk);
  // m_c rows of the subset's diagonal tiles (rows D0 .. D0 + ND - 1)
  constexpr int D0 = mf_first_diag(NT, t0), ND = mf_ndiag(NT, t0, ST);
  static_assert(ND >= 1, "every subset holds a diagonal tile");
  mf_f32x4 acc[ST], accm[ND];
#pragma unroll
  for (int t = 0; t < ST; ++t) acc[t] = mf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < ND; ++t) accm[t] = mf_f32x4{0.f, 0.f, 0.f, 0.f};
  float s0 = 0.f;
  // B operand of ones in column 0: accm[ta] column 0 collects sum_j A[a][j]
  uint4 ones = make_uint4(0u, 0u, 0u, 0u);
  if (l16 == 0) ones = make_uint4(0x3f803f80u, 0x3f803f80u, 0x3f803f80u, 0x3f803f80u);
  for (int64_t j0 = r0; j0 < r1; j0 += kMfStage) {
    __syncthreads();
    // stage u (column-major, bf16 planes) and V (column-major, FP32) of rows
    // [j0, j0 + 256); a short unroll, as the accumulators stay live across it
#pragma unroll 4
    for (int e = tid; e < kMfStage * W; e += 1024) {
      const int row = e / W, col = e % W;  // consecutive threads read one row
      const int64_t j = j0 + row;
      float u = 0.f;
      if (j < r1 && col < r) u = Store<FMT>::load(U, j * r + col, scale, col);
      const uint16_t h = mf_bf16(u);
      Us[0][col][row] = h;
      if (SPLIT) Us[SPLIT ? 1 : 0][col][row] = mf_bf16(u - mf_f32(h));
    }
    for (int e = tid; e < kMfStage * kMfK; e += 1024) {
      const int row = e / kMfK, cc = e % kMfK;
      const int64_t j = j0 + row;
      Vs[cc][row] = j < r1 ? V[j * kMfK + cc] : 0.f;
    }
    __syncthreads();
    const int rows = static_cast<int>(min<int64_t>(kMfStage, r1 - j0));
    if (sub == 0)
      for (int q = lane; q < rows; q += 64) s0 += Vs[c][q];
    for (int k0 = 0; k0 < rows; k0 += 32) {  // rows past `rows` are staged as zeros
      const int jr = k0 + 8 * g;
      const float4 va = *reinterpret_cast<const float4*>(&Vs[c][jr]);
      const float4 vb = *reinterpret_cast<const float4*>(&Vs[c][jr + 4]);
      const float vv[8] = {va.x, va.y, va.z, va.w, vb.x, vb.y, vb.z, vb.w};
      int cur_ta = -1;
      uint4 Ahi = make_uint4(0u, 0u, 0u, 0u), Alo = Ahi;
#pragma unroll
      for (int s = 0; s < ST; ++s) {
        const int t = t0 + s;
        int ta, tb;
        mf_tile<NT>(t, ta, tb);
        if (ta != cur_ta) {
          // A = u o v_c as bf16 hi + lo (the FP32 product carried to about 2^-17)
          const uint4 w = *reinterpret_cast<const uint4*>(&Us[0][ta * 16 + l16][jr]);
          uint4 z = make_uint4(0u, 0u, 0u, 0u);
          if (SPLIT) z = *reinterpret_cast<const uint4*>(&Us[SPLIT ? 1 : 0][ta * 16 + l16][jr]);
          uint16_t hi[8], lo[8];
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            float u = mf_f32(mf_elem(w, e));
            if (SPLIT) u += mf_f32(mf_elem(z, e));
            const float p = u * vv[e];
            hi[e] = mf_bf16(p);
            lo[e] = mf_bf16(p - mf_f32(hi[e]));
          }
          Ahi = mf_pack8(hi);
          Alo = mf_pack8(lo);
          cur_ta = ta;
          if (tb == ta) {  // the diagonal tile starts row ta: its m_c rows
#pragma unroll
            for (int q = 0; q < ND; ++q)
              if (q == ta - D0) {
                accm[q] = mf_mma(Ahi, ones, accm[q]);
                accm[q] = mf_mma(Alo, ones, accm[q]);
              }
          }
        }
        const uint4 B = *reinterpret_cast<const uint4*>(&Us[0][tb * 16 + l16][jr]);
        acc[s] = mf_mma(Ahi, B, acc[s]);
        acc[s] = mf_mma(Alo, B, acc[s]);
        if (SPLIT) {
          const uint4 Bl = *reinterpret_cast<const uint4*>(&Us[SPLIT ? 1 : 0][tb * 16 + l16][jr]);
          acc[s] = mf_mma(Ahi, Bl, acc[s]);
        }
      }
    }
  }
  const int64_t base = (static_cast<int64_t>(blockIdx.x) * kMfK + c) * ntiles;
#pragma unroll
  for (int s = 0; s < ST; ++s) {
    const int t = t0 + s;
    *reinterpret_cast<float4*>(&P[(base + t) * 256 + lane * 4]) = make_float4(acc[s][0], acc[s][1], acc[s][2], acc[s][3]);
  }
  float* pm = Pm + (static_cast<int64_t>(blockIdx.x) * kMfK + c) * (W + 1);
  // column 0 of the ones product: lane 16g holds rows 4g + e of tile ta
  if (l16 == 0)
#pragma unroll
    for (int q = 0; q < ND; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) pm[(D0 + q) * 16 + 4 * g + e] = accm[q][e];
  if (sub == 0) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) s0 += __shfl_xor(s0, d, 64);
    if (lane == 0) pm[W] = s0;
  }
}

template <int FMT, int WT, bool SPLIT>
__global__ __launch_bounds__(1024) void fac_mfma_reduce_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                               const float* __restrict__ scale, int64_t n, int r,
                                                               const float* __restrict__ V, int64_t rows_per_chunk,
                                                               float* __restrict__ P, float* __restrict__ Pm) {
  // factor planes: bf16(u), and with SPLIT bf16(u - bf16(u)) as well
  __shared__ uint16_t Us[SPLIT ? 2 : 1][WT * 32][kMfStride];
  __shared__ float Vs[kMfK][kMfStride];
  constexpr int NS = mf_ntiles(WT) / mf_sub_tiles(WT);
  static_assert(NS <= 4, "subset dispatch covers four subsets");
  if (blockIdx.y == 0) {
    fac_mfma_reduce_subset<FMT, WT, SPLIT, 0>(Us, Vs, U, scale, n, r, V, rows_per_chunk, P, Pm);
  } else if constexpr (NS > 1) {
    if (blockIdx.y == 1) {
      fac_mfma_reduce_subset<FMT, WT, SPLIT, 1>(Us, Vs, U, scale, n, r, V, rows_per_chunk, P, Pm);
    } else if constexpr (NS > 2) {
      if (blockIdx.y == 2) {
        fac_mfma_reduce_subset<FMT, WT, SPLIT, 2>(Us, Vs, U, scale, n, r, V, rows_per_chunk, P, Pm);
      } else if constexpr (NS > 3) {
        fac_mfma_reduce_subset<FMT, WT, SPLIT, 3>(Us, Vs, U, scale, n, r, V, rows_per_chunk, P, Pm);
      }
    }
  }
}

// S_c from the partials as the apply kernel's A operands, split into bf16
// hi + lo planes: Sb[plane][(((c * NT + tn) * WT + ka) * 64 + lane) * 8 + e]
// holds S_c[tn * 16 + lane % 16][ka * 32 + 16 (e / 4) + 4 (lane / 16) + e % 4];
// plus mv[c][0..W) = m_c and mv[c][W] = s_c in FP32.  One thread per partial
// element: the FP64 sum over the chunks in order, then the value and, off
// the diagonal tiles, its mirror.
template <int WT>
__global__ __launch_bounds__(256) void fac_mfma_finalize_kernel(const float* __restrict__ P,
                                                                const float* __restrict__ Pm, int nchunks,
                                                                uint16_t* __restrict__ Sb, float* __restrict__ mv) {
  constexpr int W = WT * 32;
  constexpr int NT = 2 * WT;
  constexpr int ntiles = mf_ntiles(WT);
  const int64_t total = static_cast<int64_t>(kMfK) * ntiles * 256;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= total) {
    const int64_t f = e - total;  // the m_c, s_c entries
    if (f >= static_cast<int64_t>(kMfK) * (W + 1)) return;
    double s = 0.0;
    for (int k = 0; k < nchunks; ++k) s += static_cast<double>(Pm[static_cast<int64_t>(k) * kMfK * (W + 1) + f]);
    mv[f] = static_cast<float>(s);
    return;
  }
  const int c = static_cast<int>(e / (ntiles * 256));
  const int64_t off = e - static_cast<int64_t>(c) * ntiles * 256;
  double s = 0.0;
  for (int k = 0; k < nchunks; ++k) s += static_cast<double>(P[(static_cast<int64_t>(k) * kMfK + c) * ntiles * 256 + off]);
  const int t = static_cast<int>(off / 256);
  const int q = static_cast<int>(off % 256), lane = q / 4, el = q % 4;
  int ta, tb;
  mf_tile<NT>(t, ta, tb);
  const int a = ta * 16 + 4 * (lane >> 4) + el, b = tb * 16 + (lane & 15);  // accumulator D[i][j] -> S[a][b]
  const float sf = static_cast<float>(s);
  const uint16_t h = mf_bf16(sf), lo = mf_bf16(sf - mf_f32(h));
  const int64_t plane = static_cast<int64_t>(kMfK) * NT * WT * 512;
  auto put = [&](int row, int col) {
    const int tn = row >> 4, l16 = row & 15, ka = col >> 5, w = col & 31;
    const int ee = (w >> 4) * 4 + (w & 3), gg = (w & 15) >> 2;
    const int64_t o = ((((static_cast<int64_t>(c) * NT + tn) * WT + ka) * 64) + l16 + 16 * gg) * 8 + ee;
    Sb[o] = h;
    Sb[plane + o] = lo;
  };
  put(a, b);
  if (ta != tb) put(b, a);
}

// Y[i][c] = u_i^T S_c u_i - 2 u_i . m_c + s_c - (1 - |u_i|^2)^2 V[i][c];
// 4 waves x RT row tiles of 16.  T^T = S_c U^T per row tile: A = S tile
// (tn, ka) from Sb, B = U^T block (ka, rows); D[j][i] = T[i][tn * 16 + j] lands
// in lane i + 16 (j / 4), element j % 4, beside the u_i element that multiplies
// it (the apply's K order), so the quadratic form's epilogue multiplies
// registers by registers and sums the 4 lanes of a row with two shuffles.
// S tiles are read straight from global memory (shared by every workgroup,
// L2-resident); no LDS.
template <int FMT, int WT, bool SPLIT, int RT>
__global__ __launch_bounds__(256) void fac_mfma_apply_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                              const float* __restrict__ scale, int64_t n, int r,
                                                              const uint16_t* __restrict__ Sb,
                                                              const float* __restrict__ mv,
                                                              const float* __restrict__ V, float* __restrict__ Y) {
  constexpr int W = WT * 32;
  constexpr int NT = 2 * WT;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int64_t w0 = (static_cast<int64_t>(blockIdx.x) * 4 + wave) * (RT * 16);
  // B operands: lane (l16, g), element e of block ka holds u[row][ka*32 + 16(e/4) + 4g + e%4]
  uint4 Bu[RT][WT], Bl[SPLIT ? RT : 1][SPLIT ? WT : 1];
  float nrm[RT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const int64_t i = w0 + rt * 16 + l16;
    float s2 = 0.f;
#pragma unroll
    for (int ka = 0; ka < WT; ++ka) {
      uint16_t h[8], lo[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int col = ka * 32 + 16 * (e >> 2) + 4 * g + (e & 3);
        float u = 0.f;
        if (i < n && col < r) u = Store<FMT>::load(U, i * r + col, scale, col);
        h[e] = mf_bf16(u);
        lo[e] = SPLIT ? mf_bf16(u - mf_f32(h[e])) : static_cast<uint16_t>(0);
        const float uv = SPLIT ? mf_f32(h[e]) + mf_f32(lo[e]) : mf_f32(h[e]);
        s2 += uv * uv;
      }
      Bu[rt][ka] = mf_pack8(h);
      if (SPLIT) Bl[SPLIT ? rt : 0][SPLIT ? ka : 0] = mf_pack8(lo);
    }
    s2 += __shfl_xor(s2, 16, 64);
    s2 += __shfl_xor(s2, 32, 64);
    nrm[rt] = s2;
  }
  float yv[RT][4];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int q = 0; q < 4; ++q) yv[rt][q] = 0.f;
  constexpr int64_t sn = static_cast<int64_t>(NT) * WT * 512;
  const int64_t plane = static_cast<int64_t>(kMfK) * sn;
#pragma unroll 1
  for (int c = 0; c < kMfK; ++c) {
    const uint16_t* Sh = Sb + static_cast<int64_t>(c) * sn;
    const float* mc = mv + static_cast<int64_t>(c) * (W + 1);
    float ys[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) ys[rt] = 0.f;
    // linear term: u_i . m_c over the lane's elements
#pragma unroll
    for (int ka = 0; ka < WT; ++ka) {
      float m8[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) m8[e] = mc[ka * 32 + 16 * (e >> 2) + 4 * g + (e & 3)];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          float uv = mf_f32(mf_elem(Bu[rt][ka], e));
          if (SPLIT) uv += mf_f32(mf_elem(Bl[SPLIT ? rt : 0][SPLIT ? ka : 0], e));
          ys[rt] -= 2.f * uv * m8[e];
        }
    }
    // one output tile at a time (the loop is kept rolled so the compiler does
    // not hoist every tile's S loads); the epilogue's u elements are picked
    // from the B registers with selects
#pragma unroll 1
    for (int tn = 0; tn < NT; ++tn) {
      uint4 Ah[WT], Al[WT];
#pragma unroll
      for (int ka = 0; ka < WT; ++ka) {
        const int64_t o = (static_cast<int64_t>(tn * WT + ka) * 64 + lane) * 8;
        Ah[ka] = *reinterpret_cast<const uint4*>(Sh + o);
        Al[ka] = *reinterpret_cast<const uint4*>(Sh + plane + o);
      }
      mf_f32x4 d[RT];
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) d[rt] = mf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ka = 0; ka < WT; ++ka)
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          d[rt] = mf_mma(Ah[ka], Bu[rt][ka], d[rt]);
          d[rt] = mf_mma(Al[ka], Bu[rt][ka], d[rt]);
          if (SPLIT) d[rt] = mf_mma(Ah[ka], Bl[SPLIT ? rt : 0][SPLIT ? ka : 0], d[rt]);
        }
      // D[j][i]: lane l16 + 16g holds T[i][tn*16 + 4g + e'], i = l16; the
      // matching u_i element is element (tn % 2) * 4 + e' of block tn / 2
      const int kb = tn >> 1;
      const bool odd = (tn & 1) != 0;
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        uint4 bs = Bu[rt][0], bl = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
        for (int ka = 1; ka < WT; ++ka)
          if (ka == kb) bs = Bu[rt][ka];
        if (SPLIT) {
          bl = Bl[SPLIT ? rt : 0][0];
#pragma unroll
          for (int ka = 1; ka < WT; ++ka)
            if (ka == kb) bl = Bl[SPLIT ? rt : 0][SPLIT ? ka : 0];
        }
        const uint32_t h01 = odd ? bs.z : bs.x, h23 = odd ? bs.w : bs.y;
        const uint32_t l01 = odd ? bl.z : bl.x, l23 = odd ? bl.w : bl.y;
        const uint32_t hw[2] = {h01, h23}, lw[2] = {l01, l23};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t hd = hw[e >> 1], ld = lw[e >> 1];
          float uv = mf_f32(static_cast<uint16_t>((e & 1) ? hd >> 16 : hd & 0xffffu));
          if (SPLIT) uv += mf_f32(static_cast<uint16_t>((e & 1) ? ld >> 16 : ld & 0xffffu));
          ys[rt] += d[rt][e] * uv;
        }
      }
    }
    // the row's 4 lanes l16 + 16g hold its partial sums; lane g = c % 4 keeps column c
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      float v = ys[rt];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      v += mc[W];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c == 4 * q + g) yv[rt][q] = v;
    }
  }
#pragma unroll
  for (int rt = 0; rt < RT; ++rt) {
    const int64_t i = w0 + rt * 16 + l16;
    if (i >= n) continue;
    const float dd = 1.f - nrm[rt];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 4 * q + g;
      Y[i * kMfK + c] = yv[rt][q] - dd * dd * V[i * kMfK + c];
    }
  }
}

}  // namespace
