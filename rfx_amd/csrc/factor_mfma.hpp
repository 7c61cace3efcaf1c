// Factor-MDS operator on the matrix cores (bf16 MFMA, FP32 accumulation).
//
// With u~_i = [u_i, +1] and u^_j = [u_j, -1] (both padded with zeros to
// NT*16 columns), u~_i . u^_j = u_i . u_j - 1, so for a block V of 16 columns
//   (E V)_ic = sum_j (1 - u_i.u_j)^2 v_jc - (1 - |u_i|^2)^2 v_ic
//            = u~_i^T Q_c u~_i - (1 - |u_i|^2)^2 v_ic,
//   Q_c = sum_j v_jc u^_j u^_j^T        (NT*16 x NT*16, symmetric).
// Both steps are GEMMs over the factor matrix:
//   reduce  Q_c = (U^ o v_c)^T U^  over row chunks (K = rows), upper tiles
//           only, FP32 partials per chunk summed in FP64 in a fixed order;
//   apply   T_c = U~ Q_c (K = NT*16), then y_ic = sum_b T_c[i][b] u~_ib with a
//           16-lane reduction in registers.
// Operands are bf16: factors stored in BF16 are exact, INT8/FP8 factors are
// dequantised (q * s_c) and rounded to bf16.  The products u * v and the
// entries of Q are carried as unevaluated bf16 pairs hi + lo (two MFMAs), so
// the operator error is set by the stored factors, not by the arithmetic.
// SPLIT (FP32 factors) also carries u as a pair hi + lo (three MFMAs per
// product: hi*hi, lo*hi, hi*lo), an FP32-accurate operator on the matrix
// cores that the refinement sweeps use.
//
// MFMA 16x16x16 bf16 (v_mfma_f32_16x16x16bf16_1k), wave64 register layout:
//   A[i][k]: lane i + 16*(k/4), element k%4
//   B[k][j]: lane j + 16*(k/4), element k%4
//   D[i][j]: lane j + 16*(i/4), element i%4
#pragma once

#include "rfx_common.hpp"

namespace {

typedef short mf_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMfK = 16;          // columns of the MDS block
constexpr int kMfStage = 256;     // rows staged in LDS per step
constexpr int kMfStride = 260;    // padded LDS row stride (bf16 elements) of the column-major stage
constexpr int kMfMaxNT = 5;       // NT*16 >= rank + 1, rank <= 64

__device__ __forceinline__ uint16_t mf_bf16(float f) {
  uint32_t b = __float_as_uint(f);
  if ((b & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  b += 0x7fffu + ((b >> 16) & 1u);
  return static_cast<uint16_t>(b >> 16);
}
__device__ __forceinline__ float mf_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

__device__ __forceinline__ mf_bf16x4 mf_pack(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
  mf_bf16x4 v;
  v[0] = static_cast<short>(a);
  v[1] = static_cast<short>(b);
  v[2] = static_cast<short>(c);
  v[3] = static_cast<short>(d);
  return v;
}

// upper tile index t -> (ta, tb), ta <= tb, row-major over the NT x NT tiles
__device__ __forceinline__ void mf_tile(int t, int NT, int& ta, int& tb) {
  ta = 0;
  while (t >= NT - ta) {
    t -= NT - ta;
    ++ta;
  }
  tb = ta + t;
}

// Partials: P[chunk][c][tile][lane*4 + e] (FP32), one workgroup per row chunk,
// 16 waves: wave c accumulates the upper tiles of Q_c.
template <int FMT, int NT, bool SPLIT>
__global__ __launch_bounds__(1024) void fac_mfma_reduce_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                               const float* __restrict__ scale, int64_t n, int r,
                                                               const float* __restrict__ V,
                                                               int64_t rows_per_chunk, float* __restrict__ P) {
  constexpr int W = NT * 16;
  constexpr int ntiles = NT * (NT + 1) / 2;
  // factor planes: bf16(u), and with SPLIT bf16(u - bf16(u)) as well
  __shared__ uint16_t Us[SPLIT ? 2 : 1][W][kMfStride];
  __shared__ float Vs[kMfK][kMfStride];
  const int tid = threadIdx.x, lane = tid & 63, c = tid >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rows_per_chunk;
  const int64_t r1 = min(n, r0 + rows_per_chunk);
  mf_f32x4 acc[ntiles];
#pragma unroll
  for (int t = 0; t < ntiles; ++t) acc[t] = mf_f32x4{0.f, 0.f, 0.f, 0.f};
  for (int64_t j0 = r0; j0 < r1; j0 += kMfStage) {
    __syncthreads();
    // stage u^ (column-major, bf16) and V (column-major, FP32) of rows [j0, j0 + 256)
    for (int e = tid; e < kMfStage * W; e += 1024) {
      const int row = e / W, col = e % W;  // consecutive threads read one row
      const int64_t j = j0 + row;
      float u = 0.f;
      if (j < r1) {
        if (col < r)
          u = Store<FMT>::load(U, j * r + col, scale, col);
        else if (col == r)
          u = -1.f;
      }
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
    for (int k0 = 0; k0 < rows; k0 += 16) {
      const int jr = k0 + 4 * g;
      const float4 vv = *reinterpret_cast<const float4*>(&Vs[c][jr]);
      // A = u^ o v_c as an unevaluated bf16 sum hi + lo (the FP32 product
      // u * v is exact to 2^-24; hi + lo carries it to about 2^-17)
      mf_bf16x4 Bop[NT], Blo[NT], Ahi[NT], Alo[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint2 w = *reinterpret_cast<const uint2*>(&Us[0][t * 16 + l16][jr]);
        const uint16_t h0 = w.x & 0xffffu, h1 = w.x >> 16, h2 = w.y & 0xffffu, h3 = w.y >> 16;
        Bop[t] = mf_pack(h0, h1, h2, h3);
        float u0 = mf_f32(h0), u1 = mf_f32(h1), u2 = mf_f32(h2), u3 = mf_f32(h3);
        if (SPLIT) {
          const uint2 z = *reinterpret_cast<const uint2*>(&Us[SPLIT ? 1 : 0][t * 16 + l16][jr]);
          const uint16_t z0 = z.x & 0xffffu, z1 = z.x >> 16, z2 = z.y & 0xffffu, z3 = z.y >> 16;
          Blo[t] = mf_pack(z0, z1, z2, z3);
          u0 += mf_f32(z0);
          u1 += mf_f32(z1);
          u2 += mf_f32(z2);
          u3 += mf_f32(z3);
        }
        const float p0 = u0 * vv.x, p1 = u1 * vv.y, p2 = u2 * vv.z, p3 = u3 * vv.w;
        const uint16_t q0 = mf_bf16(p0), q1 = mf_bf16(p1), q2 = mf_bf16(p2), q3 = mf_bf16(p3);
        Ahi[t] = mf_pack(q0, q1, q2, q3);
        Alo[t] = mf_pack(mf_bf16(p0 - mf_f32(q0)), mf_bf16(p1 - mf_f32(q1)), mf_bf16(p2 - mf_f32(q2)),
                         mf_bf16(p3 - mf_f32(q3)));
      }
#pragma unroll
      for (int ta = 0; ta < NT; ++ta)
#pragma unroll
        for (int tb = ta; tb < NT; ++tb) {
          const int t = ta * NT - ta * (ta - 1) / 2 + (tb - ta);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ahi[ta], Bop[tb], acc[t], 0, 0, 0);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Alo[ta], Bop[tb], acc[t], 0, 0, 0);
          if (SPLIT) acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ahi[ta], Blo[tb], acc[t], 0, 0, 0);
        }
    }
  }
  float* out = P + (static_cast<int64_t>(blockIdx.x) * kMfK + c) * ntiles * 256;
#pragma unroll
  for (int t = 0; t < ntiles; ++t)
    *reinterpret_cast<float4*>(&out[t * 256 + lane * 4]) = make_float4(acc[t][0], acc[t][1], acc[t][2], acc[t][3]);
}

// Q_c from the partials, written as the apply kernel's B operands, split into
// bf16 hi + lo planes: Qb[plane][((c*NT + tn)*NT + ka)*256 + lane*4 + e] holds
// Q_c[ka*16 + 4*(lane/16) + e][tn*16 + lane%16].  Each of 64 entries per
// workgroup is summed by 4 threads over interleaved chunks in FP64, then the
// 4 parts are added in a fixed order (deterministic).
constexpr int kMfFinParts = 4;
__global__ __launch_bounds__(256) void fac_mfma_finalize_kernel(const float* __restrict__ P, int nchunks, int NT,
                                                                uint16_t* __restrict__ Qb) {
  __shared__ double part[kMfFinParts][64];
  const int ntiles = NT * (NT + 1) / 2;
  const int64_t total = static_cast<int64_t>(kMfK) * ntiles * 256;
  const int sub = threadIdx.x & 63, pr = threadIdx.x >> 6;
  const int64_t e = static_cast<int64_t>(blockIdx.x) * 64 + sub;
  double s = 0.0;
  if (e < total) {
    const int c = static_cast<int>(e / (ntiles * 256));
    const int64_t off = e - static_cast<int64_t>(c) * ntiles * 256;
    for (int k = pr; k < nchunks; k += kMfFinParts)
      s += static_cast<double>(P[(static_cast<int64_t>(k) * kMfK + c) * ntiles * 256 + off]);
  }
  part[pr][sub] = s;
  __syncthreads();
  if (pr != 0 || e >= total) return;
  s = part[0][sub] + part[1][sub] + part[2][sub] + part[3][sub];
  const int c = static_cast<int>(e / (ntiles * 256));
  const int t = static_cast<int>((e / 256) % ntiles);
  const int q = static_cast<int>(e % 256), lane = q / 4, el = q % 4;
  int ta, tb;
  mf_tile(t, NT, ta, tb);
  const int a = ta * 16 + 4 * (lane >> 4) + el, b = tb * 16 + (lane & 15);
  // Q as bf16 hi + lo: plane 0 holds hi, plane 1 (offset kMfK * NT * NT * 256) lo
  const float qf = static_cast<float>(s);
  const uint16_t h = mf_bf16(qf), lo = mf_bf16(qf - mf_f32(h));
  const int64_t plane = static_cast<int64_t>(kMfK) * NT * NT * 256;
  auto put = [&](int row, int col) {
    const int ka = row >> 4, ia = row & 15, tn = col >> 4;
    const int ln = (ia >> 2) * 16 + (col & 15);
    const int64_t o = ((static_cast<int64_t>(c) * NT + tn) * NT + ka) * 256 + ln * 4 + (ia & 3);
    Qb[o] = h;
    Qb[plane + o] = lo;
  };
  put(a, b);
  if (ta != tb) put(b, a);
}

// Y[i][c] = u~_i^T Q_c u~_i - (1 - |u_i|^2)^2 V[i][c]; 4 waves x 64 rows.
//
// Computed as T^T = Q_c U~^T per 16-row tile (Q_c is stored exactly
// symmetric, so the finalize kernel's B-operand layout of a Q tile is also
// its A-operand layout): with A = Q tile (tn, ka) and B = U~^T tile (ka, rt),
// D[j][i] = T[i][tn*16 + j] lands in lane i + 16*(j/4), element j%4 -- the
// same lane and element that hold u~[i][tn*16 + j] in the B operand of k-block
// tn, so the quadratic form's epilogue multiplies registers by registers and
// sums over the 4 lanes of a row with two shuffles.  Q tiles are read straight
// from global memory (a few hundred KB shared by every workgroup, L2-resident)
// in operand layout; no LDS, so occupancy is set by registers alone.
template <int FMT, int NT, bool SPLIT>
__global__ __launch_bounds__(256) void fac_mfma_apply_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                              const float* __restrict__ scale, int64_t n, int r,
                                                              const uint16_t* __restrict__ Qb,
                                                              const float* __restrict__ V, float* __restrict__ Y) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int64_t w0 = static_cast<int64_t>(blockIdx.x) * 256 + wave * 64;
  // B operands: lane (l16, g) holds u~[row][ka*16 + 4g + e] of row w0 + rt*16 + l16
  mf_bf16x4 Bu[4][NT], Bl[SPLIT ? 4 : 1][NT];
  float nrm[4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    const int64_t i = w0 + rt * 16 + l16;
    float s2 = 0.f;
#pragma unroll
    for (int ka = 0; ka < NT; ++ka) {
      uint16_t h[4], lo[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = ka * 16 + 4 * g + e;
        float u = 0.f;
        if (i < n) {
          if (col < r)
            u = Store<FMT>::load(U, i * r + col, scale, col);
          else if (col == r)
            u = 1.f;
        }
        h[e] = mf_bf16(u);
        lo[e] = SPLIT ? mf_bf16(u - mf_f32(h[e])) : static_cast<uint16_t>(0);
        const float uv = SPLIT ? mf_f32(h[e]) + mf_f32(lo[e]) : mf_f32(h[e]);
        if (col < r) s2 += uv * uv;
      }
      Bu[rt][ka] = mf_pack(h[0], h[1], h[2], h[3]);
      if (SPLIT) Bl[SPLIT ? rt : 0][ka] = mf_pack(lo[0], lo[1], lo[2], lo[3]);
    }
    s2 += __shfl_xor(s2, 16, 64);
    s2 += __shfl_xor(s2, 32, 64);
    nrm[rt] = s2;
  }
  float yv[4][4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int q = 0; q < 4; ++q) yv[rt][q] = 0.f;
  constexpr int qn = NT * NT * 256;
  const int64_t plane = static_cast<int64_t>(kMfK) * qn;
  for (int c = 0; c < kMfK; ++c) {
    const uint16_t* Qh = Qb + static_cast<int64_t>(c) * qn;
    float ys[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int tn = 0; tn < NT; ++tn) {
      mf_bf16x4 Ah[NT], Al[NT];
#pragma unroll
      for (int ka = 0; ka < NT; ++ka) {
        const int64_t o = static_cast<int64_t>((tn * NT + ka) * 64 + lane) * 4;
        const uint2 w = *reinterpret_cast<const uint2*>(Qh + o);
        const uint2 z = *reinterpret_cast<const uint2*>(Qh + plane + o);
        Ah[ka] = mf_pack(w.x & 0xffffu, w.x >> 16, w.y & 0xffffu, w.y >> 16);
        Al[ka] = mf_pack(z.x & 0xffffu, z.x >> 16, z.y & 0xffffu, z.y >> 16);
      }
      mf_f32x4 d[4];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) d[rt] = mf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ka = 0; ka < NT; ++ka)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          d[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah[ka], Bu[rt][ka], d[rt], 0, 0, 0);
          d[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Al[ka], Bu[rt][ka], d[rt], 0, 0, 0);
          if (SPLIT)
            d[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Ah[ka], Bl[SPLIT ? rt : 0][ka], d[rt], 0, 0, 0);
        }
      // D[j][i]: lane l16 + 16g holds T[i][tn*16 + 4g + e], i = l16
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const mf_bf16x4 bu = Bu[rt][tn];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          float uv = mf_f32(static_cast<uint16_t>(bu[e]));
          if (SPLIT) uv += mf_f32(static_cast<uint16_t>(Bl[SPLIT ? rt : 0][tn][e]));
          ys[rt] += d[rt][e] * uv;
        }
      }
    }
    // the row's 4 lanes l16 + 16g hold its partial sums; lane g = c % 4 keeps column c
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      float v = ys[rt];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (c == 4 * q + g) yv[rt][q] = v;
    }
  }
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
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
