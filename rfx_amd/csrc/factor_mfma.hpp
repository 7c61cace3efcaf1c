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
template <int FMT, int NT, bool SPLIT>
__global__ __launch_bounds__(256) void fac_mfma_apply_kernel(const typename Store<FMT>::T* __restrict__ U,
                                                              const float* __restrict__ scale, int64_t n, int r,
                                                              const uint16_t* __restrict__ Qb,
                                                              const float* __restrict__ V, float* __restrict__ Y) {
  constexpr int W = NT * 16;
  constexpr int kRowStride = W + 4;
  __shared__ uint16_t Ut[SPLIT ? 2 : 1][256][kRowStride];
  __shared__ uint16_t Qs[2][NT * NT * 256];
  __shared__ float nrm[256];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, l16 = lane & 15;
  const int64_t i0 = static_cast<int64_t>(blockIdx.x) * 256;
  for (int e = tid; e < 256 * W; e += 256) {
    const int row = e / W, col = e % W;
    const int64_t i = i0 + row;
    float u = 0.f;
    if (i < n) {
      if (col < r)
        u = Store<FMT>::load(U, i * r + col, scale, col);
      else if (col == r)
        u = 1.f;
    }
    const uint16_t h = mf_bf16(u);
    Ut[0][row][col] = h;
    if (SPLIT) Ut[SPLIT ? 1 : 0][row][col] = mf_bf16(u - mf_f32(h));
  }
  __syncthreads();
  auto ut = [&](int row, int col) -> float {
    float v = mf_f32(Ut[0][row][col]);
    if (SPLIT) v += mf_f32(Ut[SPLIT ? 1 : 0][row][col]);
    return v;
  };
  {
    float s = 0.f;
    for (int col = 0; col < r; ++col) {
      const float u = ut(tid, col);
      s += u * u;
    }
    nrm[tid] = s;
  }
  // A operands of the wave's 4 row tiles: A[i][k] = u~[row][ka*16 + k]
  mf_bf16x4 Aop[4][NT], Alo[SPLIT ? 4 : 1][NT];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ka = 0; ka < NT; ++ka) {
      const uint2 w = *reinterpret_cast<const uint2*>(&Ut[0][wave * 64 + rt * 16 + l16][ka * 16 + 4 * g]);
      Aop[rt][ka] = mf_pack(w.x & 0xffffu, w.x >> 16, w.y & 0xffffu, w.y >> 16);
      if (SPLIT) {
        const uint2 z =
            *reinterpret_cast<const uint2*>(&Ut[SPLIT ? 1 : 0][wave * 64 + rt * 16 + l16][ka * 16 + 4 * g]);
        Alo[SPLIT ? rt : 0][ka] = mf_pack(z.x & 0xffffu, z.x >> 16, z.y & 0xffffu, z.y >> 16);
      }
    }
  float yv[4][4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int e = 0; e < 4; ++e) yv[rt][e] = 0.f;
  constexpr int qn = NT * NT * 256;
  for (int c = 0; c < kMfK; ++c) {
    __syncthreads();
    for (int e = tid; e < qn / 4; e += 256) {
      reinterpret_cast<uint2*>(Qs[0])[e] = reinterpret_cast<const uint2*>(Qb + static_cast<int64_t>(c) * qn)[e];
      reinterpret_cast<uint2*>(Qs[1])[e] =
          reinterpret_cast<const uint2*>(Qb + static_cast<int64_t>(kMfK) * qn + static_cast<int64_t>(c) * qn)[e];
    }
    __syncthreads();
    mf_bf16x4 Bhi[NT][NT], Blo[NT][NT];
#pragma unroll
    for (int tn = 0; tn < NT; ++tn)
#pragma unroll
      for (int ka = 0; ka < NT; ++ka) {
        const uint2 w = *reinterpret_cast<const uint2*>(&Qs[0][((tn * NT + ka) * 64 + lane) * 4]);
        Bhi[tn][ka] = mf_pack(w.x & 0xffffu, w.x >> 16, w.y & 0xffffu, w.y >> 16);
        const uint2 z = *reinterpret_cast<const uint2*>(&Qs[1][((tn * NT + ka) * 64 + lane) * 4]);
        Blo[tn][ka] = mf_pack(z.x & 0xffffu, z.x >> 16, z.y & 0xffffu, z.y >> 16);
      }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int tn = 0; tn < NT; ++tn) {
        mf_f32x4 d = mf_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ka = 0; ka < NT; ++ka) {
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Aop[rt][ka], Bhi[tn][ka], d, 0, 0, 0);
          d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Aop[rt][ka], Blo[tn][ka], d, 0, 0, 0);
          if (SPLIT) d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Alo[SPLIT ? rt : 0][ka], Bhi[tn][ka], d, 0, 0, 0);
        }
        // D[i][b]: row 4g + e of the tile, column tn*16 + l16
#pragma unroll
        for (int e = 0; e < 4; ++e) s[e] += d[e] * ut(wave * 64 + rt * 16 + 4 * g + e, tn * 16 + l16);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float v = s[e];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        if (l16 == c) yv[rt][e] = v;
      }
    }
  }
  // lane 16g + c holds rows 4g + e of every row tile for column c = l16
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int row = wave * 64 + rt * 16 + 4 * g + e;
      const int64_t i = i0 + row;
      if (i < n) {
        const float d = 1.f - nrm[row];
        Y[i * kMfK + l16] = yv[rt][e] - d * d * V[i * kMfK + l16];
      }
    }
}

}  // namespace
