// DENSE co-occurrence tile (shared by the one-problem kernel and the batched
// persistent kernel): a workgroup owns a 128 x 128 output tile; the row and
// column keys of 32 trees at a time are staged in LDS after a per-variant
// transform, and each thread compares an 8 x 8 register tile.
#pragma once

#include "gather.hpp"
#include "rfx_common.hpp"

namespace {

constexpr int kTile = 128;
constexpr int kSub = 8;
constexpr int kTreeChunk = 32;
constexpr int kThreads = 256;

template <int MODE>
__device__ __forceinline__ int32_t row_key(int32_t k) {
  if (MODE == kMode_ALL) return k >> 1;
  if (MODE == kMode_OOB_ASYM) return (k >= 0 && !(k & 1)) ? k : -1;
  if (MODE == kMode_OOB_ROW) return (k >= 0 && !(k & 1)) ? (k >> 1) : -1;
  return k;  // OOB_UPPER, PAIR (invalid stays -1)
}
template <int MODE>
__device__ __forceinline__ int32_t col_key(int32_t k) {
  if (MODE == kMode_ALL || MODE == kMode_OOB_ROW) return k < 0 ? -2 : (k >> 1);
  if (MODE == kMode_OOB_ASYM) return (k >= 0 && (k & 1)) ? (k ^ 1) : -2;
  if (MODE == kMode_OOB_UPPER) return k < 0 ? -2 : (k ^ 1);
  return k < 0 ? 0x40000000 : k;  // PAIR
}
template <int MODE>
__device__ __forceinline__ int32_t row_pad() {
  return -1;
}
template <int MODE>
__device__ __forceinline__ int32_t col_pad() {
  return MODE == kMode_PAIR ? 0x40000000 : -2;
}

// Upper-tile decode: linear k -> (bi, bj), bi <= bj, row-major over the
// upper triangle of an nb x nb tile grid.
__device__ __forceinline__ void upper_tile(int64_t k, int64_t nb, int64_t& bi, int64_t& bj) {
  const double b = 2.0 * nb + 1.0;
  int64_t r = static_cast<int64_t>((b - sqrt(b * b - 8.0 * static_cast<double>(k))) * 0.5);
  if (r < 0) r = 0;
  if (r >= nb) r = nb - 1;
  auto start = [nb](int64_t q) { return q * nb - (q * (q - 1)) / 2; };
  while (r > 0 && start(r) > k) --r;
  while (r + 1 < nb && start(r + 1) <= k) ++r;
  bi = r;
  bj = r + (k - start(r));
}

// Workgroups are dispatched round-robin over the 8 XCDs; renumber them so a
// contiguous run of tiles (sharing row keys) lands on one XCD and its L2.
__device__ __forceinline__ int64_t xcd_linear(int64_t b, int64_t nblocks) {
  const int64_t q = nblocks / 8, r = nblocks % 8;
  const int64_t x = b % 8, local = b / 8;
  return x * q + (x < r ? x : r) + local;
}

// One 128 x 128 output tile (bi, bj) of the DENSE compare.  sR / sS are the
// LDS key stages [kTreeChunk][kTile] of the calling kernel.
template <int MODE, int LAYOUT>
__device__ __forceinline__ void dense_tile_body(const int32_t* __restrict__ keys, int64_t n, int32_t ntree,
                                                int64_t bi, int64_t bj, int32_t accumulate,
                                                int32_t* __restrict__ out, int32_t* __restrict__ out2,
                                                int32_t (*sR)[kTile], int32_t (*sS)[kTile]) {
  const int64_t i0 = bi * kTile, j0 = bj * kTile;
  const int tid = threadIdx.x;
  const int fast = tid & 15, slow = tid >> 4;
  // Row-major packed output wants columns contiguous per thread; the
  // column-major full matrix wants rows contiguous.
  const int rb = (LAYOUT == RFX_LAYOUT_UPPER ? slow : fast) * kSub;
  const int cb = (LAYOUT == RFX_LAYOUT_UPPER ? fast : slow) * kSub;

  uint32_t cnt[kSub][kSub];
#pragma unroll
  for (int r = 0; r < kSub; ++r)
#pragma unroll
    for (int c = 0; c < kSub; ++c) cnt[r][c] = 0;

  for (int32_t t0 = 0; t0 < ntree; t0 += kTreeChunk) {
    const int32_t tc = min(kTreeChunk, ntree - t0);
    __syncthreads();
    for (int e = tid; e < kTreeChunk * kTile; e += kThreads) {
      const int tt = e / kTile, q = e % kTile;
      const int64_t i = i0 + q, j = j0 + q;
      const bool tv = tt < tc;
      const int64_t trow = static_cast<int64_t>(t0 + tt) * n;
      sR[tt][q] = (tv && i < n) ? row_key<MODE>(keys[trow + i]) : row_pad<MODE>();
      sS[tt][q] = (tv && j < n) ? col_key<MODE>(keys[trow + j]) : col_pad<MODE>();
    }
    __syncthreads();
    for (int tt = 0; tt < tc; ++tt) {
      int32_t a[kSub], b[kSub];
      const int4* pa = reinterpret_cast<const int4*>(&sR[tt][rb]);
      const int4* pb = reinterpret_cast<const int4*>(&sS[tt][cb]);
      const int4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
      a[0] = a0.x; a[1] = a0.y; a[2] = a0.z; a[3] = a0.w;
      a[4] = a1.x; a[5] = a1.y; a[6] = a1.z; a[7] = a1.w;
      b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w;
      b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
#pragma unroll
      for (int r = 0; r < kSub; ++r)
#pragma unroll
        for (int c = 0; c < kSub; ++c) {
          if (MODE == kMode_PAIR) {
            const uint32_t xr = static_cast<uint32_t>(a[r] ^ b[c]);
            cnt[r][c] += xr < 2u ? ((xr << 16) | 1u) : 0u;
          } else {
            cnt[r][c] += (a[r] == b[c]) ? 1u : 0u;
          }
        }
    }
  }

#pragma unroll
  for (int r = 0; r < kSub; ++r) {
    const int64_t i = i0 + rb + r;
    if (i >= n) continue;
#pragma unroll
    for (int c = 0; c < kSub; ++c) {
      const int64_t j = j0 + cb + c;
      if (j >= n) continue;
      int64_t o;
      if (LAYOUT == RFX_LAYOUT_UPPER) {
        if (j < i) continue;
        o = rfx_upper_index(i, j, n);
      } else {
        o = i + j * n;
      }
      if (MODE == kMode_PAIR) {
        const int32_t va = static_cast<int32_t>(cnt[r][c] & 0xffffu);
        const int32_t vb = static_cast<int32_t>(cnt[r][c] >> 16);
        out[o] = accumulate ? out[o] + va : va;
        out2[o] = accumulate ? out2[o] + vb : vb;
      } else {
        const int32_t v = static_cast<int32_t>(cnt[r][c]);
        out[o] = accumulate ? out[o] + v : v;
      }
    }
  }
}

}  // namespace
