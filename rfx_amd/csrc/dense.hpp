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
        rfx_emit_count(out + o, va, accumulate);
        rfx_emit_count(out2 + o, vb, accumulate);
      } else {
        rfx_emit_count(out + o, static_cast<int32_t>(cnt[r][c]), accumulate);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Packed 16-bit form for ALL and OOB_UPPER, UPPER layout, for forests whose
// trees have at most kDense16MaxNodes nodes: a tree's leaf ids then differ
// modulo 16,384, so a key shrinks to (leaf mod 16,384) << 1 | bag (15 bits)
// with bit 15 free for the pads (row 0xFFFF, column 0xFFFE, which never
// compare equal to anything).  Two trees share a 32-bit word; per word and
// output pair, d = a ^ b, min(d, 1) per 16-bit half counts the non-matching
// slots (v_xor_b32, v_pk_min_u16, v_pk_add_u16): 3 operations for two
// tree-pairs against 2 per tree-pair.  count = slots - non-matches.
constexpr int64_t kDense16MaxNodes = 16384;
// algo flag of count_problem_slice (accum.hip): DENSE in the packed 16-bit
// form, for a caller that has checked every tree's node count
constexpr int32_t kAlgoDense16 = 0x1000;
constexpr int kTreeChunk16 = 64;  // trees per LDS stage (32 words per column)
typedef unsigned short rfx_u16x2 __attribute__((ext_vector_type(2)));

template <int MODE>
__device__ __forceinline__ uint32_t row_key16(int32_t k) {
  if (k < 0) return 0xFFFFu;
  if (MODE == kMode_ALL) return (static_cast<uint32_t>(k) >> 1) & 0x3FFFu;
  return static_cast<uint32_t>(k) & 0x7FFFu;  // OOB_UPPER
}
template <int MODE>
__device__ __forceinline__ uint32_t col_key16(int32_t k) {
  if (k < 0) return 0xFFFEu;
  if (MODE == kMode_ALL) return (static_cast<uint32_t>(k) >> 1) & 0x3FFFu;
  return (static_cast<uint32_t>(k) & 0x7FFFu) ^ 1u;  // OOB_UPPER: same leaf, other bag
}

template <int MODE>
__device__ __forceinline__ void dense16_tile_body(const int32_t* __restrict__ keys, int64_t n, int32_t ntree,
                                                  int64_t bi, int64_t bj, int32_t* __restrict__ out,
                                                  uint32_t (*sR)[kTile], uint32_t (*sS)[kTile]) {
  constexpr int kWords = kTreeChunk16 / 2;
  const int64_t i0 = bi * kTile, j0 = bj * kTile;
  const int tid = threadIdx.x;
  const int rb = (tid >> 4) * kSub, cb = (tid & 15) * kSub;  // UPPER: columns contiguous per thread
  rfx_u16x2 cnt[kSub][kSub];
#pragma unroll
  for (int r = 0; r < kSub; ++r)
#pragma unroll
    for (int c = 0; c < kSub; ++c) cnt[r][c] = rfx_u16x2{0, 0};
  uint32_t slots = 0;  // tree slots compared per output pair, pads included
  const uint32_t ones = 0x00010001u;
  for (int32_t t0 = 0; t0 < ntree; t0 += kTreeChunk16) {
    const int32_t tc = min(kTreeChunk16, ntree - t0);
    const int words = (tc + 1) / 2;
    __syncthreads();
    for (int e = tid; e < kWords * kTile; e += kThreads) {
      const int w = e / kTile, q = e % kTile;
      const int64_t i = i0 + q, j = j0 + q;
      const int32_t ta = t0 + 2 * w, tb = ta + 1;
      const bool va = 2 * w < tc, vb = 2 * w + 1 < tc;
      const uint32_t ra = (va && i < n) ? row_key16<MODE>(keys[static_cast<int64_t>(ta) * n + i]) : 0xFFFFu;
      const uint32_t rb2 = (vb && i < n) ? row_key16<MODE>(keys[static_cast<int64_t>(tb) * n + i]) : 0xFFFFu;
      const uint32_t ca = (va && j < n) ? col_key16<MODE>(keys[static_cast<int64_t>(ta) * n + j]) : 0xFFFEu;
      const uint32_t cb2 = (vb && j < n) ? col_key16<MODE>(keys[static_cast<int64_t>(tb) * n + j]) : 0xFFFEu;
      sR[w][q] = ra | (rb2 << 16);
      sS[w][q] = ca | (cb2 << 16);
    }
    __syncthreads();
    for (int w = 0; w < words; ++w) {
      uint32_t a[kSub], b[kSub];
      const uint4* pa = reinterpret_cast<const uint4*>(&sR[w][rb]);
      const uint4* pb = reinterpret_cast<const uint4*>(&sS[w][cb]);
      const uint4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
      a[0] = a0.x; a[1] = a0.y; a[2] = a0.z; a[3] = a0.w;
      a[4] = a1.x; a[5] = a1.y; a[6] = a1.z; a[7] = a1.w;
      b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w;
      b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
#pragma unroll
      for (int r = 0; r < kSub; ++r)
#pragma unroll
        for (int c = 0; c < kSub; ++c) {
          // min per 16-bit half, written out: the compiler would turn the
          // vector min into two compares and selects
          uint32_t m;
          asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(a[r] ^ b[c]), "v"(ones));
          cnt[r][c] += __builtin_bit_cast(rfx_u16x2, m);
        }
    }
    slots += 2 * static_cast<uint32_t>(words);
  }
#pragma unroll
  for (int r = 0; r < kSub; ++r) {
    const int64_t i = i0 + rb + r;
    if (i >= n) continue;
#pragma unroll
    for (int c = 0; c < kSub; ++c) {
      const int64_t j = j0 + cb + c;
      if (j >= n || j < i) continue;
      const uint32_t miss = static_cast<uint32_t>(cnt[r][c].x) + static_cast<uint32_t>(cnt[r][c].y);
      out[rfx_upper_index(i, j, n)] = static_cast<int32_t>(slots - miss);
    }
  }
}

}  // namespace
