// Per-line gather of the GATHER co-occurrence algorithm, shared by the
// one-problem kernels and the batched persistent kernel.
//
// A workgroup owns one output line (row i of UPPER, column i of FULL) chunk
// [lo, hi) and counts it in LDS: every tree's member run of the line's leaf
// (the samples that follow i in the leaf, or the whole leaf for FULL),
// restricted to [lo, hi) by two binary searches, is flattened across the 64
// lanes of a wavefront in aligned 16-byte blocks of the member array (prefix
// sum of block counts, each lane finds the run of its block by a 6-step
// search over the prefix) so that every lane issues independent loads
// regardless of how ragged the runs are; matches are added with LDS atomics
// and the chunk is written (or added) once.
#pragma once

#include "rfx_common.hpp"

namespace {

constexpr int kMode_ALL = RFX_PROX_ALL;
constexpr int kMode_OOB_ASYM = RFX_PROX_OOB_ASYM;
constexpr int kMode_OOB_UPPER = RFX_PROX_OOB_UPPER;
constexpr int kMode_PAIR = 3;
constexpr int kMode_OOB_ROW = RFX_PROX_OOB_ROW;

#ifndef RFX_GATHER_CHUNK
#define RFX_GATHER_CHUNK 8192
#endif
#ifndef RFX_GATHER_LINE_MAX_N
#define RFX_GATHER_LINE_MAX_N 32768
#endif
constexpr int kGatherChunk = RFX_GATHER_CHUNK;  // LDS words per workgroup
// lines up to this length use the line kernels, longer ones the gallop kernel
constexpr int64_t kGatherLineMaxN = RFX_GATHER_LINE_MAX_N;

// Rows are spread over the 8 XCDs (each with its own L2) in groups of
// kXcdGroup consecutive rows: the dispatcher sends workgroup b to XCD b % 8,
// so workgroups b = x + 8 * local on XCD x take the rows of groups
// x, x + 8, x + 16, ...  Neighbouring rows, which share pe cache lines, then
// run on one XCD, and the long rows of the upper triangle are still spread
// evenly.  The grid is padded to a multiple of 8 * kXcdGroup rows.
constexpr int64_t kXcdGroup = 64;

inline int64_t xcd_rows_padded(int64_t n) { return (n + 8 * kXcdGroup - 1) / (8 * kXcdGroup) * (8 * kXcdGroup); }

__device__ __forceinline__ int64_t xcd_row(int64_t b) {
  const int64_t x = b & 7, local = b >> 3;
  return (8 * (local / kXcdGroup) + x) * kXcdGroup + local % kXcdGroup;
}

// first position in members[lo, hi) whose sample index is >= key; M is
// uint32_t, or uint16_t for the 16-bit copy kept when n <= kMembers16MaxN
template <class M>
__device__ __forceinline__ uint32_t member_lb(const M* __restrict__ members, uint32_t lo, uint32_t hi, uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if ((static_cast<uint32_t>(members[mid]) >> 1) < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The wavefront's 64 member runs (lane l: run [s, s + len), its sample's
// in-bag bit) flattened across the lanes in aligned 16-byte blocks of the
// member array, and each member's count added to the LDS counters
// acc[column - lo].  All lanes of the wavefront must call.  A lane's run
// [s, s + len) touches the
// blocks floor(s / P) .. floor((s + len - 1) / P) (P = 16 / sizeof(M)
// members per block); the union of all lanes' blocks is spread over the
// lanes by the same prefix-sum owner search, so the search and its
// shuffles are paid once per block, each block is one aligned 16-byte load,
// and the members of a block outside the run are masked off.  Blocks may
// reach up to P - 1 members past the end of the member array: the array's
// allocation is padded to 256 bytes.
#ifndef RFX_GATHER_PIECE_UNROLL
#define RFX_GATHER_PIECE_UNROLL 2
#endif
template <int MODE, int LAYOUT, class M>
__device__ __forceinline__ void flat_visit_blocks(uint32_t s, uint32_t len, uint32_t bag, int lane,
                                                  const M* __restrict__ members, uint32_t* acc, int64_t lo,
                                                  int64_t line) {
  constexpr uint32_t P = 16 / sizeof(M);
  constexpr int U = RFX_GATHER_PIECE_UNROLL;
  const uint32_t e = s + len;
  const uint32_t nb = len ? (e - 1) / P - s / P + 1 : 0u;
  uint32_t pre = nb;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(pre, d, 64);
    if (lane >= d) pre += y;
  }
  const uint32_t excl = pre - nb;
  const uint32_t total = __shfl(pre, 63, 64);
  for (uint32_t g0 = 0; g0 < total; g0 += 64 * U) {
    uint4 w[U];
    uint32_t blo[U], bhi[U], bb[U], bo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = g0 + u * 64 + lane;
      int o = 0;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t v = __shfl(pre, o + d - 1, 64);
        if (v <= g) o += d;
      }
      o = min(o, 63);
      const uint32_t so = __shfl(s, o, 64);
      const uint32_t eo = __shfl(e, o, 64);
      const uint32_t xo = __shfl(excl, o, 64);
      bo[u] = __shfl(bag, o, 64);
      const bool ok = g < total;
      bb[u] = (so / P + (g - xo)) * P;  // first member index of the block
      blo[u] = ok ? (so > bb[u] ? so - bb[u] : 0u) : P;
      bhi[u] = ok ? min(P, eo - bb[u]) : 0u;
      w[u] = ok ? *reinterpret_cast<const uint4*>(members + bb[u]) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t wv[4] = {w[u].x, w[u].y, w[u].z, w[u].w};
      // members blo .. bhi - 1 of the block are in the run (P <= 8)
      const uint32_t vm = ((1u << bhi[u]) - 1u) & ~((1u << blo[u]) - 1u);
#pragma unroll
      for (uint32_t k = 0; k < P; ++k) {
        if (!((vm >> k) & 1u)) continue;
        const uint32_t m = sizeof(M) == 2 ? ((wv[k / 2] >> (16 * (k % 2))) & 0xffffu) : wv[k];
        const int64_t q = m >> 1;
        const uint32_t bq = m & 1u;
        if (LAYOUT == RFX_LAYOUT_FULL && MODE != kMode_OOB_ROW && q == line) continue;
        uint32_t inc = 0;
        if (MODE == kMode_ALL) inc = 1u;
        if (MODE == kMode_OOB_UPPER) inc = bq != bo[u] ? 1u : 0u;
        if (MODE == kMode_OOB_ASYM || MODE == kMode_OOB_ROW) inc = bq == 0u ? 1u : 0u;
        if (MODE == kMode_PAIR) inc = 1u | (bq != bo[u] ? 0x10000u : 0u);
        if (inc) atomicAdd(&acc[q - lo], inc);
      }
    }
  }
}

// The visit form the kernels use.
template <int MODE, int LAYOUT, class M>
__device__ __forceinline__ void line_visits(uint32_t s, uint32_t len, uint32_t bag, int lane,
                                            const M* __restrict__ members, uint32_t* acc, int64_t lo, int64_t line) {
  flat_visit_blocks<MODE, LAYOUT>(s, len, bag, lane, members, acc, lo, line);
}

// First position p in [cur, end) whose member column is >= key: a galloping
// search from cur (the runs are sorted by column), cost O(log(p - cur)).
__device__ __forceinline__ uint32_t gallop_lb(const uint32_t* __restrict__ members, uint32_t cur, uint32_t end,
                                              uint32_t key) {
  if (cur >= end || (members[cur] >> 1) >= key) return cur;
  uint32_t lo = cur, step = 1, hi;
  for (;;) {  // invariant: column of members[lo] < key
    const uint32_t probe = lo + step;
    if (probe >= end) {
      hi = end;
      break;
    }
    if ((members[probe] >> 1) >= key) {
      hi = probe;
      break;
    }
    lo = probe;
    step <<= 1;
  }
  return member_lb(members, lo + 1, hi, key);
}

// `chunk` LDS counters per workgroup of THREADS threads (THREADS / 64
// wavefronts, each taking 64 trees at a time).
template <int MODE, int LAYOUT, class M, int THREADS = 256>
__device__ __forceinline__ void gather_line_body(const uint2* __restrict__ pe, const uint32_t* __restrict__ bstart,
                                                 const uint32_t* __restrict__ mbucket, int64_t n, int32_t ntree,
                                                 const M* __restrict__ members, int32_t accumulate,
                                                 int32_t* __restrict__ out, int32_t* __restrict__ out2, int64_t line,
                                                 int64_t chunk, uint32_t* acc, uint32_t& diag_valid,
                                                 int64_t chunk_len = kGatherChunk) {
  // UPPER rows are chunked from the diagonal: chunk y of row i covers
  // [i + y*CH, i + (y+1)*CH); FULL columns in aligned chunks.
  const int64_t lo = (LAYOUT == RFX_LAYOUT_UPPER ? line : 0) + chunk * chunk_len;
  if (lo >= n) return;
  const int64_t hi = min(n, lo + chunk_len);
  const int len_chunk = static_cast<int>(hi - lo);
  for (int q = threadIdx.x; q < len_chunk; q += THREADS) acc[q] = 0u;
  if (threadIdx.x == 0) diag_valid = 0u;
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t valid = 0;
  for (int32_t tb = wave * 64; tb < ntree; tb += THREADS) {
    const int32_t t = tb + lane;
    uint32_t s = 0, len = 0, bag = 0;
    if (t < ntree) {
      const uint2 q = pe[static_cast<int64_t>(t) * n + line];
      const uint32_t end = q.y & 0x7fffffffu;
      if (end > 0) {
        ++valid;
        bag = q.y >> 31;
        uint32_t b0, b1;
        if (LAYOUT == RFX_LAYOUT_UPPER) {
          b0 = q.x + 1;
          if (lo > line) b0 = member_lb(members, b0, end, static_cast<uint32_t>(lo));
        } else {
          const uint32_t bs = bstart[mbucket[q.x]];
          b0 = lo > 0 ? member_lb(members, bs, end, static_cast<uint32_t>(lo)) : bs;
        }
        b1 = hi < n ? member_lb(members, b0, end, static_cast<uint32_t>(hi)) : end;
        s = b0;
        len = b1 > b0 ? b1 - b0 : 0u;
      }
    }
    if (MODE == kMode_OOB_ASYM && !bag) len = 0;  // column j of P needs j in-bag
    line_visits<MODE, LAYOUT>(s, len, bag, lane, members, acc, lo, line);
  }
  if (MODE == kMode_ALL || MODE == kMode_PAIR) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) valid += __shfl_down(valid, d, 64);
    if (lane == 0 && valid) atomicAdd(&diag_valid, valid);
  }
  __syncthreads();
  if ((MODE == kMode_ALL || MODE == kMode_PAIR) && threadIdx.x == 0 && line >= lo && line < hi)
    acc[line - lo] = diag_valid;
  __syncthreads();

  const int64_t base = LAYOUT == RFX_LAYOUT_UPPER ? rfx_upper_index(line, lo, n) : lo + line * n;
  for (int q = threadIdx.x; q < len_chunk; q += THREADS) {
    const uint32_t v = acc[q];
    if (MODE == kMode_PAIR) {
      const int32_t va = static_cast<int32_t>(v & 0xffffu), vb = static_cast<int32_t>(v >> 16);
      rfx_emit_count(out + base + q, va, accumulate);
      rfx_emit_count(out2 + base + q, vb, accumulate);
    } else {
      rfx_emit_count(out + base + q, static_cast<int32_t>(v), accumulate);
    }
  }
}


// Long line in chunks of `chunk` LDS counters, in order, for a
// workgroup of THREADS threads: thread x holds the cursors of trees
// g*THREADS + x (g < G, so T <= THREADS G) into their member runs.  For every
// chunk a lane finds where each of its runs crosses the chunk end by a
// galloping search from its cursor, and the wavefront then visits the 64 run
// pieces flattened across its lanes -- a leaf of 10,000 members costs its
// wavefront no more per member than a leaf of 2.  No binary search from the
// run start is ever repeated: cursors only move forward, and every pe entry
// of the line is read once.
template <int MODE, int LAYOUT, int G, class M, int THREADS = 256>
__device__ __forceinline__ void gather_gallop_body(const uint2* __restrict__ pe, const uint32_t* __restrict__ bstart,
                                                   const uint32_t* __restrict__ mbucket, int64_t n, int32_t ntree,
                                                   const M* __restrict__ members, int32_t accumulate,
                                                   int32_t* __restrict__ out, int32_t* __restrict__ out2,
                                                   int64_t line, uint32_t* acc, uint32_t& diag_valid,
                                                   int64_t chunk = kGatherChunk) {
  const int tid = threadIdx.x, lane = tid & 63;
  uint32_t cur[G], end[G];
  uint32_t bagm = 0, valid = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t t = g * THREADS + tid;
    cur[g] = 0;
    end[g] = 0;
    if (t < ntree) {
      const uint2 q = pe[static_cast<int64_t>(t) * n + line];
      const uint32_t e = q.y & 0x7fffffffu;
      if (e > 0) {
        ++valid;
        const uint32_t bag = q.y >> 31;
        bagm |= bag << g;
        cur[g] = LAYOUT == RFX_LAYOUT_UPPER ? q.x + 1 : bstart[mbucket[q.x]];
        end[g] = (MODE == kMode_OOB_ASYM && !bag) ? cur[g] : e;
      }
    }
  }
  if (tid == 0) diag_valid = 0u;
  __syncthreads();
  if (MODE == kMode_ALL || MODE == kMode_PAIR) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) valid += __shfl_down(valid, d, 64);
    if (lane == 0 && valid) atomicAdd(&diag_valid, valid);
  }
  const int64_t first = LAYOUT == RFX_LAYOUT_UPPER ? line : 0;
  for (int64_t lo = first; lo < n; lo += chunk) {
    const int64_t hi = min(n, lo + chunk);
    const int len_chunk = static_cast<int>(hi - lo);
    __syncthreads();
    for (int q = tid; q < len_chunk; q += THREADS) acc[q] = 0u;
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t e2 = hi < n ? gallop_lb(members, cur[g], end[g], static_cast<uint32_t>(hi)) : end[g];
      const uint32_t s0 = cur[g];
      cur[g] = e2;
      line_visits<MODE, LAYOUT>(s0, e2 - s0, (bagm >> g) & 1u, lane, members, acc, lo, line);
    }
    __syncthreads();
    if ((MODE == kMode_ALL || MODE == kMode_PAIR) && tid == 0 && line >= lo && line < hi) acc[line - lo] = diag_valid;
    __syncthreads();
    const int64_t base = LAYOUT == RFX_LAYOUT_UPPER ? rfx_upper_index(line, lo, n) : lo + line * n;
    for (int q = tid; q < len_chunk; q += THREADS) {
      const uint32_t v = acc[q];
      if (MODE == kMode_PAIR) {
        const int32_t va = static_cast<int32_t>(v & 0xffffu), vb = static_cast<int32_t>(v >> 16);
        rfx_emit_count(out + base + q, va, accumulate);
        rfx_emit_count(out2 + base + q, vb, accumulate);
      } else {
        rfx_emit_count(out + base + q, static_cast<int32_t>(v), accumulate);
      }
    }
  }
}

}  // namespace
