// A2 + A3: same-leaf co-occurrence counts over all trees, bit-exact int32.
//
// Input is the key array of rfx_leaf_assign: keys[t][i] = (global node id << 1)
// | in-bag, or -1 when sample i ends on a non-terminal node of tree t.
// Three algorithms compute the same integers (GATHER is described at its
// kernel below):
//
// DENSE  (tile compare).  A workgroup owns a 128 x 128 tile of the output;
//        the row and column keys of 32 trees at a time are staged in LDS
//        after a per-variant transform, and each thread compares an 8 x 8
//        register tile (v_cmp_eq + v_addc per tree-pair), so the count
//        matrix is written exactly once, with no atomics at all.  Cost is
//        T*n^2 compares: VALU-bound, independent of the leaf structure.
//
// BUCKET (leaf buckets).  A radix sort groups the (tree, sample) slots by
//        terminal node (histogram, scan, scatter), then every thread walks
//        the bucket of one slot and adds 1 with an integer atomic for each
//        partner of larger sample index.  Cost is sum over trees and leaves
//        of |leaf|^2 plus one pass over the output, so for fully grown trees
//        on large n (expected count per pair << 1) it is far below T*n^2.
//
// The per-variant transform maps keys to a row key R and column key S with
// R_i == S_j exactly when the pair counts, so one compare kernel serves
//   ALL        R = k >> 1 (invalid -1),            S = k >> 1 (invalid -2)
//   OOB_UPPER  R = k,                              S = k ^ 1  (invalid -2)
//   OOB_ASYM   R = k if out-of-bag else -1,        S = k ^ 1 if in-bag else -2
// and the fused PAIR mode packs ALL (low 16 bits) and OOB_UPPER (high 16
// bits) counters in one register: x = R ^ S, ALL += x < 2, OOB += x == 1.
//
// Reference semantics: src/rf_proximity.cpp:179-230 (OOB_ASYM, accumulated
// per tree), cuda/rf_proximity_upper_triangle.cu:238-248 (OOB_UPPER fold).

#include <algorithm>
#include <mutex>

#include "rfx_common.hpp"
#include "bucket.hpp"
#include "gather.hpp"
#include "dense.hpp"

// The workspace build records read by RFX_ALGO_PREBUILT calls (bucket.hpp):
// a small table, most recent builds kept, guarded for concurrent callers.
namespace {
constexpr int kBuildRecords = 64;
std::mutex g_build_mu;
const void* g_build_key[kBuildRecords] = {};
BucketBuildRecord g_build_rec[kBuildRecords];
int g_build_next = 0;
}  // namespace

void note_bucket_build(const void* ws, const BucketBuildRecord& r) {
  std::lock_guard<std::mutex> lock(g_build_mu);
  for (int k = 0; k < kBuildRecords; ++k)
    if (g_build_key[k] == ws) {
      g_build_rec[k] = r;
      return;
    }
  g_build_key[g_build_next] = ws;
  g_build_rec[g_build_next] = r;
  g_build_next = (g_build_next + 1) % kBuildRecords;
}

bool last_bucket_build(const void* ws, BucketBuildRecord* r) {
  std::lock_guard<std::mutex> lock(g_build_mu);
  for (int k = 0; k < kBuildRecords; ++k)
    if (g_build_key[k] == ws) {
      *r = g_build_rec[k];
      return true;
    }
  return false;
}

namespace {


template <int MODE, int LAYOUT>
__global__ __launch_bounds__(kThreads) void dense_tile_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                              int32_t ntree, int64_t nb, int32_t accumulate,
                                                              int32_t* __restrict__ out,
                                                              int32_t* __restrict__ out2) {
  __shared__ int32_t sR[kTreeChunk][kTile];
  __shared__ int32_t sS[kTreeChunk][kTile];
  const int64_t nblocks = gridDim.x;
  const int64_t lin = xcd_linear(blockIdx.x, nblocks);
  int64_t bi, bj;
  if (LAYOUT == RFX_LAYOUT_UPPER) {
    upper_tile(lin, nb, bi, bj);
  } else {
    bi = lin % nb;
    bj = lin / nb;
  }
  dense_tile_body<MODE, LAYOUT>(keys, n, ntree, bi, bj, accumulate, out, out2, sR, sS);
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void dense16_tile_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                                int32_t ntree, int64_t nb,
                                                                int32_t* __restrict__ out) {
  __shared__ uint32_t sR[kTreeChunk16 / 2][kTile];
  __shared__ uint32_t sS[kTreeChunk16 / 2][kTile];
  const int64_t lin = xcd_linear(blockIdx.x, gridDim.x);
  int64_t bi, bj;
  upper_tile(lin, nb, bi, bj);
  dense16_tile_body<MODE>(keys, n, ntree, bi, bj, out, sR, sS);
}

// Rows [r0, r1) of the full matrix, written row-major (r - r0) * n + j.
template <int MODE>
__global__ __launch_bounds__(kThreads) void dense_rows_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                              int32_t ntree, int64_t r0, int64_t r1,
                                                              int32_t* __restrict__ out) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const int64_t i = r0 + blockIdx.y;
  if (j >= n || i >= r1) return;
  uint32_t c = 0;
  for (int32_t t = 0; t < ntree; ++t) {
    const int64_t tr = static_cast<int64_t>(t) * n;
    c += (row_key<MODE>(keys[tr + i]) == col_key<MODE>(keys[tr + j])) ? 1u : 0u;
  }
  out[(i - r0) * n + j] = static_cast<int32_t>(c);
}

// One thread per sorted slot: its partners of larger sample index are the
// rest of its leaf's member list (ascending); each pair is added once with a
// global integer atomic.
template <int MODE, int LAYOUT>
__global__ __launch_bounds__(256) void bucket_pair_kernel(const uint32_t* __restrict__ members,
                                                          const uint32_t* __restrict__ mbucket,
                                                          const uint32_t* __restrict__ bend,
                                                          const uint32_t* __restrict__ nvalid_p, int64_t n,
                                                          uint32_t total_nodes, int32_t* __restrict__ out,
                                                          int32_t* __restrict__ out2) {
  const uint32_t nvalid = *nvalid_p;
  for (int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; m < nvalid;
       m += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint32_t leaf = mbucket[m];
    if (leaf >= total_nodes) continue;  // off-leaf slot (per-tree layout: at the end of its tree's segment)
    const uint32_t me = members[m];
    const uint32_t e = bend[leaf];
    const int64_t i = me >> 1;
    const uint32_t bi = me & 1u;
    if (MODE == kMode_OOB_ROW && bi == 0) atomicAdd(&out[i + i * n], 1);  // i with itself
    for (uint32_t q = static_cast<uint32_t>(m) + 1; q < e; ++q) {
      const uint32_t o = members[q];
      const int64_t j = o >> 1;
      const uint32_t bj = o & 1u;
      if (MODE == kMode_OOB_ROW) {
        if (bi == 0) atomicAdd(&out[i + j * n], 1);
        if (bj == 0) atomicAdd(&out[j + i * n], 1);
        continue;
      }
      if (MODE == kMode_ALL || MODE == kMode_PAIR) {
        if (LAYOUT == RFX_LAYOUT_UPPER) {
          atomicAdd(&out[rfx_upper_index(i, j, n)], 1);
        } else {
          atomicAdd(&out[i + j * n], 1);
          atomicAdd(&out[j + i * n], 1);
        }
      }
      if (bi != bj) {
        int32_t* o2 = MODE == kMode_PAIR ? out2 : out;
        if (MODE == kMode_OOB_UPPER || MODE == kMode_PAIR) {
          if (LAYOUT == RFX_LAYOUT_UPPER) {
            atomicAdd(&o2[rfx_upper_index(i, j, n)], 1);
          } else {
            atomicAdd(&o2[i + j * n], 1);
            atomicAdd(&o2[j + i * n], 1);
          }
        } else if (MODE == kMode_OOB_ASYM) {
          // row = the out-of-bag sample, column = the in-bag one
          if (bi == 0)
            atomicAdd(&out[i + j * n], 1);
          else
            atomicAdd(&out[j + i * n], 1);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// GATHER algorithm: one workgroup owns one output line segment (a row of the
// packed upper triangle, or a column of the column-major full matrix) and
// collects that line's partners from the leaf buckets into LDS counters, so
// every output element is written exactly once, coalesced, with no global
// atomics and no separate zero fill.  Each wavefront takes 64 trees at a
// time (lane = tree): one load of pe gives the line's position in its leaf
// and the leaf's end, so the partners of larger index are a contiguous run
// (members are sorted); when the line is split into column chunks, binary or
// galloping searches trim the run to the chunk.  The wavefront then visits
// the union of its 64 runs in aligned 16-byte blocks of the member array
// (owner lane of a block by a search over the prefix sum of block counts),
// so a leaf of 1000 members and one of 2 cost the same per member.  PAIR
// packs ALL (low 16 bits) and OOB (high 16) per LDS word.
// ---------------------------------------------------------------------------
// Lines of at most kGatherLineMaxN <= kMembers16MaxN samples read the 16-bit
// member copy.
static_assert(kGatherLineMaxN <= kMembers16MaxN, "line kernel reads the 16-bit members");
#ifndef RFX_GATHER_ROWS
#define RFX_GATHER_ROWS 1
#endif
#ifndef RFX_LINE_MAX_LDS
#define RFX_LINE_MAX_LDS 81920
#endif
#ifndef RFX_LINE_THREADS
#define RFX_LINE_THREADS 512
#endif
#ifndef RFX_LINE_WAVES_PER_EU
#define RFX_LINE_WAVES_PER_EU 1
#endif
constexpr int kLineThreads = RFX_LINE_THREADS;
constexpr int64_t kLineMaxLds = RFX_LINE_MAX_LDS;  // bytes of LDS counters per workgroup
// the line kernels have no static LDS; their requests are at most kLineMaxLds
static_assert(kLineMaxLds <= kLdsLimitBytes, "line counters exceed the LDS");

// Counters of one line chunk in dynamic LDS: the whole line when it fits
// kLineMaxLds (no chunk searches, every pe entry read once), else chunks of
// kLineMaxLds / 4 columns.
inline int64_t line_chunk_len(int64_t n) { return std::min<int64_t>(n, kLineMaxLds / 4 - 4); }

template <int MODE, int LAYOUT>
__global__ __launch_bounds__(kLineThreads, RFX_LINE_WAVES_PER_EU) void gather_line_kernel(
    const uint2* __restrict__ pe, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ mbucket, int64_t n,
    int32_t ntree, const uint16_t* __restrict__ members, int32_t accumulate, int32_t* __restrict__ out,
    int32_t* __restrict__ out2, int64_t chunk_len) {
  extern __shared__ uint32_t line_lds[];
  const int64_t line = xcd_row(blockIdx.x);
  if (line >= n) return;
  gather_line_body<MODE, LAYOUT, uint16_t, kLineThreads>(pe, bstart, mbucket, n, ntree, members, accumulate, out,
                                                         out2, line, blockIdx.y, line_lds + 4, line_lds[0],
                                                         chunk_len);
}

// Whole lines, persistent: when a whole line's counters fit the LDS budget,
// each workgroup loops over the lines b, b + G, b + 2G, ... (G, the grid, a
// multiple of the XCD row group, so a workgroup stays on its XCD and its
// lines spread over the whole triangle).  The pe entries of the next line
// are loaded before the current line is visited, so their latency overlaps
// the member visits instead of opening every line.  No chunk searches: a
// line's runs are [position + 1, end) (UPPER) or the whole leaf (FULL).
template <int MODE, int LAYOUT, int KT>
__device__ __forceinline__ void line_whole_body(const uint2 (&q)[KT], const uint32_t* __restrict__ bstart,
                                                const uint32_t* __restrict__ mbucket, int64_t n,
                                                const uint16_t* __restrict__ members, int32_t accumulate,
                                                int32_t* __restrict__ out, int32_t* __restrict__ out2, int64_t line,
                                                uint32_t* acc, uint32_t& diag_valid) {
  const int64_t lo = LAYOUT == RFX_LAYOUT_UPPER ? line : 0;
  const int len_line = static_cast<int>(n - lo);
  for (int c = threadIdx.x; c < len_line; c += kLineThreads) acc[c] = 0u;
  if (threadIdx.x == 0) diag_valid = 0u;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  uint32_t valid = 0;
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    uint32_t s = 0, len = 0, bag = 0;
    const uint32_t end = q[k].y & 0x7fffffffu;
    if (end > 0) {
      ++valid;
      bag = q[k].y >> 31;
      s = LAYOUT == RFX_LAYOUT_UPPER ? q[k].x + 1 : bstart[mbucket[q[k].x]];
      len = end - s;
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
  if ((MODE == kMode_ALL || MODE == kMode_PAIR) && threadIdx.x == 0) acc[line - lo] = diag_valid;
  __syncthreads();
  const int64_t base = LAYOUT == RFX_LAYOUT_UPPER ? rfx_upper_index(line, line, n) : line * n;
  for (int c = threadIdx.x; c < len_line; c += kLineThreads) {
    const uint32_t v = acc[c];
    if (MODE == kMode_PAIR) {
      const int32_t va = static_cast<int32_t>(v & 0xffffu), vb = static_cast<int32_t>(v >> 16);
      rfx_emit_count(out + base + c, va, accumulate);
      rfx_emit_count(out2 + base + c, vb, accumulate);
    } else {
      rfx_emit_count(out + base + c, static_cast<int32_t>(v), accumulate);
    }
  }
}

template <int KT>
__device__ __forceinline__ void line_pe_load(const uint2* __restrict__ pe, int64_t n, int32_t ntree, int64_t line,
                                             uint2 (&q)[KT]) {
  const int32_t t0 = static_cast<int32_t>(threadIdx.x);
#pragma unroll
  for (int k = 0; k < KT; ++k) {
    const int32_t t = t0 + k * kLineThreads;
    q[k] = t < ntree ? pe[static_cast<int64_t>(t) * n + line] : make_uint2(0u, 0u);
  }
}

template <int MODE, int LAYOUT, int KT>
__global__ __launch_bounds__(kLineThreads) void gather_rows_kernel(const uint2* __restrict__ pe,
                                                                   const uint32_t* __restrict__ bstart,
                                                                   const uint32_t* __restrict__ mbucket, int64_t n,
                                                                   int32_t ntree,
                                                                   const uint16_t* __restrict__ members,
                                                                   int32_t accumulate, int32_t* __restrict__ out,
                                                                   int32_t* __restrict__ out2, int64_t rows_padded) {
  extern __shared__ uint32_t line_lds[];
  const int64_t G = gridDim.x;
  int64_t b = blockIdx.x;
  while (b < rows_padded && xcd_row(b) >= n) b += G;
  uint2 q[KT];
  if (b < rows_padded) line_pe_load<KT>(pe, n, ntree, xcd_row(b), q);
  while (b < rows_padded) {
    const int64_t line = xcd_row(b);
    uint2 cur[KT];
#pragma unroll
    for (int k = 0; k < KT; ++k) cur[k] = q[k];
    int64_t bn = b + G;
    while (bn < rows_padded && xcd_row(bn) >= n) bn += G;
    if (bn < rows_padded) line_pe_load<KT>(pe, n, ntree, xcd_row(bn), q);
    line_whole_body<MODE, LAYOUT, KT>(cur, bstart, mbucket, n, members, accumulate, out, out2, line, line_lds + 4,
                                      line_lds[0]);
    __syncthreads();  // the line's counters are read out before the next line zeroes them
    b = bn;
  }
}

#ifndef RFX_ROWS_WG_PER_CU
#define RFX_ROWS_WG_PER_CU 16
#endif

template <int MODE, int LAYOUT, int KT>
int launch_rows(hipStream_t st, const BucketWs& w, int64_t n, int32_t ntree, int32_t accumulate, int32_t* out,
                int32_t* out2) {
  const int64_t rows_padded = xcd_rows_padded(n);
  const int64_t g = std::min<int64_t>(rows_padded, 256 * RFX_ROWS_WG_PER_CU / (8 * kXcdGroup) * (8 * kXcdGroup));
  hipLaunchKernelGGL((gather_rows_kernel<MODE, LAYOUT, KT>), dim3(static_cast<unsigned>(g)), dim3(kLineThreads),
                     static_cast<unsigned>(4 * (n + 4)), st, w.pe, w.bstart, w.mbucket, n, ntree, w.members16,
                     accumulate, out, out2, rows_padded);
  return rfx_launch_status();
}

// ---------------------------------------------------------------------------
// GATHER, long rows (n > 32768): one workgroup per output line, its column
// chunks walked with per-tree cursors (gather_gallop_body in gather.hpp).
// ---------------------------------------------------------------------------
#ifndef RFX_GALLOP_CHUNK
#define RFX_GALLOP_CHUNK 16384
#endif
// LDS counters per workgroup.  At n = 200,000, T = 300: 2,048 columns 87.2 ms
// per call, 4,096 65.7, 6,144 62.8, 8,192 64.8, 16,384 76.4, 32,768 107.5;
// at T = 1000 (four cursors per lane) 6,144 takes 143.8 ms against 137.0.
constexpr int64_t kGallopChunk = RFX_GALLOP_CHUNK;
// dynamic counters + the kernel's static diag_valid word
static_assert(4 * kGallopChunk + 4 <= kLdsLimitBytes, "gallop chunk exceeds the LDS");
#ifndef RFX_GALLOP_THREADS
#define RFX_GALLOP_THREADS 1024
#endif
constexpr int kGallopThreads = RFX_GALLOP_THREADS;  // one tree cursor per thread and group
static_assert(kGallopThreads % 64 == 0 && kGallopThreads <= 1024, "gallop workgroup size");
template <int MODE, int LAYOUT, int G>
__global__ __launch_bounds__(kGallopThreads) void gather_gallop_kernel(const uint2* __restrict__ pe,
                                                            const uint32_t* __restrict__ bstart,
                                                            const uint32_t* __restrict__ mbucket, int64_t n,
                                                            int32_t ntree, const uint32_t* __restrict__ members,
                                                            int32_t accumulate, int32_t* __restrict__ out,
                                                            int32_t* __restrict__ out2, int64_t chunk) {
  extern __shared__ uint32_t gallop_acc[];
  __shared__ uint32_t diag_valid;
  const int64_t line = xcd_row(blockIdx.x);
  if (line >= n) return;
  gather_gallop_body<MODE, LAYOUT, G, uint32_t, kGallopThreads>(pe, bstart, mbucket, n, ntree, members, accumulate,
                                                                out, out2, line, gallop_acc, diag_valid, chunk);
}

// ---------------------------------------------------------------------------
// GATHER, long rows split by the chunk-boundary table (n > 32,768 with at
// most kWinMaxBounds boundaries): one workgroup of kGallopThreads threads per
// output line, thread x holding tree g * kGallopThreads + x.  At the start a
// thread loads its tree's pe entry and key (independent loads), then the
// leaf's boundary positions from wtab into registers; chunk c of the line is
// the piece [cursor, boundary c + 1) of each run, so no search is made.  The
// column chunks are aligned to kWinChunk and the members are read as 16-bit
// offsets within their 32,768-column window (members_w).  For UPPER the
// first chunk starts at the line's own column.
//
// Workgroups take the lines in the order of tree 0's member list (the lines
// grouped by their tree-0 leaf), so the lines in flight on an XCD are similar
// and share leaves and boundary rows in its L2; the per-line pe entries and
// leaves come from line-major copies (rpe, rleaf) for that reason.
//
// The LDS counter of column x of chunk [lo, hi) sits at index x - lo + sh,
// sh in [0, 4) chosen so that index and output address agree mod 16 bytes:
// the counters are cleared and flushed as aligned 16-byte vectors (a scalar
// head and tail where the output row is not 16-byte aligned; all scalar when
// out and out2 differ in alignment).
// ---------------------------------------------------------------------------
constexpr int kWinAccWords = static_cast<int>(kWinChunk) + 8;  // counters (+ shift, rounding) and the diagonal word
constexpr int kWinDiagWord = static_cast<int>(kWinChunk) + 4;

template <int MODE>
__device__ __forceinline__ void win_emit(uint32_t v, int32_t* __restrict__ out, int32_t* __restrict__ out2,
                                         int64_t o, int32_t accumulate) {
  if (MODE == kMode_PAIR) {
    const int32_t va = static_cast<int32_t>(v & 0xffffu), vb = static_cast<int32_t>(v >> 16);
    rfx_emit_count(out + o, va, accumulate);
    rfx_emit_count(out2 + o, vb, accumulate);
  } else {
    rfx_emit_count(out + o, static_cast<int32_t>(v), accumulate);
  }
}

typedef int32_t rfx_v4i __attribute__((ext_vector_type(4)));

// 16-byte form of rfx_emit_count's overwrite: a nontemporal vector store.
__device__ __forceinline__ void win_store4(int4* p, int4 v) {
  __builtin_nontemporal_store(rfx_v4i{v.x, v.y, v.z, v.w}, reinterpret_cast<rfx_v4i*>(p));
}

// Counters acc[L0 + j], j < count, to out[o0 + j] (and out2 for PAIR); the
// vector body needs (L0 + j) % 4 == 0 exactly where &out[o0 + j] is 16-byte
// aligned, which the caller's shift provides (vec false: all scalar).
template <int MODE, int THREADS>
__device__ __forceinline__ void win_flush(const uint32_t* __restrict__ acc, int L0, int count, bool vec,
                                          int32_t* __restrict__ out, int32_t* __restrict__ out2, int64_t o0,
                                          int32_t accumulate) {
  const int tid = threadIdx.x;
  const int head = vec ? min(count, static_cast<int>((4 - ((reinterpret_cast<uintptr_t>(out + o0) >> 2) & 3)) & 3))
                       : count;
  const int nvec = (count - head) >> 2;
  const int tail0 = head + 4 * nvec;
  const int nscalar = head + (count - tail0);
  for (int k = tid; k < nscalar; k += THREADS) {
    const int j = k < head ? k : tail0 + (k - head);
    win_emit<MODE>(acc[L0 + j], out, out2, o0 + j, accumulate);
  }
  const uint4* a4 = reinterpret_cast<const uint4*>(acc + L0 + head);
  int4* p4 = reinterpret_cast<int4*>(out + o0 + head);
  int4* q4 = MODE == kMode_PAIR ? reinterpret_cast<int4*>(out2 + o0 + head) : nullptr;
  for (int k = tid; k < nvec; k += THREADS) {
    const uint4 v = a4[k];
    if (MODE == kMode_PAIR) {
      int4 va = make_int4(static_cast<int32_t>(v.x & 0xffffu), static_cast<int32_t>(v.y & 0xffffu),
                          static_cast<int32_t>(v.z & 0xffffu), static_cast<int32_t>(v.w & 0xffffu));
      int4 vb = make_int4(static_cast<int32_t>(v.x >> 16), static_cast<int32_t>(v.y >> 16),
                          static_cast<int32_t>(v.z >> 16), static_cast<int32_t>(v.w >> 16));
      if (accumulate) {
        const int4 pa = p4[k], pb = q4[k];
        va = make_int4(va.x + pa.x, va.y + pa.y, va.z + pa.z, va.w + pa.w);
        vb = make_int4(vb.x + pb.x, vb.y + pb.y, vb.z + pb.z, vb.w + pb.w);
      }
      win_store4(p4 + k, va);
      win_store4(q4 + k, vb);
    } else {
      int4 va = make_int4(static_cast<int32_t>(v.x), static_cast<int32_t>(v.y), static_cast<int32_t>(v.z),
                          static_cast<int32_t>(v.w));
      if (accumulate) {
        const int4 pa = p4[k];
        va = make_int4(va.x + pa.x, va.y + pa.y, va.z + pa.z, va.w + pa.w);
      }
      win_store4(p4 + k, va);
    }
  }
}

// Two workgroups per CU (eight waves per SIMD): at most 64 registers.
template <int MODE, int LAYOUT, int G>
__global__ __launch_bounds__(kGallopThreads) __attribute__((amdgpu_waves_per_eu(8, 8))) void gather_window_kernel(
    const uint32_t* __restrict__ order, const uint2* __restrict__ rpe, const uint32_t* __restrict__ rleaf,
    const uint32_t* __restrict__ wtab, int nb, const uint32_t* __restrict__ bstart, int64_t n, int32_t ntree,
    const uint16_t* __restrict__ members, int32_t accumulate, int32_t* __restrict__ out,
    int32_t* __restrict__ out2) {
  extern __shared__ uint4 win_acc4[];  // kWinAccWords words, 16-byte aligned (no static LDS)
  uint32_t* const win_acc = reinterpret_cast<uint32_t*>(win_acc4);
  const int64_t slot = xcd_row(blockIdx.x);
  if (slot >= n) return;
  const int64_t line = static_cast<int64_t>(order[slot] >> 1);  // tree 0's slot order
  const int tid = threadIdx.x, lane = tid & 63;
  const int nch = nb + 1;
  const int c0 = LAYOUT == RFX_LAYOUT_UPPER ? static_cast<int>(line / kWinChunk) : 0;
  // per tree: cursor, end of the counted run, and the run's positions at the
  // chunk boundaries (wb[g][k]: start of chunk c0 + k + 1)
  uint32_t cur[G], end[G], wb[G][kWinMaxBounds];
  uint32_t bagm = 0, valid = 0;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int32_t t = g * kGallopThreads + tid;
    cur[g] = 0;
    end[g] = 0;
    uint32_t e = 0, bag = 0, pos = 0, leaf = 0;
    if (t < ntree) {
      const int64_t s = line * ntree + t;
      const uint2 q = rpe[s];
      leaf = rleaf[s];
      e = q.y & 0x7fffffffu;
      bag = q.y >> 31;
      pos = q.x;
    }
    if (e > 0) {
      ++valid;
      bagm |= bag << g;
      const uint32_t* tb = wtab + static_cast<int64_t>(leaf) * nb;
#pragma unroll
      for (int k = 0; k < kWinMaxBounds; ++k) wb[g][k] = c0 + k < nb ? tb[c0 + k] : e;
      cur[g] = LAYOUT == RFX_LAYOUT_UPPER ? pos + 1 : bstart[leaf];
      end[g] = (MODE == kMode_OOB_ASYM && !bag) ? cur[g] : e;
    } else {
#pragma unroll
      for (int k = 0; k < kWinMaxBounds; ++k) wb[g][k] = 0u;
    }
  }
  if (tid == 0) win_acc[kWinDiagWord] = 0u;
  __syncthreads();
  if (MODE == kMode_ALL || MODE == kMode_PAIR) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) valid += __shfl_down(valid, d, 64);
    if (lane == 0 && valid) atomicAdd(&win_acc[kWinDiagWord], valid);
  }
  const bool vec = ((reinterpret_cast<uintptr_t>(out) ^ reinterpret_cast<uintptr_t>(out2 ? out2 : out)) & 15) == 0;
  for (int c = c0; c < nch; ++c) {
    const int64_t lo = static_cast<int64_t>(c) * kWinChunk;
    const int64_t hi = min(n, lo + kWinChunk);
    const int64_t from = LAYOUT == RFX_LAYOUT_UPPER ? max(lo, line) : lo;  // first column written
    const int64_t wbase = lo & ~(kWinSpan - 1);
    const int64_t o0 = LAYOUT == RFX_LAYOUT_UPPER ? rfx_upper_index(line, from, n) : from + line * n;
    const int count = static_cast<int>(hi - from);
    // counter of column x at x - lo + sh: index and &out[o] agree mod 16 bytes
    const int sh = vec ? static_cast<int>(((reinterpret_cast<uintptr_t>(out + o0) >> 2) - (from - lo)) & 3) : 0;
    const int L0 = static_cast<int>(from - lo) + sh;
    __syncthreads();
    for (int q = (L0 >> 2) + tid; q < ((L0 + count + 3) >> 2); q += kGallopThreads)
      win_acc4[q] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const uint32_t stop = min(wb[g][0], end[g]);
      const uint32_t s0 = cur[g];
      cur[g] = max(cur[g], stop);
#pragma unroll
      for (int k = 0; k + 1 < kWinMaxBounds; ++k) wb[g][k] = wb[g][k + 1];
      wb[g][kWinMaxBounds - 1] = end[g];
      line_visits<MODE, LAYOUT>(s0, stop > s0 ? stop - s0 : 0u, (bagm >> g) & 1u, lane, members, win_acc,
                                lo - wbase - sh, line - wbase);
    }
    __syncthreads();
    if ((MODE == kMode_ALL || MODE == kMode_PAIR) && tid == 0 && line >= lo && line < hi)
      win_acc[line - lo + sh] = win_acc[kWinDiagWord];
    __syncthreads();
    win_flush<MODE, kGallopThreads>(win_acc, L0, count, vec, out, out2, o0, accumulate);
  }
}
// counters of one chunk with their shift and rounding, and the diagonal word
static_assert(4 * kWinAccWords <= kLdsLimitBytes, "window chunk exceeds the LDS");

template <int MODE, int LAYOUT, int G>
int launch_window(hipStream_t st, const BucketWs& w, int64_t n, int32_t ntree, int32_t accumulate, int32_t* out,
                  int32_t* out2) {
  const size_t lds = 4 * static_cast<size_t>(kWinAccWords);
  hipLaunchKernelGGL((gather_window_kernel<MODE, LAYOUT, G>), dim3(static_cast<unsigned>(xcd_rows_padded(n))),
                     dim3(kGallopThreads), lds, st, w.members, w.rpe, w.rleaf, w.wtab, w.nb, w.bstart, n, ntree,
                     w.members_w, accumulate, out, out2);
  return rfx_launch_status();
}

template <int MODE, int LAYOUT, int G>
int launch_gallop(hipStream_t st, const BucketWs& w, int64_t n, int32_t ntree, int32_t accumulate, int32_t* out,
                  int32_t* out2) {
  const size_t lds = 4 * static_cast<size_t>(kGallopChunk);
  hipLaunchKernelGGL((gather_gallop_kernel<MODE, LAYOUT, G>), dim3(static_cast<unsigned>(xcd_rows_padded(n))),
                     dim3(kGallopThreads), lds, st, w.pe, w.bstart, w.mbucket, n, ntree, w.members, accumulate, out,
                     out2, kGallopChunk);
  return rfx_launch_status();
}

// Diagonal of ALL: C_ii = number of trees in which i ends on a terminal node.
__global__ void all_diag_kernel(const int32_t* __restrict__ keys, int64_t n, int32_t ntree, int32_t layout,
                                int32_t* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t c = 0;
  for (int32_t t = 0; t < ntree; ++t) c += keys[static_cast<int64_t>(t) * n + i] >= 0 ? 1 : 0;
  const int64_t o = layout == RFX_LAYOUT_UPPER ? rfx_upper_index(i, i, n) : i + i * n;
  out[o] += c;
}

__global__ void sum_squares_kernel(const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
                                   int64_t len, unsigned long long* __restrict__ acc) {
  unsigned long long s = 0;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < len;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const unsigned long long c = bend[e] - bstart[e];
    s += c * c;
  }
  for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(acc, s);
}

// Packed 16-bit DENSE (ALL or OOB_UPPER, UPPER, overwrite): only for forests
// whose trees have at most kDense16MaxNodes nodes, which the caller knows.
template <int MODE>
int run_dense16(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int32_t* out) {
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t tiles = nb * (nb + 1) / 2;
  if (tiles > 0x7fffffffLL || ntree > 131070) return RFX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL((dense16_tile_kernel<MODE>), dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st, keys, n,
                     ntree, nb, out);
  return rfx_launch_status();
}

template <int MODE, int LAYOUT>
int run_dense(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int32_t accumulate, int32_t* out,
              int32_t* out2) {
  const int64_t nb = (n + kTile - 1) / kTile;
  const int64_t tiles = LAYOUT == RFX_LAYOUT_UPPER ? nb * (nb + 1) / 2 : nb * nb;
  if (tiles > 0x7fffffffLL) return RFX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL((dense_tile_kernel<MODE, LAYOUT>), dim3(static_cast<unsigned>(tiles)), dim3(kThreads), 0, st,
                     keys, n, ntree, nb, accumulate, out, out2);
  return rfx_launch_status();
}

template <int MODE, int LAYOUT>
int run_bucket(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
               int32_t accumulate, int32_t* out, int32_t* out2, void* ws, bool prebuilt) {
  const BucketWs w = bucket_ws_carve(ws, n, ntree, total_nodes, true, true);
  const int64_t slots = n * static_cast<int64_t>(ntree);
  const int64_t elems = LAYOUT == RFX_LAYOUT_UPPER ? n * (n + 1) / 2 : n * n;
  if (!accumulate) {
    int rc = fill_u32(st, out, elems, 0u);
    if (!rc && out2) rc = fill_u32(st, out2, elems, 0u);
    if (rc) return rc;
  }
  int rc = prebuilt ? RFX_OK : bucket_build(st, keys, n, ntree, total_nodes, w);
  if (rc) return rc;
  hipLaunchKernelGGL((bucket_pair_kernel<MODE, LAYOUT>), dim3(grid_for(slots, 256)), dim3(256), 0, st, w.members,
                     w.mbucket, w.bend, w.nvalid, n, static_cast<uint32_t>(total_nodes), out, out2);
  if (MODE == kMode_ALL || MODE == kMode_PAIR)
    hipLaunchKernelGGL(all_diag_kernel, dim3(rfx_ceil_div(n, 256)), dim3(256), 0, st, keys, n, ntree, LAYOUT, out);
  return rfx_launch_status();
}

template <int MODE, int LAYOUT>
int launch_gather_kernels(hipStream_t st, const BucketWs& w, int64_t n, int32_t ntree, int32_t accumulate,
                          int32_t* out, int32_t* out2);

template <int MODE, int LAYOUT>
int run_gather(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
               int32_t accumulate, int32_t* out, int32_t* out2, void* ws, bool prebuilt) {
  const BucketWs w = bucket_ws_carve(ws, n, ntree, total_nodes, true, true);
  // the line kernels (n <= kGatherLineMaxN) read no 32-bit members: a lean
  // build; for UPPER they read only pe and the 16-bit members
  const bool lean = n <= kGatherLineMaxN && w.members16 != nullptr;
  int rc = prebuilt ? RFX_OK
                    : bucket_build(st, keys, n, ntree, total_nodes, w, lean, lean && LAYOUT == RFX_LAYOUT_UPPER);
  if (rc) return rc;
  return launch_gather_kernels<MODE, LAYOUT>(st, w, n, ntree, accumulate, out, out2);
}

// The GATHER kernel for the shape: whole lines, line chunks, or the gallop
// walk (reads w.pe, w.bstart, w.mbucket, w.members, w.members16 only).
template <int MODE, int LAYOUT>
int launch_gather_kernels(hipStream_t st, const BucketWs& w, int64_t n, int32_t ntree, int32_t accumulate,
                          int32_t* out, int32_t* out2) {
  // the whole-line and line-chunk kernels read the 16-bit member copy, which
  // exists only for n <= kGatherLineMaxN (<= kMembers16MaxN)
  if (n <= kGatherLineMaxN && !w.members16) return RFX_ERR_INVALID;
  if (RFX_GATHER_ROWS && n <= kGatherLineMaxN && 4 * (n + 4) <= kLineMaxLds) {
    if (ntree <= kLineThreads) return launch_rows<MODE, LAYOUT, 1>(st, w, n, ntree, accumulate, out, out2);
    if (ntree <= 2 * kLineThreads) return launch_rows<MODE, LAYOUT, 2>(st, w, n, ntree, accumulate, out, out2);
    if (ntree <= 4 * kLineThreads) return launch_rows<MODE, LAYOUT, 4>(st, w, n, ntree, accumulate, out, out2);
    if (ntree <= 8 * kLineThreads) return launch_rows<MODE, LAYOUT, 8>(st, w, n, ntree, accumulate, out, out2);
  }
  if (n <= kGatherLineMaxN) {
    const int64_t chunk_len = line_chunk_len(n);
    dim3 grid(static_cast<unsigned>(xcd_rows_padded(n)), rfx_ceil_div(n, chunk_len));
    hipLaunchKernelGGL((gather_line_kernel<MODE, LAYOUT>), grid, dim3(kLineThreads),
                       static_cast<unsigned>(4 * (chunk_len + 4)), st, w.pe, w.bstart, w.mbucket, n, ntree,
                       w.members16, accumulate, out, out2, chunk_len);
    return rfx_launch_status();
  }
  constexpr int TH = kGallopThreads;
  if (w.wtab && w.members_w && w.rpe && w.nb == win_bounds(n)) {
    if (ntree <= TH) return launch_window<MODE, LAYOUT, 1>(st, w, n, ntree, accumulate, out, out2);
    if (ntree <= 2 * TH) return launch_window<MODE, LAYOUT, 2>(st, w, n, ntree, accumulate, out, out2);
  }
  if (ntree <= TH) return launch_gallop<MODE, LAYOUT, 1>(st, w, n, ntree, accumulate, out, out2);
  if (ntree <= 2 * TH) return launch_gallop<MODE, LAYOUT, 2>(st, w, n, ntree, accumulate, out, out2);
  if (ntree <= 4 * TH) return launch_gallop<MODE, LAYOUT, 4>(st, w, n, ntree, accumulate, out, out2);
  if (TH < 512 && ntree <= 8 * TH) return launch_gallop<MODE, LAYOUT, 8>(st, w, n, ntree, accumulate, out, out2);
  if (TH < 256 && ntree <= 16 * TH) return launch_gallop<MODE, LAYOUT, 16>(st, w, n, ntree, accumulate, out, out2);
  return RFX_ERR_UNSUPPORTED;
}

// A RFX_ALGO_PREBUILT call needs the workspace's last build to be a full
// one (rfx_prox_build_buckets, or a BUCKET / long-line GATHER count call) of
// the same shape, with the window arrays where the GATHER kernel reads them.
bool prebuilt_usable(const void* ws, int64_t n, int32_t ntree, int64_t total_nodes, int32_t algo) {
  BucketBuildRecord r;
  if (!last_bucket_build(ws, &r)) return false;
  if (!r.full || r.n != n || r.ntree != ntree || r.total_nodes != total_nodes) return false;
  return !(algo == RFX_ALGO_GATHER && win_enabled(n) && !r.windows);
}

template <int MODE>
int dispatch(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes, int32_t layout,
             int32_t algo, int32_t accumulate, int32_t* out, int32_t* out2, void* ws) {
  const bool prebuilt = (algo & RFX_ALGO_PREBUILT) != 0;
  algo &= ~RFX_ALGO_PREBUILT;
  if (prebuilt && !prebuilt_usable(ws, n, ntree, total_nodes, algo)) return RFX_ERR_INVALID;
  if (algo == RFX_ALGO_DENSE) {
    return layout == RFX_LAYOUT_UPPER ? run_dense<MODE, RFX_LAYOUT_UPPER>(st, keys, n, ntree, accumulate, out, out2)
                                      : run_dense<MODE, RFX_LAYOUT_FULL>(st, keys, n, ntree, accumulate, out, out2);
  }
  if (algo == RFX_ALGO_GATHER) {
    return layout == RFX_LAYOUT_UPPER
               ? run_gather<MODE, RFX_LAYOUT_UPPER>(st, keys, n, ntree, total_nodes, accumulate, out, out2, ws, prebuilt)
               : run_gather<MODE, RFX_LAYOUT_FULL>(st, keys, n, ntree, total_nodes, accumulate, out, out2, ws, prebuilt);
  }
  return layout == RFX_LAYOUT_UPPER
             ? run_bucket<MODE, RFX_LAYOUT_UPPER>(st, keys, n, ntree, total_nodes, accumulate, out, out2, ws, prebuilt)
             : run_bucket<MODE, RFX_LAYOUT_FULL>(st, keys, n, ntree, total_nodes, accumulate, out, out2, ws, prebuilt);
}

// AUTO: compare the compare-kernel cost T*n^2/2 (one tile compare per
// tree-pair) with the leaf-bucket cost: sum over trees and leaves of |leaf|^2
// partner visits plus one pass over the output.  GATHER (LDS row segments)
// serves T <= 4096 (the long-line kernel keeps per-tree cursors in registers);
// beyond that the global-atomic BUCKET form is used.  The per-item weights
// are MI355X
// measurements on the covtype-like workload (n = 20,000, T = 500).
constexpr double kVisitVsCompare = 8.0;
constexpr double kWriteVsCompare = 4.0;
constexpr double kAtomicVsCompare = 24.0;

int choose_algo(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes, void* ws,
                size_t ws_bytes, int32_t* algo) {
  if (!ws || ws_bytes < bucket_ws_bytes(n, ntree, total_nodes, true)) {
    *algo = RFX_ALGO_DENSE;
    return RFX_OK;
  }
  double sq = 0;
  int rc = rfx_prox_pair_stats(st, keys, n, ntree, total_nodes, ws, ws_bytes, &sq);
  if (rc) return rc;
  const double nn = static_cast<double>(n) * static_cast<double>(n);
  const double dense_cost = 0.5 * static_cast<double>(ntree) * nn;
  const bool gather = ntree <= 4096;
  const double sparse_cost =
      sq * (gather ? kVisitVsCompare : 0.5 * kAtomicVsCompare) + 0.5 * nn * kWriteVsCompare;
  *algo = sparse_cost < dense_cost ? (gather ? RFX_ALGO_GATHER : RFX_ALGO_BUCKET) : RFX_ALGO_DENSE;
  return RFX_OK;
}

int check_common(const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes, int32_t layout, int32_t algo) {
  if (!keys || n <= 0 || n > RFX_MAX_SAMPLES || ntree <= 0 || ntree > 65535) return RFX_ERR_INVALID;
  if (layout != RFX_LAYOUT_FULL && layout != RFX_LAYOUT_UPPER) return RFX_ERR_INVALID;
  if (algo & RFX_ALGO_PREBUILT) {
    const int32_t base = algo & ~RFX_ALGO_PREBUILT;
    if (base != RFX_ALGO_BUCKET && base != RFX_ALGO_GATHER) return RFX_ERR_INVALID;
    algo = base;
  }
  if (algo < RFX_ALGO_AUTO || algo > RFX_ALGO_GATHER) return RFX_ERR_INVALID;
  if (algo != RFX_ALGO_DENSE && (total_nodes <= 0 || total_nodes > RFX_MAX_TOTAL_NODES)) return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  return RFX_OK;
}

}  // namespace

// One problem's UPPER counts (not accumulated) on bucket structures that may
// be the problem's slice of a batched build: pe of the problem's slots,
// positions global to members / members16, keys of the problem.  Used by the
// batch (batch.hip) for the forests it launches on their own.
int count_problem_slice(hipStream_t st, int32_t variant, int32_t algo, const int32_t* keys, int64_t n,
                        int32_t ntree, const uint2* pe, const uint32_t* bstart, const uint32_t* mbucket,
                        const uint32_t* members, const uint16_t* members16, int32_t* out, int32_t* out2) {
  if (algo == (RFX_ALGO_DENSE | kAlgoDense16)) {
    if (variant == RFX_PROX_ALL) return run_dense16<kMode_ALL>(st, keys, n, ntree, out);
    if (variant == RFX_PROX_OOB_UPPER) return run_dense16<kMode_OOB_UPPER>(st, keys, n, ntree, out);
    algo = RFX_ALGO_DENSE;  // PAIR: the 32-bit form
  }
  if (algo == RFX_ALGO_DENSE) {
    switch (variant) {
      case RFX_PROX_ALL:
        return run_dense<kMode_ALL, RFX_LAYOUT_UPPER>(st, keys, n, ntree, 0, out, out2);
      case RFX_PROX_OOB_UPPER:
        return run_dense<kMode_OOB_UPPER, RFX_LAYOUT_UPPER>(st, keys, n, ntree, 0, out, out2);
      case RFX_PROX_PAIR:
        return run_dense<kMode_PAIR, RFX_LAYOUT_UPPER>(st, keys, n, ntree, 0, out, out2);
      default:
        return RFX_ERR_INVALID;
    }
  }
  BucketWs w{};
  w.pe = const_cast<uint2*>(pe);
  w.bstart = const_cast<uint32_t*>(bstart);
  w.mbucket = const_cast<uint32_t*>(mbucket);
  w.members = const_cast<uint32_t*>(members);
  w.members16 = const_cast<uint16_t*>(members16);
  switch (variant) {
    case RFX_PROX_ALL:
      return launch_gather_kernels<kMode_ALL, RFX_LAYOUT_UPPER>(st, w, n, ntree, 0, out, out2);
    case RFX_PROX_OOB_UPPER:
      return launch_gather_kernels<kMode_OOB_UPPER, RFX_LAYOUT_UPPER>(st, w, n, ntree, 0, out, out2);
    case RFX_PROX_PAIR:
      return launch_gather_kernels<kMode_PAIR, RFX_LAYOUT_UPPER>(st, w, n, ntree, 0, out, out2);
    default:
      return RFX_ERR_INVALID;
  }
}

extern "C" size_t rfx_prox_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes, int32_t algo) {
  if (algo == RFX_ALGO_DENSE) return 0;
  if (n <= 0 || ntree <= 0 || total_nodes <= 0) return 0;
  return bucket_ws_bytes(n, ntree, total_nodes, true);
}

extern "C" int rfx_prox_pair_stats(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                   int64_t total_nodes, void* ws, size_t ws_bytes, double* sum_sq_host) {
  if (!keys || !ws || !sum_sq_host || n <= 0 || ntree <= 0 || total_nodes <= 0 ||
      total_nodes > RFX_MAX_TOTAL_NODES)
    return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  if (ws_bytes < bucket_ws_bytes(n, ntree, total_nodes, true)) return RFX_ERR_WORKSPACE;
  const BucketWs w = bucket_ws_carve(ws, n, ntree, total_nodes, true, true);
  int rc = bucket_build(st, keys, n, ntree, total_nodes, w);
  if (rc) return rc;
  rc = fill_u32(st, w.stat, 2, 0u);
  if (rc) return rc;
  hipLaunchKernelGGL(sum_squares_kernel, dim3(grid_for(total_nodes, 256)), dim3(256), 0, st, w.bstart, w.bend,
                     total_nodes, w.stat);
  rc = rfx_launch_status();
  if (rc) return rc;
  unsigned long long h = 0;
  RFX_HIP_TRY(hipMemcpyAsync(&h, w.stat, 8, hipMemcpyDeviceToHost, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));
  *sum_sq_host = static_cast<double>(h);
  return RFX_OK;
}

extern "C" int rfx_prox_build_buckets(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                      int64_t total_nodes, void* ws, size_t ws_bytes) {
  if (!keys || !ws || n <= 0 || n > RFX_MAX_SAMPLES || ntree <= 0 || total_nodes <= 0 ||
      total_nodes > RFX_MAX_TOTAL_NODES)
    return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  if (ws_bytes < bucket_ws_bytes(n, ntree, total_nodes, true)) return RFX_ERR_WORKSPACE;
  return bucket_build(st, keys, n, ntree, total_nodes, bucket_ws_carve(ws, n, ntree, total_nodes, true, true));
}

extern "C" int rfx_prox_accumulate(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                   int64_t total_nodes, int32_t variant, int32_t layout, int32_t algo,
                                   int32_t accumulate, int32_t* out, void* ws, size_t ws_bytes) {
  int rc = check_common(keys, n, ntree, total_nodes, layout, algo);
  if (rc) return rc;
  if (!out) return RFX_ERR_INVALID;
  if ((variant == RFX_PROX_OOB_ASYM || variant == RFX_PROX_OOB_ROW) && layout != RFX_LAYOUT_FULL)
    return RFX_ERR_UNSUPPORTED;
  if (algo == RFX_ALGO_AUTO) {
    rc = choose_algo(st, keys, n, ntree, total_nodes, ws, ws_bytes, &algo);
    if (rc) return rc;
  }
  {
    const int32_t base = algo & ~RFX_ALGO_PREBUILT;
    if ((base == RFX_ALGO_BUCKET || base == RFX_ALGO_GATHER) && (!ws || ws_bytes < bucket_ws_bytes(n, ntree, total_nodes, true)))
      return RFX_ERR_WORKSPACE;
  }
  switch (variant) {
    case RFX_PROX_ALL:
      return dispatch<kMode_ALL>(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out, nullptr, ws);
    case RFX_PROX_OOB_ASYM:
      return dispatch<kMode_OOB_ASYM>(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out, nullptr, ws);
    case RFX_PROX_OOB_UPPER:
      return dispatch<kMode_OOB_UPPER>(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out, nullptr,
                                       ws);
    case RFX_PROX_OOB_ROW:
      return dispatch<kMode_OOB_ROW>(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out, nullptr, ws);
    default:
      return RFX_ERR_INVALID;
  }
}

extern "C" int rfx_prox_accumulate_pair(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                        int64_t total_nodes, int32_t layout, int32_t algo, int32_t accumulate,
                                        int32_t* out_all, int32_t* out_oob, void* ws, size_t ws_bytes) {
  int rc = check_common(keys, n, ntree, total_nodes, layout, algo);
  if (rc) return rc;
  if (!out_all || !out_oob) return RFX_ERR_INVALID;
  if (algo == RFX_ALGO_AUTO) {
    rc = choose_algo(st, keys, n, ntree, total_nodes, ws, ws_bytes, &algo);
    if (rc) return rc;
  }
  {
    const int32_t base = algo & ~RFX_ALGO_PREBUILT;
    if ((base == RFX_ALGO_BUCKET || base == RFX_ALGO_GATHER) && (!ws || ws_bytes < bucket_ws_bytes(n, ntree, total_nodes, true)))
      return RFX_ERR_WORKSPACE;
  }
  return dispatch<kMode_PAIR>(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out_all, out_oob, ws);
}

extern "C" int rfx_prox_rows(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int32_t variant,
                             int64_t row_begin, int64_t row_end, int32_t* out) {
  if (!keys || !out || n <= 0 || ntree <= 0 || row_begin < 0 || row_end > n || row_end <= row_begin ||
      row_end - row_begin > 65535)
    return RFX_ERR_INVALID;
  dim3 grid(rfx_ceil_div(n, kThreads), static_cast<unsigned>(row_end - row_begin));
  switch (variant) {
    case RFX_PROX_ALL:
      hipLaunchKernelGGL(dense_rows_kernel<kMode_ALL>, grid, dim3(kThreads), 0, st, keys, n, ntree, row_begin,
                         row_end, out);
      break;
    case RFX_PROX_OOB_ASYM:
      hipLaunchKernelGGL(dense_rows_kernel<kMode_OOB_ASYM>, grid, dim3(kThreads), 0, st, keys, n, ntree,
                         row_begin, row_end, out);
      break;
    case RFX_PROX_OOB_UPPER:
      hipLaunchKernelGGL(dense_rows_kernel<kMode_OOB_UPPER>, grid, dim3(kThreads), 0, st, keys, n, ntree,
                         row_begin, row_end, out);
      break;
    case RFX_PROX_OOB_ROW:
      hipLaunchKernelGGL(dense_rows_kernel<kMode_OOB_ROW>, grid, dim3(kThreads), 0, st, keys, n, ntree, row_begin,
                         row_end, out);
      break;
    default:
      return RFX_ERR_INVALID;
  }
  return rfx_launch_status();
}

// ---------------------------------------------------------------------------
// No-allocation form of rfx_prox_accumulate_pair: the leaf-bucket workspace
// is a file-scope device array, so the call allocates nothing.  Each GPU the
// library runs on holds its own copy of the array until the process ends.
// Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device with that device current, and only
// that address reaches kernels.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RFX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kNoAllocMaxDevices] = {};

// Device address of the current device's scratch array.
int noalloc_scratch(void** out) {
  int dev = -1;
  RFX_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return RFX_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RFX_ERR_HIP;
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return RFX_OK;
}
}  // namespace

extern "C" int rfx_prox_accumulate_pair_noalloc(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree,
                                                int64_t total_nodes, int32_t layout, int32_t algo,
                                                int32_t accumulate, int32_t* out_all, int32_t* out_oob) {
  // Every call rebuilds the buckets it reads: a prebuilt workspace would be
  // whatever the last call on this GPU left in the shared array.
  if (algo & RFX_ALGO_PREBUILT) return RFX_ERR_INVALID;
  int rc = check_common(keys, n, ntree, total_nodes, layout, algo);
  if (rc) return rc;
  if (!out_all || !out_oob) return RFX_ERR_INVALID;
  if (algo != RFX_ALGO_DENSE && bucket_ws_bytes(n, ntree, total_nodes, true) > RFX_NOALLOC_SCRATCH_BYTES)
    return RFX_ERR_WORKSPACE;
  void* ws = nullptr;
  rc = noalloc_scratch(&ws);
  if (rc) return rc;
  return rfx_prox_accumulate_pair(st, keys, n, ntree, total_nodes, layout, algo, accumulate, out_all, out_oob, ws,
                                  RFX_NOALLOC_SCRATCH_BYTES);
}
