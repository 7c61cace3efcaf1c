// Leaf buckets of the (tree, sample) slots, shared by the co-occurrence
// counts and the matrix-free proximity products of the factor path.
//
// Every slot s = t*n + i with a terminal leaf gets the sort key "global node
// id" (off-leaf slots get total_nodes, which sorts last) and the value
// (s << 1) | in-bag.  A stable LSD radix sort (8-bit digits, 4096-key tiles:
// tile histograms, one exclusive scan over the digit-major histogram, and a
// scatter whose in-tile ranks come from wavefront ballots on the digit bits)
// groups the slots by leaf with the samples of a leaf in ascending order.
// The one-problem build sorts every tree's segment of slots [t*n, (t+1)*n)
// on its own, by the node id relative to the tree's smallest leaf id: a tree
// has fewer than 2n + 1 nodes, so 2 passes instead of 3 at n = 20,000 (3
// instead of 4 at n = 200,000).  The passes needed follow from the largest
// leaf-id range of a tree, found on the device; the host launches the
// maximum and the surplus passes return at once.  Off-leaf slots (sentinel
// key) sort to the end of their tree's segment.  The batched build sorts the
// concatenated key space in one segment, off-leaf slots at the very end.
// Up to n = 32,767 the one-problem build is instead one counting-sort
// workgroup per tree in LDS (rs_tree_lds_kernel below), with the same
// results.
// From the sorted arrays:
//   bstart[b], bend[b]  the member range of leaf b (0, 0 when empty)
//   members[p]          (i << 1) | in-bag of the p-th sorted slot
//   members16[p]        the same as 16 bits, kept when n <= kMembers16MaxN
//                       (half the bytes for the gather kernels' member runs)
//   mbucket[p]          its leaf (the sorted keys)
//   pe[s]               (position of slot s in members, end of its leaf |
//                       in-bag << 31); (0, 0) off-leaf
// so a sample finds the partners that follow it in its leaf with one load.
// For long lines (n > 32,768 and at most kWinMaxBounds column-chunk
// boundaries) the one-problem build of the count calls also writes
//   members_w[p]        ((i mod 32,768) << 1) | in-bag: the member's column
//                       relative to its 32,768-column window, 16 bits
//   wtab[b * nb + k]    for leaf b and chunk boundary k + 1 (column
//                       (k + 1) * kWinChunk): the first position in the
//                       leaf's run whose column is at or past it (bend[b]
//                       when none is)
// so a line's run splits into its column chunks by table lookups instead of
// searches, and each chunk's members are read as 16-bit window offsets.
#pragma once

#include "rfx_common.hpp"

// Host record of the last bucket build in each workspace, keyed by the
// workspace's base address (defined in accum.hip): RFX_ALGO_PREBUILT calls
// refuse a workspace whose last build skipped arrays, lacks the long-line
// window arrays their kernel reads, or was made for another shape.
struct BucketBuildRecord {
  int64_t n = 0;
  int32_t ntree = 0;
  int64_t total_nodes = 0;
  bool full = false;     // every array of the per-tree layout was written
  bool windows = false;  // members_w and wtab were written too
};
void note_bucket_build(const void* ws, const BucketBuildRecord& r);
bool last_bucket_build(const void* ws, BucketBuildRecord* r);

namespace {

constexpr int64_t kMembers16MaxN = 32768;

#ifndef RFX_WIN_CHUNK
#define RFX_WIN_CHUNK 16384
#endif
#ifndef RFX_WIN_MAX_BOUNDS
#define RFX_WIN_MAX_BOUNDS 16
#endif
constexpr int64_t kWinChunk = RFX_WIN_CHUNK;      // columns per LDS chunk of a long line
constexpr int kWinMaxBounds = RFX_WIN_MAX_BOUNDS;  // chunk boundaries held per (line, tree)
constexpr int64_t kWinSpan = 32768;                // members_w offsets are relative to these windows
static_assert((kWinChunk & (kWinChunk - 1)) == 0 && kWinSpan % kWinChunk == 0, "chunks tile the windows");
inline int win_bounds(int64_t n) { return static_cast<int>((n + kWinChunk - 1) / kWinChunk - 1); }
inline bool win_enabled(int64_t n) { return n > kMembers16MaxN && win_bounds(n) <= kWinMaxBounds; }
inline int win_shift() {
  int s = 0;
  while ((int64_t{1} << s) < kWinChunk) ++s;
  return s;
}
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanThreads * kScanItems;

constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;
constexpr int kRsTile = kRsThreads * kRsRounds;  // keys per tile

// Block-wide exclusive scan of kScanItems per thread; returns the block total.
__device__ uint32_t block_exclusive_scan(uint32_t (&v)[kScanItems], uint32_t* lds_waves) {
  uint32_t s = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    const uint32_t x = v[q];
    v[q] = s;
    s += x;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) lds_waves[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < (kScanThreads / 64) ? lds_waves[lane] : 0u;
    uint32_t wi = w;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(wi, d, 64);
      if (lane >= d) wi += y;
    }
    if (lane < (kScanThreads / 64)) lds_waves[lane] = wi - w;
    if (lane == (kScanThreads / 64) - 1) lds_waves[kScanThreads / 64] = wi;
  }
  __syncthreads();
  const uint32_t base = lds_waves[wave] + (incl - s);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) v[q] += base;
  const uint32_t total = lds_waves[kScanThreads / 64];
  __syncthreads();
  return total;
}

// The scans of a radix pass return at once when `passes` (device pass count
// of the per-tree sort) says the pass is surplus.
__device__ __forceinline__ bool rs_surplus(const int32_t* passes, int pass) { return passes && pass >= *passes; }

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t len,
                                                                   uint32_t* __restrict__ part,
                                                                   const int32_t* passes = nullptr, int pass = 0) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
  if (rs_surplus(passes, pass)) return;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) v[q] = base + q < len ? in[base + q] : 0u;
  const uint32_t tot = block_exclusive_scan(v, lw);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(uint32_t* __restrict__ part, int64_t nparts,
                                                                     const int32_t* passes = nullptr, int pass = 0) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
  if (rs_surplus(passes, pass)) return;
  uint32_t carry = 0;
  for (int64_t b0 = 0; b0 < nparts; b0 += kScanChunk) {
    const int64_t base = b0 + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) v[q] = base + q < nparts ? part[base + q] : 0u;
    const uint32_t tot = block_exclusive_scan(v, lw);
#pragma unroll
    for (int q = 0; q < kScanItems; ++q)
      if (base + q < nparts) part[base + q] = v[q] + carry;
    carry += tot;
  }
  if (threadIdx.x == 0) part[nparts] = carry;
}

// Exclusive scan of in[] into out[] (in-place allowed: each element is read
// and written by the same thread).
__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* in, int64_t len,
                                                                  const uint32_t* __restrict__ part, uint32_t* out,
                                                                  const int32_t* passes = nullptr, int pass = 0) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
  if (rs_surplus(passes, pass)) return;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) v[q] = base + q < len ? in[base + q] : 0u;
  block_exclusive_scan(v, lw);
  const uint32_t off = part[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q)
    if (base + q < len) out[base + q] = v[q] + off;
}

// Tiles never straddle a segment: segment g holds seg_len slots starting at
// g * seg_len, tile j of it [j * kRsTile, min((j + 1) * kRsTile, seg_len)),
// and histograms are laid out [segment][digit][tile], so one exclusive scan
// over them keeps every segment's keys inside the segment.  The global
// layout is one segment of all N slots.
struct RsTiles {
  int64_t seg_len;  // slots per segment
  int64_t tps;      // tiles per segment
  int64_t ntiles;   // total tiles
};

__device__ __forceinline__ void rs_tile(const RsTiles& g, int64_t tile, int64_t& start, int& len, int64_t& seg,
                                        int64_t& j) {
  seg = tile / g.tps;
  j = tile - seg * g.tps;
  start = seg * g.seg_len + j * kRsTile;
  len = static_cast<int>(min<int64_t>(kRsTile, g.seg_len - j * kRsTile));
}

__device__ __forceinline__ int64_t rs_hist_index(const RsTiles& g, int64_t seg, int64_t j, uint32_t d) {
  return (seg * 256 + d) * g.tps + j;
}

// Pass p reads buffer (p + first) & 1 and writes the other, first being the
// parity of the pass count (device value `passes`, per-tree sort) or 0 (host
// pass count, global sort), so the sorted keys always end in buffer 0 for
// the per-tree sort.  With `passes` given, a pass at or beyond it returns.
__device__ __forceinline__ bool rs_pass_buffers(const int32_t* passes, int pass, int& src) {
  if (passes) {
    const int np = *passes;
    if (pass >= np) return false;
    src = (pass + (np & 1)) & 1;
  } else {
    src = pass & 1;
  }
  return true;
}

// Sort input, global layout: key = leaf (global node id) or total_nodes,
// value = slot << 1 | bag.
__global__ void rs_prep_kernel(const int32_t* __restrict__ keys, int64_t slots, uint32_t total_nodes,
                               uint32_t* __restrict__ K, uint32_t* __restrict__ V) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= slots) return;
  const int32_t k = keys[s];
  K[s] = k >= 0 ? static_cast<uint32_t>(k) >> 1 : total_nodes;
  V[s] = (static_cast<uint32_t>(s) << 1) | (k >= 0 ? (static_cast<uint32_t>(k) & 1u) : 0u);
}

// Per-tree layout, step 1: smallest and largest leaf id of each tree
// (workgroup t), and the largest range over the trees (atomic).
__global__ __launch_bounds__(256) void rs_tree_range_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                            uint32_t* __restrict__ tmin, uint32_t* __restrict__ info) {
  __shared__ uint32_t smin[4], smax[4];
  const int64_t t = blockIdx.x;
  uint32_t lo = 0xffffffffu, hi = 0u;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const int32_t k = keys[t * n + i];
    if (k >= 0) {
      const uint32_t v = static_cast<uint32_t>(k) >> 1;
      lo = min(lo, v);
      hi = max(hi, v);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = min(lo, static_cast<uint32_t>(__shfl_xor(lo, d, 64)));
    hi = max(hi, static_cast<uint32_t>(__shfl_xor(hi, d, 64)));
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    smin[wave] = lo;
    smax[wave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    lo = min(min(smin[0], smin[1]), min(smin[2], smin[3]));
    hi = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    tmin[t] = lo;
    if (lo <= hi) atomicMax(&info[0], hi - lo);
  }
}

// Step 2: the sentinel (all ones over the digits needed) and the pass count:
// info[0] = largest range, info[1] = sentinel, info[2] = passes.
__global__ void rs_tree_passes_kernel(uint32_t* __restrict__ info) {
  const uint32_t range = info[0];
  int bits = 0;
  while (bits < 32 && (1ull << bits) <= static_cast<uint64_t>(range) + 1) ++bits;  // range + 1 < 2^bits
  const int passes = max(1, (bits + 7) / 8);
  info[1] = passes >= 4 ? 0xffffffffu : static_cast<uint32_t>((1ull << (8 * passes)) - 1);
  info[2] = static_cast<uint32_t>(passes);
}

// Step 3: key = leaf id - tmin[t] or the sentinel, into the buffer the first
// pass reads.
// Tree of slot s = t * n + i without a 64-bit integer division (a long
// instruction sequence here): the double quotient s * (1 / n) is within one of
// t for s < 2^52, and one compare either way corrects it.
__device__ __forceinline__ int64_t slot_tree(int64_t s, int64_t n, double inv_n) {
  int64_t t = static_cast<int64_t>(static_cast<double>(s) * inv_n);
  if (t * n > s)
    --t;
  else if ((t + 1) * n <= s)
    ++t;
  return t;
}

__global__ void rs_prep_tree_kernel(const int32_t* __restrict__ keys, int64_t n, int64_t slots, double inv_n,
                                    const uint32_t* __restrict__ tmin, const uint32_t* __restrict__ info,
                                    uint32_t* __restrict__ K0, uint32_t* __restrict__ K1, uint32_t* __restrict__ V0,
                                    uint32_t* __restrict__ V1) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= slots) return;
  const bool odd = info[2] & 1u;
  uint32_t* K = odd ? K1 : K0;
  uint32_t* V = odd ? V1 : V0;
  const int32_t k = keys[s];
  K[s] = k >= 0 ? (static_cast<uint32_t>(k) >> 1) - tmin[slot_tree(s, n, inv_n)] : info[1];
  V[s] = (static_cast<uint32_t>(s) << 1) | (k >= 0 ? (static_cast<uint32_t>(k) & 1u) : 0u);
}

// Digit histogram of one tile.
__global__ __launch_bounds__(kRsThreads) void rs_hist_kernel(const uint32_t* __restrict__ K0,
                                                             const uint32_t* __restrict__ K1,
                                                             const int32_t* __restrict__ passes, int pass, RsTiles g,
                                                             uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[256];
  int src;
  if (!rs_pass_buffers(passes, pass, src)) return;
  const uint32_t* K = src ? K1 : K0;
  const int shift = 8 * pass;
  int64_t start, seg, j;
  int len;
  rs_tile(g, blockIdx.x, start, len, seg, j);
  cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int q = threadIdx.x; q < len; q += kRsThreads) atomicAdd(&cnt[(K[start + q] >> shift) & 255u], 1u);
  __syncthreads();
  hist[rs_hist_index(g, seg, j, threadIdx.x)] = cnt[threadIdx.x];
}

// Stable scatter of one tile of kRsTile keys.  Wave w ranks the keys
// [w * kRsTile / 4, (w + 1) * kRsTile / 4) of the tile in kRsRounds rounds of
// 64 consecutive keys: the lanes holding the same digit in a round are found
// with 8 ballots, and a wave-private LDS counter per digit carries the count
// over the rounds (no barrier inside the loop).  One block scan over the
// digits then gives every key its position in the tile sorted by digit; the
// tile is sorted in LDS and written out in position order, so the stores of
// consecutive threads fall into the same digit run (coalesced) instead of
// one scattered 4-byte store per key.
__global__ __launch_bounds__(kRsThreads) void rs_scatter_kernel(uint32_t* __restrict__ K0, uint32_t* __restrict__ K1,
                                                                uint32_t* __restrict__ V0, uint32_t* __restrict__ V1,
                                                                const int32_t* __restrict__ passes, int pass,
                                                                RsTiles g, const uint32_t* __restrict__ offs) {
  constexpr int kWaves = kRsThreads / 64;
  constexpr int kPerWave = kRsTile / kWaves;
  constexpr int kKeysPerLane = kPerWave / 64;
  __shared__ uint32_t wcnt[kWaves][256];  // per-wave digit counts, then per-wave digit bases
  __shared__ int64_t gofs[256];           // global position of the tile's digit run minus its tile start
  __shared__ uint32_t wsum[kWaves];
  __shared__ uint32_t sk[kRsTile], sv[kRsTile];
  int src;
  if (!rs_pass_buffers(passes, pass, src)) return;
  const uint32_t* Kin = src ? K1 : K0;
  const uint32_t* Vin = src ? V1 : V0;
  uint32_t* Kout = src ? K0 : K1;
  uint32_t* Vout = src ? V0 : V1;
  const int shift = 8 * pass;
  int64_t tile0, seg, jt;
  int tile_len;
  rs_tile(g, blockIdx.x, tile0, tile_len, seg, jt);
  const int64_t seg_end = seg * g.seg_len + g.seg_len;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int d = lane; d < 256; d += 64) wcnt[wave][d] = 0u;
  const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint32_t k[kKeysPerLane], v[kKeysPerLane], rk[kKeysPerLane];
#pragma unroll
  for (int r = 0; r < kKeysPerLane; ++r) {
    const int idx = wave * kPerWave + r * 64 + lane;  // index in the tile
    const bool valid = idx < tile_len;
    k[r] = valid ? Kin[tile0 + idx] : 0u;
    v[r] = valid ? Vin[tile0 + idx] : 0u;
    const uint32_t d = (k[r] >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int bit = 0; bit < 8; ++bit) {
      const bool on = (d >> bit) & 1u;
      const uint64_t m = __ballot(on);
      peers &= on ? m : ~m;
    }
    // rank among this wave's earlier keys of digit d; the wave's LDS ops run
    // in program order, so every lane reads the count before the leader adds
    const uint32_t before = valid ? wcnt[wave][d] : 0u;
    rk[r] = before + static_cast<uint32_t>(__popcll(peers & lt_mask));
    if (valid && (peers & lt_mask) == 0ull) wcnt[wave][d] = before + static_cast<uint32_t>(__popcll(peers));
  }
  __syncthreads();
  // thread d: the tile's count of digit d, per-wave prefixes, block scan over d
  const int d = tid;
  uint32_t c[kWaves], tot = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    c[w] = wcnt[w][d];
    tot += c[w];
  }
  uint32_t incl = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t start = incl - tot;
  for (int w = 0; w < wave; ++w) start += wsum[w];
  uint32_t run = start;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    wcnt[w][d] = run;
    run += c[w];
  }
  gofs[d] = static_cast<int64_t>(offs[rs_hist_index(g, seg, jt, d)]) - start;
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kKeysPerLane; ++r) {
    const int idx = wave * kPerWave + r * 64 + lane;
    if (idx < tile_len) {
      const uint32_t pos = wcnt[wave][(k[r] >> shift) & 255u] + rk[r];
      if (pos < static_cast<uint32_t>(tile_len)) {
        sk[pos] = k[r];
        sv[pos] = v[r];
      }
    }
  }
  __syncthreads();
  const int64_t seg_begin = seg * g.seg_len;
  for (int j = tid; j < tile_len; j += kRsThreads) {
    const uint32_t key = sk[j];
    const int64_t pos = gofs[(key >> shift) & 255u] + j;
    if (pos >= seg_begin && pos < seg_end) {
      Kout[pos] = key;
      Vout[pos] = sv[j];
    }
  }
}

// Leaf ranges from the sorted keys; nvalid = first off-leaf position.
__global__ void rs_bounds_kernel(const uint32_t* __restrict__ K, int64_t N, uint32_t total_nodes,
                                 uint32_t* __restrict__ bstart, uint32_t* __restrict__ bend,
                                 uint32_t* __restrict__ nvalid) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t k = K[p];
  const uint32_t prev = p > 0 ? K[p - 1] : 0xffffffffu;
  if (k == total_nodes) {
    if (prev != total_nodes) *nvalid = static_cast<uint32_t>(p);
    return;
  }
  if (prev != k) bstart[k] = static_cast<uint32_t>(p);
  const uint32_t next = p + 1 < N ? K[p + 1] : 0xffffffffu;
  if (next != k) bend[k] = static_cast<uint32_t>(p + 1);
  if (p + 1 == N) *nvalid = static_cast<uint32_t>(N);
}

// Leaf ranges for the per-tree layout: sorted local keys in K (buffer 0)
// back to global leaf ids (mbucket, total_nodes for an off-leaf slot), the
// ranges of every leaf, nvalid = N (off-leaf slots end each tree's segment).
__global__ void rs_bounds_tree_kernel(const uint32_t* __restrict__ K, int64_t N, int64_t n, double inv_n,
                                      uint32_t total_nodes, const uint32_t* __restrict__ tmin,
                                      const uint32_t* __restrict__ info, uint32_t* __restrict__ mbucket,
                                      uint32_t* __restrict__ bstart, uint32_t* __restrict__ bend,
                                      uint32_t* __restrict__ nvalid) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t sentinel = info[1];
  // slot p's tree; its neighbours are in the same tree except across a
  // segment boundary
  const int64_t t = slot_tree(p, n, inv_n);
  const auto global = [&](int64_t q, int64_t tq) -> uint32_t {
    const uint32_t k = K[q];
    return k == sentinel ? total_nodes : k + tmin[tq];
  };
  const uint32_t g = global(p, t);
  mbucket[p] = g;
  if (p == 0) *nvalid = static_cast<uint32_t>(N);
  if (g == total_nodes) return;
  if (p == 0 || global(p - 1, p == t * n ? t - 1 : t) != g) bstart[g] = static_cast<uint32_t>(p);
  if (p + 1 == N || global(p + 1, p + 1 == (t + 1) * n ? t + 1 : t) != g) bend[g] = static_cast<uint32_t>(p + 1);
}

// members[p] = (i << 1) | bag; pe[s] = (p, end of its leaf | bag << 31),
// (0, 0) off-leaf.
__global__ void rs_members_kernel(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int64_t N, uint32_t n,
                                  double inv_n, uint32_t total_nodes, const uint32_t* __restrict__ bend,
                                  uint32_t* __restrict__ members, uint16_t* __restrict__ members16,
                                  uint2* __restrict__ pe, uint16_t* __restrict__ members_w = nullptr) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t v = V[p];
  const uint32_t s = v >> 1, bag = v & 1u;
  const uint32_t i = s - static_cast<uint32_t>(slot_tree(s, n, inv_n)) * n;
  members[p] = (i << 1) | bag;
  if (members16) members16[p] = static_cast<uint16_t>((i << 1) | bag);
  if (members_w) members_w[p] = static_cast<uint16_t>(((i & (kWinSpan - 1)) << 1) | bag);
  const uint32_t k = K[p];
  pe[s] = k == total_nodes ? make_uint2(0u, 0u) : make_uint2(static_cast<uint32_t>(p), bend[k] | (bag << 31));
}

// Chunk-boundary table of the long-line counts (wtab above): slot p of leaf
// b, column j, is the first of its run at or past boundary k + 1 for every
// boundary after its predecessor's column and at most j; the last member of
// the run also fills the boundaries after it with the run end.  Every
// boundary of every non-empty leaf is written once.
__global__ void win_table_kernel(const uint32_t* __restrict__ mbucket, const uint32_t* __restrict__ members,
                                 const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend, int64_t N,
                                 uint32_t total_nodes, int nb, int shift, uint32_t* __restrict__ wtab) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t b = mbucket[p];
  if (b >= total_nodes) return;  // off-leaf slot
  const uint32_t bs = bstart[b], be = bend[b];
  const int cj = static_cast<int>((members[p] >> 1) >> shift);
  const int first = static_cast<uint32_t>(p) == bs ? 1 : static_cast<int>((members[p - 1] >> 1) >> shift) + 1;
  uint32_t* row = wtab + static_cast<int64_t>(b) * nb;
  for (int c = first; c <= cj && c <= nb; ++c) row[c - 1] = static_cast<uint32_t>(p);
  if (static_cast<uint32_t>(p) + 1 == be)
    for (int c = cj + 1; c <= nb; ++c) row[c - 1] = be;
}

// pe and the leaf ids transposed to line-major order (64 x 64 tiles through
// LDS): rpe[line * ntree + t] = pe[t * n + line], rleaf likewise from keys.
__global__ __launch_bounds__(256) void win_rows_kernel(const uint2* __restrict__ pe, const int32_t* __restrict__ keys,
                                                       int64_t n, int32_t ntree, uint2* __restrict__ rpe,
                                                       uint32_t* __restrict__ rleaf) {
  __shared__ uint2 tp[64][65];
  __shared__ uint32_t tl[64][65];
  const int64_t l0 = static_cast<int64_t>(blockIdx.x) * 64;
  const int t0 = static_cast<int>(blockIdx.y) * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int tt = e >> 6, ll = e & 63;  // consecutive threads: consecutive lines
    const int64_t line = l0 + ll;
    const int t = t0 + tt;
    if (line < n && t < ntree) {
      const int64_t s = static_cast<int64_t>(t) * n + line;
      tp[tt][ll] = pe[s];
      tl[tt][ll] = static_cast<uint32_t>(keys[s]) >> 1;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int ll = e >> 6, tt = e & 63;  // consecutive threads: consecutive trees
    const int64_t line = l0 + ll;
    const int t = t0 + tt;
    if (line < n && t < ntree) {
      rpe[line * ntree + t] = tp[tt][ll];
      rleaf[line * ntree + t] = tl[tt][ll];
    }
  }
}

// Segmented form for a batch of problems laid end to end: slots
// [seg_start[b], seg_start[b+1]) belong to problem b with seg_n[b] samples,
// so the sample index is (s - seg_start[b]) % seg_n[b].
__global__ void rs_members_seg_kernel(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int64_t N,
                                      const int64_t* __restrict__ seg_start, const int64_t* __restrict__ seg_n,
                                      int nseg, uint32_t total_nodes, const uint32_t* __restrict__ bend,
                                      uint32_t* __restrict__ members, uint16_t* __restrict__ members16,
                                      uint2* __restrict__ pe) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t v = V[p];
  const uint32_t s = v >> 1, bag = v & 1u;
  int lo = 0, hi = nseg - 1;  // last b with seg_start[b] <= s
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_start[mid] <= static_cast<int64_t>(s))
      lo = mid;
    else
      hi = mid - 1;
  }
  const uint32_t i = static_cast<uint32_t>((static_cast<int64_t>(s) - seg_start[lo]) % seg_n[lo]);
  members[p] = (i << 1) | bag;
  if (members16) members16[p] = static_cast<uint16_t>((i << 1) | bag);  // meaningful where seg_n <= 32768
  const uint32_t k = K[p];
  pe[s] = k == total_nodes ? make_uint2(0u, 0u) : make_uint2(static_cast<uint32_t>(p), bend[k] | (bag << 31));
}

// Graph-safe fill (no memset nodes in captured graphs): 16-byte stores from
// the first 16-byte boundary, single words for the unaligned head and tail.
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, int64_t len, uint32_t v) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t head = min<int64_t>(len, ((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / 4);
  const int64_t n4 = (len - head) / 4;
  uint4* p4 = reinterpret_cast<uint4*>(p + head);
  const uint4 v4 = make_uint4(v, v, v, v);
  for (int64_t e = gid; e < n4; e += stride) p4[e] = v4;
  if (gid < head) p[gid] = v;
  for (int64_t e = head + 4 * n4 + gid; e < len; e += stride) p[e] = v;
}

struct BucketWs {
  uint32_t* bstart;
  uint32_t* bend;
  uint32_t* members;
  uint16_t* members16;  // nullptr unless n <= kMembers16MaxN
  uint16_t* members_w;  // nullptr unless the long-line window arrays are carved
  uint32_t* wtab;       // chunk-boundary table, nb entries per node (or nullptr)
  uint2* rpe;           // pe by line: rpe[line * ntree + t] (with the window arrays)
  uint32_t* rleaf;      // leaf of (line, tree) by line (with the window arrays)
  int nb;               // chunk boundaries per node in wtab
  uint32_t* mbucket;  // sorted keys (one of the two key buffers)
  uint2* pe;
  uint32_t* nvalid;
  uint32_t* K[2];
  uint32_t* V[2];
  uint32_t* hist;
  uint32_t* part;
  uint32_t* tmin;  // per-tree smallest leaf id (per-tree sort)
  uint32_t* info;  // largest range, sentinel, pass count (per-tree sort)
  bool per_tree;
  unsigned long long* stat;
  int64_t ntiles;
  int64_t nparts;
  int passes;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

int radix_passes(int64_t total_nodes) {
  int bits = 0;
  while ((1LL << bits) <= total_nodes) ++bits;  // keys range over [0, total_nodes]
  return (bits + 7) / 8;
}

// Radix tiles: ntree segments of n slots each (per-tree sort; never fewer
// tiles than the global layout of the same slots needs).
int64_t rs_tiles_per_tree(int64_t n) { return (n + kRsTile - 1) / kRsTile; }

// windows: the count calls' workspace, which also holds the long-line window
// arrays (members_w, wtab) when win_enabled(n).
size_t bucket_ws_bytes(int64_t n, int32_t ntree, int64_t total_nodes, bool windows = false) {
  const int64_t slots = n * static_cast<int64_t>(ntree);
  const int64_t ntiles = ntree * rs_tiles_per_tree(n);
  const int64_t hlen = 256 * ntiles;
  const int64_t nparts = (hlen + kScanChunk - 1) / kScanChunk;
  return align256(4 * total_nodes) * 2 + align256(4 * slots) * 5 + align256(8 * slots) + align256(4 * hlen) +
         align256(4 * (nparts + 1)) + 512 + align256(4 * static_cast<int64_t>(ntree)) + 256 +
         (n <= kMembers16MaxN ? align256(2 * slots) : 0) +
         (windows && win_enabled(n) ? align256(2 * slots) + align256(4 * total_nodes * win_bounds(n)) +
                                          align256(8 * slots) + align256(4 * slots)
                                    : 0);
}

// per_tree: the one-problem layout (segments of n slots per tree); false for
// the batched build (one segment of all slots).
BucketWs bucket_ws_carve(void* ws, int64_t n, int32_t ntree, int64_t total_nodes, bool per_tree = true,
                         bool windows = false) {
  BucketWs w;
  char* p = static_cast<char*>(ws);
  const int64_t slots = n * static_cast<int64_t>(ntree);
  w.per_tree = per_tree;
  w.ntiles = per_tree ? ntree * rs_tiles_per_tree(n) : (slots + kRsTile - 1) / kRsTile;
  const int64_t hlen = 256 * w.ntiles;
  w.nparts = (hlen + kScanChunk - 1) / kScanChunk;
  w.passes = radix_passes(total_nodes);
  w.bstart = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * total_nodes);
  w.bend = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * total_nodes);
  w.members = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * slots);
  for (int b = 0; b < 2; ++b) {
    w.K[b] = reinterpret_cast<uint32_t*>(p);
    p += align256(4 * slots);
    w.V[b] = reinterpret_cast<uint32_t*>(p);
    p += align256(4 * slots);
  }
  w.pe = reinterpret_cast<uint2*>(p);
  p += align256(8 * slots);
  w.hist = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * 256 * ntree * rs_tiles_per_tree(n));
  const int64_t nparts_max = (256 * ntree * rs_tiles_per_tree(n) + kScanChunk - 1) / kScanChunk;
  w.part = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * (nparts_max + 1));
  w.nvalid = reinterpret_cast<uint32_t*>(p);
  w.stat = reinterpret_cast<unsigned long long*>(p + 256);
  p += 512;
  w.tmin = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * static_cast<int64_t>(ntree));
  w.info = reinterpret_cast<uint32_t*>(p);
  p += 256;
  w.members16 = n <= kMembers16MaxN ? reinterpret_cast<uint16_t*>(p) : nullptr;
  if (w.members16) p += align256(2 * slots);
  w.members_w = nullptr;
  w.wtab = nullptr;
  w.rpe = nullptr;
  w.rleaf = nullptr;
  w.nb = 0;
  if (per_tree && windows && win_enabled(n)) {
    w.nb = win_bounds(n);
    w.members_w = reinterpret_cast<uint16_t*>(p);
    p += align256(2 * slots);
    w.wtab = reinterpret_cast<uint32_t*>(p);
    p += align256(4 * total_nodes * w.nb);
    w.rpe = reinterpret_cast<uint2*>(p);
    p += align256(8 * slots);
    w.rleaf = reinterpret_cast<uint32_t*>(p);
  }
  // the per-tree sort ends in buffer 0 and writes the global leaf ids into K[1]
  w.mbucket = per_tree ? w.K[1] : w.K[w.passes & 1];
  return w;
}

unsigned grid_for(int64_t work, int threads) {
  int64_t g = (work + threads - 1) / threads;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

int fill_u32(hipStream_t st, void* p, int64_t len, uint32_t v) {
  if (len <= 0) return RFX_OK;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for((len + 3) / 4, 256)), dim3(256), 0, st, static_cast<uint32_t*>(p),
                     len, v);
  return rfx_launch_status();
}

// The radix passes over the prepared keys: `passes` is the device pass count
// of the per-tree sort (surplus passes return at once) or nullptr.
int rs_passes(hipStream_t st, const BucketWs& w, const RsTiles& g, const int32_t* passes) {
  const int64_t hlen = 256 * g.ntiles;
  for (int pass = 0; pass < w.passes; ++pass) {
    hipLaunchKernelGGL(rs_hist_kernel, dim3(static_cast<unsigned>(g.ntiles)), dim3(kRsThreads), 0, st, w.K[0], w.K[1],
                       passes, pass, g, w.hist);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.hist,
                       hlen, w.part, passes, pass);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, w.part, w.nparts, passes, pass);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.hist,
                       hlen, w.part, w.hist, passes, pass);
    hipLaunchKernelGGL(rs_scatter_kernel, dim3(static_cast<unsigned>(g.ntiles)), dim3(kRsThreads), 0, st, w.K[0],
                       w.K[1], w.V[0], w.V[1], passes, pass, g, w.hist);
  }
  return rfx_launch_status();
}

// Sort of N slots by leaf and the leaf ranges (everything but members/pe),
// one segment of all slots; sorted keys end in K[passes & 1] = mbucket.
int bucket_sort(hipStream_t st, const int32_t* keys, int64_t N, int64_t total_nodes, const BucketWs& w) {
  const uint32_t tn = static_cast<uint32_t>(total_nodes);
  int rc = fill_u32(st, w.bstart, total_nodes, 0u);
  if (rc) return rc;
  rc = fill_u32(st, w.bend, total_nodes, 0u);
  if (rc) return rc;
  hipLaunchKernelGGL(rs_prep_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, keys, N, tn, w.K[0], w.V[0]);
  const RsTiles g{N, (N + kRsTile - 1) / kRsTile, (N + kRsTile - 1) / kRsTile};
  rc = rs_passes(st, w, g, nullptr);
  if (rc) return rc;
  const int fin = w.passes & 1;
  hipLaunchKernelGGL(rs_bounds_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[fin], N, tn, w.bstart, w.bend,
                     w.nvalid);
  return rfx_launch_status();
}

// Per-tree sort of the keys [T][n]: sorted local keys end in K[0], V[0]; the
// global leaf ids go to mbucket (K[1]).
int bucket_sort_trees(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                      const BucketWs& w) {
  const int64_t N = n * static_cast<int64_t>(ntree);
  const uint32_t tn = static_cast<uint32_t>(total_nodes);
  int rc = fill_u32(st, w.bstart, total_nodes, 0u);
  if (!rc) rc = fill_u32(st, w.bend, total_nodes, 0u);
  if (!rc) rc = fill_u32(st, w.info, 4, 0u);
  if (rc) return rc;
  hipLaunchKernelGGL(rs_tree_range_kernel, dim3(static_cast<unsigned>(ntree)), dim3(256), 0, st, keys, n, w.tmin,
                     w.info);
  hipLaunchKernelGGL(rs_tree_passes_kernel, dim3(1), dim3(1), 0, st, w.info);
  hipLaunchKernelGGL(rs_prep_tree_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, keys, n, N,
                     1.0 / static_cast<double>(n), w.tmin, w.info,
                     w.K[0], w.K[1], w.V[0], w.V[1]);
  const RsTiles g{n, rs_tiles_per_tree(n), w.ntiles};
  rc = rs_passes(st, w, g, reinterpret_cast<const int32_t*>(w.info + 2));
  if (rc) return rc;
  hipLaunchKernelGGL(rs_bounds_tree_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[0], N, n,
                     1.0 / static_cast<double>(n), tn, w.tmin,
                     w.info, w.mbucket, w.bstart, w.bend, w.nvalid);
  return rfx_launch_status();
}

// ---------------------------------------------------------------------------
// One-workgroup-per-tree build for n <= kTreeLdsMaxN: a stable counting sort
// of the tree's n slots by local leaf id in LDS, writing the same arrays as
// the radix path (mbucket, V[0], members, members16, pe, bstart, bend,
// nvalid) bit for bit.
//
// local = leaf id - tmin for a slot on a leaf, and nleaf (one past the
// largest) for an off-leaf slot, so those end the tree's segment.  A counting
// pass sorts by a digit: the samples are cut into B contiguous blocks
// (B <= 8 wavefronts) and 16-bit counters c[digit][block] live in LDS, two to
// a 32-bit word, beside the 16-bit permutation perm[position] = slot << 1 |
// bag that the pass fills:
//   A  each wavefront counts its block with one LDS atomic per sample;
//   B  one exclusive scan over the counters in digit-major order turns them
//      into first positions (< n, so still 16-bit);
//   C  each wavefront walks its block again, 64 samples a step: the lanes
//      of a step with the same digit are found with one ballot per digit bit;
//      a slot's position is its counter plus its rank among them, and the
//      lowest such lane advances the counter.
// Samples are taken in order throughout, so a pass is stable.  The sorted
// arrays are then written along the positions (coalesced) from the
// permutation, and pe along the slots.  A tree's ids span at most its node
// count; when the counters cannot hold that many digits (trees grown on more
// samples than are assigned here), a first pass on the low 15 bits leaves
// its order in the spare key buffer and a second pass sorts by the rest.
constexpr int64_t kTreeLdsMaxN = 32767;
constexpr int kTreeLdsThreads = 512;
constexpr int kTreeLdsWaves = kTreeLdsThreads / 64;
constexpr uint32_t kTreeLdsBytes = 159 * 1024;  // counters + the 16-bit permutation of the tree's slots
// dynamic request + rs_tree_lds_kernel's static red_lo / red_hi / wsum
static_assert(kTreeLdsBytes + 3 * 4 * kTreeLdsWaves <= kLdsLimitBytes, "per-tree sort exceeds the LDS");
constexpr int kTreeLowBits = 15;                 // low digit of the two-pass form

__device__ __forceinline__ uint32_t lds_half(const uint32_t* words, uint32_t e) {
  return (words[e >> 1] >> (16 * (e & 1u))) & 0xffffu;
}

// Lanes of the wavefront (among `active`) whose digit equals this lane's,
// comparing the low nbits bits.
__device__ __forceinline__ uint64_t same_digit_lanes(uint32_t d, int nbits, uint64_t active) {
  uint64_t m = active;
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t on = __ballot(bit);
    m &= bit ? on : ~on;
  }
  return m;
}

// One stable counting pass over elements q = 0 .. n - 1: get(q) fetches an
// element's token (kTlPrefetch steps of tokens are fetched before they are
// used, so the loads overlap), digit(token) < ndig, and place(q, token, pos)
// receives its position.  On return the counter of (digit d, last block)
// holds the end of d's positions.  All threads must call.
// keys in flight per thread in the passes over a tree's keys; the full
// build at n = 20,000, T = 500 takes 0.207 ms with 4, 0.190 with 8, 0.234 with 16
#ifndef RFX_TL_PREFETCH
#define RFX_TL_PREFETCH 8
#endif
constexpr int kTlPrefetch = RFX_TL_PREFETCH;

template <class GetF, class DigitF, class PlaceF>
__device__ void tl_count_pass(uint32_t* words, uint32_t* wsum, int64_t n, uint32_t ndig, int B, GetF get,
                              DigitF digit, PlaceF place) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int nbits = 0;
  while ((1u << nbits) < ndig) ++nbits;
  const uint32_t nwords = (ndig * static_cast<uint32_t>(B) + 1) >> 1;
  const int64_t steps = (n + 63) / 64;
  const int64_t steps_per = (steps + B - 1) / B;
  const int64_t st0 = wave * steps_per, st1 = min(steps, (wave + 1) * steps_per);
  __syncthreads();
  for (uint32_t q = tid; q < nwords; q += kTreeLdsThreads) words[q] = 0u;
  __syncthreads();
  if (wave < B) {  // A
    for (int64_t sb = st0; sb < st1; sb += kTlPrefetch) {
      uint32_t tok[kTlPrefetch];
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        const int64_t q = (sb + u) * 64 + lane;
        tok[u] = (sb + u < st1 && q < n) ? get(q) : 0u;
      }
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        if (sb + u >= st1) break;  // uniform across the wavefront
        const int64_t q = (sb + u) * 64 + lane;
        if (q < n) {
          const uint32_t e = digit(tok[u]) * B + wave;
          atomicAdd(&words[e >> 1], 1u << (16 * (e & 1u)));  // counts need no ranks
        }
      }
    }
  }
  __syncthreads();
  // B: each thread owns a contiguous run of whole words
  const uint32_t per = (nwords + kTreeLdsThreads - 1) / kTreeLdsThreads;
  const uint32_t w0 = min(nwords, tid * per), w1 = min(nwords, w0 + per);
  uint32_t local = 0;
  for (uint32_t q = w0; q < w1; ++q) local += (words[q] & 0xffffu) + (words[q] >> 16);
  uint32_t incl = local;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t run = incl - local;
  for (int q = 0; q < wave; ++q) run += wsum[q];
  for (uint32_t q = w0; q < w1; ++q) {
    const uint32_t a = words[q] & 0xffffu, b = words[q] >> 16;
    words[q] = run | ((run + a) << 16);
    run += a + b;
  }
  __syncthreads();
  if (wave < B) {  // C
    for (int64_t sb = st0; sb < st1; sb += kTlPrefetch) {
      uint32_t tok[kTlPrefetch];
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        const int64_t q = (sb + u) * 64 + lane;
        tok[u] = (sb + u < st1 && q < n) ? get(q) : 0u;
      }
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        if (sb + u >= st1) break;
        const int64_t q = (sb + u) * 64 + lane;
        const bool in = q < n;
        const uint32_t d = in ? digit(tok[u]) : 0u;
        const uint64_t m = same_digit_lanes(d, nbits, __ballot(in));
        const uint32_t e = d * B + wave;
        const uint32_t first = in ? lds_half(words, e) : 0u;
        const uint32_t rank = static_cast<uint32_t>(__popcll(m & ((1ull << lane) - 1)));
        if (in && rank == 0) atomicAdd(&words[e >> 1], static_cast<uint32_t>(__popcll(m)) << (16 * (e & 1u)));
        if (in) place(q, tok[u], first + rank);
      }
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(kTreeLdsThreads) void rs_tree_lds_kernel(
    const int32_t* __restrict__ keys, int64_t n, uint32_t total_nodes, uint32_t* __restrict__ order,
    uint32_t* __restrict__ mbucket, uint32_t* __restrict__ V0, uint32_t* __restrict__ members,
    uint16_t* __restrict__ members16, uint2* __restrict__ pe, uint32_t* __restrict__ bstart,
    uint32_t* __restrict__ bend, uint32_t* __restrict__ nvalid) {
  extern __shared__ uint32_t tl_words[];
  __shared__ uint32_t red_lo[kTreeLdsWaves], red_hi[kTreeLdsWaves], wsum[kTreeLdsWaves];
  const int64_t t = blockIdx.x;
  const int64_t base = t * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int32_t* kt = keys + base;
  // LDS: perm[pos] = i << 1 | bag (16 bits, n <= 32767), then the counters
  const uint32_t perm_words = static_cast<uint32_t>((n + 1) / 2);
  uint16_t* perm = reinterpret_cast<uint16_t*>(tl_words);
  uint32_t* words = tl_words + perm_words;
  const uint32_t emax = 2 * (kTreeLdsBytes / 4 - perm_words);  // counter entries that fit
  uint32_t lo = 0xffffffffu, hi = 0u;
  // kTlPrefetch keys in flight per thread (a plain loop waits for each load)
  for (int64_t i0 = tid; i0 < n; i0 += static_cast<int64_t>(kTreeLdsThreads) * kTlPrefetch) {
    int32_t k[kTlPrefetch];
#pragma unroll
    for (int u = 0; u < kTlPrefetch; ++u) {
      const int64_t i = i0 + static_cast<int64_t>(u) * kTreeLdsThreads;
      k[u] = i < n ? kt[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kTlPrefetch; ++u) {
      if (k[u] >= 0) {
        lo = min(lo, static_cast<uint32_t>(k[u]) >> 1);
        hi = max(hi, static_cast<uint32_t>(k[u]) >> 1);
      }
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = min(lo, static_cast<uint32_t>(__shfl_xor(lo, d, 64)));
    hi = max(hi, static_cast<uint32_t>(__shfl_xor(hi, d, 64)));
  }
  if (lane == 0) {
    red_lo[wave] = lo;
    red_hi[wave] = hi;
  }
  __syncthreads();
  lo = red_lo[0];
  hi = red_hi[0];
  for (int q = 1; q < kTreeLdsWaves; ++q) {
    lo = min(lo, red_lo[q]);
    hi = max(hi, red_hi[q]);
  }
  if (t == 0 && tid == 0) *nvalid = static_cast<uint32_t>(n * gridDim.x);
  const uint32_t nleaf = lo <= hi ? hi - lo + 1 : 0u;  // local ids 0 .. nleaf - 1; nleaf = off-leaf
  const auto local_of = [&](int32_t k) -> uint32_t { return k >= 0 ? (static_cast<uint32_t>(k) >> 1) - lo : nleaf; };
  // a slot's position: pe.x now (written along the slots), the permutation in LDS
  const auto place_final = [&](uint32_t i, int32_t k, uint32_t pos) {
    if (pos >= n) return;  // cannot happen: positions come from a scan of n counts
    const uint32_t bag = k >= 0 ? (static_cast<uint32_t>(k) & 1u) : 0u;
    perm[pos] = static_cast<uint16_t>((i << 1) | bag);
    pe[base + i].x = static_cast<uint32_t>(base + pos);
  };
  const auto get_key = [&](int64_t q) -> uint32_t { return static_cast<uint32_t>(kt[q]); };
  const bool one_pass = nleaf + 1 <= emax;
  int B = 1;
  if (one_pass) {
    const uint32_t ndig = nleaf + 1;
    B = static_cast<int>(min<uint32_t>(kTreeLdsWaves, emax / ndig));
    tl_count_pass(
        words, wsum, n, ndig, B, get_key, [&](uint32_t k) { return local_of(static_cast<int32_t>(k)); },
        [&](int64_t q, uint32_t k, uint32_t pos) { place_final(static_cast<uint32_t>(q), static_cast<int32_t>(k), pos); });
  } else {
    // leaf ids spanning more values than the counters hold: the low bits,
    // with the order kept in the spare key buffer, then the rest
    uint32_t* ord = order + base;
    constexpr uint32_t kLow = 1u << kTreeLowBits;
    tl_count_pass(
        words, wsum, n, kLow, 1, get_key,
        [&](uint32_t k) { return local_of(static_cast<int32_t>(k)) & (kLow - 1); },
        [&](int64_t q, uint32_t, uint32_t pos) {
          if (pos < n) ord[pos] = static_cast<uint32_t>(q);
        });
    const uint32_t ndig2 = (nleaf >> kTreeLowBits) + 1;
    tl_count_pass(
        words, wsum, n, ndig2, static_cast<int>(min<uint32_t>(kTreeLdsWaves, emax / ndig2)),
        [&](int64_t q) -> uint32_t { return ord[q]; },
        [&](uint32_t i) { return local_of(kt[i]) >> kTreeLowBits; },
        [&](int64_t, uint32_t i, uint32_t pos) { place_final(i, kt[i], pos); });
  }
  // the sorted segment, written along the positions (kTlPrefetch positions
  // per thread at a time, so the key gathers overlap)
  for (int64_t p0 = tid; p0 < n; p0 += static_cast<int64_t>(kTreeLdsThreads) * kTlPrefetch) {
    uint32_t v[kTlPrefetch];
    int32_t k[kTlPrefetch];
#pragma unroll
    for (int u = 0; u < kTlPrefetch; ++u) {
      const int64_t p = p0 + static_cast<int64_t>(u) * kTreeLdsThreads;
      v[u] = p < n ? perm[p] : 0u;
      k[u] = (mbucket && p < n && (v[u] >> 1) < n) ? kt[v[u] >> 1] : -1;
    }
#pragma unroll
    for (int u = 0; u < kTlPrefetch; ++u) {
      const int64_t p = p0 + static_cast<int64_t>(u) * kTreeLdsThreads;
      const uint32_t i = v[u] >> 1, bag = v[u] & 1u;
      if (p >= n || i >= n) continue;  // i < n always: every position is filled by the pass
      if (mbucket) mbucket[base + p] = k[u] >= 0 ? (static_cast<uint32_t>(k[u]) >> 1) : total_nodes;
      if (V0) V0[base + p] = (static_cast<uint32_t>(base + i) << 1) | bag;
      if (members) members[base + p] = v[u];
      if (members16) members16[base + p] = static_cast<uint16_t>(v[u]);
    }
  }
  if (one_pass) {
    // leaf d holds [end(d - 1), end(d)), end(d) = the counter of (d, last block)
    for (uint32_t d = tid; d < nleaf; d += kTreeLdsThreads) {
      const uint32_t s0 = d > 0 ? lds_half(words, (d - 1) * B + (B - 1)) : 0u;
      const uint32_t s1 = lds_half(words, d * B + (B - 1));
      if (s1 > s0 && bstart) {
        bstart[lo + d] = static_cast<uint32_t>(base + s0);
        bend[lo + d] = static_cast<uint32_t>(base + s1);
      }
    }
    for (int64_t i0 = tid; i0 < n; i0 += static_cast<int64_t>(kTreeLdsThreads) * kTlPrefetch) {
      int32_t k[kTlPrefetch];
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        const int64_t i = i0 + static_cast<int64_t>(u) * kTreeLdsThreads;
        k[u] = i < n ? kt[i] : -1;
      }
#pragma unroll
      for (int u = 0; u < kTlPrefetch; ++u) {
        const int64_t i = i0 + static_cast<int64_t>(u) * kTreeLdsThreads;
        if (i >= n) continue;
        if (k[u] < 0)
          pe[base + i] = make_uint2(0u, 0u);
        else
          pe[base + i].y = static_cast<uint32_t>(base + lds_half(words, local_of(k[u]) * B + (B - 1))) |
                           ((static_cast<uint32_t>(k[u]) & 1u) << 31);
      }
    }
    return;
  }
  // two-pass form: leaf ranges and pe ends from the sorted positions
  const auto leaf_at = [&](int64_t p) -> int64_t {
    const uint32_t i = perm[p] >> 1u;
    if (i >= n) return -1;  // cannot happen, as above
    const int32_t k = kt[i];
    return k >= 0 ? static_cast<int64_t>(static_cast<uint32_t>(k) >> 1) : -1;
  };
  for (int64_t p = tid; p < n; p += kTreeLdsThreads) {
    const int64_t g = leaf_at(p);
    if (g < 0) continue;
    if (bstart && (p == 0 || leaf_at(p - 1) != g)) bstart[g] = static_cast<uint32_t>(base + p);
    if (p + 1 == n || leaf_at(p + 1) != g) bend[g] = static_cast<uint32_t>(base + p + 1);
  }
  __syncthreads();
  for (int64_t i = tid; i < n; i += kTreeLdsThreads) {
    const int32_t k = kt[i];
    if (k < 0)
      pe[base + i] = make_uint2(0u, 0u);
    else
      pe[base + i].y = bend[static_cast<uint32_t>(k) >> 1] | ((static_cast<uint32_t>(k) & 1u) << 31);
  }
}

// Full bucket build from the leaf keys [T][n] (see the file comment).
// lean: for a caller that reads only pe, bstart, mbucket and members16 (the
// whole-line and line-chunk GATHER kernels), the per-tree build skips the
// 32-bit members and the sorted slot values.  rows_upper (with lean): for the
// UPPER line kernels, which read only pe and members16, it also skips the
// leaf bounds (and their clears) and mbucket, with the key gather that feeds
// it; bend stays the two-pass form's scratch, written for every leaf present
// before it is read.  The skipped arrays are left as they were.
int bucket_build(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                 const BucketWs& w, bool lean = false, bool rows_upper = false) {
  const int64_t N = n * static_cast<int64_t>(ntree);
  if (!w.per_tree) return RFX_ERR_INVALID;
  if (n <= kTreeLdsMaxN) {
    if (lean && !w.members16) return RFX_ERR_INVALID;
    rows_upper = rows_upper && lean;
    if (!rows_upper) {
      int rc = fill_u32(st, w.bstart, total_nodes, 0u);
      if (!rc) rc = fill_u32(st, w.bend, total_nodes, 0u);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(rs_tree_lds_kernel, dim3(static_cast<unsigned>(ntree)), dim3(kTreeLdsThreads), kTreeLdsBytes,
                       st, keys, n, static_cast<uint32_t>(total_nodes), w.K[0], rows_upper ? nullptr : w.mbucket,
                       lean ? nullptr : w.V[0], lean ? nullptr : w.members, w.members16, w.pe,
                       rows_upper ? nullptr : w.bstart, w.bend, w.nvalid);
    const int rc = rfx_launch_status();
    note_bucket_build(w.bstart, BucketBuildRecord{n, ntree, total_nodes, rc == RFX_OK && !lean, false});
    return rc;
  }
  int rc = bucket_sort_trees(st, keys, n, ntree, total_nodes, w);
  if (rc) return rc;
  hipLaunchKernelGGL(rs_members_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.mbucket, w.V[0], N,
                     static_cast<uint32_t>(n), 1.0 / static_cast<double>(n), static_cast<uint32_t>(total_nodes), w.bend,
                     w.members, w.members16, w.pe, w.members_w);
  if (w.wtab)
    hipLaunchKernelGGL(win_table_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.mbucket, w.members,
                       w.bstart, w.bend, N, static_cast<uint32_t>(total_nodes), w.nb, win_shift(), w.wtab);
  if (w.rpe)
    hipLaunchKernelGGL(win_rows_kernel, dim3(rfx_ceil_div(n, 64), rfx_ceil_div(ntree, 64)), dim3(256), 0, st, w.pe,
                       keys, n, ntree, w.rpe, w.rleaf);
  rc = rfx_launch_status();
  note_bucket_build(w.bstart, BucketBuildRecord{n, ntree, total_nodes, rc == RFX_OK, rc == RFX_OK && w.wtab});
  return rc;
}

// Bucket build over a batch of problems whose key arrays lie end to end
// (problem b: slots [seg_start[b], seg_start[b+1]), keys [T_b][seg_n[b]],
// node ids already offset so that all leaves are distinct).
int bucket_build_segmented(hipStream_t st, const int32_t* keys, int64_t N, int64_t total_nodes,
                           const int64_t* seg_start_dev, const int64_t* seg_n_dev, int nseg, const BucketWs& w,
                           uint16_t* members16 = nullptr) {
  if (w.per_tree) return RFX_ERR_INVALID;
  int rc = bucket_sort(st, keys, N, total_nodes, w);
  if (rc) return rc;
  const int fin = w.passes & 1;
  hipLaunchKernelGGL(rs_members_seg_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[fin], w.V[fin], N,
                     seg_start_dev, seg_n_dev, nseg, static_cast<uint32_t>(total_nodes), w.bend, w.members, members16,
                     w.pe);
  return rfx_launch_status();
}

}  // namespace
