// Leaf buckets of the (tree, sample) slots, shared by the co-occurrence
// counts and the matrix-free proximity products of the factor path.
//
// Every slot s = t*n + i with a terminal leaf gets the sort key "global node
// id" (off-leaf slots get total_nodes, which sorts last) and the value
// (s << 1) | in-bag.  A stable LSD radix sort (8-bit digits, 4096-key tiles:
// tile histograms, one exclusive scan over the digit-major histogram, and a
// scatter whose in-tile ranks come from wavefront ballots on the digit bits)
// groups the slots by leaf with the samples of a leaf in ascending order.
// From the sorted arrays:
//   bstart[b], bend[b]  the member range of leaf b (0, 0 when empty)
//   members[p]          (i << 1) | in-bag of the p-th sorted slot
//   members16[p]        the same as 16 bits, kept when n <= kMembers16MaxN
//                       (half the bytes for the gather kernels' member runs)
//   mbucket[p]          its leaf (the sorted keys)
//   pe[s]               (position of slot s in members, end of its leaf |
//                       in-bag << 31); (0, 0) off-leaf
// so a sample finds the partners that follow it in its leaf with one load.
#pragma once

#include "rfx_common.hpp"

namespace {

constexpr int64_t kMembers16MaxN = 32768;
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

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t len,
                                                                   uint32_t* __restrict__ part) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) v[q] = base + q < len ? in[base + q] : 0u;
  const uint32_t tot = block_exclusive_scan(v, lw);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanThreads) void scan_partials_kernel(uint32_t* __restrict__ part, int64_t nparts) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
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
                                                                  const uint32_t* __restrict__ part, uint32_t* out) {
  __shared__ uint32_t lw[kScanThreads / 64 + 1];
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

// Sort input: key = leaf (global node id) or total_nodes, value = slot<<1|bag.
__global__ void rs_prep_kernel(const int32_t* __restrict__ keys, int64_t slots, uint32_t total_nodes,
                               uint32_t* __restrict__ K, uint32_t* __restrict__ V) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= slots) return;
  const int32_t k = keys[s];
  K[s] = k >= 0 ? static_cast<uint32_t>(k) >> 1 : total_nodes;
  V[s] = (static_cast<uint32_t>(s) << 1) | (k >= 0 ? (static_cast<uint32_t>(k) & 1u) : 0u);
}

// Digit histogram of one tile, written digit-major: hist[d * ntiles + tile].
__global__ __launch_bounds__(kRsThreads) void rs_hist_kernel(const uint32_t* __restrict__ K, int64_t N, int shift,
                                                             int64_t ntiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  for (int q = 0; q < kRsRounds; ++q) {
    const int64_t g = base + q * kRsThreads + threadIdx.x;
    if (g < N) atomicAdd(&cnt[(K[g] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[static_cast<int64_t>(threadIdx.x) * ntiles + blockIdx.x] = cnt[threadIdx.x];
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
__global__ __launch_bounds__(kRsThreads) void rs_scatter_kernel(const uint32_t* __restrict__ Kin,
                                                                const uint32_t* __restrict__ Vin, int64_t N, int shift,
                                                                int64_t ntiles, const uint32_t* __restrict__ offs,
                                                                uint32_t* __restrict__ Kout,
                                                                uint32_t* __restrict__ Vout) {
  constexpr int kWaves = kRsThreads / 64;
  constexpr int kPerWave = kRsTile / kWaves;
  constexpr int kKeysPerLane = kPerWave / 64;
  __shared__ uint32_t wcnt[kWaves][256];  // per-wave digit counts, then per-wave digit bases
  __shared__ int64_t gofs[256];           // global position of the tile's digit run minus its tile start
  __shared__ uint32_t wsum[kWaves];
  __shared__ uint32_t sk[kRsTile], sv[kRsTile];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRsTile;
  const int tile_len = static_cast<int>(min<int64_t>(kRsTile, N - tile0));
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
  gofs[d] = static_cast<int64_t>(offs[static_cast<int64_t>(d) * ntiles + blockIdx.x]) - start;
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
  for (int j = tid; j < tile_len; j += kRsThreads) {
    const uint32_t key = sk[j];
    const int64_t pos = gofs[(key >> shift) & 255u] + j;
    if (pos >= 0 && pos < N) {
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

// members[p] = (i << 1) | bag; pe[s] = (p, end of its leaf | bag << 31),
// (0, 0) off-leaf.
__global__ void rs_members_kernel(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int64_t N, uint32_t n,
                                  uint32_t total_nodes, const uint32_t* __restrict__ bend,
                                  uint32_t* __restrict__ members, uint16_t* __restrict__ members16,
                                  uint2* __restrict__ pe) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const uint32_t v = V[p];
  const uint32_t s = v >> 1, bag = v & 1u;
  const uint32_t i = s % n;
  members[p] = (i << 1) | bag;
  if (members16) members16[p] = static_cast<uint16_t>((i << 1) | bag);
  const uint32_t k = K[p];
  pe[s] = k == total_nodes ? make_uint2(0u, 0u) : make_uint2(static_cast<uint32_t>(p), bend[k] | (bag << 31));
}

// Segmented form for a batch of problems laid end to end: slots
// [seg_start[b], seg_start[b+1]) belong to problem b with seg_n[b] samples,
// so the sample index is (s - seg_start[b]) % seg_n[b].
__global__ void rs_members_seg_kernel(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V, int64_t N,
                                      const int64_t* __restrict__ seg_start, const int64_t* __restrict__ seg_n,
                                      int nseg, uint32_t total_nodes, const uint32_t* __restrict__ bend,
                                      uint32_t* __restrict__ members, uint2* __restrict__ pe) {
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
  uint32_t* mbucket;  // sorted keys (one of the two key buffers)
  uint2* pe;
  uint32_t* nvalid;
  uint32_t* K[2];
  uint32_t* V[2];
  uint32_t* hist;
  uint32_t* part;
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

size_t bucket_ws_bytes(int64_t n, int32_t ntree, int64_t total_nodes) {
  const int64_t slots = n * static_cast<int64_t>(ntree);
  const int64_t ntiles = (slots + kRsTile - 1) / kRsTile;
  const int64_t hlen = 256 * ntiles;
  const int64_t nparts = (hlen + kScanChunk - 1) / kScanChunk;
  return align256(4 * total_nodes) * 2 + align256(4 * slots) * 5 + align256(8 * slots) + align256(4 * hlen) +
         align256(4 * (nparts + 1)) + 512 + (n <= kMembers16MaxN ? align256(2 * slots) : 0);
}

BucketWs bucket_ws_carve(void* ws, int64_t n, int32_t ntree, int64_t total_nodes) {
  BucketWs w;
  char* p = static_cast<char*>(ws);
  const int64_t slots = n * static_cast<int64_t>(ntree);
  w.ntiles = (slots + kRsTile - 1) / kRsTile;
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
  p += align256(4 * hlen);
  w.part = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * (w.nparts + 1));
  w.nvalid = reinterpret_cast<uint32_t*>(p);
  w.stat = reinterpret_cast<unsigned long long*>(p + 256);
  p += 512;
  w.members16 = n <= kMembers16MaxN ? reinterpret_cast<uint16_t*>(p) : nullptr;
  w.mbucket = w.K[w.passes & 1];
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

// Sort of N slots by leaf and the leaf ranges (everything but members/pe).
int bucket_sort(hipStream_t st, const int32_t* keys, int64_t N, int64_t total_nodes, const BucketWs& w) {
  const uint32_t tn = static_cast<uint32_t>(total_nodes);
  int rc = fill_u32(st, w.bstart, total_nodes, 0u);
  if (rc) return rc;
  rc = fill_u32(st, w.bend, total_nodes, 0u);
  if (rc) return rc;
  hipLaunchKernelGGL(rs_prep_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, keys, N, tn, w.K[0], w.V[0]);
  const int64_t hlen = 256 * w.ntiles;
  for (int pass = 0; pass < w.passes; ++pass) {
    const int src = pass & 1, dst = src ^ 1;
    hipLaunchKernelGGL(rs_hist_kernel, dim3(static_cast<unsigned>(w.ntiles)), dim3(kRsThreads), 0, st, w.K[src], N,
                       8 * pass, w.ntiles, w.hist);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.hist,
                       hlen, w.part);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, w.part, w.nparts);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.hist,
                       hlen, w.part, w.hist);
    hipLaunchKernelGGL(rs_scatter_kernel, dim3(static_cast<unsigned>(w.ntiles)), dim3(kRsThreads), 0, st, w.K[src],
                       w.V[src], N, 8 * pass, w.ntiles, w.hist, w.K[dst], w.V[dst]);
  }
  const int fin = w.passes & 1;
  hipLaunchKernelGGL(rs_bounds_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[fin], N, tn, w.bstart, w.bend,
                     w.nvalid);
  return rfx_launch_status();
}

// Full bucket build from the leaf keys [T][n] (see the file comment).
int bucket_build(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                 const BucketWs& w) {
  const int64_t N = n * static_cast<int64_t>(ntree);
  int rc = bucket_sort(st, keys, N, total_nodes, w);
  if (rc) return rc;
  const int fin = w.passes & 1;
  hipLaunchKernelGGL(rs_members_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[fin], w.V[fin], N,
                     static_cast<uint32_t>(n), static_cast<uint32_t>(total_nodes), w.bend, w.members, w.members16,
                     w.pe);
  return rfx_launch_status();
}

// Bucket build over a batch of problems whose key arrays lie end to end
// (problem b: slots [seg_start[b], seg_start[b+1]), keys [T_b][seg_n[b]],
// node ids already offset so that all leaves are distinct).
int bucket_build_segmented(hipStream_t st, const int32_t* keys, int64_t N, int64_t total_nodes,
                           const int64_t* seg_start_dev, const int64_t* seg_n_dev, int nseg, const BucketWs& w) {
  int rc = bucket_sort(st, keys, N, total_nodes, w);
  if (rc) return rc;
  const int fin = w.passes & 1;
  hipLaunchKernelGGL(rs_members_seg_kernel, dim3(rfx_ceil_div(N, 256)), dim3(256), 0, st, w.K[fin], w.V[fin], N,
                     seg_start_dev, seg_n_dev, nseg, static_cast<uint32_t>(total_nodes), w.bend, w.members, w.pe);
  return rfx_launch_status();
}

}  // namespace
