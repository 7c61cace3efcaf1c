// Leaf buckets of the (tree, sample) slots: histogram over terminal nodes,
// exclusive scan, stable scatter.  Shared by the co-occurrence counts and
// the matrix-free proximity products of the factor path.
#pragma once

#include "rfx_common.hpp"

namespace {

// ---------------------------------------------------------------------------
// BUCKET algorithm
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanChunk = kScanThreads * kScanItems;

__global__ void bucket_hist_kernel(const int32_t* __restrict__ keys, int64_t total, uint32_t* __restrict__ cnt) {
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int32_t k = keys[e];
    if (k >= 0) atomicAdd(&cnt[k >> 1], 1u);
  }
}

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

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(const uint32_t* __restrict__ in, int64_t len,
                                                                  const uint32_t* __restrict__ part,
                                                                  uint32_t* __restrict__ out) {
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

// Stable scatter of the (tree, sample) slots into their buckets: one
// workgroup per tree walks its samples in chunks of 256 in index order; a
// sample's position is the bucket cursor plus the number of earlier samples
// of the chunk in the same bucket, so every bucket lists its members in
// ascending sample order and the layout is deterministic.  cursor starts as
// the exclusive scan and ends as the bucket end offsets.
__global__ __launch_bounds__(256) void bucket_scatter_kernel(const int32_t* __restrict__ keys, int64_t n,
                                                             uint32_t* __restrict__ cursor,
                                                             uint32_t* __restrict__ members,
                                                             uint32_t* __restrict__ mbucket) {
  __shared__ uint32_t sb[256];
  const int64_t t = blockIdx.x;
  const int tid = threadIdx.x;
  for (int64_t i0 = 0; i0 < n; i0 += 256) {
    const int64_t i = i0 + tid;
    const int32_t k = i < n ? keys[t * n + i] : -1;
    const uint32_t b = k >= 0 ? static_cast<uint32_t>(k) >> 1 : 0xffffffffu;
    sb[tid] = b;
    __syncthreads();
    uint32_t rank = 0;
    bool last = true;
    for (int q = 0; q < 256; ++q) {
      const uint32_t bq = sb[q];
      if (bq == b) {
        rank += q < tid ? 1u : 0u;
        last = last && !(q > tid);
      }
    }
    uint32_t base = 0;
    if (k >= 0) {
      base = cursor[b];
      const uint32_t pos = base + rank;
      members[pos] = (static_cast<uint32_t>(i) << 1) | (static_cast<uint32_t>(k) & 1u);
      mbucket[pos] = b;
    }
    __syncthreads();
    if (k >= 0 && last) cursor[b] = base + rank + 1;
    __syncthreads();
  }
}

struct BucketWs {
  uint32_t* cnt;
  uint32_t* bend;
  uint32_t* part;
  uint32_t* members;
  uint32_t* mbucket;
  unsigned long long* stat;
  int64_t nparts;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t bucket_ws_bytes(int64_t n, int32_t ntree, int64_t total_nodes) {
  const int64_t nparts = (total_nodes + kScanChunk - 1) / kScanChunk;
  const int64_t slots = n * static_cast<int64_t>(ntree);
  return align256(4 * total_nodes) * 2 + align256(4 * (nparts + 1)) + align256(4 * slots) * 2 + 256;
}

BucketWs bucket_ws_carve(void* ws, int64_t n, int32_t ntree, int64_t total_nodes) {
  BucketWs w;
  char* p = static_cast<char*>(ws);
  w.nparts = (total_nodes + kScanChunk - 1) / kScanChunk;
  const int64_t slots = n * static_cast<int64_t>(ntree);
  w.cnt = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * total_nodes);
  w.bend = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * total_nodes);
  w.part = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * (w.nparts + 1));
  w.members = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * slots);
  w.mbucket = reinterpret_cast<uint32_t*>(p);
  p += align256(4 * slots);
  w.stat = reinterpret_cast<unsigned long long*>(p);
  return w;
}

unsigned grid_for(int64_t work, int threads) {
  int64_t g = (work + threads - 1) / threads;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

// Histogram + scan: cnt and bend (as bucket start) ready, part[nparts] = total.
int bucket_build(hipStream_t st, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                 const BucketWs& w) {
  const int64_t slots = n * static_cast<int64_t>(ntree);
  RFX_HIP_TRY(hipMemsetAsync(w.cnt, 0, 4 * total_nodes, st));
  hipLaunchKernelGGL(bucket_hist_kernel, dim3(grid_for(slots, 256)), dim3(256), 0, st, keys, slots, w.cnt);
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.cnt,
                     total_nodes, w.part);
  hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(kScanThreads), 0, st, w.part, w.nparts);
  hipLaunchKernelGGL(scan_apply_kernel, dim3(static_cast<unsigned>(w.nparts)), dim3(kScanThreads), 0, st, w.cnt,
                     total_nodes, w.part, w.bend);
  return rfx_launch_status();
}

}  // namespace
