// A10: co-occurrence counts for a batch of independent forests in a fixed
// number of launches, replayable as one hipGraph.
//
// The reference grows and accumulates one forest at a time, tree by tree,
// from the host (fit_batch_gpu, src/rf_random_forest.cpp:1364-1570, with the
// tree batch size of get_recommended_batch_size,
// src/rf_cuda_config.cpp:225-295); many small problems then leave the GPU
// mostly idle between launches.  Here the whole batch is one problem:
//
//  1. one leaf-assignment launch over every (problem, sample block, tree
//     group), writing keys whose node ids are offset per problem, so that the
//     batch's key arrays, laid end to end, form one key space;
//  2. one segmented leaf-bucket build (radix sort by leaf) over all slots of
//     all problems;
//  3. one persistent launch: a grid sized to the machine pops work items
//     from a queue in device memory, most expensive problems first.  An item
//     is either a run of output rows of a GATHER problem (counted in LDS by
//     the one-problem GATHER line body) or a run of 128 x 128 upper tiles of
//     a DENSE problem (the one-problem DENSE tile body).  The queue head is
//     reset by a one-thread kernel, so the whole sequence is capturable.
//
// The autotuner measures every problem's leaf statistics (sum over trees and
// leaves of |leaf|^2) on the caller's data, chooses GATHER or DENSE per
// problem with the one-problem cost model, and times the tunables (rows per
// work item, workgroups per CU).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "bucket.hpp"
#include "dense.hpp"
#include "gather.hpp"
#include "leaf.hpp"
#include "rfx_common.hpp"

// One problem's UPPER counts on its slice of a batched bucket build (accum.hip).
int count_problem_slice(hipStream_t st, int32_t variant, int32_t algo, const int32_t* keys, int64_t n,
                        int32_t ntree, const uint2* pe, const uint32_t* bstart, const uint32_t* mbucket,
                        const uint32_t* members, const uint16_t* members16, int32_t* out, int32_t* out2);

namespace {

struct LeafDesc {
  const float* x;
  const uint2* nodes;
  const int64_t* node_off;
  const uint32_t* inbag;
  int32_t* keys;
  int64_t n;
  int64_t node_base;
  int64_t bx_count;
  int32_t ntree;
};

struct GatherDesc {
  const uint2* pe;
  const int32_t* keys;  // DENSE items read the keys
  int32_t* out;
  int32_t* out2;
  int64_t n;
  int32_t ntree;
};

constexpr uint32_t kItemGather = 0, kItemDense = 1;

__global__ __launch_bounds__(kBlock) void batch_leaf_kernel(const LeafDesc* __restrict__ desc,
                                                            const int64_t* __restrict__ block_start, int nprob) {
  const int64_t blk = blockIdx.x;
  int lo = 0, hi = nprob - 1;  // last problem whose first block <= blk
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (block_start[mid] <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  const LeafDesc d = desc[lo];
  const int64_t local = blk - block_start[lo];
  leaf_numeric_body(d.x, d.n, d.nodes, d.node_off, d.ntree, d.inbag, nullptr, d.keys, d.node_base,
                    local % d.bx_count, local / d.bx_count);
}

__global__ void batch_reset_kernel(uint32_t* counter) { *counter = 0u; }

// Work item: (problem, first, count, kind) -- rows [first, first + count)
// of a GATHER problem or upper tiles [first, first + count) of a DENSE
// problem.  Every workgroup pops items until the queue is empty; the queue is
// finite, so every wave exits.
// G > 0: GATHER lines walk their column chunks with per-tree cursors
// (gather_gallop_body, T <= 256 G); G = 0: chunk by chunk with binary
// searches (gather_line_body, any T).
template <int MODE, int G>
__global__ __launch_bounds__(256) void batch_gather_kernel(const GatherDesc* __restrict__ desc,
                                                           const uint4* __restrict__ items, uint32_t nitems,
                                                           uint32_t* __restrict__ counter,
                                                           const uint32_t* __restrict__ bstart,
                                                           const uint32_t* __restrict__ mbucket,
                                                           const uint32_t* __restrict__ members) {
  // 32 KiB: the GATHER line counters or the two DENSE key stages
  __shared__ uint32_t smem[kGatherChunk];
  __shared__ uint32_t diag_valid;
  __shared__ uint32_t item_s;
  static_assert(2 * kTreeChunk * kTile <= kGatherChunk, "DENSE stages must fit the shared buffer");
  int32_t(*sR)[kTile] = reinterpret_cast<int32_t(*)[kTile]>(smem);
  int32_t(*sS)[kTile] = reinterpret_cast<int32_t(*)[kTile]>(smem + kTreeChunk * kTile);
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t it = item_s;
    if (it >= nitems) break;
    const uint4 w = items[it];
    const GatherDesc d = desc[w.x];
    if (w.w == kItemGather) {
      for (uint32_t l = 0; l < w.z; ++l) {
        const int64_t line = static_cast<int64_t>(w.y) + l;
        if constexpr (G > 0) {
          gather_gallop_body<MODE, RFX_LAYOUT_UPPER, G>(d.pe, bstart, mbucket, d.n, d.ntree, members, 0, d.out,
                                                        d.out2, line, smem, diag_valid);
          __syncthreads();
        } else {
          const int64_t nch = (d.n - line + kGatherChunk - 1) / kGatherChunk;
          for (int64_t c = 0; c < nch; ++c) {
            gather_line_body<MODE, RFX_LAYOUT_UPPER>(d.pe, bstart, mbucket, d.n, d.ntree, members, 0, d.out, d.out2,
                                                     line, c, smem, diag_valid);
            __syncthreads();
          }
        }
      }
    } else {
      const int64_t nb = (d.n + kTile - 1) / kTile;
      for (uint32_t q = 0; q < w.z; ++q) {
        int64_t bi, bj;
        upper_tile(static_cast<int64_t>(w.y) + q, nb, bi, bj);
        dense_tile_body<MODE, RFX_LAYOUT_UPPER>(d.keys, d.n, d.ntree, bi, bj, 0, d.out, d.out2, sR, sS);
        __syncthreads();
      }
    }
    __syncthreads();  // every thread has read item_s before it is rewritten
  }
}

// sum over the leaves [lo, hi) of |leaf|^2
__global__ void batch_sumsq_kernel(const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend, int64_t lo,
                                   int64_t hi, unsigned long long* __restrict__ acc) {
  unsigned long long s = 0;
  for (int64_t e = lo + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < hi;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const unsigned long long c = bend[e] - bstart[e];
    s += c * c;
  }
  for (int d = 32; d > 0; d >>= 1) s += __shfl_down(s, d, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(acc, s);
}

}  // namespace

struct rfx_batch {
  std::vector<rfx_batch_problem> prob;
  int32_t variant = RFX_PROX_OOB_UPPER;
  int64_t slots = 0, total_nodes = 0;
  int32_t max_ntree = 0;  // largest T: cursor groups of the GATHER lines
  std::vector<int64_t> seg_start, seg_n, node_base, block_start;
  std::vector<uint32_t> order;  // problems by decreasing work
  std::vector<int32_t> algo;    // RFX_ALGO_GATHER or RFX_ALGO_DENSE per problem
  std::vector<double> sumsq;    // measured sum of |leaf|^2 per problem (0 until measured)
  std::vector<uint8_t> direct;  // 1: counted by its own one-problem kernel launch, not the queue
  std::vector<uint8_t> dense16; // 1: every tree has <= kDense16MaxNodes nodes (packed DENSE form)
  int32_t lines_per_item = 4;
  int32_t blocks_per_cu = 0;  // 0: as many as fit
  // device layout inside the workspace
  size_t off_leaf = 0, off_gather = 0, off_block = 0, off_segs = 0, off_segn = 0, off_items = 0, off_counter = 0,
         off_keys = 0, off_bucket = 0, off_m16 = 0, total_bytes = 0;
  uint32_t nitems = 0;
  std::vector<uint4> items;
  void* uploaded_ws = nullptr;
  int32_t uploaded_lpi = -1;
  uint64_t uploaded_algo_sig = 0;
  // second stream of the launch sequence (own-launch DENSE problems, which
  // need only the leaf keys, overlap the bucket build and the GATHER work)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // graph
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t exec = nullptr;
  void* graph_ws = nullptr;
  int32_t graph_lpi = -1, graph_bpc = -1;
  uint64_t graph_algo_sig = 0;
  int cu_count = 0, max_bpc = 0;
};

namespace {

// Relative costs of the one-problem AUTO rule (one tile compare = 1).
constexpr double kBatchVisit = 8.0, kBatchWrite = 4.0;
// a problem whose one-problem kernel takes at least this is launched on its own
constexpr float kDirectMinMs = 0.05f;

int64_t dense_tiles(int64_t n) {
  const int64_t nb = (n + kTile - 1) / kTile;
  return nb * (nb + 1) / 2;
}

double problem_cost(const rfx_batch* b, size_t i) {
  const double n = static_cast<double>(b->prob[i].n), T = b->prob[i].ntree;
  if (b->algo[i] == RFX_ALGO_DENSE) return 0.5 * T * n * n;
  const double sq = b->sumsq[i] > 0 ? b->sumsq[i] : T * n * n;
  return 0.5 * sq * kBatchVisit + 0.5 * n * n * kBatchWrite;
}

void order_problems(rfx_batch* b) {
  b->order.resize(b->prob.size());
  std::iota(b->order.begin(), b->order.end(), 0u);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [b](uint32_t a, uint32_t c) { return problem_cost(b, a) > problem_cost(b, c); });
}

uint64_t algo_signature(const rfx_batch* b) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < b->algo.size(); ++i)
    h = (h ^ static_cast<uint64_t>(b->algo[i] + 1 + 16 * b->direct[i])) * 1099511628211ull;
  return h;
}

void build_items(rfx_batch* b) {
  b->items.clear();
  const uint32_t L = static_cast<uint32_t>(std::max(1, b->lines_per_item));
  for (uint32_t pi : b->order) {
    const int64_t n = b->prob[pi].n;
    if (b->direct[pi]) continue;
    if (b->algo[pi] == RFX_ALGO_DENSE) {
      const int64_t tiles = dense_tiles(n);
      for (int64_t t = 0; t < tiles; ++t) b->items.push_back(make_uint4(pi, static_cast<uint32_t>(t), 1u, kItemDense));
    } else {
      for (int64_t r = 0; r < n; r += L)
        b->items.push_back(make_uint4(pi, static_cast<uint32_t>(r),
                                      static_cast<uint32_t>(std::min<int64_t>(L, n - r)), kItemGather));
    }
  }
  b->nitems = static_cast<uint32_t>(b->items.size());
}

size_t layout_bytes(rfx_batch* b) {
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    const size_t at = o;
    o += align256(bytes);
    return at;
  };
  const size_t B = b->prob.size();
  // items at the finest granularity (1 row or 1 tile per item) bound every tuning
  int64_t max_items = 0;
  for (const auto& p : b->prob) max_items += std::max(p.n, dense_tiles(p.n));
  b->off_leaf = take(sizeof(LeafDesc) * B);
  b->off_gather = take(sizeof(GatherDesc) * B);
  b->off_block = take(8 * (B + 1));
  b->off_segs = take(8 * (B + 1));
  b->off_segn = take(8 * B);
  b->off_items = take(sizeof(uint4) * static_cast<size_t>(max_items));
  b->off_counter = take(256);
  b->off_keys = take(4 * static_cast<size_t>(b->slots));
  b->off_bucket = take(bucket_ws_bytes(b->slots, 1, b->total_nodes));
  b->off_m16 = take(2 * static_cast<size_t>(b->slots));
  return o;
}

int upload(hipStream_t st, rfx_batch* b, char* ws) {
  const size_t B = b->prob.size();
  const BucketWs w = bucket_ws_carve(ws + b->off_bucket, b->slots, 1, b->total_nodes, false);
  std::vector<LeafDesc> ld(B);
  std::vector<GatherDesc> gd(B);
  int32_t* keys = reinterpret_cast<int32_t*>(ws + b->off_keys);
  for (size_t i = 0; i < B; ++i) {
    const auto& p = b->prob[i];
    ld[i] = LeafDesc{p.x_col, reinterpret_cast<const uint2*>(p.nodes), p.node_off, p.inbag, keys + b->seg_start[i],
                     p.n, b->node_base[i], rfx_ceil_div(p.n, kBlock), p.ntree};
    gd[i] = GatherDesc{w.pe + b->seg_start[i], keys + b->seg_start[i], p.out, p.out2, p.n, p.ntree};
  }
  build_items(b);
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_leaf, ld.data(), sizeof(LeafDesc) * B, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_gather, gd.data(), sizeof(GatherDesc) * B, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_block, b->block_start.data(), 8 * (B + 1), hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_segs, b->seg_start.data(), 8 * (B + 1), hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_segn, b->seg_n.data(), 8 * B, hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipMemcpyAsync(ws + b->off_items, b->items.data(), sizeof(uint4) * b->items.size(),
                             hipMemcpyHostToDevice, st));
  RFX_HIP_TRY(hipStreamSynchronize(st));  // host vectors are local
  b->uploaded_ws = ws;
  b->uploaded_lpi = b->lines_per_item;
  b->uploaded_algo_sig = algo_signature(b);
  return RFX_OK;
}

template <int MODE, int G>
int launch_gather_g(hipStream_t st, const rfx_batch* b, char* ws, const BucketWs& w, unsigned grid) {
  hipLaunchKernelGGL((batch_gather_kernel<MODE, G>), dim3(grid), dim3(256), 0, st,
                     reinterpret_cast<const GatherDesc*>(ws + b->off_gather),
                     reinterpret_cast<const uint4*>(ws + b->off_items), b->nitems,
                     reinterpret_cast<uint32_t*>(ws + b->off_counter), w.bstart, w.mbucket, w.members);
  return rfx_launch_status();
}

// cursor groups of the GATHER lines: T <= 256 G, or 0 (chunked form) past 4096
int gallop_groups(int32_t max_ntree) {
  for (int g = 1; g <= 16; g *= 2)
    if (max_ntree <= 256 * g) return g;
  return 0;
}

template <int MODE>
int launch_gather(hipStream_t st, const rfx_batch* b, char* ws, const BucketWs& w, unsigned grid) {
  switch (gallop_groups(b->max_ntree)) {
    case 1:
      return launch_gather_g<MODE, 1>(st, b, ws, w, grid);
    case 2:
      return launch_gather_g<MODE, 2>(st, b, ws, w, grid);
    case 4:
      return launch_gather_g<MODE, 4>(st, b, ws, w, grid);
    case 8:
      return launch_gather_g<MODE, 8>(st, b, ws, w, grid);
    case 16:
      return launch_gather_g<MODE, 16>(st, b, ws, w, grid);
    default:
      return launch_gather_g<MODE, 0>(st, b, ws, w, grid);
  }
}

template <int MODE>
int occupancy(int g) {
  int nb = 0;
  hipError_t e;
  switch (g) {
    case 1:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 1>, 256, 0);
      break;
    case 2:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 2>, 256, 0);
      break;
    case 4:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 4>, 256, 0);
      break;
    case 8:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 8>, 256, 0);
      break;
    case 16:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 16>, 256, 0);
      break;
    default:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_gather_kernel<MODE, 0>, 256, 0);
  }
  return e == hipSuccess ? nb : 0;
}

int max_blocks_per_cu(int32_t variant, int32_t max_ntree) {
  const int g = gallop_groups(max_ntree);
  int nb = variant == RFX_PROX_ALL ? occupancy<kMode_ALL>(g)
           : variant == RFX_PROX_OOB_UPPER ? occupancy<kMode_OOB_UPPER>(g)
                                           : occupancy<kMode_PAIR>(g);
  return nb < 1 ? 1 : nb;
}

// Problem i counted by its own one-problem kernel (algo DENSE or GATHER) on
// its slice of the batched buckets.
int launch_direct(hipStream_t st, const rfx_batch* b, char* ws, const BucketWs& w, size_t i, int32_t algo) {
  const auto& p = b->prob[i];
  const int32_t* keys = reinterpret_cast<const int32_t*>(ws + b->off_keys) + b->seg_start[i];
  if (algo == RFX_ALGO_DENSE && b->dense16[i]) algo |= kAlgoDense16;
  return count_problem_slice(st, b->variant, algo, keys, p.n, p.ntree, w.pe + b->seg_start[i], w.bstart, w.mbucket,
                             w.members, reinterpret_cast<const uint16_t*>(ws + b->off_m16), p.out, p.out2);
}

// join: `st` waits for the side stream's work
int join_side(hipStream_t st, rfx_batch* b, bool side_used) {
  if (!side_used) return RFX_OK;
  RFX_HIP_TRY(hipEventRecord(b->ev_join, b->side));
  RFX_HIP_TRY(hipStreamWaitEvent(st, b->ev_join, 0));
  return RFX_OK;
}

// The launch sequence (eager, or recorded into a graph by the caller): the
// leaf keys of all problems, one segmented bucket build, the problems
// counted on their own, then the persistent queue for the rest.
int enqueue(hipStream_t st, rfx_batch* b, char* ws) {
  const int B = static_cast<int>(b->prob.size());
  const BucketWs w = bucket_ws_carve(ws + b->off_bucket, b->slots, 1, b->total_nodes, false);
  const int64_t nblocks = b->block_start[B];
  hipLaunchKernelGGL(batch_leaf_kernel, dim3(static_cast<unsigned>(nblocks)), dim3(kBlock), 0, st,
                     reinterpret_cast<const LeafDesc*>(ws + b->off_leaf),
                     reinterpret_cast<const int64_t*>(ws + b->off_block), B);
  int rc = rfx_launch_status();
  if (rc) return rc;
  // fork: the own-launch DENSE problems on the side stream, right after the leaves
  bool side_used = false;
  for (uint32_t pi : b->order) {
    if (!b->direct[pi] || b->algo[pi] != RFX_ALGO_DENSE) continue;
    if (!side_used) {
      RFX_HIP_TRY(hipEventRecord(b->ev_fork, st));
      RFX_HIP_TRY(hipStreamWaitEvent(b->side, b->ev_fork, 0));
      side_used = true;
    }
    rc = launch_direct(b->side, b, ws, w, pi, RFX_ALGO_DENSE);
    if (rc) return rc;
  }
  rc = bucket_build_segmented(st, reinterpret_cast<const int32_t*>(ws + b->off_keys), b->slots, b->total_nodes,
                              reinterpret_cast<const int64_t*>(ws + b->off_segs),
                              reinterpret_cast<const int64_t*>(ws + b->off_segn), B, w,
                              reinterpret_cast<uint16_t*>(ws + b->off_m16));
  if (rc) return rc;
  for (uint32_t pi : b->order)
    if (b->direct[pi] && b->algo[pi] == RFX_ALGO_GATHER) {
      rc = launch_direct(st, b, ws, w, pi, RFX_ALGO_GATHER);
      if (rc) return rc;
    }
  if (b->nitems == 0) return join_side(st, b, side_used);
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, st, reinterpret_cast<uint32_t*>(ws + b->off_counter));
  rc = rfx_launch_status();
  if (rc) return rc;
  const int bpc = b->blocks_per_cu > 0 ? std::min(b->blocks_per_cu, b->max_bpc) : b->max_bpc;
  const unsigned grid = static_cast<unsigned>(std::max<int64_t>(
      1, std::min<int64_t>(static_cast<int64_t>(b->cu_count) * bpc, static_cast<int64_t>(b->nitems))));
  switch (b->variant) {
    case RFX_PROX_ALL:
      rc = launch_gather<kMode_ALL>(st, b, ws, w, grid);
      break;
    case RFX_PROX_OOB_UPPER:
      rc = launch_gather<kMode_OOB_UPPER>(st, b, ws, w, grid);
      break;
    default:
      rc = launch_gather<kMode_PAIR>(st, b, ws, w, grid);
  }
  if (rc) return rc;
  return join_side(st, b, side_used);
}

void drop_graph(rfx_batch* b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->graph_ws = nullptr;
}

}  // namespace

extern "C" int rfx_batch_create(const rfx_batch_problem* problems, int32_t count, int32_t variant,
                                rfx_batch** out) {
  if (!problems || !out || count <= 0) return RFX_ERR_INVALID;
  if (variant != RFX_PROX_ALL && variant != RFX_PROX_OOB_UPPER && variant != RFX_PROX_PAIR) return RFX_ERR_INVALID;
  *out = nullptr;
  auto* b = new (std::nothrow) rfx_batch();
  if (!b) return RFX_ERR_ALLOC;
  b->prob.assign(problems, problems + count);
  b->variant = variant;
  b->seg_start.assign(count + 1, 0);
  b->seg_n.assign(count, 0);
  b->node_base.assign(count, 0);
  b->block_start.assign(count + 1, 0);
  for (int32_t i = 0; i < count; ++i) {
    const auto& p = b->prob[i];
    const bool bad = !p.x_col || !p.nodes || !p.node_off || !p.inbag || !p.out || p.n <= 0 ||
                     p.n > RFX_MAX_SAMPLES || p.p <= 0 || p.p > RFX_MAX_FEATURES || p.ntree <= 0 ||
                     p.ntree > 65535 || p.total_nodes <= 0 || (variant == RFX_PROX_PAIR && !p.out2);
    if (bad) {
      delete b;
      return RFX_ERR_INVALID;
    }
    b->seg_start[i + 1] = b->seg_start[i] + p.n * static_cast<int64_t>(p.ntree);
    b->seg_n[i] = p.n;
    b->node_base[i] = b->total_nodes;
    b->total_nodes += p.total_nodes;
    b->max_ntree = std::max(b->max_ntree, p.ntree);
    b->block_start[i + 1] = b->block_start[i] + rfx_ceil_div(p.n, kBlock) * rfx_ceil_div(p.ntree, kFastTrees);
  }
  b->slots = b->seg_start[count];
  if (b->slots >= (1LL << 31) || b->total_nodes > RFX_MAX_TOTAL_NODES || b->block_start[count] >= (1LL << 31)) {
    delete b;
    return RFX_ERR_UNSUPPORTED;
  }
  b->algo.assign(count, RFX_ALGO_GATHER);
  b->sumsq.assign(count, 0.0);
  b->direct.assign(count, 0);
  // node counts per tree (planning reads the device offsets once)
  b->dense16.assign(count, 0);
  for (int32_t i = 0; i < count; ++i) {
    const auto& p = b->prob[i];
    std::vector<int64_t> off(static_cast<size_t>(p.ntree) + 1);
    if (hipMemcpy(off.data(), p.node_off, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost) != hipSuccess) {
      delete b;
      return RFX_ERR_HIP;
    }
    int64_t most = 0;
    for (int32_t t = 0; t < p.ntree; ++t) most = std::max(most, off[t + 1] - off[t]);
    b->dense16[i] = most <= kDense16MaxNodes ? 1 : 0;
  }
  order_problems(b);
  b->total_bytes = layout_bytes(b);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&b->cu_count, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    delete b;
    return RFX_ERR_HIP;
  }
  b->max_bpc = max_blocks_per_cu(variant, b->max_ntree);
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
    rfx_batch_free(b);
    return RFX_ERR_HIP;
  }
  *out = b;
  return RFX_OK;
}

extern "C" size_t rfx_batch_workspace_bytes(const rfx_batch* b) { return b ? b->total_bytes : 0; }

extern "C" int rfx_batch_set_tuning(rfx_batch* b, int32_t lines_per_item, int32_t blocks_per_cu) {
  if (!b || lines_per_item < 1 || lines_per_item > 4096 || blocks_per_cu < 0) return RFX_ERR_INVALID;
  b->lines_per_item = lines_per_item;
  b->blocks_per_cu = blocks_per_cu;
  return RFX_OK;
}

extern "C" int rfx_batch_get_tuning(const rfx_batch* b, int32_t* lines_per_item, int32_t* blocks_per_cu) {
  if (!b) return RFX_ERR_INVALID;
  if (lines_per_item) *lines_per_item = b->lines_per_item;
  if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu > 0 ? std::min(b->blocks_per_cu, b->max_bpc) : b->max_bpc;
  return RFX_OK;
}

extern "C" int rfx_batch_run(hipStream_t st, rfx_batch* b, void* workspace, size_t workspace_bytes,
                             int32_t use_graph) {
  if (!b || !workspace) return RFX_ERR_INVALID;
  if (workspace_bytes < b->total_bytes) return RFX_ERR_WORKSPACE;
  char* ws = static_cast<char*>(workspace);
  if (b->uploaded_ws != ws || b->uploaded_lpi != b->lines_per_item || b->uploaded_algo_sig != algo_signature(b)) {
    int rc = upload(st, b, ws);
    if (rc) return rc;
  }
  if (!use_graph) return enqueue(st, b, ws);
  if (!b->exec || b->graph_ws != ws || b->graph_lpi != b->lines_per_item || b->graph_bpc != b->blocks_per_cu ||
      b->graph_algo_sig != algo_signature(b)) {
    drop_graph(b);
    if (!b->cap_stream) RFX_HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    RFX_HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(b->cap_stream, b, ws);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return RFX_ERR_HIP;
    const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
      b->exec = nullptr;
      return RFX_ERR_HIP;
    }
    b->graph_ws = ws;
    b->graph_lpi = b->lines_per_item;
    b->graph_bpc = b->blocks_per_cu;
    b->graph_algo_sig = algo_signature(b);
  }
  RFX_HIP_TRY(hipGraphLaunch(b->exec, st));
  return RFX_OK;
}

// Leaf statistics of every problem from the buckets of the last run (the
// workspace holds them after any run), then GATHER or DENSE per problem and
// queue or own launch by timing (below).
int measure_and_choose(hipStream_t st, rfx_batch* b, char* ws) {
  const BucketWs w = bucket_ws_carve(ws + b->off_bucket, b->slots, 1, b->total_nodes, false);
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(w.stat);
  const size_t B = b->prob.size();
  std::vector<double> sq(B);
  for (size_t i = 0; i < B; ++i) {
    int rc = fill_u32(st, acc, 2, 0u);
    if (rc) return rc;
    const int64_t lo = b->node_base[i], hi = lo + b->prob[i].total_nodes;
    hipLaunchKernelGGL(batch_sumsq_kernel, dim3(grid_for(hi - lo, 256)), dim3(256), 0, st, w.bstart, w.bend, lo, hi,
                       acc);
    rc = rfx_launch_status();
    if (rc) return rc;
    unsigned long long h = 0;
    RFX_HIP_TRY(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, st));
    RFX_HIP_TRY(hipStreamSynchronize(st));
    sq[i] = static_cast<double>(h);
  }
  for (size_t i = 0; i < B; ++i) b->sumsq[i] = sq[i];
  // Each problem's one-problem DENSE and GATHER kernels timed on its slice of
  // the buckets: the faster decides the algorithm; problems that take at
  // least kDirectMinMs are launched on their own, the rest go to the queue.
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) hipEventDestroy(e0);
    return RFX_ERR_HIP;
  }
  int rc = RFX_OK;
  for (size_t i = 0; i < B && !rc; ++i) {
    float t[2] = {0.f, 0.f};
    const int32_t algos[2] = {RFX_ALGO_DENSE, RFX_ALGO_GATHER};
    for (int a = 0; a < 2 && !rc; ++a) {
      rc = launch_direct(st, b, ws, w, i, algos[a]);  // warm
      if (!rc && hipEventRecord(e0, st) != hipSuccess) rc = RFX_ERR_HIP;
      if (!rc) rc = launch_direct(st, b, ws, w, i, algos[a]);
      if (!rc && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                  hipEventElapsedTime(&t[a], e0, e1) != hipSuccess))
        rc = RFX_ERR_HIP;
    }
    b->algo[i] = t[0] <= t[1] ? RFX_ALGO_DENSE : RFX_ALGO_GATHER;
    b->direct[i] = std::min(t[0], t[1]) >= kDirectMinMs ? 1 : 0;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  order_problems(b);
  return RFX_OK;
}

extern "C" int rfx_batch_set_algos(rfx_batch* b, const int32_t* algos) {
  if (!b || !algos) return RFX_ERR_INVALID;
  for (size_t i = 0; i < b->prob.size(); ++i)
    if (algos[i] != RFX_ALGO_GATHER && algos[i] != RFX_ALGO_DENSE) return RFX_ERR_INVALID;
  b->algo.assign(algos, algos + b->prob.size());
  order_problems(b);
  return RFX_OK;
}

extern "C" int rfx_batch_get_stats(const rfx_batch* b, int32_t* algos, double* sum_leaf_sq) {
  if (!b) return RFX_ERR_INVALID;
  for (size_t i = 0; i < b->prob.size(); ++i) {
    if (algos) algos[i] = b->algo[i];
    if (sum_leaf_sq) sum_leaf_sq[i] = b->sumsq[i];
  }
  return RFX_OK;
}

extern "C" int rfx_batch_set_own_launch(rfx_batch* b, const int32_t* flags) {
  if (!b || !flags) return RFX_ERR_INVALID;
  for (size_t i = 0; i < b->prob.size(); ++i) b->direct[i] = flags[i] ? 1 : 0;
  order_problems(b);
  return RFX_OK;
}

extern "C" int rfx_batch_get_own_launch(const rfx_batch* b, int32_t* flags) {
  if (!b || !flags) return RFX_ERR_INVALID;
  for (size_t i = 0; i < b->prob.size(); ++i) flags[i] = b->direct[i];
  return RFX_OK;
}

extern "C" int rfx_batch_autotune(hipStream_t st, rfx_batch* b, void* workspace, size_t workspace_bytes,
                                  int32_t reps, double* best_ms) {
  if (!b || !workspace || reps < 1) return RFX_ERR_INVALID;
  if (workspace_bytes < b->total_bytes) return RFX_ERR_WORKSPACE;
  {
    int rc = rfx_batch_run(st, b, workspace, workspace_bytes, 0);  // buckets for the statistics
    if (rc) return rc;
    rc = measure_and_choose(st, b, static_cast<char*>(workspace));
    if (rc) return rc;
  }
  static const int32_t kLines[] = {1, 4, 16};
  std::vector<int32_t> bpcs;
  for (int v = 1; v <= b->max_bpc; v *= 2) bpcs.push_back(v);
  if (bpcs.back() != b->max_bpc) bpcs.push_back(b->max_bpc);
  hipEvent_t e0, e1;
  RFX_HIP_TRY(hipEventCreate(&e0));
  RFX_HIP_TRY(hipEventCreate(&e1));
  double best = 1e300;
  int32_t best_l = b->lines_per_item, best_b = b->blocks_per_cu;
  int rc = RFX_OK;
  for (int32_t L : kLines) {
    for (int32_t bpc : bpcs) {
      b->lines_per_item = L;
      b->blocks_per_cu = bpc;
      rc = rfx_batch_run(st, b, workspace, workspace_bytes, 0);  // upload + warm
      if (rc) break;
      if (hipEventRecord(e0, st) != hipSuccess) {
        rc = RFX_ERR_HIP;
        break;
      }
      for (int r = 0; r < reps && !rc; ++r) rc = rfx_batch_run(st, b, workspace, workspace_bytes, 0);
      if (rc) break;
      float ms = 0.f;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = RFX_ERR_HIP;
        break;
      }
      if (ms / reps < best) {
        best = ms / reps;
        best_l = L;
        best_b = bpc;
      }
    }
    if (rc) break;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->lines_per_item = best_l;
  b->blocks_per_cu = best_b;
  if (rc) return rc;
  if (best_ms) *best_ms = best;
  return RFX_OK;
}

extern "C" void rfx_batch_free(rfx_batch* b) {
  if (!b) return;
  drop_graph(b);
  if (b->cap_stream) hipStreamDestroy(b->cap_stream);
  if (b->side) hipStreamDestroy(b->side);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_join) hipEventDestroy(b->ev_join);
  delete b;
}
