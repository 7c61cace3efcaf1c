// A1: leaf assignment of every sample through every tree on gfx950.
//
// One thread follows one sample down TREES_PER_BLOCK consecutive trees.
// Node records are 8 bytes (threshold bits, left << 10 | feature; -1 marks a
// terminal node) so a traversal step is one 8-byte gather plus one 4-byte
// gather of the sample's feature from the column-major matrix; the top
// levels of every tree are shared by the whole wavefront and stay in L1/L2.
//
// Semantics follow the reference CPU path exactly:
//  * out-of-bag samples (src/rf_testreebag.cpp:51-122): descend while the node
//    is split, at most 999 descents, a node that is not terminal afterwards
//    is reported as node 0; categorical features use (int)x with a 0-based
//    catgoleft lookup, which the reference grower never fills, so they always
//    go right;
//  * in-bag samples (src/rf_random_forest.cpp:2194-2244): categorical
//    features use (int)(x + 0.5); a value in [1, ncat] goes right (catgoleft
//    is all zero) and anything else stops at node 0.
// With all features numeric both rules reduce to x <= threshold -> left.

#include <algorithm>

#include "rfx_common.hpp"
#include "chains.hpp"
#include "leaf.hpp"

namespace {

constexpr int kTreesPerBlock = 8;

__global__ __launch_bounds__(kBlock) void leaf_assign_kernel(const float* __restrict__ x, int64_t n,
                                                             const uint2* __restrict__ nodes,
                                                             const int64_t* __restrict__ node_off, int32_t ntree,
                                                             const uint32_t* __restrict__ inbag,
                                                             const int32_t* __restrict__ cat,
                                                             int32_t* __restrict__ nodextr,
                                                             int32_t* __restrict__ keys) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int64_t words = (n + 31) / 32;
  const int32_t t0 = blockIdx.y * kTreesPerBlock;
  const int32_t t1 = min(t0 + kTreesPerBlock, ntree);
  for (int32_t t = t0; t < t1; ++t) {
    const int64_t base = node_off[t];
    const uint2* tn = nodes + base;
    const bool bag = (inbag[t * words + (i >> 5)] >> (i & 31)) & 1u;
    uint32_t kt = 0;
    uint2 nd = tn[0];
    if (cat == nullptr) {
      // All-numeric fast path: both traversal rules are x <= thr -> left.
      int32_t steps = 0;
      while (nd.y != 0xffffffffu && (bag || steps < kMaxOobDescents)) {
        const uint32_t f = nd.y & 1023u, l = nd.y >> 10;
        const float v = x[static_cast<int64_t>(f) * n + i];
        kt = v <= __uint_as_float(nd.x) ? l : l + 1;
        nd = tn[kt];
        ++steps;
      }
      if (nd.y != 0xffffffffu) kt = 0;  // out-of-bag depth cap reached
    } else if (bag) {
      while (nd.y != 0xffffffffu) {
        const uint32_t f = nd.y & 1023u, l = nd.y >> 10;
        const float v = x[static_cast<int64_t>(f) * n + i];
        const int32_t nc = cat[f];
        if (nc == 1) {
          kt = v <= __uint_as_float(nd.x) ? l : l + 1;
        } else {
          const int32_t jcat = static_cast<int32_t>(static_cast<double>(v) + 0.5);
          if (jcat >= 1 && jcat <= nc) {
            kt = l + 1;
          } else {
            kt = 0;
            nd = tn[0];
            break;
          }
        }
        nd = tn[kt];
      }
    } else {
      int32_t steps = 0;
      while (nd.y != 0xffffffffu && steps < kMaxOobDescents) {
        const uint32_t f = nd.y & 1023u, l = nd.y >> 10;
        const float v = x[static_cast<int64_t>(f) * n + i];
        kt = (cat[f] != 1 || v > __uint_as_float(nd.x)) ? l + 1 : l;
        nd = tn[kt];
        ++steps;
      }
      if (nd.y != 0xffffffffu) {
        kt = 0;
        nd = tn[0];
      }
    }
    const int64_t o = static_cast<int64_t>(t) * n + i;
    if (nodextr) nodextr[o] = static_cast<int32_t>(kt);
    if (keys) {
      const bool terminal = nd.y == 0xffffffffu;
      keys[o] = terminal ? static_cast<int32_t>(((base + kt) << 1) | (bag ? 1 : 0)) : -1;
    }
  }
}

// All-numeric fast path (body in leaf.hpp).
//
// One-dimensional grid laid out for the 8 XCDs.  The dispatcher sends
// workgroup b to XCD b % 8, and each XCD has its own L2, so XCD x owns the
// contiguous sample blocks [x * nsbx, (x + 1) * nsbx) and walks them through
// every tree: its share of the feature matrix stays in its L2 instead of the
// whole matrix cycling through all eight.  Within an XCD the sample blocks
// are taken in chunks of `sc` blocks (a feature footprint of at most
// RFX_LEAF_CHUNK_BYTES); per chunk the tree groups run in order with the
// chunk's sample blocks innermost, so a tree group's nodes are fetched once
// per chunk.  The default takes an XCD's whole sample range as one chunk:
// covtype-like n = 200,000, T = 1,000 (5.4 MiB of features per XCD) runs in
// 12.5 ms against 14.1 ms with 1.5 MiB chunks and 14.7 ms with 0.75 MiB --
// the node records, re-fetched once per chunk, are the larger stream.
// Workgroups past the last sample block exit.
#ifndef RFX_LEAF_CHUNK_BYTES
#define RFX_LEAF_CHUNK_BYTES (int64_t(1) << 40)
#endif
constexpr int64_t kLeafChunkBytes = RFX_LEAF_CHUNK_BYTES;

struct LeafGrid {
  int64_t nsb, ntg, nsbx, sc;
};

__device__ __forceinline__ bool leaf_grid_map(const LeafGrid& g, int64_t b, int64_t& sb, int64_t& tg) {
  const int64_t xcd = b & 7, local = b >> 3;
  const int64_t per_chunk = g.sc * g.ntg;
  const int64_t c = local / per_chunk, rem = local - c * per_chunk;
  const int64_t csize = min(g.sc, g.nsbx - c * g.sc);
  tg = rem / csize;
  sb = xcd * g.nsbx + c * g.sc + (rem - tg * csize);
  return sb < g.nsb;
}

LeafGrid leaf_grid(int64_t n, int32_t p, int32_t ntree, int block, int trees) {
  LeafGrid g;
  g.nsb = rfx_ceil_div(n, block);
  g.ntg = rfx_ceil_div(ntree, trees);
  g.nsbx = (g.nsb + 7) / 8;
  const int64_t block_bytes = static_cast<int64_t>(block) * p * 4;
  g.sc = std::max<int64_t>(1, std::min<int64_t>(g.nsbx, kLeafChunkBytes / std::max<int64_t>(block_bytes, 1)));
  return g;
}

__global__ __launch_bounds__(kBlock) void leaf_assign_numeric_kernel(const float* __restrict__ x, int64_t n,
                                                                     const uint2* __restrict__ nodes,
                                                                     const int64_t* __restrict__ node_off,
                                                                     int32_t ntree,
                                                                     const uint32_t* __restrict__ inbag,
                                                                     int32_t* __restrict__ nodextr,
                                                                     int32_t* __restrict__ keys, LeafGrid g) {
  int64_t sb, tg;
  if (!leaf_grid_map(g, blockIdx.x, sb, tg)) return;
  leaf_numeric_body(x, n, nodes, node_off, ntree, inbag, nodextr, keys, 0, sb, tg);
}

// LDS-staged form: the workgroup first copies its samples' feature rows
// into LDS ([p][LB] floats, coalesced column reads), then every traversal
// step reads the feature from LDS, so the divergent global gathers of a step
// are the node records alone.  Used while p * LB * 4 bytes fit kStageMaxLds.
// At n = 20,000, p = 54, T = 500: 8 wavefronts of 4 trees each (4 in
// lockstep) per tile take 0.31 ms, 8 of 8 trees 0.33 ms, 16 of 8 trees
// 0.34 ms, one wavefront of 16 trees (8 in lockstep) per tile 0.51 ms -- the
// single-wavefront tiles bounded occupancy by LDS.
// At n = 200,000 (about 54,000 nodes per tree, 1.5 MiB chunks) 4 wavefronts
// of 2 trees take 14.1 ms per 1,000 trees, 4 x 4 trees 13.8 ms, 8 x 2 14.2 ms,
// 4 x 1 19.5 ms, one wavefront of 16 trees 17.4 ms and 8 x 8 trees 23.7 ms:
// the node records of a tree group must stay in L2.  With one chunk per XCD
// (the default): 4 x 2 trees 12.5 ms, 2 x 4 12.6, 4 x 4 13.1, 8 x 2 13.6,
// 4 x 1 16.7, 2 x 2 17.6.
#ifndef RFX_LEAF_STAGE
#define RFX_LEAF_STAGE 1
#endif
#ifndef RFX_LEAF_STAGE_LB
#define RFX_LEAF_STAGE_LB 64
#endif
constexpr int kStageLb = RFX_LEAF_STAGE_LB;
constexpr int64_t kStageMaxLds = 64 * 1024;
static_assert(kStageMaxLds <= kLdsLimitBytes, "staged feature tile exceeds the LDS");
// Wavefronts per workgroup W share the workgroup's LDS feature tile and take
// consecutive groups of TPW trees each, ILP of them in lockstep.  Small trees
// (the forest grown on up to kStageSmallN samples) use 8 x 4 trees per
// workgroup; larger trees fewer per workgroup, so the node records that the
// sample blocks of an XCD chunk share stay within its L2.
#ifndef RFX_LEAF_SMALL_CFG
#define RFX_LEAF_SMALL_CFG 8, 4, 4
#endif
#ifndef RFX_LEAF_BIG_CFG
#define RFX_LEAF_BIG_CFG 4, 2, 2
#endif
#ifndef RFX_LEAF_SMALL_N
#define RFX_LEAF_SMALL_N 65536
#endif
constexpr int64_t kStageSmallN = RFX_LEAF_SMALL_N;

template <int W, int TPW, int ILP, bool CAP>
__global__ __launch_bounds__(kStageLb * W) void leaf_assign_staged_kernel(const float* __restrict__ x, int64_t n,
                                                                          int32_t p, const uint2* __restrict__ nodes,
                                                                          const int64_t* __restrict__ node_off,
                                                                          int32_t ntree,
                                                                          const uint32_t* __restrict__ inbag,
                                                                          int32_t* __restrict__ nodextr,
                                                                          int32_t* __restrict__ keys, LeafGrid g,
                                                                          LeafQueue q) {
  extern __shared__ float xs[];
  int64_t sb, tg;
  if (!leaf_grid_map(g, blockIdx.x, sb, tg)) return;  // uniform over the workgroup
  const int tid = threadIdx.x % kStageLb, wave = threadIdx.x / kStageLb;
  const int64_t i = sb * kStageLb + tid;
  const bool in = i < n;
  for (int32_t f = wave; f < p; f += W) xs[f * kStageLb + tid] = in ? x[static_cast<int64_t>(f) * n + i] : 0.0f;
  __syncthreads();
  const int32_t tb0 = (static_cast<int32_t>(tg) * W + wave) * TPW;
  const int32_t tb1 = min(tb0 + TPW, ntree);
  if (!in || tb0 >= tb1) return;
  const float* xr = xs + tid;
  leaf_walk_trees<ILP, CAP>([&](uint32_t f) { return xr[f * kStageLb]; }, i, n, nodes, node_off, tb0, tb1, inbag,
                            nodextr, keys, 0, q);
}

// Drain of the walk queue.  Persistent one-wavefront workgroups: a lane whose
// walk has ended takes the next queued walk at the wavefront's next refill
// (one atomic per refill, every kDrainSteps steps), so the lanes stay busy
// whatever the spread of the remaining depths.  With DRAIN_LDS the sample's
// feature row is copied into the lane's LDS column at refill, and a step is
// one node gather plus one LDS read.
//
// A walk that stands on a node of a long heavy path (the chain
// index's labels) jumps to its exit from the path: the whole wavefront
// serves one such lane at a time, one lane per (feature, side) record list
// of the path, each a binary search for the first record whose exit test the
// sample passes, then a minimum over the wavefront.  The walk continues at
// the exit's light child (or the path's end node) with its step count
// advanced by the levels skipped, so the out-of-bag 999-descent rule is
// applied to the depth reached: beyond it the walk ends off-leaf at node 0,
// as the one-pass walk would.
//
// Every walk ends: node records only point to higher node indices within
// the tree (checked by rfx_pack_nodes), and out-of-bag walks stop after
// kMaxOobDescents descents.  A wavefront exits once the queue is exhausted
// and its lanes are idle.
constexpr int kDrainBlock = 64;
constexpr int kDrainSteps = 16;
constexpr int32_t kDrainLdsMaxP = 64;
static_assert(static_cast<int64_t>(kDrainLdsMaxP) * kDrainBlock * 4 <= kLdsLimitBytes,
              "drain kernel feature rows exceed the LDS");
#ifndef RFX_LEAF_CAP_STEPS
#define RFX_LEAF_CAP_STEPS 128
#endif
constexpr int32_t kLeafCapSteps = RFX_LEAF_CAP_STEPS;
constexpr size_t kLeafQueueHeader = 256;
constexpr int64_t kLeafQueueMinItems = 1024;

struct ChainView {
  const int64_t* tree_lab;
  const int32_t* tree_path0;
  const uint32_t* labels;
  const rfx_chain::ChainPath* paths;
  const rfx_chain::ChainList* lists;
  const uint2* recs;
  const uint32_t* exits;
  int32_t ntree, npaths;
};

__device__ __forceinline__ ChainView chain_view(const unsigned char* c) {
  const rfx_chain::ChainHeader* h = reinterpret_cast<const rfx_chain::ChainHeader*>(c);
  ChainView v;
  v.tree_lab = reinterpret_cast<const int64_t*>(c + h->tree_lab);
  v.tree_path0 = reinterpret_cast<const int32_t*>(c + h->tree_path0);
  v.labels = reinterpret_cast<const uint32_t*>(c + h->labels);
  v.paths = reinterpret_cast<const rfx_chain::ChainPath*>(c + h->paths);
  v.lists = reinterpret_cast<const rfx_chain::ChainList*>(c + h->lists);
  v.recs = reinterpret_cast<const uint2*>(c + h->recs);
  v.exits = reinterpret_cast<const uint32_t*>(c + h->exits);
  v.ntree = h->ntree;
  v.npaths = h->npaths;
  return v;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = min(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), off)));
  return v;
}

// Trees [tree_base, tree_base + items' tree ids) of the chain index: queue
// items carry tree ids relative to the launch, the chain index is global.
template <bool DRAIN_LDS>
__global__ __launch_bounds__(kDrainBlock) void leaf_drain_kernel(const float* __restrict__ x, int64_t n, int32_t p,
                                                                 const uint2* __restrict__ nodes,
                                                                 const int64_t* __restrict__ node_off, LeafQueue q,
                                                                 int32_t* __restrict__ nodextr,
                                                                 int32_t* __restrict__ keys,
                                                                 const unsigned char* __restrict__ chains,
                                                                 int32_t tree_base) {
  extern __shared__ float xrow[];  // [p][kDrainBlock] with DRAIN_LDS
  const unsigned long long offered = *q.tail;
  const unsigned long long count = offered < q.capacity ? offered : q.capacity;
  if (count == 0) return;
  const ChainView cv = chain_view(chains);
  const int lane = __lane_id();
  bool active = false, drained = false, bag = false;
  uint32_t t = 0, i = 0, kt = 0;
  int32_t steps = 0;
  int64_t base = 0, lab0 = -1;
  int32_t path0 = 0;
  uint32_t nnode = 0;
  uint2 nd = make_uint2(0u, 0xffffffffu);
  for (;;) {
    if (!drained) {
      const bool need = !active;
      const unsigned long long mask = __ballot(need);
      if (mask != 0) {
        const int leader = __ffsll(mask) - 1;
        const unsigned long long cnt = __popcll(mask);
        unsigned long long b = 0;
        if (lane == leader) b = atomicAdd(q.head, cnt);
        b = __shfl(b, leader);
        if (b + cnt >= count) drained = true;
        const unsigned long long slot = b + __popcll(mask & ((1ull << lane) - 1ull));
        if (need && slot < count) {
          const uint4 it = q.items[slot];
          t = it.x;
          i = it.y;
          kt = it.z;
          steps = static_cast<int32_t>(it.w & 0x7fffffffu);
          bag = (it.w >> 31) != 0;
          base = node_off[t];
          nd = nodes[base + kt];
          {
            nnode = static_cast<uint32_t>(node_off[t + 1] - base);
            const int32_t tg = tree_base + static_cast<int32_t>(t);
            lab0 = tg < cv.ntree ? cv.tree_lab[tg] : -1;
            path0 = lab0 >= 0 ? cv.tree_path0[tg] : 0;
          }
          if (DRAIN_LDS)
            for (int32_t f = 0; f < p; ++f) xrow[f * kDrainBlock + lane] = x[static_cast<int64_t>(f) * n + i];
          active = true;
        }
      }
    }
    if (__ballot(active) == 0) break;
    for (int r = 0; r < kDrainSteps; ++r) {
      uint32_t lab = rfx_chain::kNoLabel;
      if (active) {
        if (nd.y != 0xffffffffu && (bag || steps < kMaxOobDescents)) {
          if (lab0 >= 0) lab = cv.labels[lab0 + kt];
          if (lab == rfx_chain::kNoLabel) {
            const uint32_t f = nd.y & 1023u, l = nd.y >> 10;
            const float v = DRAIN_LDS ? xrow[f * kDrainBlock + lane] : x[static_cast<int64_t>(f) * n + i];
            kt = v <= __uint_as_float(nd.x) ? l : l + 1;
            ++steps;
            nd = nodes[base + kt];
          }
        } else {
          if (nd.y != 0xffffffffu) kt = 0;  // out-of-bag depth cap reached
          const int64_t o = static_cast<int64_t>(t) * n + i;
          if (nodextr) nodextr[o] = static_cast<int32_t>(kt);
          if (keys) keys[o] = nd.y == 0xffffffffu ? static_cast<int32_t>(((base + kt) << 1) | (bag ? 1 : 0)) : -1;
          active = false;
        }
      }
      {
        unsigned long long jm = __ballot(lab != rfx_chain::kNoLabel);
        while (jm != 0) {
          const int src = __ffsll(jm) - 1;
          jm &= jm - 1ull;
          const uint32_t slab = static_cast<uint32_t>(__shfl(static_cast<int>(lab), src));
          const int32_t gp = min(__shfl(path0, src) + static_cast<int32_t>(slab >> rfx_chain::kPosBits), cv.npaths - 1);
          const uint32_t p0 = slab & rfx_chain::kPosMask;
          const uint32_t si = static_cast<uint32_t>(__shfl(static_cast<int>(i), src));
          const rfx_chain::ChainPath P = cv.paths[gp];
          uint32_t e = 0xffffffffu;
          for (uint32_t li = static_cast<uint32_t>(lane); li < P.nlists; li += kDrainBlock) {
            const rfx_chain::ChainList L = cv.lists[P.list0 + li];
            const float v = DRAIN_LDS ? xrow[static_cast<int32_t>(L.feature) * kDrainBlock + src]
                                      : x[static_cast<int64_t>(L.feature) * n + si];
            const uint2* rec = cv.recs + L.rec0;
            uint32_t lo = 0, hi = L.count;
            while (lo < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              const bool le = v <= __uint_as_float(rec[mid].x);
              if (L.right_heavy ? le : !le)
                hi = mid;
              else
                lo = mid + 1;
            }
            if (lo < L.count) e = min(e, rec[lo].y);
          }
          e = wave_min_u32(e);
          if (lane == src) {
            // levels from position p0 to the node reached
            const uint32_t land = e == 0xffffffffu ? P.end_node : cv.exits[P.exit0 + e];
            const int64_t depth = static_cast<int64_t>(steps) + (e == 0xffffffffu ? P.length - p0 : e - p0 + 1);
            if (land <= kt || land >= nnode) {
              // only an index of other trees lands here: end off-leaf (every
              // move goes to a higher node index, so no walk can cycle)
              bag = false;
              steps = kMaxOobDescents;
              nd = make_uint2(0u, 0u);
            } else if (!bag && depth > kMaxOobDescents) {
              steps = kMaxOobDescents;  // the walk would stop on the path: node 0, off-leaf
              nd = make_uint2(0u, 0u);
            } else {
              kt = land;
              steps = static_cast<int32_t>(depth);
              nd = nodes[base + kt];
            }
          }
        }
      }
    }
  }
}

__global__ void leaf_queue_reset_kernel(unsigned long long* hdr) {
  if (threadIdx.x < 2) hdr[threadIdx.x] = 0ull;
}

template <int W, int TPW, int ILP>
int launch_staged(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                  const int64_t* node_off, int32_t ntree, const uint32_t* inbag, int32_t* nodextr, int32_t* keys,
                  const LeafQueue* q, const unsigned char* chains) {
  const int64_t lds = static_cast<int64_t>(p) * kStageLb * 4;
  const uint2* nd = reinterpret_cast<const uint2*>(nodes);
  if (q == nullptr || chains == nullptr) {
    const LeafGrid g = leaf_grid(n, p, ntree, kStageLb, TPW * W);
    const int64_t blocks = 8 * g.nsbx * g.ntg;
    if (blocks >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((leaf_assign_staged_kernel<W, TPW, ILP, false>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kStageLb * W), static_cast<unsigned>(lds), stream, x_col, n, p, nd, node_off, ntree,
                       inbag, nodextr, keys, g, LeafQueue{});
    return rfx_launch_status();
  }
  // Trees in chunks whose walks the queue can hold twice over: walks that
  // pass the first-pass cap go to the queue, refused ones stay in the pass.
  const int64_t words = (n + 31) / 32;
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(ntree, static_cast<int64_t>(2 * q->capacity) / n));
  for (int64_t t0 = 0; t0 < ntree; t0 += chunk) {
    const int32_t tc = static_cast<int32_t>(std::min<int64_t>(chunk, ntree - t0));
    const LeafGrid g = leaf_grid(n, p, tc, kStageLb, TPW * W);
    const int64_t blocks = 8 * g.nsbx * g.ntg;
    if (blocks >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
    const int64_t* off = node_off + t0;
    const uint32_t* bag = inbag + t0 * words;
    int32_t* ndx = nodextr ? nodextr + t0 * n : nullptr;
    int32_t* ky = keys ? keys + t0 * n : nullptr;
    hipLaunchKernelGGL(leaf_queue_reset_kernel, dim3(1), dim3(64), 0, stream, q->tail);
    hipLaunchKernelGGL((leaf_assign_staged_kernel<W, TPW, ILP, true>), dim3(static_cast<unsigned>(blocks)),
                       dim3(kStageLb * W), static_cast<unsigned>(lds), stream, x_col, n, p, nd, off, tc, bag, ndx, ky,
                       g, *q);
    // enough one-wavefront workgroups for every CU's SIMDs, the rest exit at once
    const unsigned drain_blocks = 256 * 16;
    if (p <= kDrainLdsMaxP)
      hipLaunchKernelGGL((leaf_drain_kernel<true>), dim3(drain_blocks), dim3(kDrainBlock),
                         static_cast<unsigned>(static_cast<size_t>(p) * kDrainBlock * 4), stream, x_col, n, p, nd,
                         off, *q, ndx, ky, chains, static_cast<int32_t>(t0));
    else
      hipLaunchKernelGGL((leaf_drain_kernel<false>), dim3(drain_blocks), dim3(kDrainBlock), 0, stream, x_col,
                         n, p, nd, off, *q, ndx, ky, chains, static_cast<int32_t>(t0));
    const int rc = rfx_launch_status();
    if (rc != RFX_OK) return rc;
  }
  return RFX_OK;
}

int leaf_assign(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat, int32_t* nodextr,
                int32_t* keys, const LeafQueue* q, const unsigned char* chains) {
  if (!x_col || !nodes || !node_off || !inbag || n <= 0 || n > RFX_MAX_SAMPLES || p <= 0 || p > RFX_MAX_FEATURES ||
      ntree <= 0)
    return RFX_ERR_INVALID;
  if (!nodextr && !keys) return RFX_OK;
  if (cat == nullptr) {
    const int64_t lds = static_cast<int64_t>(p) * kStageLb * 4;
    if (RFX_LEAF_STAGE && lds <= kStageMaxLds) {
      if (n <= kStageSmallN)
        return launch_staged<RFX_LEAF_SMALL_CFG>(stream, x_col, n, p, nodes, node_off, ntree, inbag, nodextr, keys,
                                                 q, chains);
      return launch_staged<RFX_LEAF_BIG_CFG>(stream, x_col, n, p, nodes, node_off, ntree, inbag, nodextr, keys, q,
                                             chains);
    }
    const LeafGrid g = leaf_grid(n, p, ntree, kBlock, kFastTrees);
    const int64_t blocks = 8 * g.nsbx * g.ntg;
    if (blocks >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(leaf_assign_numeric_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, stream, x_col,
                       n, reinterpret_cast<const uint2*>(nodes), node_off, ntree, inbag, nodextr, keys, g);
    return rfx_launch_status();
  }
  dim3 grid(rfx_ceil_div(n, kBlock), rfx_ceil_div(ntree, kTreesPerBlock));
  hipLaunchKernelGGL(leaf_assign_kernel, grid, dim3(kBlock), 0, stream, x_col, n,
                     reinterpret_cast<const uint2*>(nodes), node_off, ntree, inbag, cat, nodextr, keys);
  return rfx_launch_status();
}

}  // namespace

extern "C" int rfx_leaf_assign(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                               const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat,
                               int32_t* nodextr, int32_t* keys) {
  return leaf_assign(stream, x_col, n, p, nodes, node_off, ntree, inbag, cat, nodextr, keys, nullptr, nullptr);
}

extern "C" size_t rfx_leaf_workspace_bytes(int64_t n, int32_t ntree) {
  if (n <= 0 || ntree <= 0) return 0;
  const int64_t walks = n * ntree;
  const int64_t items = std::max<int64_t>(
      kLeafQueueMinItems,
      std::min<int64_t>(walks, std::max<int64_t>(int64_t(1) << 16, std::min<int64_t>(walks / 16, int64_t(1) << 26))));
  return kLeafQueueHeader + static_cast<size_t>(items) * sizeof(uint4);
}

extern "C" int rfx_leaf_assign_ws(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                                  const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat,
                                  int32_t* nodextr, int32_t* keys, const void* chains, void* ws, size_t ws_bytes) {
  if (ws == nullptr || ws_bytes < kLeafQueueHeader + kLeafQueueMinItems * sizeof(uint4)) return RFX_ERR_WORKSPACE;
  if (reinterpret_cast<uintptr_t>(ws) % 16 != 0) return RFX_ERR_INVALID;
  LeafQueue q;
  q.tail = static_cast<unsigned long long*>(ws);
  q.head = q.tail + 1;
  q.items = reinterpret_cast<uint4*>(static_cast<char*>(ws) + kLeafQueueHeader);
  q.capacity = (ws_bytes - kLeafQueueHeader) / sizeof(uint4);
  q.cap_steps = kLeafCapSteps;
  return leaf_assign(stream, x_col, n, p, nodes, node_off, ntree, inbag, cat, nodextr, keys, &q,
                     static_cast<const unsigned char*>(chains));
}
