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
// are taken in chunks of `sc` blocks (a feature footprint of about 1.5 MiB);
// per chunk the tree groups run in order with the chunk's sample blocks
// innermost, so a tree group's nodes are fetched once per chunk.  Workgroups
// past the last sample block exit.
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
  g.sc = std::max<int64_t>(1, std::min<int64_t>(g.nsbx, (int64_t(3) << 19) / std::max<int64_t>(block_bytes, 1)));
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
// At n = 200,000 (about 54,000 nodes per tree) 4 wavefronts of 2 trees take
// 13.1 ms per 1,000 trees, one wavefront of 16 trees 17.4 ms and 8 x 8 trees
// 23.7 ms: there the node records of a tree group must stay in L2.
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

template <int W, int TPW, int ILP>
__global__ __launch_bounds__(kStageLb * W) void leaf_assign_staged_kernel(const float* __restrict__ x, int64_t n,
                                                                          int32_t p, const uint2* __restrict__ nodes,
                                                                          const int64_t* __restrict__ node_off,
                                                                          int32_t ntree,
                                                                          const uint32_t* __restrict__ inbag,
                                                                          int32_t* __restrict__ nodextr,
                                                                          int32_t* __restrict__ keys, LeafGrid g) {
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
  leaf_walk_trees<ILP>([&](uint32_t f) { return xr[f * kStageLb]; }, i, n, nodes, node_off, tb0, tb1, inbag, nodextr,
                       keys, 0);
}

template <int W, int TPW, int ILP>
int launch_staged(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                  const int64_t* node_off, int32_t ntree, const uint32_t* inbag, int32_t* nodextr, int32_t* keys) {
  const int64_t lds = static_cast<int64_t>(p) * kStageLb * 4;
  const LeafGrid g = leaf_grid(n, p, ntree, kStageLb, TPW * W);
  const int64_t blocks = 8 * g.nsbx * g.ntg;
  if (blocks >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL((leaf_assign_staged_kernel<W, TPW, ILP>), dim3(static_cast<unsigned>(blocks)),
                     dim3(kStageLb * W), static_cast<unsigned>(lds), stream, x_col, n, p,
                     reinterpret_cast<const uint2*>(nodes), node_off, ntree, inbag, nodextr, keys, g);
  return rfx_launch_status();
}

}  // namespace

extern "C" int rfx_leaf_assign(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                               const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat,
                               int32_t* nodextr, int32_t* keys) {
  if (!x_col || !nodes || !node_off || !inbag || n <= 0 || n > RFX_MAX_SAMPLES || p <= 0 || p > RFX_MAX_FEATURES ||
      ntree <= 0)
    return RFX_ERR_INVALID;
  if (!nodextr && !keys) return RFX_OK;
  if (cat == nullptr) {
    const int64_t lds = static_cast<int64_t>(p) * kStageLb * 4;
    if (RFX_LEAF_STAGE && lds <= kStageMaxLds) {
      if (n <= kStageSmallN)
        return launch_staged<RFX_LEAF_SMALL_CFG>(stream, x_col, n, p, nodes, node_off, ntree, inbag, nodextr, keys);
      return launch_staged<RFX_LEAF_BIG_CFG>(stream, x_col, n, p, nodes, node_off, ntree, inbag, nodextr, keys);
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
