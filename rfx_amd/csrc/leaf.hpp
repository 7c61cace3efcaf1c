// All-numeric leaf assignment body, shared by the one-forest kernel and the
// batched kernel.
//
// One thread follows one sample down kFastTrees consecutive trees, kIlp
// traversals of different trees in lockstep, so every step issues kIlp
// independent node gathers instead of one dependent node + feature chain.
// Node records are 8 bytes (threshold bits, left << 10 | feature;
// 0xffffffff in the second word marks a terminal node).  With all features
// numeric the reference's out-of-bag rule (src/rf_testreebag.cpp:51-122, at
// most 999 descents, else node 0) and in-bag rule
// (src/rf_random_forest.cpp:2194-2244) both reduce to x <= threshold -> left.
#pragma once

#include "rfx_common.hpp"

namespace {

constexpr int kBlock = 256;
constexpr int kMaxOobDescents = 999;
#ifndef RFX_LEAF_ILP
#define RFX_LEAF_ILP 2
#endif
#ifndef RFX_LEAF_TREES
#define RFX_LEAF_TREES 4
#endif
constexpr int kIlp = RFX_LEAF_ILP;
constexpr int kFastTrees = RFX_LEAF_TREES;  // trees per workgroup

// Queue of the walks that the capped first pass leaves unfinished: 16-byte
// items {tree, sample, node reached (local to the tree), steps | in-bag << 31}.
// `tail` counts the walks offered (it may pass `capacity`; the offers past it
// are refused and those walks go on in the first pass), `head` the items the
// drain kernel has taken.
struct LeafQueue {
  uint4* items;
  unsigned long long* tail;
  unsigned long long* head;
  uint64_t capacity;
  int32_t cap_steps;  // lockstep iterations of the first pass before offering
};

// Wave-aggregated offer: one atomic per wavefront, slots in lane order.
// Returns whether this lane's walk was queued.
__device__ __forceinline__ bool leaf_queue_offer(const LeafQueue& q, bool want, uint32_t t, uint32_t i, uint32_t kt,
                                                 int32_t steps, bool bag) {
  const unsigned long long mask = __ballot(want);
  if (mask == 0) return false;
  const int lane = __lane_id();
  const int leader = __ffsll(mask) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(q.tail, static_cast<unsigned long long>(__popcll(mask)));
  base = __shfl(base, leader);
  const unsigned long long slot = base + __popcll(mask & ((1ull << lane) - 1ull));
  if (!want || slot >= q.capacity) return false;
  q.items[slot] = make_uint4(t, i, kt, static_cast<uint32_t>(steps) | (bag ? 0x80000000u : 0u));
  return true;
}

// Walk of sample i through trees [tb0, tb1), ILP traversals in lockstep;
// feat(f) returns the sample's value of feature f (from global memory or an
// LDS tile).  Keys are ((key_node_base + node_off[t] + node) << 1) | in-bag,
// -1 off-leaf.  With CAP, the walks still going after q.cap_steps lockstep
// iterations are offered to the queue and left to the drain kernel, so a
// wavefront does not wait on the few samples that follow a long chain of
// single-sample splits; refused offers go on here.
template <int ILP, bool CAP = false, class Feat>
__device__ __forceinline__ void leaf_walk_trees(const Feat& feat, int64_t i, int64_t n,
                                                const uint2* __restrict__ nodes,
                                                const int64_t* __restrict__ node_off, int32_t tb0, int32_t tb1,
                                                const uint32_t* __restrict__ inbag, int32_t* __restrict__ nodextr,
                                                int32_t* __restrict__ keys, int64_t key_node_base,
                                                const LeafQueue& q = LeafQueue{}) {
  const int64_t words = (n + 31) / 32;
  for (int32_t t0 = tb0; t0 < tb1; t0 += ILP) {
    int64_t base[ILP];
    uint32_t kt[ILP];
    uint2 nd[ILP];
    bool bag[ILP], live[ILP];
    int steps[ILP];
#pragma unroll
    for (int u = 0; u < ILP; ++u) {
      const int32_t t = t0 + u;
      live[u] = t < tb1;
      base[u] = live[u] ? node_off[t] : 0;
      bag[u] = live[u] ? ((inbag[t * words + (i >> 5)] >> (i & 31)) & 1u) : true;
      kt[u] = 0;
      steps[u] = 0;
      nd[u] = live[u] ? nodes[base[u]] : make_uint2(0u, 0xffffffffu);
    }
    int32_t iter = CAP ? 0 : -1;
    bool any = true;
    while (any) {
      if (CAP && iter >= 0 && iter++ == q.cap_steps) {
        // every lane still in the loop is at this iteration
#pragma unroll
        for (int u = 0; u < ILP; ++u) {
          const bool want = live[u] && nd[u].y != 0xffffffffu && (bag[u] || steps[u] < kMaxOobDescents);
          if (leaf_queue_offer(q, want, static_cast<uint32_t>(t0 + u), static_cast<uint32_t>(i), kt[u], steps[u],
                               bag[u])) {
            live[u] = false;
            nd[u].y = 0xffffffffu;
          }
        }
        iter = -1;
      }
      any = false;
      bool go[ILP];
#pragma unroll
      for (int u = 0; u < ILP; ++u) {
        go[u] = nd[u].y != 0xffffffffu && (bag[u] || steps[u] < kMaxOobDescents);
        if (go[u]) {
          const uint32_t f = nd[u].y & 1023u, l = nd[u].y >> 10;
          const float v = feat(f);
          kt[u] = v <= __uint_as_float(nd[u].x) ? l : l + 1;
          ++steps[u];
        }
        any = any || go[u];
      }
      // the ILP node gathers of this step are independent
#pragma unroll
      for (int u = 0; u < ILP; ++u)
        if (go[u]) nd[u] = nodes[base[u] + kt[u]];
    }
#pragma unroll
    for (int u = 0; u < ILP; ++u) {
      if (!live[u]) continue;
      if (nd[u].y != 0xffffffffu) kt[u] = 0;  // out-of-bag depth cap reached
      const int64_t o = static_cast<int64_t>(t0 + u) * n + i;
      if (nodextr) nodextr[o] = static_cast<int32_t>(kt[u]);
      if (keys)
        keys[o] = nd[u].y == 0xffffffffu
                      ? static_cast<int32_t>(((key_node_base + base[u] + kt[u]) << 1) | (bag[u] ? 1 : 0))
                      : -1;
    }
  }
}

// Sample block bx (kBlock samples), tree group by (kFastTrees trees), the
// features read from the column-major matrix in global memory.
__device__ __forceinline__ void leaf_numeric_body(const float* __restrict__ x, int64_t n,
                                                  const uint2* __restrict__ nodes,
                                                  const int64_t* __restrict__ node_off, int32_t ntree,
                                                  const uint32_t* __restrict__ inbag, int32_t* __restrict__ nodextr,
                                                  int32_t* __restrict__ keys, int64_t key_node_base, int64_t bx,
                                                  int64_t by) {
  const int64_t i = bx * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t tb0 = static_cast<int32_t>(by) * kFastTrees;
  const int32_t tb1 = min(tb0 + kFastTrees, ntree);
  const float* xi = x + i;
  leaf_walk_trees<kIlp>([&](uint32_t f) { return xi[static_cast<int64_t>(f) * n]; }, i, n, nodes, node_off, tb0, tb1,
                        inbag, nodextr, keys, key_node_base);
}

}  // namespace
