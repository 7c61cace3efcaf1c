// Chain index of a packed forest: the long heavy paths of its trees, laid out
// for the leaf-assignment drain kernel (device) and built on the host.
//
// Trees grown to small nodes on large noisy data carry chains of splits that
// each peel one sample off the extreme of a feature: on the 16-component
// subspace mixture at n = 1,000,000 (minndsize 5) one path from the root
// holds 29,000 of a tree's 125,000 nodes, nearly every light child a single
// leaf, and the mean walk is 1,800 to 3,300 levels deep.  A sample walks
// such a chain one dependent node gather per level.  Heavy-path
// decomposition (heavy child = the child with more nodes below it, ties to
// the left) puts every chain on one path.  On a path, a sample leaves at the
// first node whose test sends it to the light child.  Per (feature, heavy
// side), the exit test of a node is !(x <= thr) when the heavy child is the
// left one and x <= thr when it is the right one; a node can only be the
// first exit for its (feature, side) if its threshold is a new strict
// minimum (left-heavy) or maximum (right-heavy) of that list along the
// path, so the list keeps only those "record" nodes, and along it the exit
// test is monotone: false on a prefix, true after.  The first exit of a
// sample on the path is then the smallest position that a binary search of
// each list returns.  A sample that has walked to position p0 of the path
// fails every record test before p0, so the search from anywhere on the
// path gives its exit at or after p0.
//
// Buffer layout (offsets in bytes from the start, sections 16-byte aligned):
//   ChainHeader
//   int64  tree_lab[ntree]    first label of tree t, -1 if it has no long path
//   int32  tree_path0[ntree]  global id of tree t's first long path
//   uint32 labels[]           per node of the labelled trees: (local path id
//                             << 22) | position, kNoLabel off long paths
//   ChainPath paths[npaths]
//   ChainList lists[nlists]
//   uint32 recs[2 * nrecs]    {threshold bits, position} by list, in path order
//   uint32 exits[]            per path position: the light child (local id)
#pragma once

#include <stdint.h>

namespace rfx_chain {

constexpr uint32_t kMagic = 0x43584652u;  // "RFXC"
constexpr uint32_t kNoLabel = 0xffffffffu;
constexpr uint32_t kPosBits = 22;
constexpr uint32_t kPosMask = (1u << kPosBits) - 1u;
constexpr uint32_t kMaxPathsPerTree = 1023;  // local ids 0..1022 keep labels below kNoLabel

struct ChainHeader {
  uint32_t magic;
  int32_t ntree;
  int32_t npaths;
  int32_t min_len;
  uint64_t tree_lab, tree_path0, labels, paths, lists, recs, exits, total;
};

struct ChainPath {
  uint64_t exit0;     // first entry of the path in exits
  uint32_t list0;     // first list of the path
  uint32_t nlists;
  uint32_t length;    // internal nodes on the path (positions 0 .. length-1)
  uint32_t end_node;  // terminal node that ends the path (position length)
};

struct ChainList {
  uint64_t rec0;      // first record (index into the {threshold, position} pairs)
  uint32_t count;
  uint16_t feature;
  uint16_t right_heavy;  // 0: exit when !(x <= thr); 1: exit when x <= thr
};

static_assert(sizeof(ChainHeader) == 80, "chain header layout");
static_assert(sizeof(ChainPath) == 24, "chain path layout");
static_assert(sizeof(ChainList) == 16, "chain list layout");

}  // namespace rfx_chain
