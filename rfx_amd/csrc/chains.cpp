// Host build of the chain index (layout and rationale in chains.hpp): heavy-
// path decomposition of every packed tree, the long paths' record lists,
// per-position light children and node labels.  O(nodes) per tree, trees
// in parallel.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "chains.hpp"
#include "rfx_prox.h"

namespace {

using namespace rfx_chain;

constexpr uint32_t kTerminal = 0xffffffffu;

struct TreeChains {
  std::vector<ChainPath> paths;  // exit0 / list0 / rec0 relative to this tree
  std::vector<ChainList> lists;
  std::vector<uint32_t> recs;    // {threshold bits, position} pairs
  std::vector<uint32_t> exits;
  std::vector<uint32_t> labels;  // empty when the tree has no long path
};

void tree_chains(const uint32_t* nd, int64_t nn, int32_t min_len, TreeChains& out) {
  std::vector<int64_t> size(static_cast<size_t>(nn), 1);
  for (int64_t k = nn - 1; k >= 0; --k) {
    const uint32_t y = nd[2 * k + 1];
    if (y == kTerminal) continue;
    const int64_t l = y >> 10;
    size[k] = 1 + size[l] + size[l + 1];
  }
  auto heavy_is_right = [&](int64_t k) {
    const int64_t l = nd[2 * k + 1] >> 10;
    return size[l + 1] > size[l];
  };
  // path tops: the root and every light child that is internal, in order
  std::vector<int64_t> tops;
  if (nd[1] != kTerminal) tops.push_back(0);
  for (int64_t k = 0; k < nn; ++k) {
    const uint32_t y = nd[2 * k + 1];
    if (y == kTerminal) continue;
    const int64_t l = y >> 10;
    const int64_t light = heavy_is_right(k) ? l : l + 1;
    if (nd[2 * light + 1] != kTerminal) tops.push_back(light);
  }
  std::sort(tops.begin(), tops.end());
  std::vector<int64_t> path;
  // per (feature, side) scratch: records of the current path
  std::vector<std::vector<uint32_t>> by_key;
  std::vector<float> best;
  std::vector<int32_t> touched;
  for (const int64_t top : tops) {
    if (out.paths.size() >= kMaxPathsPerTree) break;
    path.clear();
    int64_t k = top;
    while (nd[2 * k + 1] != kTerminal) {
      path.push_back(k);
      const int64_t l = nd[2 * k + 1] >> 10;
      k = heavy_is_right(k) ? l + 1 : l;
    }
    if (static_cast<int64_t>(path.size()) < min_len || path.size() > kPosMask) continue;
    if (out.labels.empty()) out.labels.assign(static_cast<size_t>(nn), kNoLabel);
    const uint32_t local = static_cast<uint32_t>(out.paths.size());
    ChainPath P;
    P.exit0 = out.exits.size();
    P.list0 = static_cast<uint32_t>(out.lists.size());
    P.length = static_cast<uint32_t>(path.size());
    P.end_node = static_cast<uint32_t>(k);
    touched.clear();
    for (size_t pos = 0; pos < path.size(); ++pos) {
      const int64_t node = path[pos];
      const uint32_t y = nd[2 * node + 1];
      const uint32_t f = y & 1023u;
      const int64_t l = y >> 10;
      const bool right = heavy_is_right(node);
      out.labels[node] = (local << kPosBits) | static_cast<uint32_t>(pos);
      out.exits.push_back(static_cast<uint32_t>(right ? l : l + 1));
      const size_t key = 2 * static_cast<size_t>(f) + (right ? 1 : 0);
      if (key >= by_key.size()) {
        by_key.resize(key + 1);
        best.resize(key + 1);
      }
      float thr;
      memcpy(&thr, &nd[2 * node], 4);
      std::vector<uint32_t>& r = by_key[key];
      // a record only where the threshold is a new strict minimum
      // (left-heavy) or maximum (right-heavy) of the list so far
      const bool rec = r.empty() || (right ? thr > best[key] : thr < best[key]);
      if (!rec) continue;
      if (r.empty()) touched.push_back(static_cast<int32_t>(key));
      best[key] = thr;
      r.push_back(nd[2 * node]);
      r.push_back(static_cast<uint32_t>(pos));
    }
    std::sort(touched.begin(), touched.end());
    for (const int32_t key : touched) {
      std::vector<uint32_t>& r = by_key[key];
      ChainList L;
      L.rec0 = out.recs.size() / 2;
      L.count = static_cast<uint32_t>(r.size() / 2);
      L.feature = static_cast<uint16_t>(key >> 1);
      L.right_heavy = static_cast<uint16_t>(key & 1);
      out.recs.insert(out.recs.end(), r.begin(), r.end());
      out.lists.push_back(L);
      r.clear();
    }
    P.nlists = static_cast<uint32_t>(out.lists.size()) - P.list0;
    out.paths.push_back(P);
  }
}

uint64_t align16(uint64_t v) { return (v + 15) & ~uint64_t(15); }

}  // namespace

extern "C" int rfx_chain_index(const int64_t* node_offset, int32_t ntree, const uint32_t* packed, int32_t min_len,
                               void* out, size_t* bytes) {
  if (!node_offset || !packed || !bytes || ntree <= 0 || min_len < 2) return RFX_ERR_INVALID;
  for (int32_t t = 0; t < ntree; ++t) {
    const int64_t nn = node_offset[t + 1] - node_offset[t];
    if (nn <= 0 || nn > RFX_MAX_NODES_PER_TREE) return RFX_ERR_INVALID;
  }
  std::vector<TreeChains> tc(static_cast<size_t>(ntree));
  {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::atomic<int32_t> next(0);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < hw; ++w)
      pool.emplace_back([&]() {
        for (int32_t t = next++; t < ntree; t = next++)
          tree_chains(packed + 2 * node_offset[t], node_offset[t + 1] - node_offset[t], min_len, tc[t]);
      });
    for (auto& th : pool) th.join();
  }
  uint64_t npaths = 0, nlists = 0, nrecs = 0, nexits = 0, nlabels = 0;
  for (const TreeChains& c : tc) {
    npaths += c.paths.size();
    nlists += c.lists.size();
    nrecs += c.recs.size() / 2;
    nexits += c.exits.size();
    nlabels += c.labels.size();
  }
  if (npaths == 0) {
    *bytes = 0;
    return RFX_OK;
  }
  if (npaths > 0x7fffffffull || nlists > 0xffffffffull) return RFX_ERR_UNSUPPORTED;
  ChainHeader h;
  memset(&h, 0, sizeof(h));
  h.magic = kMagic;
  h.ntree = ntree;
  h.npaths = static_cast<int32_t>(npaths);
  h.min_len = min_len;
  uint64_t o = align16(sizeof(ChainHeader));
  h.tree_lab = o;
  o = align16(o + 8ull * ntree);
  h.tree_path0 = o;
  o = align16(o + 4ull * ntree);
  h.labels = o;
  o = align16(o + 4ull * nlabels);
  h.paths = o;
  o = align16(o + sizeof(ChainPath) * npaths);
  h.lists = o;
  o = align16(o + sizeof(ChainList) * nlists);
  h.recs = o;
  o = align16(o + 8ull * nrecs);
  h.exits = o;
  o = align16(o + 4ull * nexits);
  h.total = o;
  if (out == nullptr) {
    *bytes = static_cast<size_t>(o);
    return RFX_OK;
  }
  if (*bytes < o) return RFX_ERR_WORKSPACE;
  char* b = static_cast<char*>(out);
  memset(b, 0, static_cast<size_t>(o));
  memcpy(b, &h, sizeof(h));
  int64_t* tree_lab = reinterpret_cast<int64_t*>(b + h.tree_lab);
  int32_t* tree_path0 = reinterpret_cast<int32_t*>(b + h.tree_path0);
  uint32_t* labels = reinterpret_cast<uint32_t*>(b + h.labels);
  ChainPath* paths = reinterpret_cast<ChainPath*>(b + h.paths);
  ChainList* lists = reinterpret_cast<ChainList*>(b + h.lists);
  uint32_t* recs = reinterpret_cast<uint32_t*>(b + h.recs);
  uint32_t* exits = reinterpret_cast<uint32_t*>(b + h.exits);
  uint64_t pa = 0, li = 0, re = 0, ex = 0, la = 0;
  for (int32_t t = 0; t < ntree; ++t) {
    const TreeChains& c = tc[t];
    tree_path0[t] = static_cast<int32_t>(pa);
    tree_lab[t] = c.labels.empty() ? -1 : static_cast<int64_t>(la);
    if (!c.labels.empty()) memcpy(labels + la, c.labels.data(), 4 * c.labels.size());
    for (ChainPath P : c.paths) {
      P.exit0 += ex;
      P.list0 += static_cast<uint32_t>(li);
      paths[pa++] = P;
    }
    for (ChainList L : c.lists) {
      L.rec0 += re;
      lists[li++] = L;
    }
    if (!c.recs.empty()) memcpy(recs + 2 * re, c.recs.data(), 4 * c.recs.size());
    if (!c.exits.empty()) memcpy(exits + ex, c.exits.data(), 4 * c.exits.size());
    re += c.recs.size() / 2;
    ex += c.exits.size();
    la += c.labels.size();
  }
  *bytes = static_cast<size_t>(o);
  return RFX_OK;
}
