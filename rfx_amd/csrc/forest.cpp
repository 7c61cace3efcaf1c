// Host-side input contract of the proximity path: the MT19937 stream, the
// bootstrap and the classification tree grower whose forests the GPU path
// consumes, plus the deterministic synthetic data generators.
//
// The grower reproduces the reference CPU grower tree-for-tree
// (src/rf_growtree.cpp:382-942 driven by src/rf_random_forest.cpp:928-1028),
// but replaces its per-node re-traversal of every in-bag sample from the root
// (src/rf_growtree.cpp:430-487) by an in-place partition of the parent's
// sample range.  Node membership is identical (children are numbered after
// their parents, so the reference traversal reaches node k exactly through
// the splits this partition applies), and the split search sees the node's
// samples through a (value, index) sort, whose order is total, so the float
// Gini arithmetic runs in the same sequence as the reference's.  Two growers
// produce these trees: grow_tree sorts each node's samples per drawn
// variable; grow_tree_presorted (the default) keeps every column of the tree
// sorted for the whole tree and is what makes n = 1,000,000 forests feasible.
//
// Everything here is plain host C++ (no HIP): the forest is an input of the
// path, grown once and exported; it is not part of the timed GPU work.

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <utility>
#include <vector>

#include "rfx_prox.h"

namespace {

// ---------------------------------------------------------------------------
// MT19937 (Matsumoto & Nishimura 1998, init_genrand seeding).  The reference
// draws doubles as y / 0xffffffff (include/rf_utils.hpp:104-135) and floats as
// the float rounding of that double (include/rf_utils.hpp:138-140).
// ---------------------------------------------------------------------------
class Mt19937 {
 public:
  explicit Mt19937(uint32_t seed) { seed_state(seed); }

  uint32_t next_u32() {
    if (pos_ >= kN) regenerate();
    uint32_t y = state_[pos_++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  double next_closed01() { return static_cast<double>(next_u32()) / 4294967295.0; }
  float next_float() { return static_cast<float>(next_closed01()); }

 private:
  static constexpr int kN = 624;
  static constexpr int kM = 397;

  void seed_state(uint32_t seed) {
    state_[0] = seed;
    for (int i = 1; i < kN; ++i) {
      state_[i] = 1812433253u * (state_[i - 1] ^ (state_[i - 1] >> 30)) + static_cast<uint32_t>(i);
    }
    pos_ = kN;
  }
  void regenerate() {
    for (int k = 0; k < kN; ++k) {
      uint32_t y = (state_[k] & 0x80000000u) | (state_[(k + 1) % kN] & 0x7fffffffu);
      uint32_t v = state_[(k + kM) % kN] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      state_[k] = v;
    }
    pos_ = 0;
  }

  uint32_t state_[kN];
  int pos_;
};

// ---------------------------------------------------------------------------
// Bootstrap with replacement (src/rf_bootstrap.cpp:14-82): n draws of
// i = int(float(u) * float(n)), clamped to n-1, multiplicities into nin.
// ---------------------------------------------------------------------------
void bootstrap_tree(int64_t n, uint32_t seed, int32_t* nin) {
  std::fill(nin, nin + n, 0);
  Mt19937 rng(seed);
  const float fn = static_cast<float>(n);
  for (int64_t d = 0; d < n; ++d) {
    float u = rng.next_float();
    int64_t i = static_cast<int64_t>(static_cast<int32_t>(u * fn));
    if (i >= n) i = n - 1;
    if (i < 0) i = 0;
    nin[i] += 1;
  }
}

struct Tree {
  std::vector<int32_t> status;
  std::vector<int32_t> bestvar;
  std::vector<float> xsplit;
  std::vector<int32_t> treemap;
  std::vector<int32_t> nodeclass;
  int32_t nnode = 0;
};

struct GrowParams {
  const float* x;  // row-major n x p
  const int32_t* cl;
  int64_t n;
  int32_t p;
  int32_t nclass;
  int32_t maxnode;
  int32_t minndsize;
  int32_t mtry;
  bool fma_gini;
};

// Per-thread scratch reused across trees.
struct GrowScratch {
  std::vector<int32_t> idx, tmp, start, end;
  std::vector<std::pair<float, int32_t>> sorted;
  std::vector<float> cc, wl, wr;
};

// One classification tree, following src/rf_growtree.cpp:382-942.
void grow_tree(const GrowParams& gp, const int32_t* nin, uint32_t seed, GrowScratch& s, Tree& t) {
  const int32_t maxnode = gp.maxnode;
  const int32_t nclass = gp.nclass;
  const int32_t p = gp.p;
  Mt19937 rng(seed);

  std::vector<int32_t> status(maxnode, 0), bestvar(maxnode, 0), treemap(2 * static_cast<size_t>(maxnode), 0),
      nodeclass(maxnode, 0);
  std::vector<float> xsplit(maxnode, 0.0f);
  s.start.assign(maxnode, 0);
  s.end.assign(maxnode, 0);

  // Unique in-bag samples in ascending order: jinbag (src/rf_bootstrap.cpp:66-78).
  s.idx.clear();
  for (int64_t i = 0; i < gp.n; ++i)
    if (nin[i] > 0) s.idx.push_back(static_cast<int32_t>(i));
  const int32_t ninbag = static_cast<int32_t>(s.idx.size());
  s.tmp.resize(ninbag);
  s.cc.assign(nclass, 0.0f);
  s.wl.assign(nclass, 0.0f);
  s.wr.assign(nclass, 0.0f);

  status[0] = 2;
  s.start[0] = 0;
  s.end[0] = ninbag;
  int32_t nnode = 1;

  for (int32_t k = 0; k < nnode; ++k) {
    if (status[k] != 2) continue;
    const int32_t b = s.start[k], e = s.end[k];
    const int32_t m = e - b;

    // Class counts of the node's unique in-bag samples (float, +1 each).
    std::fill(s.cc.begin(), s.cc.end(), 0.0f);
    for (int32_t q = b; q < e; ++q) {
      int32_t c = gp.cl[s.idx[q]];
      if (c >= 0 && c < nclass) s.cc[c] += 1.0f;
    }
    int32_t nonzero = 0, majority = 0;
    float maxc = 0.0f;
    for (int32_t j = 0; j < nclass; ++j) {
      if (s.cc[j] > 0.0f) {
        ++nonzero;
        if (s.cc[j] > maxc) {
          maxc = s.cc[j];
          majority = j;
        }
      }
    }
    nodeclass[k] = majority;
    // Too small (src/rf_growtree.cpp:494-499) or pure (511-569): terminal.
    if (m < gp.minndsize || nonzero <= 1) {
      status[k] = -1;
      continue;
    }

    // Gini numerator/denominator of the node.  The reference build
    // (CMakeLists.txt:54, -O3 -march=native, GNU C++17) contracts
    // pno += c*c into a fused multiply-add on FMA hosts; fma_gini mirrors it.
    float pno = 0.0f, pdo = 0.0f;
    for (int32_t j = 0; j < nclass; ++j) {
      if (gp.fma_gini)
        pno = std::fmaf(s.cc[j], s.cc[j], pno);
      else
        pno = pno + s.cc[j] * s.cc[j];
      pdo += s.cc[j];
    }

    float best_crit = -2.0f;
    int32_t best_var = 0;
    float best_split = 0.0f;
    for (int32_t mv = 0; mv < gp.mtry; ++mv) {
      float u = rng.next_float();
      int32_t mvar = static_cast<int32_t>(u * static_cast<float>(p));
      if (mvar >= p) mvar = p - 1;

      s.sorted.resize(m);
      for (int32_t q = 0; q < m; ++q) {
        int32_t si = s.idx[b + q];
        s.sorted[q] = {gp.x[static_cast<int64_t>(si) * p + mvar], si};
      }
      std::sort(s.sorted.begin(), s.sorted.end());

      float rrn = pno, rrd = pdo, rln = 0.0f, rld = 0.0f;
      for (int32_t j = 0; j < nclass; ++j) {
        s.wl[j] = 0.0f;
        s.wr[j] = s.cc[j];
      }
      // Split scan (src/rf_growtree.cpp:656-734).  A tied position is skipped
      // before the sample is moved to the left side, exactly as the reference
      // does; the float operations below run in its order.
      for (int32_t ii = 0; ii + 1 < m; ++ii) {
        const float v1 = s.sorted[ii].first, v2 = s.sorted[ii + 1].first;
        if (v1 >= v2) continue;
        const int32_t c = gp.cl[s.sorted[ii].second];
        rln = rln + (1.0f + 2.0f * s.wl[c]);
        rrn = rrn + (1.0f - 2.0f * s.wr[c]);
        rld = rld + 1.0f;
        rrd = rrd - 1.0f;
        s.wl[c] += 1.0f;
        s.wr[c] -= 1.0f;
        if (std::min(rrd, rld) > 1.0e-5f) {
          const float crit = (rln / rld) + (rrn / rrd);
          if (crit > best_crit) {
            best_crit = crit;
            best_var = mvar;
            best_split = v1;
          }
        }
      }
    }

    if (best_crit > -1.0f && nnode + 2 <= maxnode) {
      const int32_t lc = nnode, rc = nnode + 1;
      status[k] = 1;
      bestvar[k] = best_var;
      xsplit[k] = best_split;
      treemap[2 * static_cast<size_t>(k)] = lc;
      treemap[2 * static_cast<size_t>(k) + 1] = rc;
      status[lc] = 2;
      status[rc] = 2;
      // Stable partition of [b, e) by x <= split: the rule every traversal uses.
      int32_t nl = 0, nr = 0;
      for (int32_t q = b; q < e; ++q) {
        int32_t si = s.idx[q];
        if (gp.x[static_cast<int64_t>(si) * p + best_var] <= best_split)
          s.idx[b + nl++] = si;
        else
          s.tmp[nr++] = si;
      }
      std::copy(s.tmp.begin(), s.tmp.begin() + nr, s.idx.begin() + b + nl);
      s.start[lc] = b;
      s.end[lc] = b + nl;
      s.start[rc] = b + nl;
      s.end[rc] = e;
      nnode += 2;
    } else {
      status[k] = -1;
    }
  }

  status.resize(nnode);
  bestvar.resize(nnode);
  xsplit.resize(nnode);
  treemap.resize(2 * static_cast<size_t>(nnode));
  nodeclass.resize(nnode);
  t.status = std::move(status);
  t.bestvar = std::move(bestvar);
  t.xsplit = std::move(xsplit);
  t.treemap = std::move(treemap);
  t.nodeclass = std::move(nodeclass);
  t.nnode = nnode;
}

// ---------------------------------------------------------------------------
// Presorted grower: the same trees as grow_tree, without a sort per node.
//
// grow_tree sorts the node's samples by (value, index) once per drawn
// variable, so a node of m samples costs mtry sorts of m gathered values.
// On data whose trees peel one sample off a large impure node per split
// (the 16-component mixture with 24 noise columns and minndsize 5 has chains
// over a thousand nodes deep), the sum of node sizes reaches n x 100 and the
// sorts dominate.  Here every feature column of the tree's in-bag samples is
// kept sorted by (value, index) for the whole tree, each node owning the same
// range [b, e) in every column: the global (value, index) order is filtered
// to the in-bag samples once per tree, and after a split every column's range
// is stably partitioned by the side each sample went to.  A stable partition
// of a (value, index)-sorted range leaves both halves (value, index)-sorted,
// and the split variable's own range is already partitioned (its left part is
// the prefix of values <= the split value), so the scans see the same
// sequences, and do the same float operations, as grow_tree's.
//
// A column record holds the value and (local in-bag id << 8) | class, so the
// fast path needs nclass <= 256, labels in [0, nclass), fewer than 2^24
// in-bag samples and no NaN; rfx_forest_grow uses grow_tree otherwise.
// ---------------------------------------------------------------------------
struct ColRec {
  float v;
  uint32_t lc;  // (local id << 8) | class
};

struct PresortedData {
  // Global order of every column: (value, sample) pairs sorted by value then
  // sample, column-major p x n.
  std::vector<ColRec> order;
  bool avx2 = false;
  bool lanes = false;  // AVX2 and a column array whose indices fit int32
};

struct PeelEvent {
  int32_t b, e;    // the split node's range
  int32_t s;       // records on the small side
  bool few_left;   // the small side is the left child
  int32_t lid_at;  // offset of the small side's local ids in the lid pool
};

struct PeelLog {
  std::vector<PeelEvent> events;
  std::vector<int32_t> lids;
  std::vector<std::vector<int32_t>> pend;  // per column, event ids, outermost first
  std::vector<std::pair<int32_t, int32_t>> rem;  // (position, destination)
  std::vector<int32_t> gpos;
  std::vector<ColRec> saved;

  void reset(int32_t p) {
    events.clear();
    lids.clear();
    pend.resize(p);
    for (auto& v : pend) v.clear();
  }
  static int32_t big_b(const PeelEvent& ev) { return ev.few_left ? ev.b + ev.s : ev.b; }
  static int32_t big_e(const PeelEvent& ev) { return ev.few_left ? ev.e : ev.e - ev.s; }
};

// Pass-1 output of the split scan: numerators and positions.
struct ScanBuf {
  std::vector<float> ln, rn;
  std::vector<int32_t> at;
};

struct FastScratch {
  std::vector<int32_t> lid;   // n: local in-bag id, -1 out of bag
  std::vector<int32_t> sid;   // m0: sample of each local id
  std::vector<ColRec> col;    // p x m0, column-major
  std::vector<ColRec> tmp;    // m0
  std::vector<uint8_t> left;  // m0
  std::vector<int32_t> start, end;
  std::vector<int32_t> wl, wr;  // class counts as integers
  std::vector<float> cc;
  ScanBuf sb;
  std::vector<int32_t> vars;
  std::vector<int32_t> lane_wl, lane_wr;  // 8 x nclass: class counts per lane
  std::vector<int32_t> clean_col, cc_at;
  std::vector<float> cc_pool;
  PeelLog peels;
};

// Smallest node scanned eight columns at a time.
constexpr int32_t kLaneScanMin = 256;

void presort_columns(const GrowParams& gp, int nthreads, PresortedData& d) {
  const int64_t n = gp.n;
  const int32_t p = gp.p;
  d.order.resize(static_cast<size_t>(p) * n);
  d.avx2 = __builtin_cpu_supports("avx2");
  d.lanes = d.avx2 && static_cast<int64_t>(p) * n < (int64_t{1} << 31);
  std::atomic<int32_t> next{0};
  auto worker = [&]() {
    std::vector<std::pair<float, int32_t>> s(n);
    for (;;) {
      const int32_t v = next.fetch_add(1);
      if (v >= p) break;
      for (int64_t i = 0; i < n; ++i) s[i] = {gp.x[i * p + v], static_cast<int32_t>(i)};
      std::sort(s.begin(), s.end());
      ColRec* o = d.order.data() + static_cast<size_t>(v) * n;
      for (int64_t i = 0; i < n; ++i) o[i] = {s[i].first, static_cast<uint32_t>(s[i].second)};
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < nthreads && i < p; ++i) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
}

// Split scan of one (value, index)-sorted column range: grow_tree's float
// arithmetic in the same order, in two passes.
//
// Pass 1 runs the serial part: for every position that is not tied with the
// next, the class moves to the left side and the Gini numerators advance,
//   rln = rln + (1 + 2 wl[c]),  rrn = rrn + (1 - 2 wr[c]),
// and the pair (rln, rrn) is stored with the position.  The increments are
// exact integers in float (|.| <= 2m + 1 < 2^24), so the counts are kept as
// integers.  The left and right sizes rld, rrd are the integers cnt and m - cnt
// (pdo is the node size, exact in float), so min(rrd, rld) > 1e-5 always holds.
// Pass 2 evaluates crit = rln / cnt + rrn / (m - cnt) for all stored
// positions, eight at a time (IEEE division and addition, the same roundings
// as the scalar code), and keeps the first position of the largest value;
// the scalar loop's "crit > best" updates end on exactly that position.
struct ScanBest {
  float crit;
  int32_t var;
  float split;
  int32_t nl;
};

inline int32_t scan_prefix(const ColRec* c, int32_t m, float pno, int32_t nclass, const float* cc, int32_t* wl,
                           int32_t* wr, float* ln, float* rn, int32_t* at) {
  for (int32_t j = 0; j < nclass; ++j) {
    wl[j] = 0;
    wr[j] = static_cast<int32_t>(cc[j]);
  }
  int32_t k = 0;
  float rln = 0.0f, rrn = pno;
  float v1 = c[0].v;
  for (int32_t ii = 0; ii + 1 < m; ++ii) {
    const float v2 = c[ii + 1].v;
    if (v1 < v2) {
      const uint32_t cls = c[ii].lc & 0xffu;
      const int32_t a = wl[cls], bb = wr[cls];
      wl[cls] = a + 1;
      wr[cls] = bb - 1;
      rln = rln + static_cast<float>(1 + 2 * a);
      rrn = rrn + static_cast<float>(1 - 2 * bb);
      ln[k] = rln;
      rn[k] = rrn;
      at[k] = ii;
      ++k;
    }
    v1 = v2;
  }
  return k;
}

// First index of the largest crit over positions [0, K); returns -1 if none
// exceeds `floor`.
int32_t scan_best_scalar(const float* ln, const float* rn, int32_t K, int32_t m, float floor, float* out) {
  float best = floor;
  int32_t arg = -1;
  for (int32_t k = 0; k < K; ++k) {
    const float crit = (ln[k] / static_cast<float>(k + 1)) + (rn[k] / static_cast<float>(m - k - 1));
    if (crit > best) {
      best = crit;
      arg = k;
    }
  }
  *out = best;
  return arg;
}

__attribute__((target("avx2"))) int32_t scan_best_avx2(const float* ln, const float* rn, int32_t K, int32_t m,
                                                         float floor, float* out) {
  const int32_t K8 = K & ~7;
  __m256 vbest = _mm256_set1_ps(-3.0e38f);
  __m256i varg = _mm256_set1_epi32(-1);
  __m256i vk = _mm256_setr_epi32(0, 1, 2, 3, 4, 5, 6, 7);
  const __m256i one = _mm256_set1_epi32(1), eight = _mm256_set1_epi32(8), vm = _mm256_set1_epi32(m);
  for (int32_t k = 0; k < K8; k += 8) {
    const __m256 l = _mm256_loadu_ps(ln + k), r = _mm256_loadu_ps(rn + k);
    const __m256 dl = _mm256_cvtepi32_ps(_mm256_add_epi32(vk, one));
    const __m256 dr = _mm256_cvtepi32_ps(_mm256_sub_epi32(_mm256_sub_epi32(vm, vk), one));
    const __m256 crit = _mm256_add_ps(_mm256_div_ps(l, dl), _mm256_div_ps(r, dr));
    const __m256 gt = _mm256_cmp_ps(crit, vbest, _CMP_GT_OQ);
    vbest = _mm256_blendv_ps(vbest, crit, gt);
    varg = _mm256_blendv_epi8(varg, vk, _mm256_castps_si256(gt));
    vk = _mm256_add_epi32(vk, eight);
  }
  alignas(32) float bl[8];
  alignas(32) int32_t al[8];
  _mm256_store_ps(bl, vbest);
  _mm256_store_si256(reinterpret_cast<__m256i*>(al), varg);
  // Lane maxima hold their lane's first maximum; the overall first maximum
  // is the smallest index among the lanes with the largest value.
  float best = floor;
  int32_t arg = -1;
  for (int q = 0; q < 8; ++q) {
    if (al[q] < 0) continue;
    if (bl[q] > best || (bl[q] == best && arg >= 0 && al[q] < arg)) {
      best = bl[q];
      arg = al[q];
    }
  }
  for (int32_t k = K8; k < K; ++k) {
    const float crit = (ln[k] / static_cast<float>(k + 1)) + (rn[k] / static_cast<float>(m - k - 1));
    if (crit > best) {
      best = crit;
      arg = k;
    }
  }
  *out = best;
  return arg;
}

inline void scan_column(const ColRec* c, int32_t m, int32_t mvar, float pno, int32_t nclass, const float* cc,
                        int32_t* wl, int32_t* wr, ScanBuf& sb, bool avx2, ScanBest& best) {
  const int32_t K = scan_prefix(c, m, pno, nclass, cc, wl, wr, sb.ln.data(), sb.rn.data(), sb.at.data());
  float crit;
  const int32_t arg = avx2 ? scan_best_avx2(sb.ln.data(), sb.rn.data(), K, m, best.crit, &crit)
                           : scan_best_scalar(sb.ln.data(), sb.rn.data(), K, m, best.crit, &crit);
  if (arg >= 0) {
    const int32_t ii = sb.at[arg];
    best.crit = crit;
    best.var = mvar;
    best.split = c[ii].v;
    best.nl = ii + 1;
  }
}

// Up to eight columns scanned together, one per AVX2 lane, for nodes where one
// class dominates.  Each lane runs the scalar scan's float sequence on its
// own column: rln and rrn advance by the exact integer increments of the
// lane's class, and crit = rln / cnt + rrn / (m - cnt) is formed with the same
// IEEE single divisions and addition.  The counts of the lane's current
// class are kept in registers and written back to the lane's table when its
// class changes (rare when one class dominates).  Each lane keeps the first
// position of its largest crit; the caller merges lanes in draw order with
// the scalar rule, so the result is that of the columns scanned one by one.
struct LaneBest {
  float crit[8];
  int32_t at[8];
};

__attribute__((target("avx2"))) inline void gather_lanes(const ColRec* col, __m128i ilo, __m128i ihi, __m256& v,
                                                          __m256i& c) {
  const long long* base = reinterpret_cast<const long long*>(col);
  const __m256 lo = _mm256_castsi256_ps(_mm256_i32gather_epi64(base, ilo, 8));
  const __m256 hi = _mm256_castsi256_ps(_mm256_i32gather_epi64(base, ihi, 8));
  // lo = v0 l0 v1 l1 | v2 l2 v3 l3, hi = v4 l4 v5 l5 | v6 l6 v7 l7
  const __m256d vv = _mm256_castps_pd(_mm256_shuffle_ps(lo, hi, _MM_SHUFFLE(2, 0, 2, 0)));
  const __m256d ll = _mm256_castps_pd(_mm256_shuffle_ps(lo, hi, _MM_SHUFFLE(3, 1, 3, 1)));
  v = _mm256_castpd_ps(_mm256_permute4x64_pd(vv, _MM_SHUFFLE(3, 1, 2, 0)));
  c = _mm256_and_si256(_mm256_castpd_si256(_mm256_permute4x64_pd(ll, _MM_SHUFFLE(3, 1, 2, 0))),
                       _mm256_set1_epi32(0xff));
}

// base[q]: index in col of lane q's first record; every lane reads m records.
__attribute__((target("avx2"))) void scan_lanes_avx2(const ColRec* col, const int32_t* base, int32_t m, float pno,
                                                     int32_t nclass, const float* cc, int32_t* wl, int32_t* wr,
                                                     LaneBest& out) {
  alignas(32) int32_t a_[8], b_[8], cur_[8];
  __m128i ilo = _mm_loadu_si128(reinterpret_cast<const __m128i*>(base));
  __m128i ihi = _mm_loadu_si128(reinterpret_cast<const __m128i*>(base + 4));
  const __m128i one4 = _mm_set1_epi32(1);
  __m256 v1;
  __m256i c;
  gather_lanes(col, ilo, ihi, v1, c);
  _mm256_store_si256(reinterpret_cast<__m256i*>(cur_), c);
  for (int q = 0; q < 8; ++q) {
    for (int32_t j = 0; j < nclass; ++j) {
      wl[q * nclass + j] = 0;
      wr[q * nclass + j] = static_cast<int32_t>(cc[j]);
    }
    a_[q] = 0;
    b_[q] = static_cast<int32_t>(cc[cur_[q]]);
  }
  __m256i a = _mm256_load_si256(reinterpret_cast<const __m256i*>(a_));
  __m256i bb = _mm256_load_si256(reinterpret_cast<const __m256i*>(b_));
  __m256i cur = c;
  __m256i cnt = _mm256_setzero_si256();
  __m256 rln = _mm256_setzero_ps(), rrn = _mm256_set1_ps(pno);
  __m256 best = _mm256_set1_ps(-3.0e38f);
  __m256i bestat = _mm256_set1_epi32(-1);
  const __m256i one = _mm256_set1_epi32(1), vm = _mm256_set1_epi32(m);
  for (int32_t ii = 0; ii + 1 < m; ++ii) {
    if ((ii & 7) == 0) {
      // The gathers walk eight streams; fetch each a few lines ahead.
      for (int q = 0; q < 8; ++q) _mm_prefetch(reinterpret_cast<const char*>(col + base[q] + ii + 64), _MM_HINT_T0);
    }
    ilo = _mm_add_epi32(ilo, one4);
    ihi = _mm_add_epi32(ihi, one4);
    __m256 v2;
    __m256i c2;
    gather_lanes(col, ilo, ihi, v2, c2);
    const __m256 moved = _mm256_cmp_ps(v1, v2, _CMP_LT_OQ);
    const __m256i sw = _mm256_andnot_si256(_mm256_cmpeq_epi32(c, cur), _mm256_castps_si256(moved));
    if (__builtin_expect(!_mm256_testz_si256(sw, sw), 0)) {
      _mm256_store_si256(reinterpret_cast<__m256i*>(a_), a);
      _mm256_store_si256(reinterpret_cast<__m256i*>(b_), bb);
      _mm256_store_si256(reinterpret_cast<__m256i*>(cur_), cur);
      alignas(32) int32_t c_[8], sw_[8];
      _mm256_store_si256(reinterpret_cast<__m256i*>(c_), c);
      _mm256_store_si256(reinterpret_cast<__m256i*>(sw_), sw);
      for (int q = 0; q < 8; ++q) {
        if (!sw_[q]) continue;
        wl[q * nclass + cur_[q]] = a_[q];
        wr[q * nclass + cur_[q]] = b_[q];
        cur_[q] = c_[q];
        a_[q] = wl[q * nclass + c_[q]];
        b_[q] = wr[q * nclass + c_[q]];
      }
      a = _mm256_load_si256(reinterpret_cast<const __m256i*>(a_));
      bb = _mm256_load_si256(reinterpret_cast<const __m256i*>(b_));
      cur = _mm256_load_si256(reinterpret_cast<const __m256i*>(cur_));
    }
    const __m256i mv = _mm256_castps_si256(moved);
    const __m256 t1 = _mm256_cvtepi32_ps(_mm256_add_epi32(_mm256_add_epi32(a, a), one));
    const __m256 t2 = _mm256_cvtepi32_ps(_mm256_sub_epi32(one, _mm256_add_epi32(bb, bb)));
    rln = _mm256_blendv_ps(rln, _mm256_add_ps(rln, t1), moved);
    rrn = _mm256_blendv_ps(rrn, _mm256_add_ps(rrn, t2), moved);
    a = _mm256_sub_epi32(a, mv);
    bb = _mm256_add_epi32(bb, mv);
    cnt = _mm256_sub_epi32(cnt, mv);
    const __m256 dl = _mm256_cvtepi32_ps(cnt), dr = _mm256_cvtepi32_ps(_mm256_sub_epi32(vm, cnt));
    const __m256 crit = _mm256_add_ps(_mm256_div_ps(rln, dl), _mm256_div_ps(rrn, dr));
    const __m256 better = _mm256_and_ps(moved, _mm256_cmp_ps(crit, best, _CMP_GT_OQ));
    best = _mm256_blendv_ps(best, crit, better);
    bestat = _mm256_blendv_epi8(bestat, _mm256_set1_epi32(ii), _mm256_castps_si256(better));
    v1 = v2;
    c = c2;
  }
  _mm256_storeu_ps(out.crit, best);
  _mm256_storeu_si256(reinterpret_cast<__m256i*>(out.at), bestat);
}

// Stable partition of one column range by the side of every record.
inline void partition_flags(ColRec* c, int32_t m, const uint8_t* left, ColRec* tmp) {
  int32_t nl = 0, nr = 0;
  for (int32_t q = 0; q < m; ++q) {
    const ColRec x = c[q];
    const int32_t f = left[x.lc >> 8];
    c[nl] = x;
    tmp[nr] = x;
    nl += f;
    nr += 1 - f;
  }
  std::copy(tmp, tmp + nr, c + nl);
}

// A split that sends only a few records to one side ("peel": on the
// 16-component data, chains of tens of thousands of nodes each lose one
// sample) is applied to the split column at once -- its range is already in
// order -- but to every other column only when that column is next read.
// A column's pending peels are nested: each one's node is the larger child of
// the previous one.  They are applied in one pass over the outermost range:
// the peeled records are found by binary search on (value, local id), which
// is the range's order before any of them was applied (local ids ascend with
// the sample index), and the records between them move by whole segments to
// their final places.  Each record ends where applying the peels one by one,
// as stable partitions, would have put it.
inline int32_t find_record(const ColRec* c, int32_t lo, int32_t hi, float v, uint32_t l) {
  while (lo < hi) {
    const int32_t mid = (lo + hi) >> 1;
    const ColRec r = c[mid];
    if (r.v < v || (!(v < r.v) && (r.lc >> 8) < l))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Apply column v's pending peels.  c = the column, xv = &x[v] (row stride p).
void apply_peels(PeelLog& pl, int32_t v, ColRec* c, const float* xv, int32_t p, const int32_t* sid) {
  std::vector<int32_t>& ids = pl.pend[v];
  if (ids.empty()) return;
  const PeelEvent& outer = pl.events[ids.front()];
  const PeelEvent& inner = pl.events[ids.back()];
  const int32_t B = outer.b, E = outer.e;
  pl.rem.clear();
  for (int32_t id : ids) {
    const PeelEvent& ev = pl.events[id];
    pl.gpos.resize(ev.s);
    for (int32_t k = 0; k < ev.s; ++k) {
      const uint32_t l = static_cast<uint32_t>(pl.lids[ev.lid_at + k]);
      pl.gpos[k] = find_record(c, B, E, xv[static_cast<int64_t>(sid[l]) * p], l);
    }
    std::sort(pl.gpos.begin(), pl.gpos.end());
    const int32_t d0 = ev.few_left ? ev.b : ev.e - ev.s;
    for (int32_t k = 0; k < ev.s; ++k) pl.rem.push_back({pl.gpos[k], d0 + k});
  }
  std::sort(pl.rem.begin(), pl.rem.end());
  const int32_t R = static_cast<int32_t>(pl.rem.size());
  pl.saved.resize(R);
  for (int32_t k = 0; k < R; ++k) pl.saved[k] = c[pl.rem[k].first];
  // Segment j holds the records between removed positions j - 1 and j; its
  // records move by L - j, L = records peeled to the left.
  const int32_t L = PeelLog::big_b(inner) - B;
  auto seg = [&](int32_t j, int32_t& lo, int32_t& hi) {
    lo = j == 0 ? B : pl.rem[j - 1].first + 1;
    hi = j == R ? E : pl.rem[j].first;
  };
  for (int32_t j = L - 1; j >= 0; --j) {
    int32_t lo, hi;
    seg(j, lo, hi);
    if (hi > lo) std::memmove(c + lo + (L - j), c + lo, sizeof(ColRec) * static_cast<size_t>(hi - lo));
  }
  for (int32_t j = L + 1; j <= R; ++j) {
    int32_t lo, hi;
    seg(j, lo, hi);
    if (hi > lo) std::memmove(c + lo - (j - L), c + lo, sizeof(ColRec) * static_cast<size_t>(hi - lo));
  }
  for (int32_t k = 0; k < R; ++k) c[pl.rem[k].second] = pl.saved[k];
  ids.clear();
}

void grow_tree_presorted(const GrowParams& gp, const PresortedData& d, const int32_t* nin, uint32_t seed,
                         FastScratch& s, Tree& t) {
  const int32_t maxnode = gp.maxnode;
  const int32_t nclass = gp.nclass;
  const int32_t p = gp.p;
  const int64_t n = gp.n;
  Mt19937 rng(seed);

  std::vector<int32_t> status(maxnode, 0), bestvar(maxnode, 0), treemap(2 * static_cast<size_t>(maxnode), 0),
      nodeclass(maxnode, 0);
  std::vector<float> xsplit(maxnode, 0.0f);
  s.start.assign(maxnode, 0);
  s.end.assign(maxnode, 0);
  // A column that is in order over each node's range when the node is
  // reached (its parent's split column), and class counts handed down by
  // peel splits (offset into cc_pool, or -1).
  s.clean_col.assign(maxnode, 0);
  s.cc_at.assign(maxnode, -1);
  s.cc_pool.clear();
  s.peels.reset(p);

  // Unique in-bag samples, numbered in ascending sample order (jinbag).
  s.lid.resize(n);
  s.sid.clear();
  for (int64_t i = 0; i < n; ++i) {
    s.lid[i] = nin[i] > 0 ? static_cast<int32_t>(s.sid.size()) : -1;
    if (nin[i] > 0) s.sid.push_back(static_cast<int32_t>(i));
  }
  const int32_t m0 = static_cast<int32_t>(s.sid.size());
  s.col.resize(static_cast<size_t>(p) * m0);
  s.tmp.resize(m0);
  s.left.resize(m0);
  for (int32_t v = 0; v < p; ++v) {
    const ColRec* o = d.order.data() + static_cast<size_t>(v) * n;
    ColRec* c = s.col.data() + static_cast<size_t>(v) * m0;
    int32_t q = 0;
    for (int64_t r = 0; r < n; ++r) {
      const uint32_t i = o[r].lc;
      const int32_t l = s.lid[i];
      if (l >= 0) c[q++] = {o[r].v, (static_cast<uint32_t>(l) << 8) | static_cast<uint32_t>(gp.cl[i])};
    }
  }
  s.cc.assign(nclass, 0.0f);
  s.wl.assign(nclass, 0);
  s.wr.assign(nclass, 0);
  if (d.lanes) {
    s.lane_wl.resize(static_cast<size_t>(nclass) * 8);
    s.lane_wr.resize(static_cast<size_t>(nclass) * 8);
  }
  s.sb.ln.resize(m0);
  s.sb.rn.resize(m0);
  s.sb.at.resize(m0);
  auto column = [&](int32_t v) { return s.col.data() + static_cast<size_t>(v) * m0; };
  auto make_clean = [&](int32_t v) { apply_peels(s.peels, v, column(v), gp.x + v, p, s.sid.data()); };

  status[0] = 2;
  s.start[0] = 0;
  s.end[0] = m0;
  int32_t nnode = 1;

  for (int32_t k = 0; k < nnode; ++k) {
    if (status[k] != 2) continue;
    const int32_t b = s.start[k], e = s.end[k];
    const int32_t m = e - b;

    if (s.cc_at[k] >= 0) {
      std::copy(s.cc_pool.begin() + s.cc_at[k], s.cc_pool.begin() + s.cc_at[k] + nclass, s.cc.begin());
    } else {
      std::fill(s.cc.begin(), s.cc.end(), 0.0f);
      const ColRec* c0 = column(s.clean_col[k]) + b;
      for (int32_t q = 0; q < m; ++q) s.cc[c0[q].lc & 0xffu] += 1.0f;
    }
    int32_t nonzero = 0, majority = 0;
    float maxc = 0.0f;
    for (int32_t j = 0; j < nclass; ++j) {
      if (s.cc[j] > 0.0f) {
        ++nonzero;
        if (s.cc[j] > maxc) {
          maxc = s.cc[j];
          majority = j;
        }
      }
    }
    nodeclass[k] = majority;
    if (m < gp.minndsize || nonzero <= 1) {
      status[k] = -1;
      continue;
    }

    float pno = 0.0f;
    for (int32_t j = 0; j < nclass; ++j) {
      if (gp.fma_gini)
        pno = std::fmaf(s.cc[j], s.cc[j], pno);
      else
        pno = pno + s.cc[j] * s.cc[j];
    }

    ScanBest best{-2.0f, 0, 0.0f, 0};
    // The drawn variables do not depend on the scans, so they are drawn first.
    s.vars.resize(gp.mtry);
    for (int32_t mv = 0; mv < gp.mtry; ++mv) {
      float u = rng.next_float();
      int32_t mvar = static_cast<int32_t>(u * static_cast<float>(p));
      if (mvar >= p) mvar = p - 1;
      s.vars[mv] = mvar;
      make_clean(mvar);
    }
    if (d.lanes && m >= kLaneScanMin && maxc >= 0.9f * static_cast<float>(m)) {
      // Columns in groups of eight lanes, merged in draw order.
      for (int32_t g0 = 0; g0 < gp.mtry; g0 += 8) {
        const int32_t gn = std::min(8, gp.mtry - g0);
        alignas(16) int32_t base[8];
        for (int q = 0; q < 8; ++q) base[q] = s.vars[g0 + (q < gn ? q : 0)] * m0 + b;
        LaneBest lb;
        scan_lanes_avx2(s.col.data(), base, m, pno, nclass, s.cc.data(), s.lane_wl.data(), s.lane_wr.data(), lb);
        for (int32_t q = 0; q < gn; ++q) {
          if (lb.at[q] >= 0 && lb.crit[q] > best.crit) {
            best.crit = lb.crit[q];
            best.var = s.vars[g0 + q];
            best.split = s.col[static_cast<size_t>(base[q]) + lb.at[q]].v;
            best.nl = lb.at[q] + 1;
          }
        }
      }
    } else {
      for (int32_t mv = 0; mv < gp.mtry; ++mv)
        scan_column(column(s.vars[mv]) + b, m, s.vars[mv], pno, nclass, s.cc.data(), s.wl.data(), s.wr.data(),
                    s.sb, d.avx2, best);
    }

    if (best.crit > -1.0f && nnode + 2 <= maxnode) {
      const int32_t lc = nnode, rc = nnode + 1;
      status[k] = 1;
      bestvar[k] = best.var;
      xsplit[k] = best.split;
      treemap[2 * static_cast<size_t>(k)] = lc;
      treemap[2 * static_cast<size_t>(k) + 1] = rc;
      status[lc] = 2;
      status[rc] = 2;
      s.clean_col[lc] = best.var;
      s.clean_col[rc] = best.var;
      // Sides: the split column's left part is its first best.nl records.
      const ColRec* cb = column(best.var) + b;
      const int32_t nl = best.nl, nr = m - nl;
      const int32_t few = std::min(nl, nr);
      if (static_cast<int64_t>(few) * 64 <= m) {
        PeelEvent ev{b, e, few, nl <= nr, static_cast<int32_t>(s.peels.lids.size())};
        const ColRec* fc = ev.few_left ? cb : cb + nl;
        // Class counts of both children, from the parent's and the few records'.
        const int32_t at_few = static_cast<int32_t>(s.cc_pool.size());
        s.cc_pool.resize(at_few + 2 * static_cast<size_t>(nclass), 0.0f);
        float* ccf = s.cc_pool.data() + at_few;
        float* ccb = ccf + nclass;
        for (int32_t q = 0; q < few; ++q) {
          s.peels.lids.push_back(static_cast<int32_t>(fc[q].lc >> 8));
          ccf[fc[q].lc & 0xffu] += 1.0f;
        }
        for (int32_t j = 0; j < nclass; ++j) ccb[j] = s.cc[j] - ccf[j];
        s.cc_at[ev.few_left ? lc : rc] = at_few;
        s.cc_at[ev.few_left ? rc : lc] = at_few + nclass;
        const int32_t id = static_cast<int32_t>(s.peels.events.size());
        s.peels.events.push_back(ev);
        for (int32_t v = 0; v < p; ++v) {
          if (v == best.var) continue;
          std::vector<int32_t>& pv = s.peels.pend[v];
          if (!pv.empty()) {
            const PeelEvent& last = s.peels.events[pv.back()];
            if (b < PeelLog::big_b(last) || e > PeelLog::big_e(last)) make_clean(v);
          }
          pv.push_back(id);
        }
      } else {
        for (int32_t v = 0; v < p; ++v) make_clean(v);
        for (int32_t q = 0; q < m; ++q) s.left[cb[q].lc >> 8] = q < nl ? 1 : 0;
        for (int32_t v = 0; v < p; ++v) {
          if (v == best.var) continue;
          partition_flags(column(v) + b, m, s.left.data(), s.tmp.data());
        }
      }
      s.start[lc] = b;
      s.end[lc] = b + nl;
      s.start[rc] = b + nl;
      s.end[rc] = e;
      nnode += 2;
    } else {
      status[k] = -1;
    }
  }

  status.resize(nnode);
  bestvar.resize(nnode);
  xsplit.resize(nnode);
  treemap.resize(2 * static_cast<size_t>(nnode));
  nodeclass.resize(nnode);
  t.status = std::move(status);
  t.bestvar = std::move(bestvar);
  t.xsplit = std::move(xsplit);
  t.treemap = std::move(treemap);
  t.nodeclass = std::move(nodeclass);
  t.nnode = nnode;
}

// ---------------------------------------------------------------------------
// Counter-based generator for the synthetic datasets: SplitMix64 finaliser of
// (seed, stream, index).  Identical on every host because it is pure integer
// arithmetic followed by libm calls of the same image.
// ---------------------------------------------------------------------------
inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
struct Crng {
  uint64_t key;
  Crng(uint64_t seed, uint64_t stream) : key(mix64(seed * 0x100000001b3ull ^ mix64(stream + 0x632be59bd9b4e019ull))) {}
  uint64_t u64(uint64_t i) const { return mix64(key ^ mix64(i)); }
  double uniform(uint64_t i) const { return (static_cast<double>(u64(i) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal(uint64_t i) const {  // Box-Muller, one variate per index
    double u1 = uniform(2 * i), u2 = uniform(2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int64_t below(uint64_t i, int64_t m) const { return static_cast<int64_t>(u64(i) % static_cast<uint64_t>(m)); }
};

// 3-class isotropic Gaussian mixture: means ~ N(0, 4I), unit variance.
void make_gauss_mixture(int64_t n, int32_t p, int32_t nclass, double mean_sd, uint64_t seed, float* x, int32_t* y) {
  Crng rm(seed, 1), rc(seed, 2), rx(seed, 3);
  std::vector<double> mu(static_cast<size_t>(nclass) * p);
  for (size_t q = 0; q < mu.size(); ++q) mu[q] = mean_sd * rm.normal(q);
  for (int64_t i = 0; i < n; ++i) {
    int32_t c = static_cast<int32_t>(rc.below(i, nclass));
    y[i] = c;
    for (int32_t m = 0; m < p; ++m)
      x[i * p + m] = static_cast<float>(mu[static_cast<size_t>(c) * p + m] + rx.normal(static_cast<uint64_t>(i) * p + m));
  }
}

// Covertype-like: 10 lognormal(0,1) columns min-max rescaled to [0, 3000],
// a 4-way and a 40-way one-hot block (44 binary columns), 7 imbalanced classes
// from a random depth-8 axis-aligned tree plus 5% label noise.
void make_covtype_like(int64_t n, uint64_t seed, float* x, int32_t* y) {
  const int32_t p = 54;
  Crng rx(seed, 11), rcat(seed, 12), rtree(seed, 13), rnoise(seed, 14);
  for (int m = 0; m < 10; ++m) {
    double lo = 1e300, hi = -1e300;
    std::vector<double> col(n);
    for (int64_t i = 0; i < n; ++i) {
      col[i] = std::exp(rx.normal(static_cast<uint64_t>(m) * n + i));
      lo = std::min(lo, col[i]);
      hi = std::max(hi, col[i]);
    }
    const double sc = hi > lo ? 3000.0 / (hi - lo) : 0.0;
    for (int64_t i = 0; i < n; ++i) x[i * p + m] = static_cast<float>((col[i] - lo) * sc);
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t a = rcat.below(2 * i, 4), b2 = rcat.below(2 * i + 1, 40);
    for (int m = 10; m < p; ++m) x[i * p + m] = 0.0f;
    x[i * p + 10 + a] = 1.0f;
    x[i * p + 14 + b2] = 1.0f;
  }
  // Depth-8 label tree: heap-ordered internal nodes 0..254, leaves 255..510.
  const double pri[7] = {0.36, 0.49, 0.06, 0.005, 0.016, 0.03, 0.035};
  int32_t fvar[255];
  float fthr[255];
  int32_t leafcls[256];
  uint64_t q = 0;
  for (int v = 0; v < 255; ++v) {
    fvar[v] = static_cast<int32_t>(rtree.below(q++, p));
    int64_t who = rtree.below(q++, n);
    fthr[v] = x[who * p + fvar[v]];
    if (fvar[v] >= 10) fthr[v] = 0.5f;  // binary column: split 0 | 1
  }
  // Leaf of every sample, then classes given to leaves in a random order so
  // that the class masses follow the priors: each leaf takes the class with
  // the largest remaining deficit prior_c * n - assigned_c.
  std::vector<int32_t> leaf(n);
  std::vector<int64_t> mass(256, 0);
  for (int64_t i = 0; i < n; ++i) {
    int v = 0;
    while (v < 255) v = 2 * v + 1 + (x[i * p + fvar[v]] <= fthr[v] ? 0 : 1);
    leaf[i] = v - 255;
    mass[v - 255] += 1;
  }
  int order[256];
  for (int l = 0; l < 256; ++l) order[l] = l;
  for (int l = 255; l > 0; --l) std::swap(order[l], order[static_cast<int>(rtree.below(q++, l + 1))]);
  double assigned[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int r = 0; r < 256; ++r) {
    const int l = order[r];
    int best = 0;
    double bd = -1e300;
    for (int c = 0; c < 7; ++c) {
      const double d = pri[c] * static_cast<double>(n) - assigned[c];
      if (d > bd) {
        bd = d;
        best = c;
      }
    }
    leafcls[l] = best;
    assigned[best] += static_cast<double>(mass[l]);
  }
  for (int64_t i = 0; i < n; ++i) {
    int32_t c = leafcls[leaf[i]];
    if (rnoise.uniform(2 * i) < 0.05) {
      int32_t d = static_cast<int32_t>(rnoise.below(2 * i + 1, 6));
      c = d >= c ? d + 1 : d;
    }
    y[i] = c;
  }
}

// 16-component Gaussian mixture in a random 8-dim subspace of the first 8
// columns plus (p - 8) N(0,1) noise columns; 10 balanced classes, classes
// 0..5 spread over two components each.
void make_subspace_mixture(int64_t n, int32_t p, uint64_t seed, float* x, int32_t* y) {
  Crng rm(seed, 21), rc(seed, 22), rx(seed, 23), rq(seed, 24);
  const int d = 8, K = 16;
  double mu[K][d];
  for (int k = 0; k < K; ++k)
    for (int j = 0; j < d; ++j) mu[k][j] = 3.0 * rm.normal(static_cast<uint64_t>(k) * d + j);
  // Random rotation by Gram-Schmidt of a Gaussian 8x8.
  double Q[d][d];
  for (int a = 0; a < d; ++a)
    for (int b = 0; b < d; ++b) Q[a][b] = rq.normal(static_cast<uint64_t>(a) * d + b);
  for (int a = 0; a < d; ++a) {
    for (int c = 0; c < a; ++c) {
      double dot = 0;
      for (int b = 0; b < d; ++b) dot += Q[a][b] * Q[c][b];
      for (int b = 0; b < d; ++b) Q[a][b] -= dot * Q[c][b];
    }
    double nr = 0;
    for (int b = 0; b < d; ++b) nr += Q[a][b] * Q[a][b];
    nr = std::sqrt(nr);
    for (int b = 0; b < d; ++b) Q[a][b] /= nr;
  }
  for (int64_t i = 0; i < n; ++i) {
    int32_t c = static_cast<int32_t>(rc.below(2 * i, 10));
    int k = c;
    if (c < 6 && rc.below(2 * i + 1, 2) == 1) k = c + 10;
    y[i] = c;
    double z[d];
    for (int j = 0; j < d; ++j) z[j] = mu[k][j] + rx.normal(static_cast<uint64_t>(i) * p + j);
    for (int a = 0; a < d; ++a) {
      double v = 0;
      for (int b = 0; b < d; ++b) v += Q[a][b] * z[b];
      x[i * p + a] = static_cast<float>(v);
    }
    for (int m = d; m < p; ++m) x[i * p + m] = static_cast<float>(rx.normal(static_cast<uint64_t>(i) * p + m));
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

struct rfx_forest {
  int64_t n = 0;
  int32_t p = 0;
  std::vector<Tree> trees;
  std::vector<int32_t> nin;  // [T][n]
};

extern "C" {

uint32_t rfx_mt19937_u32(uint32_t seed, int64_t index) {
  Mt19937 r(seed);
  uint32_t v = 0;
  for (int64_t i = 0; i <= index; ++i) v = r.next_u32();
  return v;
}

int rfx_bootstrap(int64_t n, uint32_t seed, int32_t* nin) {
  if (n <= 0 || !nin) return RFX_ERR_INVALID;
  bootstrap_tree(n, seed, nin);
  return RFX_OK;
}

int rfx_make_dataset(int32_t kind, int64_t n, int32_t p, int32_t nclass, uint64_t seed, float* x, int32_t* y) {
  if (n <= 0 || !x || !y) return RFX_ERR_INVALID;
  switch (kind) {
    case RFX_DATA_GAUSS_MIXTURE:
      if (p <= 0 || nclass < 2) return RFX_ERR_INVALID;
      make_gauss_mixture(n, p, nclass, 2.0, seed, x, y);
      return RFX_OK;
    case RFX_DATA_COVTYPE_LIKE:
      if (p != 54) return RFX_ERR_INVALID;
      make_covtype_like(n, seed, x, y);
      return RFX_OK;
    case RFX_DATA_SUBSPACE_MIXTURE:
      if (p < 8) return RFX_ERR_INVALID;
      make_subspace_mixture(n, p, seed, x, y);
      return RFX_OK;
    case RFX_DATA_OVERLAP_MIXTURE:
      if (p <= 0 || nclass < 2) return RFX_ERR_INVALID;
      make_gauss_mixture(n, p, nclass, 0.35, seed, x, y);
      return RFX_OK;
    default:
      return RFX_ERR_INVALID;
  }
}

int rfx_forest_grow(const float* x, const int32_t* y, int64_t n, int32_t p, int32_t nclass, int32_t ntree,
                    int32_t mtry, int32_t minndsize, int32_t maxnode, int32_t iseed, int32_t boot_seed,
                    int32_t flags, int32_t nthreads, rfx_forest** out) {
  return rfx_forest_grow_progress(x, y, n, p, nclass, ntree, mtry, minndsize, maxnode, iseed, boot_seed, flags,
                                  nthreads, nullptr, out);
}

int rfx_forest_grow_progress(const float* x, const int32_t* y, int64_t n, int32_t p, int32_t nclass, int32_t ntree,
                             int32_t mtry, int32_t minndsize, int32_t maxnode, int32_t iseed, int32_t boot_seed,
                             int32_t flags, int32_t nthreads, int32_t* trees_done, rfx_forest** out) {
  if (trees_done) __atomic_store_n(trees_done, 0, __ATOMIC_RELAXED);
  if (!x || !y || !out || n <= 0 || n > RFX_MAX_SAMPLES || p <= 0 || p > RFX_MAX_FEATURES || nclass < 2 ||
      ntree <= 0 || mtry <= 0 || minndsize < 1)
    return RFX_ERR_INVALID;
  if (maxnode <= 0) {
    int64_t mn = 2 * n + 1;
    maxnode = static_cast<int32_t>(mn);
  }
  rfx_forest* f = new (std::nothrow) rfx_forest;
  if (!f) return RFX_ERR_ALLOC;
  f->n = n;
  f->p = p;
  try {
    f->trees.resize(ntree);
    f->nin.assign(static_cast<size_t>(ntree) * n, 0);
  } catch (...) {
    delete f;
    return RFX_ERR_ALLOC;
  }
  GrowParams gp{x, y, n, p, nclass, maxnode, minndsize, mtry, (flags & RFX_GROW_NO_FMA) == 0};
  int nt = nthreads > 0 ? nthreads : static_cast<int>(std::thread::hardware_concurrency());
  if (nt < 1) nt = 1;
  // The presorted grower needs byte-sized valid labels and NaN-free features.
  bool presorted = (flags & RFX_GROW_PLAIN) == 0 && nclass <= 256;
  for (int64_t i = 0; presorted && i < n; ++i) presorted = y[i] >= 0 && y[i] < nclass;
  for (int64_t i = 0; presorted && i < n * p; ++i) presorted = !std::isnan(x[i]);
  PresortedData pd;
  if (presorted) {
    try {
      presort_columns(gp, nt, pd);
    } catch (...) {
      delete f;
      return RFX_ERR_ALLOC;
    }
  }
  if (nt > ntree) nt = ntree;
  std::atomic<int32_t> next{0};
  std::atomic<int> failed{0};
  auto worker = [&]() {
    GrowScratch s;
    FastScratch fs;
    for (;;) {
      int32_t t = next.fetch_add(1);
      if (t >= ntree) break;
      try {
        int32_t* nin = f->nin.data() + static_cast<size_t>(t) * n;
        // Bootstrap seed 12345 + t and grower seed iseed + t: the reference
        // seeds its bootstrap from the never-assigned g_config.iseed = 12345
        // (src/rf_bootstrap.cpp:24, include/rf_config.hpp:61) and its grower
        // from the user's iseed (src/rf_random_forest.cpp:515).
        bootstrap_tree(n, static_cast<uint32_t>(boot_seed + t), nin);
        if (presorted)
          grow_tree_presorted(gp, pd, nin, static_cast<uint32_t>(iseed + t), fs, f->trees[t]);
        else
          grow_tree(gp, nin, static_cast<uint32_t>(iseed + t), s, f->trees[t]);
        if (trees_done) __atomic_add_fetch(trees_done, 1, __ATOMIC_RELAXED);
      } catch (...) {
        failed = 1;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int i = 1; i < nt; ++i) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  if (failed) {
    delete f;
    return RFX_ERR_ALLOC;
  }
  *out = f;
  return RFX_OK;
}

void rfx_forest_free(rfx_forest* f) { delete f; }

int32_t rfx_forest_ntree(const rfx_forest* f) { return f ? static_cast<int32_t>(f->trees.size()) : 0; }

int32_t rfx_forest_nnode(const rfx_forest* f, int32_t t) {
  if (!f || t < 0 || t >= static_cast<int32_t>(f->trees.size())) return 0;
  return f->trees[t].nnode;
}

int rfx_forest_export(const rfx_forest* f, int64_t* node_offset, int32_t* nodestatus, int32_t* bestvar,
                      float* xbestsplit, int32_t* treemap, int32_t* nodeclass, int32_t* nin) {
  if (!f || !node_offset) return RFX_ERR_INVALID;
  const int32_t T = static_cast<int32_t>(f->trees.size());
  int64_t off = 0;
  for (int32_t t = 0; t < T; ++t) {
    node_offset[t] = off;
    const Tree& tr = f->trees[t];
    if (nodestatus) std::copy(tr.status.begin(), tr.status.end(), nodestatus + off);
    if (bestvar) std::copy(tr.bestvar.begin(), tr.bestvar.end(), bestvar + off);
    if (xbestsplit) std::copy(tr.xsplit.begin(), tr.xsplit.end(), xbestsplit + off);
    if (treemap) std::copy(tr.treemap.begin(), tr.treemap.end(), treemap + 2 * off);
    if (nodeclass) std::copy(tr.nodeclass.begin(), tr.nodeclass.end(), nodeclass + off);
    off += tr.nnode;
  }
  node_offset[T] = off;
  if (nin) std::copy(f->nin.begin(), f->nin.end(), nin);
  return RFX_OK;
}

// Pack a forest given as the reference's per-node arrays (flattened over
// trees by node_offset) into the 8-byte node records the leaf kernel reads:
// .x = split threshold bits, .y = (left child << 10) | split feature, or -1
// for a terminal node.  The right child is always left + 1
// (src/rf_growtree.cpp:755-777).
int rfx_pack_nodes(const int64_t* node_offset, int32_t ntree, const int32_t* nodestatus, const int32_t* bestvar,
                   const float* xbestsplit, const int32_t* treemap, int32_t p, uint32_t* packed) {
  if (!node_offset || !nodestatus || !bestvar || !xbestsplit || !treemap || !packed || ntree <= 0 || p <= 0 ||
      p > RFX_MAX_FEATURES)
    return RFX_ERR_INVALID;
  for (int32_t t = 0; t < ntree; ++t) {
    const int64_t b = node_offset[t], e = node_offset[t + 1];
    const int64_t nn = e - b;
    if (nn <= 0 || nn > RFX_MAX_NODES_PER_TREE) return RFX_ERR_INVALID;
    for (int64_t k = 0; k < nn; ++k) {
      const int64_t g = b + k;
      uint32_t thr;
      std::memcpy(&thr, &xbestsplit[g], 4);
      packed[2 * g] = thr;
      if (nodestatus[g] == -1) {
        packed[2 * g + 1] = 0xffffffffu;
      } else if (nodestatus[g] == 1) {
        const int32_t l = treemap[2 * g], r = treemap[2 * g + 1], v = bestvar[g];
        if (l <= k || r != l + 1 || r >= nn || v < 0 || v >= p) return RFX_ERR_INVALID;
        packed[2 * g + 1] = (static_cast<uint32_t>(l) << 10) | static_cast<uint32_t>(v);
      } else {
        return RFX_ERR_INVALID;
      }
    }
  }
  return RFX_OK;
}

// OOB counts per sample: o_i = #{t : nin_t(i) = 0}, times `mult` (the
// reference CPU path counts each OOB event twice: src/rf_bootstrap.cpp:61 and
// src/rf_random_forest.cpp:1108-1110, so it passes mult = 2).
int rfx_oob_counts(const int32_t* nin, int64_t n, int32_t ntree, int32_t mult, int32_t* nout) {
  if (!nin || !nout || n <= 0 || ntree <= 0) return RFX_ERR_INVALID;
  std::fill(nout, nout + n, 0);
  for (int32_t t = 0; t < ntree; ++t)
    for (int64_t i = 0; i < n; ++i)
      if (nin[static_cast<size_t>(t) * n + i] == 0) nout[i] += mult;
  return RFX_OK;
}

}  // extern "C"

extern "C" int rfx_pack_inbag_bits(const int32_t* nin, int64_t n, int32_t ntree, uint32_t* bits) {
  if (!nin || !bits || n <= 0 || ntree <= 0) return RFX_ERR_INVALID;
  const int64_t W = (n + 31) / 32;
  std::fill(bits, bits + W * ntree, 0u);
  for (int32_t t = 0; t < ntree; ++t)
    for (int64_t i = 0; i < n; ++i)
      if (nin[static_cast<size_t>(t) * n + i] > 0) bits[t * W + i / 32] |= 1u << (i % 32);
  return RFX_OK;
}
