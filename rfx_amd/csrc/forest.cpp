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
// Gini arithmetic runs in the same sequence as the reference's.
//
// Everything here is plain host C++ (no HIP): the forest is an input of the
// path, grown once and exported; it is not part of the timed GPU work.

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
  if (nt > ntree) nt = ntree;
  std::atomic<int32_t> next{0};
  std::atomic<int> failed{0};
  auto worker = [&]() {
    GrowScratch s;
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
        grow_tree(gp, nin, static_cast<uint32_t>(iseed + t), s, f->trees[t]);
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
