// ref_driver.cpp -- a thin driver over the reference's own CPU sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile (target `ref`) compiles this
// file together with the reference sources it calls (read in place from the
// reference tree, never copied) into oracle/_ref/librfxref.so.
//
//   ref_run_forest  repeats what RandomForest::grow_tree_single does per tree
//                   (src/rf_random_forest.cpp:928-1308) to expose the per-tree
//                   arrays the class keeps private: rf::cpu_bootstrap,
//                   rf::cpu_growtree (through rf::cpu_growtree_wrapper),
//                   rf::cpu_testreebag for out-of-bag leaves, an in-bag
//                   traversal written here after RandomForest::find_terminal_node
//                   (src/rf_random_forest.cpp:2194-2244), the second nout
//                   increment (src/rf_random_forest.cpp:1108-1110), and
//                   rf::cpu_proximity.  The tests pin that in-bag traversal
//                   against the class itself (ref_rf_fit gives the same
//                   proximity only if every in-bag leaf agrees).
//   ref_rf_fit      the reference's RandomForest class end to end on the CPU
//                   (fit -> grow_tree_single per tree -> finalize_training):
//                   finished proximity, nout and, for the unsupervised task,
//                   RandomForest::compute_outlier_scores.
//   ref_mds_3d      rf::compute_mds_3d_cpu (src/rf_mds_cpu.cpp:88-341).  Its
//                   LAPACK dsyevd_ is bound at link time to the LAPACK in this
//                   image (scipy's bundled OpenBLAS, scipy_dsyevd_) by renaming
//                   the undefined symbol of the compiled object.
//   ref_finishprox, ref_proximity_rfgap  rf::cpu_finishprox and
//                   rf::cpu_proximity_rfgap.
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "rf_bootstrap.hpp"
#include "rf_config.hpp"
#include "rf_finishprox.hpp"
#include "rf_growtree.hpp"
#include "rf_mds_cpu.hpp"
#include "rf_proximity.hpp"
#include "rf_random_forest.hpp"
#include "rf_testreebag.hpp"

using rf::integer_t;
using rf::real_t;

namespace {

integer_t inbag_leaf(const real_t* xi, const integer_t* st, const integer_t* bv, const real_t* xs, const integer_t* tm,
                     const integer_t* catgoleft, const integer_t* cat, integer_t mdim, integer_t maxnode,
                     integer_t maxcat) {
  integer_t k = 0;
  for (integer_t it = 0; it < maxnode; ++it) {
    if (st[k] == -1) return k;
    const integer_t v = bv[k];
    if (v < 0 || v >= mdim) return 0;
    const real_t f = xi[v];
    if (cat[v] == 1) {
      k = f <= xs[k] ? tm[2 * k] : tm[2 * k + 1];
    } else {
      const integer_t jcat = static_cast<integer_t>(f + 0.5);
      if (jcat >= 1 && jcat <= cat[v])
        k = catgoleft[k * maxcat + (jcat - 1)] == 1 ? tm[2 * k] : tm[2 * k + 1];
      else
        return 0;
    }
  }
  return 0;
}

}  // namespace

extern "C" int ref_run_forest(const float* x, const int32_t* y, int32_t n, int32_t p, int32_t nclass, int32_t ntree,
                              int32_t mtry, int32_t minndsize, int32_t iseed, const int32_t* cat_in, int32_t maxcat,
                              int32_t do_prox, int32_t* nnode_out, int32_t* nodestatus, int32_t* bestvar,
                              float* xbestsplit, int32_t* treemap, int32_t* nin_out, int32_t* nodextr_out,
                              int32_t* nout, double* prox, int32_t casewise) {
  const integer_t maxnode = 2 * n + 1;
  if (maxcat < 1) maxcat = 1;
  rf::g_config.use_casewise = casewise != 0;
  rf::g_config.use_sparse = false;
  std::vector<integer_t> cat(p, 1);
  if (cat_in)
    for (int m = 0; m < p; ++m) cat[m] = cat_in[m];
  const size_t big = static_cast<size_t>(n) + maxnode + 16;
  std::vector<real_t> weight(n, 1.0f), win(n), tnodewt(maxnode), tcat(big), classpop(big * nclass), wr(big),
      wl(big), tclasscat(big * nclass), tclasspop(big), tmpclass(big);
  std::vector<integer_t> jinbag(n), joobag(n), amat(static_cast<size_t>(p) * n + 16), ties(big), jtr(n),
      nodextr_ws(n), idmove(big), igoleft(big), igl(big), nodestart(big), nodestop(big), itemp(big), itmp(big),
      icat(big), catgoleft(static_cast<size_t>(maxcat) * maxnode), nodeclass(maxnode);
  std::vector<integer_t> nod(maxnode), ncount(n), ncn(n), nodexb(n), ndbegin(maxnode + 1), npcase(n);
  for (integer_t i = 0; i < n; ++i) nout[i] = 0;
  for (integer_t t = 0; t < ntree; ++t) {
    integer_t* nin = nin_out + static_cast<size_t>(t) * n;
    integer_t* st = nodestatus + static_cast<size_t>(t) * maxnode;
    integer_t* bv = bestvar + static_cast<size_t>(t) * maxnode;
    real_t* xs = xbestsplit + static_cast<size_t>(t) * maxnode;
    integer_t* tm = treemap + static_cast<size_t>(t) * 2 * maxnode;
    integer_t* ndx = nodextr_out + static_cast<size_t>(t) * n;
    integer_t ninbag = 0, noobag = 0, nnode = 0;
    rf::cpu_bootstrap(weight.data(), n, win.data(), nin, nout, jinbag.data(), joobag.data(), ninbag, noobag, t);
    for (integer_t k = 0; k < maxnode; ++k) {
      st[k] = 0;
      bv[k] = 0;
      xs[k] = 0.0f;
      tm[2 * k] = 0;
      tm[2 * k + 1] = 0;
    }
    rf::cpu_growtree_wrapper(x, win.data(), y, cat.data(), ties.data(), nin, p, n, nclass, maxnode, minndsize, mtry,
                             ninbag, maxcat, iseed + t, 25, 25, amat.data(), jinbag.data(), tnodewt.data(), xs, st, bv,
                             tm, catgoleft.data(), nodeclass.data(), jtr.data(), nodextr_ws.data(), nnode,
                             idmove.data(), igoleft.data(), igl.data(), nodestart.data(), nodestop.data(),
                             itemp.data(), itmp.data(), icat.data(), tcat.data(), classpop.data(), wr.data(), wl.data(),
                             tclasscat.data(), tclasspop.data(), tmpclass.data());
    nnode_out[t] = nnode;
    rf::cpu_testreebag(x, xs, nin, tm, bv, nodeclass.data(), cat.data(), st, catgoleft.data(), n, p, nnode, maxcat,
                       jtr.data(), ndx);
    for (integer_t i = 0; i < n; ++i) {
      if (nin[i] == 0) {
        nout[i] += 1;
      } else {
        ndx[i] = inbag_leaf(x + static_cast<size_t>(i) * p, st, bv, xs, tm, catgoleft.data(), cat.data(), p, maxnode,
                            maxcat);
      }
    }
    if (do_prox && prox)
      rf::cpu_proximity(st, ndx, nin, n, nnode, prox, nod.data(), ncount.data(), ncn.data(), nodexb.data(),
                        ndbegin.data(), npcase.data());
  }
  return 0;
}

extern "C" int ref_finishprox(int32_t n, const int32_t* nout, double* prox, double* proxsym) {
  rf::cpu_finishprox(n, nout, prox, proxsym);
  return 0;
}

// rf::cpu_proximity_rfgap on flattened [tree * n + sample] arrays (its
// per-tree vectors are built here, as RandomForest keeps them).
extern "C" int ref_proximity_rfgap(int32_t ntree, int32_t n, const int32_t* nin, const int32_t* nodextr, double* prox) {
  std::vector<std::vector<integer_t>> tn(ntree), tx(ntree);
  for (int32_t t = 0; t < ntree; ++t) {
    tn[t].assign(nin + static_cast<size_t>(t) * n, nin + static_cast<size_t>(t + 1) * n);
    tx[t].assign(nodextr + static_cast<size_t>(t) * n, nodextr + static_cast<size_t>(t + 1) * n);
  }
  rf::cpu_proximity_rfgap(ntree, n, tn, tx, prox);
  return 0;
}

// RandomForest fit on the CPU.  task: 0 classification (y int32 labels),
// 2 unsupervised (y ignored).  cat: per-feature category counts or NULL (all
// numeric), set through RandomForest::set_categorical_features as the
// classifier does.  Outputs (any may be NULL): prox n*n float64 as
// get_proximity_matrix returns it, nout [n] (get_nout_ptr), outlier [n]
// float32 (get_outlier_scores, unsupervised only).
extern "C" int ref_rf_fit(const float* x, const int32_t* y, int32_t n, int32_t p, int32_t nclass, int32_t ntree,
                          int32_t mtry, int32_t minndsize, int32_t iseed, const int32_t* cat, int32_t maxcat,
                          int32_t task, int32_t casewise, double* prox_out, int32_t* nout_out, float* outlier_out) {
  try {
    rf::RandomForestConfig cfg;
    cfg.nsample = n;
    cfg.ntree = ntree;
    cfg.mdim = p;
    cfg.nclass = nclass;
    cfg.maxcat = maxcat < 1 ? 1 : maxcat;
    cfg.mtry = mtry;
    cfg.maxnode = 2 * n + 1;
    cfg.minndsize = minndsize;
    cfg.iseed = iseed;
    cfg.task_type = task == 2 ? rf::TaskType::UNSUPERVISED : rf::TaskType::CLASSIFICATION;
    cfg.compute_proximity = true;
    cfg.compute_importance = false;
    cfg.use_casewise = casewise != 0;
    cfg.use_gpu = false;
    cfg.execution_mode = "cpu";
    rf::RandomForest forest(cfg);
    if (cat) forest.set_categorical_features(cat, p);
    if (task == 2)
      forest.fit_unsupervised(x);
    else
      forest.fit(x, y);
    const rf::dp_t* pm = forest.get_proximity_matrix();
    if (prox_out) {
      if (!pm) return -2;
      std::memcpy(prox_out, pm, sizeof(double) * static_cast<size_t>(n) * n);
    }
    if (nout_out) std::memcpy(nout_out, forest.get_nout_ptr(), sizeof(int32_t) * n);
    if (outlier_out) std::memcpy(outlier_out, forest.get_outlier_scores(), sizeof(float) * n);
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

// rf::compute_mds_3d_cpu(prox, n, false, nout, use_rfgap) -> coords n x 3
// row-major.  prox is column-major float64 as the pybind wrapper passes it.
extern "C" int ref_mds_3d(const double* prox, int32_t n, const int32_t* nout, int32_t use_rfgap, double* coords) {
  try {
    const std::vector<double> c = rf::compute_mds_3d_cpu(prox, n, false, nout, use_rfgap != 0);
    if (c.size() != static_cast<size_t>(n) * 3) return -2;
    std::memcpy(coords, c.data(), sizeof(double) * c.size());
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}
