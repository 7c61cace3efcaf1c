/*
 * oracle.c -- CPU restatement of the reference proximity path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (rfx_amd/) loads this
 * library: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use it as the checker and as the timed CPU port of the reference's OpenMP
 * path.  Built by oracle/Makefile with -ffp-contract=off.
 *
 * Forests are given as the reference's own per-node arrays, flattened over
 * trees by node_offset[T+1] (nodestatus, bestvar, xbestsplit, treemap with
 * two entries per node), plus nin[T][n] bootstrap multiplicities and the
 * row-major sample matrix x[i*p + m].
 *
 * Functions and the reference code each follows:
 *   orc_leaf_assign     src/rf_testreebag.cpp:13-127 (out-of-bag samples) and
 *                       src/rf_random_forest.cpp:2194-2244 (in-bag samples),
 *                       combined as at src/rf_random_forest.cpp:1043, 1295-1303
 *   orc_prox_v1         src/rf_proximity.cpp:99-231, called once per tree as at
 *                       src/rf_random_forest.cpp:1305-1308 (OpenMP parallel-for
 *                       over out-of-bag rows with an atomic FP64 +=)
 *   orc_prox_v1_casewise the same with use_casewise (src/rf_proximity.cpp:
 *                       186-223): each in-bag partner kk of an out-of-bag i
 *                       adds nin(kk) / sum over the leaf's in-bag cases of nin
 *   orc_prox_rfgap      src/rf_proximity.cpp:238-391 (cpu_proximity_rfgap):
 *                       nin(j) * (1 / node draws) summed per (i, j) in tree
 *                       order, rows scaled by 1 / |S_i|, then the symmetric
 *                       average; cases matched by equal reached node
 *   orc_outlier_scores  src/rf_random_forest.cpp:2489-2520: float row sums
 *                       of squares (each step float(fma(p, p, (double)sum)),
 *                       the contraction the reference's GNU C++ -O3 build
 *                       makes), 1/sum, divided by the (n/2)-th smallest
 *   orc_counts_bucket   the ALL / OOB_ASYM / OOB_UPPER integer counts by leaf
 *                       buckets (ALL has no reference implementation; it is
 *                       pinned by brute force and closed-form invariants)
 *   orc_counts_brute    the same by the O(T n^2) definition (tiny n only);
 *                       OOB_ROW (variant 4) follows cuda/rf_proximity.cu:32-67
 *   orc_finishprox      src/rf_finishprox.cpp:12-74
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#pragma STDC FP_CONTRACT OFF

#define ORC_ALL 0
#define ORC_OOB_ASYM 1
#define ORC_OOB_UPPER 2
#define ORC_OOB_ROW 4 /* cuda/rf_proximity.cu:32-67: i out-of-bag, any j of its leaf, i itself included */

/* Terminal node reached by sample i in tree t (node id local to the tree). */
static int32_t orc_traverse(const float* xi, int32_t inbag, const int32_t* st, const int32_t* bv, const float* xs,
                            const int32_t* tm, int32_t nnode, const int32_t* cat) {
  if (inbag) {
    /* find_terminal_node: loop bounded by maxnode; categorical (int)(x+0.5)
       in [1, ncat] -> catgoleft (all zero) -> right, else root. */
    int32_t k = 0;
    for (int64_t guard = 0; guard < 2 * (int64_t)nnode + 2; ++guard) {
      if (st[k] == -1) return k;
      int32_t v = bv[k];
      float f = xi[v];
      if (!cat || cat[v] == 1) {
        k = (f <= xs[k]) ? tm[2 * k] : tm[2 * k + 1];
      } else {
        int32_t jcat = (int32_t)((double)f + 0.5);
        if (jcat >= 1 && jcat <= cat[v])
          k = tm[2 * k + 1];
        else
          return 0;
      }
    }
    return 0;
  } else {
    /* cpu_testreebag: descend while split, count < 1000 guard. */
    int32_t kt = 0;
    int32_t count = 0;
    while (st[kt] == 1 && count < 1000) {
      count++;
      if (count >= 1000) break;
      int32_t v = bv[kt];
      float f = xi[v];
      if (!cat || cat[v] == 1) {
        kt = (f <= xs[kt]) ? tm[2 * kt] : tm[2 * kt + 1];
      } else {
        int32_t cv = (int32_t)f;
        /* catgoleft is never written by the grower: every category goes right */
        (void)cv;
        kt = tm[2 * kt + 1];
      }
      if (kt < 0 || kt >= nnode) {
        kt = 0;
        break;
      }
      if (st[kt] != 1) break;
    }
    if (st[kt] == -1 || st[kt] == 2) return kt;
    return 0;
  }
}

int orc_leaf_assign(const float* x, int64_t n, int32_t p, int32_t ntree, const int64_t* node_offset,
                    const int32_t* nodestatus, const int32_t* bestvar, const float* xbestsplit, const int32_t* treemap,
                    const int32_t* nin, const int32_t* cat, int32_t* nodextr) {
  if (!x || !node_offset || !nodestatus || !bestvar || !xbestsplit || !treemap || !nin || !nodextr) return -1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int32_t t = 0; t < ntree; ++t) {
    const int64_t b = node_offset[t];
    const int32_t nn = (int32_t)(node_offset[t + 1] - b);
    for (int64_t i = 0; i < n; ++i) {
      nodextr[(int64_t)t * n + i] = orc_traverse(x + i * p, nin[(int64_t)t * n + i] > 0, nodestatus + b, bestvar + b,
                                                 xbestsplit + b, treemap + 2 * b, nn, cat);
    }
  }
  return 0;
}

/* One tree of the reference's accumulating dense branch.  ncount is indexed
 * by terminal ordinal, so it holds max(n, nterm) entries: the reference sizes
 * it n (src/rf_proximity.cpp:108-171), enough for a tree grown on the n
 * samples it assigns, but a tree assigned only a subset of its samples can
 * have more leaves than samples. */
static void v1_one_tree(const int32_t* st, const int32_t* ndx, const int32_t* nin, int64_t n, int32_t nnode,
                        double* prox, int32_t* nod, int32_t* ncount, int32_t* ncn, int32_t* nodexb, int32_t* ndbegin,
                        int32_t* npcase, int casewise) {
  int32_t nterm = 0;
  for (int32_t k = 0; k < nnode; ++k) nod[k] = -1;
  for (int32_t k = 0; k < nnode; ++k)
    if (st[k] == -1) nod[k] = nterm++;
  for (int32_t k = 0; k < nterm; ++k) ncount[k] = 0;
  for (int64_t i = 0; i < n; ++i) {
    ncn[i] = 0;
    nodexb[i] = -1;
  }
  for (int64_t i = 0; i < n; ++i) {
    int32_t node = ndx[i];
    if (node >= 0 && node < nnode) {
      int32_t k = nod[node];
      if (k >= 0 && k < nterm) {
        ncount[k]++;
        ncn[i] = ncount[k];
        nodexb[i] = k;
      }
    }
  }
  ndbegin[0] = 0;
  for (int32_t k = 1; k <= nterm; ++k) ndbegin[k] = ndbegin[k - 1] + ncount[k - 1];
  for (int64_t i = 0; i < n; ++i) npcase[i] = -1;
  for (int64_t i = 0; i < n; ++i) {
    int32_t k = nodexb[i];
    if (k >= 0 && k < nterm) npcase[ndbegin[k] + ncn[i] - 1] = (int32_t)i;
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    if (nin[i] != 0) continue;
    int32_t k = nodexb[i];
    if (k < 0 || k >= nterm) continue;
    int64_t nodesize = 0; /* bootstrap draws that landed in the leaf */
    for (int32_t j = ndbegin[k]; j < ndbegin[k + 1]; ++j) {
      int32_t kk = npcase[j];
      if (kk >= 0 && nin[kk] > 0) nodesize += nin[kk];
    }
    if (casewise && nodesize == 0) continue;
    for (int32_t j = ndbegin[k]; j < ndbegin[k + 1]; ++j) {
      int32_t kk = npcase[j];
      if (kk >= 0 && nin[kk] > 0) {
        const double w = casewise ? (double)nin[kk] / (double)nodesize : 1.0;
#pragma omp atomic
        prox[i + (int64_t)kk * n] += w;
      }
    }
  }
}

/* Accumulate all trees into prox (column-major n x n FP64, caller-zeroed). */
static int prox_v1_all(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr, const int32_t* nin,
                int64_t n, int32_t ntree, double* prox, int casewise) {
  int32_t maxnn = 0;
  for (int32_t t = 0; t < ntree; ++t) {
    int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    if (nn > maxnn) maxnn = nn;
  }
  int32_t* nod = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 1));
  int32_t* ws = (int32_t*)malloc(sizeof(int32_t) * (size_t)(3 * n));
  int32_t* ncount = (int32_t*)malloc(sizeof(int32_t) * (size_t)((int64_t)maxnn + 1 > n ? (int64_t)maxnn + 1 : n));
  int32_t* ndbegin = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 2));
  if (!nod || !ws || !ncount || !ndbegin) {
    free(nod);
    free(ws);
    free(ncount);
    free(ndbegin);
    return -2;
  }
  for (int32_t t = 0; t < ntree; ++t) {
    v1_one_tree(nodestatus + node_offset[t], nodextr + (int64_t)t * n, nin + (int64_t)t * n, n,
                (int32_t)(node_offset[t + 1] - node_offset[t]), prox, nod, ncount, ws, ws + n, ndbegin, ws + 2 * n,
                casewise);
  }
  free(nod);
  free(ws);
  free(ncount);
  free(ndbegin);
  return 0;
}

int orc_prox_v1(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr, const int32_t* nin,
                int64_t n, int32_t ntree, double* prox) {
  return prox_v1_all(nodestatus, node_offset, nodextr, nin, n, ntree, prox, 0);
}

int orc_prox_v1_casewise(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr,
                         const int32_t* nin, int64_t n, int32_t ntree, double* prox) {
  return prox_v1_all(nodestatus, node_offset, nodextr, nin, n, ntree, prox, 1);
}

/* RF-GAP: the reference visits, for each case i, the trees in which i is out
 * of bag in order and adds nin_t(j) * (1 / M) to prox[i + j n] for every
 * in-bag j with nodextr_t(j) == nodextr_t(i) >= 0.  Here the tree loop is
 * outermost and the in-bag cases of a node are grouped once per tree; every
 * (i, j) entry still receives its terms in tree order, so the sums are the
 * reference's bit for bit.  prox: n x n FP64 column-major, overwritten. */
int orc_prox_rfgap(const int32_t* nin, const int32_t* nodextr, int64_t n, int32_t ntree, double* prox) {
  memset(prox, 0, sizeof(double) * (size_t)(n * n));
  int32_t* oobn = (int32_t*)calloc((size_t)n, sizeof(int32_t));
  int32_t* start = NULL;
  int32_t* memb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int64_t* draws = NULL;
  if (!oobn || !memb) {
    free(oobn);
    free(memb);
    return -2;
  }
  for (int32_t t = 0; t < ntree; ++t)
    for (int64_t i = 0; i < n; ++i) oobn[i] += nin[(int64_t)t * n + i] == 0;
  for (int32_t t = 0; t < ntree; ++t) {
    const int32_t* nt = nin + (int64_t)t * n;
    const int32_t* xt = nodextr + (int64_t)t * n;
    int32_t range = 0;
    for (int64_t i = 0; i < n; ++i)
      if (xt[i] + 1 > range) range = xt[i] + 1;
    free(start);
    free(draws);
    start = (int32_t*)calloc((size_t)range + 2, sizeof(int32_t));
    draws = (int64_t*)calloc((size_t)range + 1, sizeof(int64_t));
    if (!start || !draws) {
      free(oobn);
      free(memb);
      free(start);
      free(draws);
      return -2;
    }
    /* in-bag cases grouped by node (counting sort), draws per node */
    for (int64_t j = 0; j < n; ++j)
      if (nt[j] > 0 && xt[j] >= 0) {
        start[xt[j] + 1]++;
        draws[xt[j]] += nt[j];
      }
    for (int32_t k = 0; k < range; ++k) start[k + 1] += start[k];
    for (int64_t j = 0; j < n; ++j)
      if (nt[j] > 0 && xt[j] >= 0) memb[start[xt[j]]++] = (int32_t)j;
    for (int32_t k = range; k > 0; --k) start[k] = start[k - 1];
    start[0] = 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
      if (nt[i] != 0) continue;
      const int32_t k = xt[i];
      if (k < 0 || draws[k] == 0) continue;
      const double inv = 1.0 / (double)draws[k];
      for (int32_t q = start[k]; q < start[k + 1]; ++q) {
        const int32_t j = memb[q];
        prox[i + (int64_t)j * n] += (double)nt[j] * inv;
      }
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    if (oobn[i] == 0) {
      prox[i + i * n] = 1.0;
      continue;
    }
    const double f = 1.0 / (double)oobn[i];
    for (int64_t j = 0; j < n; ++j) prox[i + j * n] *= f;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = i + 1; j < n; ++j) {
      const double avg = (prox[i + j * n] + prox[j + i * n]) / 2.0;
      prox[i + j * n] = avg;
      prox[j + i * n] = avg;
    }
  free(oobn);
  free(memb);
  free(start);
  free(draws);
  return 0;
}

static int cmp_float(const void* a, const void* b) {
  const float x = *(const float*)a, y = *(const float*)b;
  return (x > y) - (x < y);
}

/* Outlier scores of a finished n x n FP64 proximity, read as the reference
 * reads it (prox[i * n + j]); out: n floats. */
int orc_outlier_scores(const double* prox, int64_t n, float* out) {
  float* sorted = (float*)malloc(sizeof(float) * (size_t)n);
  if (!sorted) return -2;
  for (int64_t i = 0; i < n; ++i) {
    float sum = 0.0f;
    for (int64_t j = 0; j < n; ++j) {
      const double v = prox[i * n + j];
      sum = (float)fma(v, v, (double)sum);
    }
    out[i] = sum > 0.0f ? 1.0f / sum : 0.0f;
    sorted[i] = out[i];
  }
  qsort(sorted, (size_t)n, sizeof(float), cmp_float);
  const float median = sorted[n / 2];
  if (median > 0.0f)
    for (int64_t i = 0; i < n; ++i) out[i] /= median;
  free(sorted);
  return 0;
}

/* Output index helpers: FULL column-major, UPPER packed row-major. */
static inline int64_t up_idx(int64_t i, int64_t j, int64_t n) { return i * n - (i * (i - 1)) / 2 + (j - i); }

static int valid_leaf(const int32_t* st, int32_t node, int32_t nnode) {
  return node >= 0 && node < nnode && st[node] == -1;
}

/* Integer counts by leaf buckets.  layout 0 = FULL, 1 = UPPER. */
int orc_counts_bucket(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr,
                      const int32_t* nin, int64_t n, int32_t ntree, int32_t variant, int32_t layout, int32_t* out) {
  if ((variant == ORC_OOB_ASYM || variant == ORC_OOB_ROW) && layout != 0) return -1;
  int32_t maxnn = 0;
  for (int32_t t = 0; t < ntree; ++t) {
    int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    if (nn > maxnn) maxnn = nn;
  }
  int32_t* head = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 1));
  int32_t* memb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* cnt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 1));
  if (!head || !memb || !cnt) {
    free(head);
    free(memb);
    free(cnt);
    return -2;
  }
  for (int32_t t = 0; t < ntree; ++t) {
    const int32_t* st = nodestatus + node_offset[t];
    const int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    const int32_t* ndx = nodextr + (int64_t)t * n;
    const int32_t* bag = nin + (int64_t)t * n;
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nn + 1));
    for (int64_t i = 0; i < n; ++i)
      if (valid_leaf(st, ndx[i], nn)) cnt[ndx[i] + 1]++;
    head[0] = 0;
    for (int32_t k = 0; k < nn; ++k) head[k + 1] = head[k] + cnt[k + 1];
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nn + 1));
    for (int64_t i = 0; i < n; ++i)
      if (valid_leaf(st, ndx[i], nn)) memb[head[ndx[i]] + cnt[ndx[i]]++] = (int32_t)i; /* ascending */
#pragma omp parallel for schedule(dynamic, 16)
    for (int32_t k = 0; k < nn; ++k) {
      for (int32_t a = head[k]; a < head[k + 1]; ++a) {
        const int64_t i = memb[a];
        if (variant == ORC_OOB_ROW) {
          const int bi = bag[i] > 0;
          if (!bi) out[i + i * n] += 1;
          for (int32_t b2 = a + 1; b2 < head[k + 1]; ++b2) {
            const int64_t j = memb[b2];
            if (!bi) out[i + j * n] += 1;
            if (!(bag[j] > 0)) out[j + i * n] += 1;
          }
          continue;
        }
        if (variant == ORC_ALL) {
          /* diagonal and pairs j > i; rows are owned by distinct i within a leaf,
             but different leaves never share (i, j), so no race */
          int64_t o = layout ? up_idx(i, i, n) : i + i * n;
          out[o] += 1;
        }
        for (int32_t b2 = a + 1; b2 < head[k + 1]; ++b2) {
          const int64_t j = memb[b2];
          const int bi = bag[i] > 0, bj = bag[j] > 0;
          if (variant == ORC_ALL) {
            if (layout) {
              out[up_idx(i, j, n)] += 1;
            } else {
              out[i + j * n] += 1;
              out[j + i * n] += 1;
            }
          } else if (bi != bj) {
            if (variant == ORC_OOB_UPPER) {
              if (layout) {
                out[up_idx(i, j, n)] += 1;
              } else {
                out[i + j * n] += 1;
                out[j + i * n] += 1;
              }
            } else { /* OOB_ASYM: row = out-of-bag sample */
              if (!bi)
                out[i + j * n] += 1;
              else
                out[j + i * n] += 1;
            }
          }
        }
      }
    }
  }
  free(head);
  free(memb);
  free(cnt);
  return 0;
}

/* The definition, O(T n^2): out FULL column-major. */
int orc_counts_brute(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr, const int32_t* nin,
                     int64_t n, int32_t ntree, int32_t variant, int32_t* out) {
  for (int32_t t = 0; t < ntree; ++t) {
    const int32_t* st = nodestatus + node_offset[t];
    const int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    const int32_t* ndx = nodextr + (int64_t)t * n;
    const int32_t* bag = nin + (int64_t)t * n;
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) {
        if (!valid_leaf(st, ndx[i], nn) || !valid_leaf(st, ndx[j], nn) || ndx[i] != ndx[j]) continue;
        int hit;
        if (variant == ORC_ALL)
          hit = 1;
        else if (variant == ORC_OOB_ASYM)
          hit = bag[i] == 0 && bag[j] > 0;
        else if (variant == ORC_OOB_ROW)
          hit = bag[i] == 0;
        else
          hit = (bag[i] == 0) != (bag[j] == 0);
        out[i + j * n] += hit;
      }
  }
  return 0;
}

/* rf::cpu_finishprox, step for step (column-major n x n FP64). */
int orc_finishprox(int64_t n, const int32_t* nout, const double* prox, double* proxsym) {
  for (int64_t q = 0; q < n * n; ++q) proxsym[q] = prox[q];
  for (int64_t i = 0; i < n; ++i) proxsym[i + i * n] = (double)nout[i];
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = i + 1; j < n; ++j) {
      double a = (prox[i + j * n] + prox[j + i * n]) / 2.0;
      proxsym[i + j * n] = a;
      proxsym[j + i * n] = a;
    }
  for (int64_t i = 0; i < n; ++i)
    if (nout[i] > 0) {
      double r = 1.0 / (double)nout[i];
      for (int64_t j = 0; j < n; ++j) proxsym[i + j * n] *= r;
    }
  for (int64_t i = 0; i < n; ++i) {
    proxsym[i + i * n] = 1.0;
    for (int64_t j = i + 1; j < n; ++j) {
      double a = 0.5 * (proxsym[i + j * n] + proxsym[j + i * n]);
      proxsym[i + j * n] = a;
      proxsym[j + i * n] = a;
    }
  }
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Full rows of the ALL and OOB_UPPER count matrices for the rows listed in
 * rows[nrows] (any order): all_out[r * n + j] = C[rows[r], j] (with the
 * diagonal C_ii = number of trees in which i ends on a leaf) and
 * oob_out[r * n + j] = U[rows[r], j] = P[i, j] + P[j, i] (diagonal 0), the
 * sums over trees of the reference's same-leaf rule (src/rf_proximity.cpp:
 * 179-230 for P).  Per tree: a counting sort of the samples by leaf, then
 * each listed row walks its leaf's members, O(T (n + nrows * leaf)) in all,
 * so it runs at sizes where the whole matrix does not fit the host. */
int orc_counts_rows(const int32_t* nodestatus, const int64_t* node_offset, const int32_t* nodextr,
                    const int32_t* nin, int64_t n, int32_t ntree, const int64_t* rows, int32_t nrows,
                    int32_t* all_out, int32_t* oob_out) {
  int32_t maxnn = 0;
  for (int32_t t = 0; t < ntree; ++t) {
    int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    if (nn > maxnn) maxnn = nn;
  }
  for (int32_t r = 0; r < nrows; ++r)
    if (rows[r] < 0 || rows[r] >= n) return -1;
  int32_t* head = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 1));
  int32_t* memb = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  int32_t* cnt = (int32_t*)malloc(sizeof(int32_t) * (size_t)(maxnn + 1));
  if (!head || !memb || !cnt) {
    free(head);
    free(memb);
    free(cnt);
    return -2;
  }
  memset(all_out, 0, sizeof(int32_t) * (size_t)nrows * (size_t)n);
  memset(oob_out, 0, sizeof(int32_t) * (size_t)nrows * (size_t)n);
  for (int32_t t = 0; t < ntree; ++t) {
    const int32_t* st = nodestatus + node_offset[t];
    const int32_t nn = (int32_t)(node_offset[t + 1] - node_offset[t]);
    const int32_t* ndx = nodextr + (int64_t)t * n;
    const int32_t* bag = nin + (int64_t)t * n;
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nn + 1));
    for (int64_t i = 0; i < n; ++i)
      if (valid_leaf(st, ndx[i], nn)) cnt[ndx[i] + 1]++;
    head[0] = 0;
    for (int32_t k = 0; k < nn; ++k) head[k + 1] = head[k] + cnt[k + 1];
    memset(cnt, 0, sizeof(int32_t) * (size_t)(nn + 1));
    for (int64_t i = 0; i < n; ++i)
      if (valid_leaf(st, ndx[i], nn)) memb[head[ndx[i]] + cnt[ndx[i]]++] = (int32_t)i;
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t r = 0; r < nrows; ++r) {
      const int64_t i = rows[r];
      const int32_t k = ndx[i];
      if (!valid_leaf(st, k, nn)) continue;
      const int bi = bag[i] > 0;
      int32_t* ar = all_out + (int64_t)r * n;
      int32_t* orow = oob_out + (int64_t)r * n;
      for (int32_t a = head[k]; a < head[k + 1]; ++a) {
        const int64_t j = memb[a];
        ar[j] += 1; /* j == i gives the diagonal */
        if ((bag[j] > 0) != bi) orow[j] += 1;
      }
    }
  }
  free(head);
  free(memb);
  free(cnt);
  return 0;
}
