/*
 * rfx_prox.h -- C ABI of the MI355X random-forest proximity path.
 *
 * Plain pointers, sizes and status codes; no C++ or torch types.  Every
 * device buffer is owned by the caller (allocated with hipMalloc or taken from
 * any HIP allocator); every GPU entry point enqueues work on the stream it is
 * given and returns without synchronising unless stated.  Errors are returned
 * as negative status codes, never swallowed.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository).
 */
#ifndef RFX_PROX_H
#define RFX_PROX_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define RFX_OK 0
#define RFX_ERR_INVALID (-1)     /* bad argument, shape or forest            */
#define RFX_ERR_ALLOC (-2)       /* host allocation failed                   */
#define RFX_ERR_HIP (-3)         /* a HIP call failed                        */
#define RFX_ERR_WORKSPACE (-4)   /* workspace too small                      */
#define RFX_ERR_UNSUPPORTED (-5) /* combination not implemented             */

/* ---- limits ------------------------------------------------------------- */
#define RFX_MAX_SAMPLES 4000000LL
#define RFX_MAX_FEATURES 1023
#define RFX_MAX_NODES_PER_TREE 4194303LL /* left child index fits 22 bits   */
#define RFX_MAX_TOTAL_NODES 1073741823LL /* (global node id << 1) | bag fits int32 */

/* ---- synthetic datasets (rfx_make_dataset) ------------------------------ */
#define RFX_DATA_GAUSS_MIXTURE 0    /* class means ~ N(0, 4I), x ~ N(mu, I)  */
#define RFX_DATA_COVTYPE_LIKE 1     /* p = 54, 7 imbalanced classes          */
#define RFX_DATA_SUBSPACE_MIXTURE 2 /* 16 comps in 8-dim subspace + noise    */
#define RFX_DATA_OVERLAP_MIXTURE 3  /* class means ~ N(0, 0.35^2 I), x ~ N(mu, I) */

/* ---- grower flags -------------------------------------------------------- */
#define RFX_GROW_NO_FMA 1 /* Gini pno sum without fused multiply-add        */
#define RFX_GROW_PLAIN 2  /* per-node sorts instead of presorted columns     */

/* ---- proximity variants -------------------------------------------------
 * ALL       C[i,j] = #{t : leaf_t(i) = leaf_t(j)}                (symmetric)
 * OOB_ASYM  P[i,j] = #{t : i out-of-bag, j in-bag, same leaf}   (reference
 *           src/rf_proximity.cpp:179-230, non-case-wise)
 * OOB_UPPER U[i,j] = P[i,j] + P[j,i]                            (reference
 *           cuda/rf_proximity_upper_triangle.cu:149-306 fold, diagonal 0)
 * A sample whose traversal ends on a non-terminal node takes part in no pair.
 */
#define RFX_PROX_ALL 0
#define RFX_PROX_OOB_ASYM 1
#define RFX_PROX_OOB_UPPER 2
/* OOB_ROW  P4[i,j] = #{t : i out-of-bag, same leaf as j}, any j including i
 *          (cuda/rf_proximity.cu:32-67, non-case-wise); FULL layout only. */
#define RFX_PROX_OOB_ROW 4

/* layouts: FULL = n x n, element (i,j) at i + j*n (column-major, as the
 * reference's prox[n + kk*nsample]); UPPER = packed upper triangle with the
 * diagonal, row-major: (i,j), i <= j, at i*n - i*(i-1)/2 + (j-i), 64-bit. */
#define RFX_LAYOUT_FULL 0
#define RFX_LAYOUT_UPPER 1

/* accumulation algorithms */
#define RFX_ALGO_AUTO 0   /* choose from the forest's leaf-size statistics  */
#define RFX_ALGO_DENSE 1  /* LDS-staged key tiles, register-tiled compare   */
#define RFX_ALGO_BUCKET 2 /* leaf buckets, integer atomics per matching pair */
#define RFX_ALGO_GATHER 3 /* leaf buckets gathered per output line in LDS  */
/* Flag or-ed into a BUCKET or GATHER algo: the workspace already holds the
 * leaf buckets of these keys (rfx_prox_build_buckets), so only the counting
 * kernel runs.  A call without it builds in the workspace only what its own
 * kernel reads and leaves the rest of the buckets unspecified: build again
 * with rfx_prox_build_buckets before the next RFX_ALGO_PREBUILT call.  The
 * library records, per workspace address, the shape of its last build and
 * whether it was complete; a PREBUILT call whose workspace was last built
 * partially, for another (n, T, total_nodes), or not at all returns
 * RFX_ERR_INVALID.  It cannot see whether the keys changed since. */
#define RFX_ALGO_PREBUILT 0x100

/* ========================================================================
 * Host side: forest input contract (not timed)
 * ======================================================================== */
typedef struct rfx_forest rfx_forest;

/* MT19937 output #index (0-based) after init_genrand(seed); the stream of
 * include/rf_utils.hpp:74-141. */
uint32_t rfx_mt19937_u32(uint32_t seed, int64_t index);

/* Bootstrap multiplicities of one tree; replaces rf::cpu_bootstrap's nin
 * output (src/rf_bootstrap.cpp:14-82) for a given seed. */
int rfx_bootstrap(int64_t n, uint32_t seed, int32_t* nin);

/* Deterministic synthetic data, x row-major n x p float32, y int32 classes. */
int rfx_make_dataset(int32_t kind, int64_t n, int32_t p, int32_t nclass, uint64_t seed, float* x, int32_t* y);

/* Grow a classification forest exactly as the reference CPU path would:
 * bootstrap seed boot_seed + t (src/rf_bootstrap.cpp:24), grower seed
 * iseed + t (src/rf_random_forest.cpp:515), rf::cpu_growtree semantics
 * (src/rf_growtree.cpp:382-942).  maxnode <= 0 means 2n+1.  nthreads <= 0
 * means all hardware threads (trees are independent). */
int rfx_forest_grow(const float* x, const int32_t* y, int64_t n, int32_t p, int32_t nclass, int32_t ntree,
                    int32_t mtry, int32_t minndsize, int32_t maxnode, int32_t iseed, int32_t boot_seed,
                    int32_t flags, int32_t nthreads, rfx_forest** out);
/* The same, counting finished trees into *trees_done (atomically, from the
 * worker threads) so that a caller on another thread can report progress:
 * the counterpart of the reference's per-tree progress_callback_(itree + 1,
 * ntree) (src/rf_random_forest.cpp:518-521), polled instead of called back
 * because the trees grow on several threads. */
int rfx_forest_grow_progress(const float* x, const int32_t* y, int64_t n, int32_t p, int32_t nclass, int32_t ntree,
                             int32_t mtry, int32_t minndsize, int32_t maxnode, int32_t iseed, int32_t boot_seed,
                             int32_t flags, int32_t nthreads, int32_t* trees_done, rfx_forest** out);
void rfx_forest_free(rfx_forest* f);
int32_t rfx_forest_ntree(const rfx_forest* f);
int32_t rfx_forest_nnode(const rfx_forest* f, int32_t t);
/* Export the reference's per-tree arrays flattened by node_offset[T+1]:
 * nodestatus/bestvar/xbestsplit/nodeclass [total], treemap [2*total],
 * nin [T*n].  Any output pointer except node_offset may be NULL. */
int rfx_forest_export(const rfx_forest* f, int64_t* node_offset, int32_t* nodestatus, int32_t* bestvar,
                      float* xbestsplit, int32_t* treemap, int32_t* nodeclass, int32_t* nin);
/* Pack reference node arrays into the device node records (2 x uint32 per
 * node) read by rfx_leaf_assign. */
int rfx_pack_nodes(const int64_t* node_offset, int32_t ntree, const int32_t* nodestatus, const int32_t* bestvar,
                   const float* xbestsplit, const int32_t* treemap, int32_t p, uint32_t* packed);
/* nout[i] = mult * #{t : nin_t(i) = 0}; the reference CPU path has mult = 2
 * (src/rf_bootstrap.cpp:61 and src/rf_random_forest.cpp:1108-1110). */
int rfx_oob_counts(const int32_t* nin, int64_t n, int32_t ntree, int32_t mult, int32_t* nout);
/* In-bag bitmask: bit (i % 32) of word [t * ceil(n/32) + i/32] = (nin_t(i) > 0). */
int rfx_pack_inbag_bits(const int32_t* nin, int64_t n, int32_t ntree, uint32_t* bits);

/* ========================================================================
 * GPU side
 * ======================================================================== */

/* A1: leaf assignment of every sample through every tree.  Replaces
 * rf::cpu_testreebag (src/rf_testreebag.cpp:13-127, out-of-bag samples) and
 * RandomForest::find_terminal_node (src/rf_random_forest.cpp:2194-2244,
 * in-bag samples), as called per tree at src/rf_random_forest.cpp:1043 and
 * 1295-1303; GPU twin gpu_compute_nodextr_inbag_kernel
 * (cuda/rf_growtree.cu:1763-1850).
 *   x_col      device, p x n float32 column-major (feature m of sample i at m*n+i)
 *   nodes      device, packed node records (rfx_pack_nodes)
 *   node_off   device, int64 [T+1]
 *   inbag      device, in-bag bitmask (rfx_pack_inbag_bits)
 *   cat        device int32 [p] category counts (1 = numeric) or NULL (all numeric)
 *   nodextr    device int32 [T][n] node reached (reference semantics) or NULL
 *   keys       device int32 [T][n]: (global node id << 1 | in-bag) when the
 *              node is terminal, else -1; or NULL */
int rfx_leaf_assign(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                    const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat,
                    int32_t* nodextr, int32_t* keys);

/* Chain index of packed trees (host): the heavy paths of at least min_len
 * internal nodes, for rfx_leaf_assign_ws.  Trees grown to small nodes on
 * large noisy data hold chains of splits that each peel one sample off a
 * feature's extreme (at n = 1,000,000, mean walks of thousands of levels);
 * on such a path the exit of a sample is found by binary searches instead of
 * a walk of one dependent node gather per level.  Two calls: out == NULL sets
 * *bytes to the size needed (0 when no tree has such a path: pass chains =
 * NULL then); else out (host memory of *bytes) receives the index, which is
 * position independent and goes to the device as is.
 *   node_offset  int64 [T+1];  packed  rfx_pack_nodes output */
int rfx_chain_index(const int64_t* node_offset, int32_t ntree, const uint32_t* packed, int32_t min_len, void* out,
                    size_t* bytes);

/* Workspace bytes rfx_leaf_assign_ws takes for (n, T): a 256-byte header and
 * a queue of max(1024, min(n*T, max(65536, min(n*T/16, 2^26)))) 16-byte walk
 * records. */
size_t rfx_leaf_workspace_bytes(int64_t n, int32_t ntree);

/* rfx_leaf_assign for forests with chains (same outputs, bit for bit, as
 * rfx_leaf_assign; same reference interfaces).  With all features numeric
 * and a chain index (device copy of rfx_chain_index's output), the walks
 * still going after 128 lockstep levels are queued in ws and finished by a
 * kernel that refills idle lanes and jumps along the indexed paths; trees
 * go in chunks whose walks the queue can hold twice over, and a full queue
 * only sends the refused walks back to the first pass.  chains == NULL or a
 * categorical forest: the one-pass kernel.  ws: device memory of ws_bytes
 * (16-byte aligned, at least 256 + 16 * 1024 bytes, else
 * RFX_ERR_WORKSPACE); its contents need no initialisation.  Calls that share
 * ws race. */
int rfx_leaf_assign_ws(hipStream_t stream, const float* x_col, int64_t n, int32_t p, const uint32_t* nodes,
                       const int64_t* node_off, int32_t ntree, const uint32_t* inbag, const int32_t* cat,
                       int32_t* nodextr, int32_t* keys, const void* chains, void* ws, size_t ws_bytes);

/* Workspace bytes rfx_prox_accumulate needs for (n, T, total_nodes, algo). */
size_t rfx_prox_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes, int32_t algo);

/* A2+A3: co-occurrence counts from leaf keys.  Replaces the per-tree
 * rf::cpu_proximity (include/rf_proximity.hpp:10-21, src/rf_proximity.cpp:
 * 99-231) and rf::cuda::gpu_proximity_upper_triangle_fp16
 * (include/rf_proximity_upper_triangle.hpp:42-55) with one call over all
 * trees; counts are int32 and exact.
 *   keys        device int32 [T][n] from rfx_leaf_assign
 *   out         device int32, FULL (n*n) or UPPER (n(n+1)/2)
 *   accumulate  0: out is overwritten; 1: counts are added to out
 *   workspace   device, >= rfx_prox_workspace_bytes(...) (may be NULL for DENSE) */
int rfx_prox_accumulate(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                        int32_t variant, int32_t layout, int32_t algo, int32_t accumulate, int32_t* out,
                        void* workspace, size_t workspace_bytes);

/* A2 alone: the leaf buckets (the reference's per-tree nod/ncount/ndbegin/
 * npcase CSR of src/rf_proximity.cpp:108-171, for all trees at once) of the
 * keys into the workspace, for calls with RFX_ALGO_PREBUILT. */
int rfx_prox_build_buckets(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                           void* workspace, size_t workspace_bytes);

/* A3, case-wise weights: the use_casewise branch of the reference's
 * accumulating proximity (rf::cpu_proximity, src/rf_proximity.cpp:186-223,
 * weight nin(kk) / nodesize at 214-216).  For every tree, each out-of-bag
 * case i and in-bag case j of the same leaf add nin_t(j) / (sum of nin_t over
 * the leaf's in-bag cases) to out[i + j n] (FP64 n x n column-major, row =
 * the out-of-bag case, the reference's prox layout); accumulate = 0
 * overwrites, 1 adds.  nin: device int32 [T][n] bootstrap multiplicities (the
 * in-bag bits of keys must agree with nin > 0).  The weights of different
 * trees are summed in no fixed order: entries may differ from the
 * reference's tree-ordered FP64 sums in the last bits. */
size_t rfx_prox_casewise_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes);
int rfx_prox_casewise(hipStream_t stream, const int32_t* keys, const int32_t* nin, int64_t n, int32_t ntree,
                      int64_t total_nodes, int32_t accumulate, double* out, void* workspace, size_t workspace_bytes);

/* RF-GAP proximity.  Replaces rf::gpu_proximity_rfgap (include/
 * rf_proximity.hpp:42, cuda/rf_proximity.cu:541-786) and
 * rf::cpu_proximity_rfgap (include/rf_proximity.hpp:34, src/rf_proximity.cpp:
 * 238-391), called from the forest's finalisation (src/rf_random_forest.cpp:
 * 1702, 1707, 1798) with the per-tree nin and reached nodes, flattened
 * [tree * n + sample].  For every case i, p_ij = (1 / |S_i|) * sum over the
 * trees t in which i is out of bag (nin_t(i) = 0) of nin_t(j) * (1 / M_t)
 * for each in-bag j that reached the same node (M_t = that node's sum of
 * nin_t over in-bag cases); then out[i + j n] = out[j + i n] =
 * (p_ij + p_ji) / 2, and out[i + i n] = 1 for a case that is never out of
 * bag.  Cases are matched by equal node index, terminal or not, as in the
 * reference.
 *   nin       device int32 [T][n] bootstrap multiplicities
 *   nodextr   device int32 [T][n] node reached (rfx_leaf_assign's nodextr;
 *             negative entries take part in no pair)
 *   node_off  device int64 [T+1] node offsets (total_nodes = node_off[T])
 *   out       device FP64 n x n, overwritten, symmetric
 * Sums run in no fixed tree order: entries agree with the reference to about
 * 1e-16 relative. */
size_t rfx_prox_rfgap_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes);
int rfx_prox_rfgap(hipStream_t stream, const int32_t* nin, const int32_t* nodextr, const int64_t* node_off, int64_t n,
                   int32_t ntree, int64_t total_nodes, double* out, void* workspace, size_t workspace_bytes);

/* ALL and OOB_UPPER in one pass (same layout for both outputs). */
int rfx_prox_accumulate_pair(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree,
                             int64_t total_nodes, int32_t layout, int32_t algo, int32_t accumulate,
                             int32_t* out_all, int32_t* out_oob, void* workspace, size_t workspace_bytes);

/* No-allocation form of rfx_prox_accumulate_pair, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference.
 * The leaf buckets live in a static device array of RFX_NOALLOC_SCRATCH_BYTES
 * per GPU, held from library load until the process ends; the call allocates
 * nothing and takes its device address once per device (the current device
 * must be the stream's).  Problems whose bucket workspace
 * (rfx_prox_workspace_bytes with a BUCKET or GATHER algo) exceeds the array
 * return RFX_ERR_WORKSPACE; DENSE needs none.  AUTO, DENSE, BUCKET and GATHER
 * as in rfx_prox_accumulate_pair; RFX_ALGO_PREBUILT is refused
 * (RFX_ERR_INVALID): every call rebuilds its buckets.  All calls on one GPU
 * share the array, which keeps the last call's contents: calls that overlap
 * on one GPU (concurrent streams or host threads) race and must be
 * serialised by the caller. */
#define RFX_NOALLOC_SCRATCH_BYTES (64LL << 20)
int rfx_prox_accumulate_pair_noalloc(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree,
                                     int64_t total_nodes, int32_t layout, int32_t algo, int32_t accumulate,
                                     int32_t* out_all, int32_t* out_oob);

/* Rows [row_begin, row_end) of the FULL matrix of one variant, written
 * densely as (row_end-row_begin) x n int32 row-major (row r, column j at
 * (r-row_begin)*n + j).  Used to check large problems block by block. */
int rfx_prox_rows(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree, int32_t variant,
                  int64_t row_begin, int64_t row_end, int32_t* out);

/* Leaf-size statistics of a key array: sum over trees and leaves of
 * |leaf|^2 (matching ordered pairs incl. self), as a double.  Feeds the
 * AUTO choice and the bench's work accounting. */
int rfx_prox_pair_stats(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                        void* workspace, size_t workspace_bytes, double* sum_sq_host);

/* A5: finishprox.  Replaces rf::cpu_finishprox (src/rf_finishprox.cpp:
 * 12-74): S_ii = 1; for i != j, a = (P_ij + P_ji)/2, S_ij = 0.5*(a*r_i + a*r_j)
 * with r_k = 1/nout_k (r_k = 1 when nout_k = 0), in float64, operation for
 * operation as the reference, so the result is bit-identical.
 *   counts      device int32, OOB_ASYM FULL (in_layout FULL) or OOB_UPPER
 *               UPPER (in_layout UPPER)
 *   nout        device int32 [n]
 *   proxsym     device float64, FULL column-major (out_layout FULL) or UPPER */
int rfx_finishprox(hipStream_t stream, const int32_t* counts, int32_t in_layout, const int32_t* nout, int64_t n,
                   double* proxsym, int32_t out_layout);

/* A5 on an FP64 accumulated matrix, as rf::cpu_finishprox receives it
 * (include/rf_finishprox.hpp:9-12): prox FULL column-major float64 (any
 * values, e.g. case-wise weights), same arithmetic, bit-identical. */
int rfx_finishprox_f64(hipStream_t stream, const double* prox, const int32_t* nout, int64_t n, double* proxsym,
                       int32_t out_layout);

/* Outlier scores of a finished proximity matrix.  Replaces
 * RandomForest::compute_outlier_scores (src/rf_random_forest.cpp:2489-2520;
 * results read by get_outlier_scores, include/rf_random_forest.hpp:210):
 * s_i = 1 / sum_j prox[i n + j]^2 (0 if the sum is not positive), summed in
 * float in j order with each step rounding fma(p, p, (double)sum) to float, as
 * the reference's GNU C++ -O3 build evaluates it; then scores_i = s_i / m with
 * m = the (n/2)-th smallest s (no division when m <= 0).  prox: device FP64
 * n x n (the finishprox output or the RF-GAP matrix); scores: device float32
 * [n]; median_host may be NULL.  Bit-identical to the reference.  Synchronises
 * the stream (the median is selected on the host). */
int rfx_outlier_scores(hipStream_t stream, const double* prox, int64_t n, float* scores, float* median_host);

/* ---- MDS ---------------------------------------------------------------- */
#define RFX_MDS_COUNTS 0    /* M_ij = c_ij * scale (e.g. ALL counts / T)       */
#define RFX_MDS_REFERENCE 1 /* OOB_UPPER counts -> finishprox -> float32 -> nout renormalisation */
#define RFX_MDS_MATRIX 2    /* a given FP64 proximity matrix (rfx_mds_dense_matrix) */
#define RFX_MDS_RFGAP 3     /* a given FP64 RF-GAP matrix (rfx_mds_dense_rfgap) */

#define RFX_PREC_FP32 0
#define RFX_PREC_FP64 1
#define RFX_PREC_BF16 2
#define RFX_PREC_INT8 3
#define RFX_PREC_FP8 4 /* FP8 e4m3 (OCP) */
#define RFX_PREC_FP16 5 /* IEEE half, only as a count-conversion target */

/* A6+A7: classical 3-D MDS of a count-matrix proximity.  Replaces
 * rf::compute_mds_3d_cpu (include/rf_mds_cpu.hpp:27-33, src/rf_mds_cpu.cpp:
 * 88-341) as called by the classifier's compute_mds_3d_cpu
 * (python/randomforest_py.cpp:1221-1252): D = max(M) - M (zero diagonal),
 * B = -1/2 J D^2 J with FP64 row means, top three eigenpairs with
 * lambda > 1e-10 in descending order, X = V sqrt(lambda).  Eigenpairs come
 * from block subspace iteration (16 vectors) with FP64 CholeskyQR2 and
 * Rayleigh-Ritz; the D^2 operator runs in `precision` (FP32 or FP64).
 *   counts      device int32, symmetric counts (COUNTS; UPPER or FULL) or
 *               OOB_UPPER counts (REFERENCE; UPPER)
 *   scale       COUNTS: M_ij = c_ij * scale
 *   nout        device int32 [n] (REFERENCE), the reference's doubled nout
 *   coords      device float64 n x 3 row-major (reference output layout)
 *   eigvals_host host float64 [16] Ritz values (descending), may be NULL
 *   iters_host  host int32, iterations taken, may be NULL
 * Synchronises the stream (the k x k problems run on the host).  Returns
 * RFX_ERR_INVALID when fewer than three eigenvalues exceed 1e-10, as the
 * reference throws. */
size_t rfx_mds_dense_workspace_bytes(int64_t n, int32_t precision);
int rfx_mds_dense(hipStream_t stream, const int32_t* counts, int32_t layout, int64_t n, int32_t mode, double scale,
                  const int32_t* nout, int32_t precision, int32_t max_iter, double tol, void* workspace,
                  size_t workspace_bytes, double* coords, double* eigvals_host, int32_t* iters_host);

/* A6+A7 on a given FP64 proximity: replaces rf::compute_mds_3d_cpu
 * (include/rf_mds_cpu.hpp:27-33) for its non-RF-GAP inputs.  prox is a
 * device FP64 n x n column-major matrix (the finishprox output S, or any
 * symmetric proximity).  With nout (device int32 [n]) the reference's
 * renormalisation applies (src/rf_mds_cpu.cpp:154-186): M_ij = 1/2
 * (S_ij/nout_i + S_ji/nout_j), rows with nout_i = 0 undivided, M_ii = 1;
 * with nout = NULL, M = S.  Then as rfx_mds_dense (same workspace size). */
int rfx_mds_dense_matrix(hipStream_t stream, const double* prox, int64_t n, const int32_t* nout, int32_t precision,
                         int32_t max_iter, double tol, void* workspace, size_t workspace_bytes, double* coords,
                         double* eigvals_host, int32_t* iters_host);

/* A6+A7 on a given FP64 RF-GAP proximity: rf::compute_mds_3d_cpu with
 * use_rfgap and OOB counts (src/rf_mds_cpu.cpp:125-153), as the classifier
 * calls it (python/randomforest_py.cpp:1236-1244, nout = the reference's
 * doubled OOB counts): M_ij = 1/2 (P_ij/nout_i + P_ji/nout_j) off the
 * diagonal and M_ii = P_ii/nout_i, rows with nout_i = 0 undivided (no unit
 * diagonal is imposed).  nout is required.  Then as rfx_mds_dense (same
 * workspace size). */
int rfx_mds_dense_rfgap(hipStream_t stream, const double* prox, int64_t n, const int32_t* nout, int32_t precision,
                        int32_t max_iter, double tol, void* workspace, size_t workspace_bytes, double* coords,
                        double* eigvals_host, int32_t* iters_host);

/* Element-wise conversion of int32 counts to FP64, FP32 or FP16 (IEEE
 * round-to-nearest-even; exact up to 2048 for FP16).  The reference keeps
 * counts as FP64 (src/rf_proximity.cpp:179-230) or FP16
 * (cuda/rf_proximity_upper_triangle.cu:149-306). */
int rfx_counts_convert(hipStream_t stream, const int32_t* counts, int64_t len, int32_t format, void* out);

/* A8: low-rank factors of the ALL-count proximity P = C/T (PSD): U (n x rank,
 * FP32 row-major) with P ~ U U^T from randomized subspace iteration whose
 * products C V are matrix-free through the leaf buckets (nothing n x n is
 * formed).  Replaces the factor accumulation of
 * rf::cuda::LowRankProximityMatrix (include/rf_proximity_lowrank.hpp:27-420,
 * cuda/rf_proximity_lowrank.cu:863-1299) and get_factors (1753-1779).
 *   node_offset_host  host int64 [T+1] (tree chunking of the leaf sums)
 *   eigvals_host      host float64 [rank]: top eigenvalues of C/T, may be NULL
 * Synchronises the stream. */
size_t rfx_lowrank_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes, int32_t rank, int32_t oversample);
int rfx_lowrank_factors(hipStream_t stream, const int32_t* keys, int64_t n, int32_t ntree, int64_t total_nodes,
                        const int64_t* node_offset_host, int32_t rank, int32_t oversample, int32_t power_iters,
                        uint64_t seed, void* workspace, size_t workspace_bytes, float* U, double* eigvals_host);

/* Factor storage: format FP32 (copy), BF16, INT8 (q = round(u/s_c), |q| <= 127,
 * s_c = max|u_c|/127) or FP8 (e4m3, s_c = max|u_c|/448); scales FP32 [rank]
 * (1 for FP32/BF16).  Replaces quantize_factors
 * (cuda/rf_proximity_lowrank.cu:1412-1553). */
int rfx_quantize_factors(hipStream_t stream, const float* U, int64_t n, int32_t rank, int32_t format, void* Uq,
                         float* scales);

/* P = U U^T: the dense FP64 n x n column-major proximity of factors stored
 * in `format` (FP32 factors need no scales), each entry the exact FP64 sum
 * of the dequantised products in column order.  Replaces
 * LowRankProximityMatrix::reconstruct_full_matrix
 * (include/rf_proximity_lowrank.hpp:142, cuda/rf_proximity_lowrank.cu:
 * 1307-1367, P = A B^T with A = B here).  rank <= 128, n <= 4,194,240. */
int rfx_reconstruct_factors(hipStream_t stream, const void* Uq, const float* scales, int32_t format, int64_t n,
                            int32_t rank, double* out);

/* A9: classical 3-D MDS from factors: E_ij = (1 - u_i.u_j)^2 off the diagonal,
 * 0 on it, double-centred with FP64 row means; top three eigenpairs by the
 * same subspace iteration as rfx_mds_dense, operator arithmetic in
 * `precision` (FP32 or FP64) over factors stored in `format`.  Replaces
 * LowRankProximityMatrix::compute_mds_from_factors
 * (cuda/rf_proximity_lowrank.cu:1935-2531).  rank <= 128.  Synchronises. */
size_t rfx_mds_factors_workspace_bytes(int64_t n, int32_t rank, int32_t precision);
int rfx_mds_factors(hipStream_t stream, const void* Uq, const float* scales, int32_t format, int64_t n, int32_t rank,
                    int32_t precision, int32_t max_iter, double tol, void* workspace, size_t workspace_bytes,
                    double* coords, double* eigvals_host, int32_t* iters_host);

/* A9 on the matrix cores: the same MDS from factors stored in `format`, with
 * the E V products as bf16 MFMA GEMMs (u * v and Q carried as bf16 hi + lo
 * pairs, FP32 accumulation, FP64 partial sums) until the Ritz residual is
 * below tol, then, when U32 (device FP32 n x rank row-major, the factors the
 * stored copy was made from) is given, up to refine_max refinement sweeps
 * from the converged block until refine_tol, with an FP32-accurate operator
 * on U32 (the same GEMMs with u carried as a bf16 pair hi + lo).  FP64 row
 * means, CholeskyQR2, Grams and Rayleigh-Ritz throughout.  iters_host[0] =
 * matrix-core iterations, iters_host[1] = refinement sweeps.  rank <= 128.
 * Synchronises.
 * Replaces LowRankProximityMatrix::compute_mds_from_factors
 * (cuda/rf_proximity_lowrank.cu:1935-2531). */
size_t rfx_mds_factors_mixed_workspace_bytes(int64_t n, int32_t rank);
int rfx_mds_factors_mixed(hipStream_t stream, const void* Uq, const float* scales, int32_t format, const float* U32,
                          int64_t n, int32_t rank, int32_t max_iter, double tol, int32_t refine_max,
                          double refine_tol, void* workspace, size_t workspace_bytes, double* coords,
                          double* eigvals_host, int32_t* iters_host);

/* ========================================================================
 * A10: batches of independent forests
 * ======================================================================== */
/* Batch-only variant: ALL counts into out and OOB_UPPER counts into out2. */
#define RFX_PROX_PAIR 3

/* One problem of a batch: an all-numeric forest (categorical features are
 * not supported here) and its UPPER int32 output(s), all device pointers. */
typedef struct rfx_batch_problem {
  const float* x_col;      /* p x n float32 column-major */
  int64_t n;
  int32_t p;
  int32_t ntree;
  const uint32_t* nodes;   /* packed node records (rfx_pack_nodes) */
  const int64_t* node_off; /* int64 [T+1], offsets within this forest */
  int64_t total_nodes;     /* node_off[T] */
  const uint32_t* inbag;   /* in-bag bitmask (rfx_pack_inbag_bits) */
  int32_t* out;            /* UPPER n(n+1)/2 int32, overwritten */
  int32_t* out2;           /* UPPER, OOB_UPPER counts for RFX_PROX_PAIR, else NULL */
} rfx_batch_problem;

typedef struct rfx_batch rfx_batch;

/* Plan a batch: leaf assignment and co-occurrence counts (variant ALL,
 * OOB_UPPER or PAIR; UPPER layout) of `count` forests in one leaf launch,
 * one leaf-bucket sort over all of them, the large problems' own count
 * kernels (after rfx_batch_autotune) and one persistent gather launch fed
 * by a device work queue (largest problems first) for the rest.  The reference has
 * no multi-forest batching; it runs each forest tree by tree from the host
 * (fit_batch_gpu, src/rf_random_forest.cpp:1364-1570; batch sizing
 * get_recommended_batch_size, src/rf_cuda_config.cpp:225-295).  The plan
 * copies the descriptors and reads each forest's node offsets once (a
 * synchronous copy: forests whose trees have at most 16,384 nodes get the
 * packed 16-bit DENSE compare); the device buffers they name must outlive it.
 * Needs the current device to be the one the batch runs on. */
int rfx_batch_create(const rfx_batch_problem* problems, int32_t count, int32_t variant, rfx_batch** out);
size_t rfx_batch_workspace_bytes(const rfx_batch* b);
/* Enqueue the batch on `stream`.  use_graph = 1 records the launch sequence
 * into a hipGraph on the first call (re-recorded when the workspace or the
 * tuning changes) and replays it.  The first call with a workspace uploads
 * the descriptors and synchronises; later calls do not.  Two runs of one
 * plan must not overlap (they share the workspace). */
int rfx_batch_run(hipStream_t stream, rfx_batch* b, void* workspace, size_t workspace_bytes, int32_t use_graph);
/* Measure every problem on the batch's own data: its sum of |leaf|^2 and
 * the time of its one-problem DENSE and GATHER kernels on its slice of the
 * batched leaf buckets.  The faster algorithm is kept; a problem whose
 * kernel takes 0.05 ms or more is from then on launched on its own (inside
 * the same launch sequence or graph), the others stay in the persistent
 * queue.  Then every (rows per work item, workgroups per CU) candidate of
 * the queue is timed (reps eager runs each) and the fastest kept;
 * synchronises.  Before autotune every problem goes to the queue, GATHER
 * unless rfx_batch_set_algos says otherwise. */
int rfx_batch_autotune(hipStream_t stream, rfx_batch* b, void* workspace, size_t workspace_bytes, int32_t reps,
                       double* best_ms);
int rfx_batch_set_tuning(rfx_batch* b, int32_t lines_per_item, int32_t blocks_per_cu);
/* Per-problem algorithm (RFX_ALGO_GATHER or RFX_ALGO_DENSE, `count`
 * entries); the autotuner sets them from measured leaf statistics. */
int rfx_batch_set_algos(rfx_batch* b, const int32_t* algos);
/* Per-problem algorithm and measured sum over trees and leaves of |leaf|^2
 * (0 before the autotuner ran); either pointer may be NULL. */
int rfx_batch_get_stats(const rfx_batch* b, int32_t* algos, double* sum_leaf_sq);
/* flags[i] = 1 when problem i is counted by its own kernel launch (set by
 * rfx_batch_autotune), 0 when it goes through the persistent queue.  No
 * counterpart in the reference, which has no multi-forest batching. */
int rfx_batch_get_own_launch(const rfx_batch* b, int32_t* flags);
/* Override those flags (after rfx_batch_autotune): all zeros sends every
 * problem through the persistent queue, the comparison that justifies the
 * autotuner's own launches. */
int rfx_batch_set_own_launch(rfx_batch* b, const int32_t* flags);
int rfx_batch_get_tuning(const rfx_batch* b, int32_t* lines_per_item, int32_t* blocks_per_cu);
void rfx_batch_free(rfx_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* RFX_PROX_H */
