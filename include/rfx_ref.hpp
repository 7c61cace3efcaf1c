// rfx_ref.hpp -- C++ entry points with the reference's signatures, running
// on the MI355X through the C ABI of rfx_prox.h.
//
// A host program written against the reference's headers switches by
// including this header and calling rfx::ref:: instead of rf:: /
// rf::cuda::.  Argument meaning, array layouts and ownership follow the
// reference; where the reference is silent on failure (void functions) these
// throw std::runtime_error instead of returning quietly, and the helper
// functions that the reference documents as returning false / an empty
// vector do exactly that.  Every call synchronises (the reference calls are
// synchronous) and uses the null stream of the current device.
//
// The per-tree functions exist for drop-in compatibility with callers that
// loop over trees; the all-trees path (rfx_leaf_assign +
// rfx_prox_accumulate over every tree in one call) is the fast one.
#ifndef RFX_REF_HPP
#define RFX_REF_HPP

#include <hip/hip_fp16.h>

#include <cstdint>
#include <vector>

#include "rfx_prox.h"

namespace rfx {
namespace ref {

using integer_t = int32_t;
using dp_t = double;

// The process-wide switch the reference's per-tree proximity reads from its
// global configuration (rf::g_config.use_casewise, include/rf_config.hpp:39,
// read at src/rf_proximity.cpp:34 and cuda/rf_proximity.cu:345, 411; set by
// RandomForest::fit from its config, src/rf_random_forest.cpp:442).  A caller
// that sets rf::g_config.use_casewise sets rfx::ref::g_config.use_casewise.
struct Config {
  bool use_casewise = false;  // weights nin(j) / (sum of nin over the leaf) instead of 1
};
extern Config g_config;

// Per-tree OOB proximity, accumulated into the caller's column-major n x n
// float64 prox.  Replaces rf::cpu_proximity (include/rf_proximity.hpp:10-21,
// src/rf_proximity.cpp:99-231): for i out-of-bag and j in-bag in the same
// terminal node (nodestatus -1), prox[i + j*n] += 1, or with
// g_config.use_casewise += nin(j) / (sum of nin over the node's in-bag cases)
// (src/rf_proximity.cpp:186-223).  This is the reference's accumulating
// dense branch at every n; its n > 5000 block-sparse branch
// (src/rf_proximity.cpp:37-98), which overwrites prox with the last tree
// alone, is not reproduced.  nod/ncount/ncn/nodexb/ndbegin/npcase are the
// reference's caller-owned scratch arrays; they are accepted and not touched.
void cpu_proximity(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t nsample,
                   integer_t nnode, dp_t* prox, integer_t* nod, integer_t* ncount, integer_t* ncn, integer_t* nodexb,
                   integer_t* ndbegin, integer_t* npcase);
// Replaces rf::gpu_proximity (cuda/rf_proximity.cu:257-527) with its
// dispatch: when nsample > 5000 and nnode > 10 it applies the rule of
// cuda_proximity_kernel (cuda/rf_proximity.cu:32-67): for i out-of-bag,
// prox[i + j*n] += 1 for every j in the same terminal node, in-bag or not,
// i itself included, with no case-wise weights (that kernel ignores them);
// otherwise the rule of cpu_proximity above (cuda_batch_proximity_kernel,
// cuda/rf_proximity.cu:72-127, case-wise per g_config).  The reference
// launches the first kernel on a one-dimensional grid although it indexes
// rows by the second grid dimension, so it only ever fills row 0; here every
// out-of-bag row gets the kernel's rule.  Accumulates into prox, which the
// reference's caller does with a zeroed per-tree matrix
// (cuda/rf_growtree.cu:4096-4118).
void gpu_proximity(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t nsample,
                   integer_t nnode, dp_t* prox, integer_t* nod, integer_t* ncount, integer_t* ncn, integer_t* nodexb,
                   integer_t* ndbegin, integer_t* npcase);

// One tree's OOB fold U[i<=j] = P_ij + P_ji into a device packed upper
// triangle of n(n+1)/2 IEEE half values, overwritten (exact: counts <= 2048).
// Replaces rf::cuda::gpu_proximity_upper_triangle_fp16
// (include/rf_proximity_upper_triangle.hpp:42-55,
// cuda/rf_proximity_upper_triangle.cu:149-306, zeroed at 948).
void gpu_proximity_upper_triangle_fp16(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin,
                                       integer_t nsample, integer_t nnode, __half* prox_upper_fp16,
                                       integer_t* nod, integer_t* ncount, integer_t* ncn, integer_t* nodexb,
                                       integer_t* ndbegin, integer_t* npcase);

// Replaces rf::cpu_finishprox / rf::gpu_finishprox
// (include/rf_finishprox.hpp:9-12, src/rf_finishprox.cpp:12-74): host
// float64 n x n column-major prox in, proxsym out, bit-identical.
void cpu_finishprox(integer_t nsample, const integer_t* nout, dp_t* prox, dp_t* proxsym);
void gpu_finishprox(integer_t nsample, const integer_t* nout, dp_t* prox, dp_t* proxsym);

// Replaces rf::compute_mds_3d_cpu (include/rf_mds_cpu.hpp:27-33,
// src/rf_mds_cpu.cpp:88-341): n x 3 row-major coordinates of the classical
// MDS of the host float64 column-major proximity.  With oob_counts the
// reference's renormalisation is applied: by nout with a unit diagonal
// (src/rf_mds_cpu.cpp:154-186), or with use_rfgap the RF-GAP form that keeps
// the diagonal (125-153).  Throws std::runtime_error when fewer than three
// eigenvalues exceed 1e-10 (as the reference).  memory_check is accepted;
// the GPU form needs O(n) memory besides the input.
std::vector<double> compute_mds_3d_cpu(const dp_t* proximity_matrix, integer_t n_samples, bool memory_check = true,
                                       const integer_t* oob_counts_rfgap = nullptr, bool use_rfgap = false);

// RF-GAP proximity over all trees into the host float64 n x n prox
// (overwritten, symmetric).  Replaces rf::cpu_proximity_rfgap and
// rf::gpu_proximity_rfgap (include/rf_proximity.hpp:34-46,
// src/rf_proximity.cpp:238-391, cuda/rf_proximity.cu:541-786): per-tree
// bootstrap multiplicities and reached nodes, nested or flattened
// [tree * nsample + sample].  Runs rfx_prox_rfgap; node offsets are formed
// from each tree's largest node index.
void cpu_proximity_rfgap(integer_t ntree, integer_t nsample, const std::vector<std::vector<integer_t>>& tree_nin,
                         const std::vector<std::vector<integer_t>>& tree_nodextr, dp_t* prox);
void gpu_proximity_rfgap(integer_t ntree, integer_t nsample, const integer_t* tree_nin_flat,
                         const integer_t* tree_nodextr_flat, dp_t* prox);

// RandomForest::compute_outlier_scores (src/rf_random_forest.cpp:2489-2520)
// on a host float64 n x n finished proximity: the n float scores, bit-identical.
std::vector<float> compute_outlier_scores(const dp_t* proximity_matrix, integer_t nsample);

// Storage levels of the factors (the reference's QuantizationLevel names
// FP32/FP16/INT8/NF4; here the stored formats of rfx_quantize_factors).
enum class FactorFormat : int32_t { FP32 = RFX_PREC_FP32, BF16 = RFX_PREC_BF16, INT8 = RFX_PREC_INT8,
                                    FP8 = RFX_PREC_FP8 };

// Low-rank proximity factors, the interface of
// rf::cuda::LowRankProximityMatrix (include/rf_proximity_lowrank.hpp:27-420)
// over this build's factor path: trees are added as the reference's per-tree
// arrays, finalize_accumulation forms rank-r factors U of P = C/T (ALL
// counts / trees, PSD) matrix-free from the leaf buckets, and get_factors
// returns A = B = U (device float64, n x r column-major, object-owned), as the
// reference's get_factors (cuda/rf_proximity_lowrank.cu:1753-1779).
class LowRankProximityMatrix {
 public:
  LowRankProximityMatrix(integer_t n_samples, integer_t initial_rank, FactorFormat format = FactorFormat::BF16,
                         integer_t max_rank = 0);
  ~LowRankProximityMatrix();
  LowRankProximityMatrix(const LowRankProximityMatrix&) = delete;
  LowRankProximityMatrix& operator=(const LowRankProximityMatrix&) = delete;

  bool initialize();
  // One tree (host arrays: nodestatus[nnode], nodextr[n], nin[n]).
  void add_tree_contribution(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin,
                             integer_t nnode);
  // The reference's per-tree entry for a precomputed FP16 OOB triangle
  // (include/rf_proximity_lowrank.hpp:106).  A packed count triangle carries
  // no leaf structure, from which this build's matrix-free factors of C / T
  // are formed, so the call throws std::runtime_error naming
  // add_tree_contribution, the entry that takes the tree itself.
  void add_tree_contribution_incremental_upper_triangle_fp16(const __half* tree_proximity_upper_fp16,
                                                             integer_t nsample);
  // Builds the factors once all trees are in (final_call true).
  void finalize_accumulation(integer_t n_trees_processed, bool final_call = true);
  void get_factors(dp_t** A_dev, dp_t** B_dev, integer_t* rank);
  // n x k row-major MDS coordinates from the factors (k <= 3).
  std::vector<double> compute_mds_from_factors(integer_t k = 3);
  // P = A B^T (= U U^T) into the device float64 n x n column-major
  // output_prox, from the factors stored in `level` (FP32 uses the FP32
  // factors; another level than the object's storage quantises a copy).
  // Replaces reconstruct_full_matrix (include/rf_proximity_lowrank.hpp:142,
  // cuda/rf_proximity_lowrank.cu:1307-1367).
  void reconstruct_full_matrix(dp_t* output_prox, integer_t nsample, FactorFormat level = FactorFormat::FP32);
  integer_t get_rank() const { return rank_; }
  integer_t get_n_samples() const { return n_; }
  std::vector<double> top_eigenvalues() const { return eig_; }

 private:
  void reserve_trees(int64_t trees);
  integer_t n_, rank_;
  FactorFormat format_;
  int64_t trees_ = 0, cap_trees_ = 0, total_nodes_ = 0;
  std::vector<int64_t> node_offset_{0};
  int32_t* keys_ = nullptr;      // device [T][n]
  float* U_ = nullptr;           // device n x r float32 row-major
  void* Uq_ = nullptr;           // device stored factors
  float* scales_ = nullptr;      // device [r]
  double* A_ = nullptr;          // device n x r float64 column-major
  std::vector<double> eig_;
  bool ready_ = false;
};

// Helpers of include/rf_lowrank_helpers.hpp:12-24: host copies allocated
// with new[] (the caller delete[]s them); false / an empty vector on failure.
bool get_lowrank_factors_host(void* lowrank_proximity_ptr, dp_t** A_host, dp_t** B_host, integer_t* rank,
                              integer_t nsamples);
std::vector<double> compute_mds_from_factors_host(void* lowrank_proximity_ptr, integer_t k);
inline std::vector<double> compute_mds_3d_from_factors_host(void* lowrank_proximity_ptr) {
  return compute_mds_from_factors_host(lowrank_proximity_ptr, 3);
}
// include/rf_lowrank_helpers.hpp:29-34 (cuda/rf_lowrank_helpers.cu:73-130):
// the full float64 n x n column-major P = U U^T (FP32 factors) in a host
// array allocated with new[] that the caller delete[]s; false on failure and,
// as the reference, for nsamples > 50,000.
bool reconstruct_proximity_matrix_host(void* lowrank_proximity_ptr, dp_t** output_prox, integer_t nsamples);

}  // namespace ref
}  // namespace rfx

#endif  // RFX_REF_HPP
