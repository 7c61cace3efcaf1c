// C++ entry points with the reference's signatures (include/rfx_ref.hpp),
// host code over the C ABI of rfx_prox.h.  Each call stages its host inputs
// in device memory, runs the HIP path on the null stream of the current
// device and synchronises, as the reference's calls do.

#include "rfx_ref.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace rfx {
namespace ref {

Config g_config;

namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void rfx_ok(int rc, const char* what) {
  if (rc != RFX_OK) throw std::runtime_error(std::string(what) + " failed with status " + std::to_string(rc));
}

// Device buffer owned for the duration of a call.
struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(size_t bytes) {
    if (bytes) hip_ok(hipMalloc(&p, bytes), "hipMalloc");
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <typename T>
void upload(void* dst, const T* src, size_t count) {
  hip_ok(hipMemcpy(dst, src, sizeof(T) * count, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

template <typename T>
void download(T* dst, const void* src, size_t count) {
  hip_ok(hipMemcpy(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

// Leaf keys of one tree from the reference's arrays: (node << 1) | in-bag when
// the node reached is terminal (nodestatus -1), else -1.  node_base offsets
// the node ids when several trees share one key space.
void tree_keys(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t n,
               integer_t nnode, int64_t node_base, int32_t* keys) {
  for (integer_t i = 0; i < n; ++i) {
    const integer_t k = nodextr[i];
    const bool term = k >= 0 && k < nnode && nodestatus[k] == -1;
    keys[i] = term ? static_cast<int32_t>(((node_base + k) << 1) | (nin[i] > 0 ? 1 : 0)) : -1;
  }
}

void check_tree_args(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t n,
                     integer_t nnode) {
  if (!nodestatus || !nodextr || !nin || n <= 0 || nnode <= 0 || nnode > RFX_MAX_NODES_PER_TREE)
    throw std::runtime_error("invalid tree arrays");
}

// One tree's counts of `variant` in `layout` into a device int32 buffer.
void one_tree_counts(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t n,
                     integer_t nnode, int32_t variant, int32_t layout, int32_t* out_dev) {
  std::vector<int32_t> keys(n);
  tree_keys(nodestatus, nodextr, nin, n, nnode, 0, keys.data());
  DevBuf kd(4 * static_cast<size_t>(n));
  upload(kd.p, keys.data(), keys.size());
  const size_t wsb = rfx_prox_workspace_bytes(n, 1, nnode, RFX_ALGO_GATHER);
  DevBuf ws(wsb);
  rfx_ok(rfx_prox_accumulate(nullptr, kd.as<int32_t>(), n, 1, nnode, variant, layout, RFX_ALGO_GATHER, 0, out_dev,
                             ws.p, wsb),
         "rfx_prox_accumulate");
  hip_ok(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

// One tree's integer counts of `variant` (FULL), added to the host prox.
void add_tree_counts(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t nsample,
                     integer_t nnode, int32_t variant, dp_t* prox) {
  const size_t nn = static_cast<size_t>(nsample) * nsample;
  DevBuf out(4 * nn);
  one_tree_counts(nodestatus, nodextr, nin, nsample, nnode, variant, RFX_LAYOUT_FULL, out.as<int32_t>());
  std::vector<int32_t> h(nn);
  download(h.data(), out.p, nn);
  for (size_t e = 0; e < nn; ++e) prox[e] += static_cast<dp_t>(h[e]);
}

// One tree's case-wise weights (rfx_prox_casewise over a single tree), added
// to the host prox.
void add_tree_casewise(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t n,
                       integer_t nnode, dp_t* prox) {
  std::vector<int32_t> keys(n);
  tree_keys(nodestatus, nodextr, nin, n, nnode, 0, keys.data());
  const size_t nn = static_cast<size_t>(n) * n;
  DevBuf kd(4 * static_cast<size_t>(n)), nd(4 * static_cast<size_t>(n)), pd(8 * nn);
  upload(kd.p, keys.data(), keys.size());
  upload(nd.p, nin, n);
  upload(pd.p, prox, nn);
  const size_t wsb = rfx_prox_casewise_workspace_bytes(n, 1, nnode);
  DevBuf ws(wsb);
  rfx_ok(rfx_prox_casewise(nullptr, kd.as<int32_t>(), nd.as<int32_t>(), n, 1, nnode, 1, pd.as<double>(), ws.p, wsb),
         "rfx_prox_casewise");
  download(prox, pd.p, nn);
}

}  // namespace

void cpu_proximity(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t nsample,
                   integer_t nnode, dp_t* prox, integer_t*, integer_t*, integer_t*, integer_t*, integer_t*,
                   integer_t*) {
  check_tree_args(nodestatus, nodextr, nin, nsample, nnode);
  if (!prox) throw std::runtime_error("prox is null");
  if (g_config.use_casewise)
    add_tree_casewise(nodestatus, nodextr, nin, nsample, nnode, prox);
  else
    add_tree_counts(nodestatus, nodextr, nin, nsample, nnode, RFX_PROX_OOB_ASYM, prox);
}

void gpu_proximity(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin, integer_t nsample,
                   integer_t nnode, dp_t* prox, integer_t* nod, integer_t* ncount, integer_t* ncn, integer_t* nodexb,
                   integer_t* ndbegin, integer_t* npcase) {
  check_tree_args(nodestatus, nodextr, nin, nsample, nnode);
  if (!prox) throw std::runtime_error("prox is null");
  if (nsample > 5000 && nnode > 10)  // the reference's use_persistent branch
    add_tree_counts(nodestatus, nodextr, nin, nsample, nnode, RFX_PROX_OOB_ROW, prox);
  else
    cpu_proximity(nodestatus, nodextr, nin, nsample, nnode, prox, nod, ncount, ncn, nodexb, ndbegin, npcase);
}

void gpu_proximity_upper_triangle_fp16(const integer_t* nodestatus, const integer_t* nodextr, const integer_t* nin,
                                       integer_t nsample, integer_t nnode, __half* prox_upper_fp16, integer_t*,
                                       integer_t*, integer_t*, integer_t*, integer_t*, integer_t*) {
  check_tree_args(nodestatus, nodextr, nin, nsample, nnode);
  hipPointerAttribute_t attr;
  if (!prox_upper_fp16 || hipPointerGetAttributes(&attr, prox_upper_fp16) != hipSuccess ||
      attr.type != hipMemoryTypeDevice)
    throw std::runtime_error("prox_upper_fp16 must be device memory");
  const int64_t len = static_cast<int64_t>(nsample) * (nsample + 1) / 2;
  DevBuf out(4 * static_cast<size_t>(len));
  one_tree_counts(nodestatus, nodextr, nin, nsample, nnode, RFX_PROX_OOB_UPPER, RFX_LAYOUT_UPPER,
                  out.as<int32_t>());
  rfx_ok(rfx_counts_convert(nullptr, out.as<int32_t>(), len, RFX_PREC_FP16, static_cast<void*>(prox_upper_fp16)),
         "rfx_counts_convert");
  hip_ok(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

void gpu_finishprox(integer_t nsample, const integer_t* nout, dp_t* prox, dp_t* proxsym) {
  if (nsample <= 0 || !nout || !prox || !proxsym) throw std::runtime_error("invalid finishprox arguments");
  const size_t nn = static_cast<size_t>(nsample) * nsample;
  DevBuf pd(8 * nn), sd(8 * nn), od(4 * static_cast<size_t>(nsample));
  upload(pd.p, prox, nn);
  upload(od.p, nout, nsample);
  rfx_ok(rfx_finishprox_f64(nullptr, pd.as<double>(), od.as<int32_t>(), nsample, sd.as<double>(), RFX_LAYOUT_FULL),
         "rfx_finishprox_f64");
  download(proxsym, sd.p, nn);
}

void cpu_finishprox(integer_t nsample, const integer_t* nout, dp_t* prox, dp_t* proxsym) {
  gpu_finishprox(nsample, nout, prox, proxsym);
}

std::vector<double> compute_mds_3d_cpu(const dp_t* proximity_matrix, integer_t n_samples, bool,
                                       const integer_t* oob_counts_rfgap, bool use_rfgap) {
  if (!proximity_matrix || n_samples < 4) throw std::runtime_error("MDS needs a proximity matrix with n >= 4");
  const int64_t n = n_samples;
  const size_t nn = static_cast<size_t>(n) * n;
  DevBuf pd(8 * nn), od(oob_counts_rfgap ? 4 * static_cast<size_t>(n) : 0), cd(8 * 3 * static_cast<size_t>(n));
  upload(pd.p, proximity_matrix, nn);
  if (oob_counts_rfgap) upload(od.p, oob_counts_rfgap, n);
  const size_t wsb = rfx_mds_dense_workspace_bytes(n, RFX_PREC_FP64);
  DevBuf ws(wsb);
  double eig[16];
  int32_t iters = 0;
  const int rc = (use_rfgap && oob_counts_rfgap)
                     ? rfx_mds_dense_rfgap(nullptr, pd.as<double>(), n, od.as<int32_t>(), RFX_PREC_FP64, 1000, 1e-10,
                                           ws.p, wsb, cd.as<double>(), eig, &iters)
                     : rfx_mds_dense_matrix(nullptr, pd.as<double>(), n, od.as<int32_t>(), RFX_PREC_FP64, 1000, 1e-10,
                                            ws.p, wsb, cd.as<double>(), eig, &iters);
  if (rc == RFX_ERR_INVALID)
    throw std::runtime_error("MDS failed: fewer than 3 positive eigenvalues");
  rfx_ok(rc, "rfx_mds_dense");
  std::vector<double> coords(3 * static_cast<size_t>(n));
  download(coords.data(), cd.p, coords.size());
  return coords;
}

void gpu_proximity_rfgap(integer_t ntree, integer_t nsample, const integer_t* tree_nin_flat,
                         const integer_t* tree_nodextr_flat, dp_t* prox) {
  if (ntree <= 0 || nsample <= 0 || !tree_nin_flat || !tree_nodextr_flat || !prox)
    throw std::runtime_error("invalid RF-GAP arguments");
  const size_t slots = static_cast<size_t>(ntree) * nsample, nn = static_cast<size_t>(nsample) * nsample;
  // node offsets from each tree's largest reached node
  std::vector<int64_t> off(static_cast<size_t>(ntree) + 1, 0);
  for (integer_t t = 0; t < ntree; ++t) {
    integer_t mx = -1;
    for (integer_t i = 0; i < nsample; ++i) mx = std::max(mx, tree_nodextr_flat[static_cast<size_t>(t) * nsample + i]);
    off[t + 1] = off[t] + std::max<int64_t>(mx + 1, 1);
  }
  DevBuf nd(4 * slots), xd(4 * slots), od(8 * off.size()), pd(8 * nn);
  upload(nd.p, tree_nin_flat, slots);
  upload(xd.p, tree_nodextr_flat, slots);
  upload(od.p, off.data(), off.size());
  const size_t wsb = rfx_prox_rfgap_workspace_bytes(nsample, ntree, off.back());
  DevBuf ws(wsb);
  rfx_ok(rfx_prox_rfgap(nullptr, nd.as<int32_t>(), xd.as<int32_t>(), od.as<int64_t>(), nsample, ntree, off.back(),
                        pd.as<double>(), ws.p, wsb),
         "rfx_prox_rfgap");
  download(prox, pd.p, nn);
}

void cpu_proximity_rfgap(integer_t ntree, integer_t nsample, const std::vector<std::vector<integer_t>>& tree_nin,
                         const std::vector<std::vector<integer_t>>& tree_nodextr, dp_t* prox) {
  if (ntree <= 0 || nsample <= 0 || tree_nin.size() != static_cast<size_t>(ntree) ||
      tree_nodextr.size() != static_cast<size_t>(ntree))
    throw std::runtime_error("RF-GAP: per-tree arrays do not match ntree");
  const size_t n = static_cast<size_t>(nsample);
  std::vector<integer_t> nin(n * ntree), ndx(n * ntree);
  for (integer_t t = 0; t < ntree; ++t) {
    if (tree_nin[t].size() != n || tree_nodextr[t].size() != n)
      throw std::runtime_error("RF-GAP: per-tree arrays do not match nsample");
    std::copy(tree_nin[t].begin(), tree_nin[t].end(), nin.begin() + static_cast<size_t>(t) * n);
    std::copy(tree_nodextr[t].begin(), tree_nodextr[t].end(), ndx.begin() + static_cast<size_t>(t) * n);
  }
  gpu_proximity_rfgap(ntree, nsample, nin.data(), ndx.data(), prox);
}

std::vector<float> compute_outlier_scores(const dp_t* proximity_matrix, integer_t nsample) {
  if (!proximity_matrix || nsample <= 0) throw std::runtime_error("invalid outlier-score arguments");
  const size_t nn = static_cast<size_t>(nsample) * nsample;
  DevBuf pd(8 * nn), sd(4 * static_cast<size_t>(nsample));
  upload(pd.p, proximity_matrix, nn);
  rfx_ok(rfx_outlier_scores(nullptr, pd.as<double>(), nsample, sd.as<float>(), nullptr), "rfx_outlier_scores");
  std::vector<float> out(static_cast<size_t>(nsample));
  download(out.data(), sd.p, out.size());
  return out;
}

// ---------------------------------------------------------------------------
// LowRankProximityMatrix
// ---------------------------------------------------------------------------
LowRankProximityMatrix::LowRankProximityMatrix(integer_t n_samples, integer_t initial_rank, FactorFormat format,
                                               integer_t max_rank)
    : n_(n_samples), rank_(max_rank > 0 ? max_rank : initial_rank), format_(format) {
  if (n_samples < 4 || rank_ < 1 || rank_ > 128) throw std::runtime_error("LowRankProximityMatrix: bad size/rank");
}

LowRankProximityMatrix::~LowRankProximityMatrix() {
  for (void* p : {static_cast<void*>(keys_), static_cast<void*>(U_), Uq_, static_cast<void*>(scales_),
                  static_cast<void*>(A_)})
    if (p) (void)hipFree(p);
}

bool LowRankProximityMatrix::initialize() {
  try {
    reserve_trees(64);
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

void LowRankProximityMatrix::reserve_trees(int64_t trees) {
  if (trees <= cap_trees_) return;
  int64_t cap = cap_trees_ ? cap_trees_ : 64;
  while (cap < trees) cap *= 2;
  int32_t* nk = nullptr;
  hip_ok(hipMalloc(&nk, 4 * static_cast<size_t>(cap) * n_), "hipMalloc keys");
  if (keys_) {
    (void)hipMemcpy(nk, keys_, 4 * static_cast<size_t>(trees_) * n_, hipMemcpyDeviceToDevice);
    (void)hipFree(keys_);
  }
  keys_ = nk;
  cap_trees_ = cap;
}

void LowRankProximityMatrix::add_tree_contribution(const integer_t* nodestatus, const integer_t* nodextr,
                                                   const integer_t* nin, integer_t nnode) {
  check_tree_args(nodestatus, nodextr, nin, n_, nnode);
  if (trees_ >= 65535) throw std::runtime_error("LowRankProximityMatrix: too many trees");
  if (total_nodes_ + nnode > RFX_MAX_TOTAL_NODES) throw std::runtime_error("LowRankProximityMatrix: too many nodes");
  reserve_trees(trees_ + 1);
  std::vector<int32_t> k(n_);
  tree_keys(nodestatus, nodextr, nin, n_, nnode, total_nodes_, k.data());
  upload(keys_ + trees_ * n_, k.data(), k.size());
  ++trees_;
  total_nodes_ += nnode;
  node_offset_.push_back(total_nodes_);
  ready_ = false;
}

void LowRankProximityMatrix::add_tree_contribution_incremental_upper_triangle_fp16(const __half*, integer_t) {
  throw std::runtime_error(
      "LowRankProximityMatrix::add_tree_contribution_incremental_upper_triangle_fp16: a per-tree FP16 proximity "
      "triangle carries no leaf structure; this build forms the factors of C/T from the trees' leaves -- add each "
      "tree with add_tree_contribution(nodestatus, nodextr, nin, nnode)");
}

void LowRankProximityMatrix::finalize_accumulation(integer_t, bool final_call) {
  if (!final_call) return;
  if (trees_ == 0) throw std::runtime_error("LowRankProximityMatrix: no trees added");
  const int32_t over = 16;
  const size_t wsb = rfx_lowrank_workspace_bytes(n_, static_cast<int32_t>(trees_), total_nodes_, rank_, over);
  DevBuf ws(wsb);
  if (!U_) hip_ok(hipMalloc(&U_, 4 * static_cast<size_t>(n_) * rank_), "hipMalloc U");
  eig_.assign(rank_, 0.0);
  rfx_ok(rfx_lowrank_factors(nullptr, keys_, n_, static_cast<int32_t>(trees_), total_nodes_, node_offset_.data(),
                             rank_, over, 2, 1, ws.p, wsb, U_, eig_.data()),
         "rfx_lowrank_factors");
  const size_t elt = format_ == FactorFormat::FP32 ? 4 : format_ == FactorFormat::BF16 ? 2 : 1;
  if (!Uq_) hip_ok(hipMalloc(&Uq_, elt * n_ * rank_), "hipMalloc Uq");
  if (!scales_) hip_ok(hipMalloc(&scales_, 4 * static_cast<size_t>(rank_)), "hipMalloc scales");
  rfx_ok(rfx_quantize_factors(nullptr, U_, n_, rank_, static_cast<int32_t>(format_), Uq_, scales_),
         "rfx_quantize_factors");
  // A = B = U as float64 column-major
  std::vector<float> uh(static_cast<size_t>(n_) * rank_);
  download(uh.data(), U_, uh.size());
  std::vector<double> ah(uh.size());
  for (int64_t i = 0; i < n_; ++i)
    for (int32_t c = 0; c < rank_; ++c) ah[i + static_cast<size_t>(c) * n_] = uh[i * rank_ + c];
  if (!A_) hip_ok(hipMalloc(&A_, 8 * ah.size()), "hipMalloc A");
  upload(A_, ah.data(), ah.size());
  ready_ = true;
}

void LowRankProximityMatrix::get_factors(dp_t** A_dev, dp_t** B_dev, integer_t* rank) {
  if (!ready_) throw std::runtime_error("LowRankProximityMatrix: factors not finalized");
  if (A_dev) *A_dev = A_;
  if (B_dev) *B_dev = A_;
  if (rank) *rank = rank_;
}

std::vector<double> LowRankProximityMatrix::compute_mds_from_factors(integer_t k) {
  if (!ready_) throw std::runtime_error("LowRankProximityMatrix: factors not finalized");
  if (k < 1 || k > 3) throw std::runtime_error("compute_mds_from_factors supports k <= 3");
  if (rank_ > 128) throw std::runtime_error("factor MDS supports rank <= 128");
  // matrix cores on the stored factors, refined with the FP32 factors
  const size_t wsb = rfx_mds_factors_mixed_workspace_bytes(n_, rank_);
  DevBuf ws(wsb), cd(8 * 3 * static_cast<size_t>(n_));
  double eig[16];
  int32_t iters[2] = {0, 0};
  rfx_ok(rfx_mds_factors_mixed(nullptr, Uq_, scales_, static_cast<int32_t>(format_), U_, n_, rank_, 500, 1e-4, 60,
                               5e-5, ws.p, wsb, cd.as<double>(), eig, iters),
         "rfx_mds_factors_mixed");
  std::vector<double> c3(3 * static_cast<size_t>(n_));
  download(c3.data(), cd.p, c3.size());
  std::vector<double> out(static_cast<size_t>(k) * n_);
  for (int64_t i = 0; i < n_; ++i)
    for (int32_t c = 0; c < k; ++c) out[i * k + c] = c3[i * 3 + c];
  return out;
}

void LowRankProximityMatrix::reconstruct_full_matrix(dp_t* output_prox, integer_t nsample, FactorFormat level) {
  if (!ready_) throw std::runtime_error("LowRankProximityMatrix: factors not finalized");
  if (!output_prox || nsample != n_) throw std::runtime_error("reconstruct_full_matrix: bad output or nsample");
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, output_prox) != hipSuccess || attr.type != hipMemoryTypeDevice)
    throw std::runtime_error("reconstruct_full_matrix: output_prox must be device memory");
  if (level == FactorFormat::FP32) {
    rfx_ok(rfx_reconstruct_factors(nullptr, U_, nullptr, RFX_PREC_FP32, n_, rank_, output_prox),
           "rfx_reconstruct_factors");
  } else if (level == format_) {
    rfx_ok(rfx_reconstruct_factors(nullptr, Uq_, scales_, static_cast<int32_t>(format_), n_, rank_, output_prox),
           "rfx_reconstruct_factors");
  } else {
    const size_t elt = level == FactorFormat::BF16 ? 2 : 1;
    DevBuf q(elt * static_cast<size_t>(n_) * rank_), s(4 * static_cast<size_t>(rank_));
    rfx_ok(rfx_quantize_factors(nullptr, U_, n_, rank_, static_cast<int32_t>(level), q.p, s.as<float>()),
           "rfx_quantize_factors");
    rfx_ok(rfx_reconstruct_factors(nullptr, q.p, s.as<float>(), static_cast<int32_t>(level), n_, rank_, output_prox),
           "rfx_reconstruct_factors");
  }
  hip_ok(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}

bool reconstruct_proximity_matrix_host(void* obj, dp_t** output_prox, integer_t nsamples) {
  try {
    auto* m = static_cast<LowRankProximityMatrix*>(obj);
    if (!m || !output_prox || nsamples != m->get_n_samples() || nsamples > 50000) return false;
    const size_t nn = static_cast<size_t>(nsamples) * nsamples;
    DevBuf d(8 * nn);
    m->reconstruct_full_matrix(d.as<dp_t>(), nsamples, FactorFormat::FP32);
    auto* h = new (std::nothrow) dp_t[nn];
    if (!h) return false;
    if (hipMemcpy(h, d.p, 8 * nn, hipMemcpyDeviceToHost) != hipSuccess) {
      delete[] h;
      return false;
    }
    *output_prox = h;
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

bool get_lowrank_factors_host(void* obj, dp_t** A_host, dp_t** B_host, integer_t* rank, integer_t nsamples) {
  try {
    auto* m = static_cast<LowRankProximityMatrix*>(obj);
    if (!m || !A_host || !B_host || !rank || nsamples != m->get_n_samples()) return false;
    dp_t *A = nullptr, *B = nullptr;
    integer_t r = 0;
    m->get_factors(&A, &B, &r);
    const size_t cnt = static_cast<size_t>(nsamples) * r;
    auto* ah = new (std::nothrow) dp_t[cnt];
    auto* bh = new (std::nothrow) dp_t[cnt];
    if (!ah || !bh) {
      delete[] ah;
      delete[] bh;
      return false;
    }
    if (hipMemcpy(ah, A, 8 * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(bh, B, 8 * cnt, hipMemcpyDeviceToHost) != hipSuccess) {
      delete[] ah;
      delete[] bh;
      return false;
    }
    *A_host = ah;
    *B_host = bh;
    *rank = r;
    return true;
  } catch (const std::exception&) {
    return false;
  }
}

std::vector<double> compute_mds_from_factors_host(void* obj, integer_t k) {
  try {
    auto* m = static_cast<LowRankProximityMatrix*>(obj);
    if (!m) return {};
    return m->compute_mds_from_factors(k);
  } catch (const std::exception&) {
    return {};
  }
}

}  // namespace ref
}  // namespace rfx
