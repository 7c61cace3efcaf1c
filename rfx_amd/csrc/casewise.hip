// Case-wise (bootstrap-weighted) proximity, the reference's use_casewise
// branch of the accumulating V1 proximity (src/rf_proximity.cpp:186-223):
// for every tree t, an out-of-bag case i and an in-bag case j of the same
// leaf add
//     P_w[i, j] += nin_t(j) / sum over the leaf's in-bag cases k of nin_t(k)
// into the FP64 n x n column-major matrix (row = the out-of-bag case), the
// layout of the reference's prox.  The weights of one out-of-bag case over
// its leaf sum to 1.
//
// Built on the leaf buckets of the integer counts: the leaf's bootstrap draws
// sum_k nin_t(k) come from one integer-atomic pass over the slots, then one
// workgroup per (output column j, row chunk of at most 20,480 rows; in-bag j
// only contributes) gathers the out-of-bag members of j's leaf in every tree
// into FP64 LDS counters, with j's weight nin_t(j) / draws, and writes the
// chunk of the column once.  The member runs of a wavefront's 64 trees are
// flattened across its lanes in 16-byte blocks of the member array (16-bit
// members up to 32,768 cases), as in the integer gather.  FP64 LDS atomics
// add the weights of different trees in no fixed order, so an entry may differ
// from the reference's tree-ordered sum in its last bits (about 1e-16
// relative).
//
// RF-GAP (src/rf_proximity.cpp:238-391, cuda/rf_proximity.cu:541-786) is the
// same leaf sum with the weight formed as nin_t(j) * (1 / draws), followed by
// a row scaling by 1 / |S_i| (S_i = the trees in which i is out of bag), the
// symmetric average (p_ij + p_ji) / 2 and a unit diagonal for cases that are
// never out of bag.  The scaling and the average run in one in-place pass
// over pairs of mirrored 64 x 64 tiles.
#include <algorithm>

#include "rfx_common.hpp"
#include "bucket.hpp"
#include "dense.hpp"
#include "gather.hpp"

namespace {

constexpr int kCwUnroll = 4;

// draws[leaf] = sum of nin over the leaf's in-bag slots, read from the sorted
// slots of the bucket build (position p: leaf id K[p], V[p] = slot << 1 |
// in-bag).  The positions of one leaf are contiguous, so a segmented scan
// across the wavefront sums each leaf's run and the run's last lane adds it
// with one atomic (two when a run crosses a wavefront boundary).  draws must
// be zeroed; the off-leaf sentinel id (total_nodes) is skipped.
__global__ void cw_leaf_draws_kernel(const uint32_t* __restrict__ K, const uint32_t* __restrict__ V,
                                     const int32_t* __restrict__ nin, int64_t slots, uint32_t total_nodes,
                                     uint32_t* __restrict__ draws) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool in = p < slots;
  const uint32_t k = in ? K[p] : total_nodes;
  uint32_t sum = 0;
  if (in) {
    const uint32_t v = V[p];
    if (v & 1u) sum = static_cast<uint32_t>(nin[v >> 1]);
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(sum, d, 64);
    const uint32_t ky = __shfl_up(k, d, 64);
    if (lane >= d && ky == k) sum += y;
  }
  const uint32_t knext = __shfl_down(k, 1, 64);
  const bool last = lane == 63 || knext != k;
  if (in && last && k < total_nodes && sum) atomicAdd(&draws[k], sum);
}

#ifndef RFX_CW_MAX_ROWS
#define RFX_CW_MAX_ROWS 20480  // FP64 LDS counters per workgroup: up to 160 KiB, a whole column to n = 20,480
#endif
#ifndef RFX_CW_THREADS
#define RFX_CW_THREADS 512
#endif
constexpr int kCwThreads = RFX_CW_THREADS;
constexpr int64_t kCwMaxRows = RFX_CW_MAX_ROWS;
// cw_line_kernel has no static LDS: its dynamic request is 8 * chunk_len <= 8 * kCwMaxRows.
static_assert(8 * kCwMaxRows <= kLdsLimitBytes, "case-wise line counters exceed the LDS");

// The wavefront's 64 member runs (lane l: run [s, s + len) of its tree, with
// that tree's weight w) flattened across the lanes in aligned 16-byte blocks
// of the member array: one owner search and one 16-byte load per block, and
// w added to acc[row - lo] for each out-of-bag member of the block inside the
// run.  All lanes of the wavefront must call.
template <class M>
__device__ __forceinline__ void cw_visit(uint32_t s, uint32_t len, double w, int lane, const M* __restrict__ members,
                                         double* acc, int64_t lo) {
  constexpr uint32_t P = 16 / sizeof(M);
  constexpr int U = kCwUnroll;
  const uint32_t e = s + len;
  const uint32_t nb = len ? (e - 1) / P - s / P + 1 : 0u;
  uint32_t pre = nb;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(pre, d, 64);
    if (lane >= d) pre += y;
  }
  const uint32_t excl = pre - nb;
  const uint32_t total = __shfl(pre, 63, 64);
  for (uint32_t g0 = 0; g0 < total; g0 += 64 * U) {
    uint4 blk[U];
    uint32_t blo[U], bhi[U];
    double wv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t g = g0 + u * 64 + lane;
      int o = 0;
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t v = __shfl(pre, o + d - 1, 64);
        if (v <= g) o += d;
      }
      o = min(o, 63);
      const uint32_t so = __shfl(s, o, 64), eo = __shfl(e, o, 64), xo = __shfl(excl, o, 64);
      wv[u] = __shfl(w, o, 64);
      const bool ok = g < total;
      const uint32_t bb = (so / P + (g - xo)) * P;  // first member index of the block
      blo[u] = ok ? (so > bb ? so - bb : 0u) : P;
      bhi[u] = ok ? min(P, eo - bb) : 0u;
      blk[u] = ok ? *reinterpret_cast<const uint4*>(members + bb) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t bw[4] = {blk[u].x, blk[u].y, blk[u].z, blk[u].w};
#pragma unroll
      for (uint32_t k = 0; k < P; ++k) {
        if (k < blo[u] || k >= bhi[u]) continue;
        const uint32_t m = sizeof(M) == 2 ? ((bw[k / 2] >> (16 * (k % 2))) & 0xffffu) : bw[k];
        if (!(m & 1u)) atomicAdd(&acc[static_cast<int64_t>(m >> 1) - lo], wv[u]);
      }
    }
  }
}

// One workgroup per (in-bag line j, row chunk): the out-of-bag members of j's
// leaf in every tree where j is in bag add j's weight in that tree to the FP64
// LDS counters of the chunk's rows; the chunk of column j is written once.
// kRecip: the weight as nin * (1 / draws) (RF-GAP) rather than nin / draws.
template <bool kRecip, class M>
__global__ __launch_bounds__(kCwThreads) void cw_line_kernel(const uint2* __restrict__ pe,
                                                             const uint32_t* __restrict__ bstart,
                                                             const uint32_t* __restrict__ mbucket,
                                                             const M* __restrict__ members,
                                                             const uint32_t* __restrict__ draws,
                                                             const int32_t* __restrict__ nin, int64_t n, int32_t ntree,
                                                             int64_t chunk_len, int32_t accumulate,
                                                             double* __restrict__ out) {
  extern __shared__ double cw_lds[];
  double* acc = cw_lds;
  const int64_t j = xcd_row(blockIdx.x);
  const int64_t lo = static_cast<int64_t>(blockIdx.y) * chunk_len;
  if (j >= n || lo >= n) return;
  const int64_t hi = min(n, lo + chunk_len);
  const int len_chunk = static_cast<int>(hi - lo);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int q = tid; q < len_chunk; q += kCwThreads) acc[q] = 0.0;
  __syncthreads();
  for (int32_t tb = wave * 64; tb < ntree; tb += kCwThreads) {
    const int32_t t = tb + lane;
    uint32_t s = 0, len = 0;
    double w = 0.0;
    if (t < ntree) {
      const uint2 q = pe[static_cast<int64_t>(t) * n + j];
      const uint32_t end = q.y & 0x7fffffffu;
      if (end > 0 && (q.y >> 31)) {  // j in bag in tree t
        const uint32_t leaf = mbucket[q.x];
        const double c = static_cast<double>(nin[static_cast<int64_t>(t) * n + j]);
        const double m = static_cast<double>(draws[leaf]);
        w = kRecip ? c * (1.0 / m) : c / m;
        const uint32_t bs = bstart[leaf];
        const uint32_t b0 = lo > 0 ? member_lb(members, bs, end, static_cast<uint32_t>(lo)) : bs;
        const uint32_t b1 = hi < n ? member_lb(members, b0, end, static_cast<uint32_t>(hi)) : end;
        s = b0;
        len = b1 > b0 ? b1 - b0 : 0u;
      }
    }
    cw_visit(s, len, w, lane, members, acc, lo);
  }
  __syncthreads();
  double* col = out + j * n + lo;
  for (int q = tid; q < len_chunk; q += kCwThreads) col[q] = accumulate ? col[q] + acc[q] : acc[q];
}

// RF-GAP keys from the node each case reached, terminal or not: the
// reference matches cases by equal node index (src/rf_proximity.cpp:
// 320-355), so a case whose traversal stopped early still meets the cases
// that stopped at the same node.  key = (node_off[t] + node) << 1 | in-bag,
// -1 for a negative node index, which the reference skips.  The same pass
// counts the trees in which each case is out of bag (nin == 0) into oob[i]
// (zeroed beforehand): thread (i, tree group) covers kRgTreeGroup trees.
constexpr int kRgTreeGroup = 16;

__global__ void rg_keys_kernel(const int32_t* __restrict__ nin, const int32_t* __restrict__ nodextr,
                               const int64_t* __restrict__ node_off, int64_t n, int32_t ntree,
                               int32_t* __restrict__ keys, uint32_t* __restrict__ oob) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t t0 = static_cast<int32_t>(blockIdx.y) * kRgTreeGroup;
  const int32_t t1 = min(ntree, t0 + kRgTreeGroup);
  uint32_t zeros = 0;
  for (int32_t t = t0; t < t1; ++t) {
    const int64_t s = static_cast<int64_t>(t) * n + i;
    const int32_t k = nodextr[s], c = nin[s];
    const int64_t nn = node_off[t + 1] - node_off[t];
    keys[s] = (k >= 0 && k < nn) ? static_cast<int32_t>(((node_off[t] + k) << 1) | (c > 0 ? 1 : 0)) : -1;
    zeros += c == 0 ? 1u : 0u;
  }
  if (zeros) atomicAdd(&oob[i], zeros);
}

constexpr int kRgTile = 64;

// In place on the column-major n x n matrix W (W[i + j n], row = the
// out-of-bag case): p_ij = W_ij * (1 / S_i) (rows with S_i = 0 are all zero
// and stay so), then P_ij = P_ji = (p_ij + p_ji) / 2 off the diagonal, and
// P_ii = 1 where S_i = 0, else p_ii.  Workgroup (I, J), I <= J, owns the
// mirrored tiles (I, J) and (J, I): the (J, I) tile is read into LDS
// transposed, the (I, J) tile is read, averaged and written directly, and
// the averages are written back to (J, I) from LDS.  Both global passes run
// along the contiguous (row) direction.
// The finishing pass reads and writes every element once: nontemporal loads
// and stores leave the L2 alone (1,292 -> 1,260 us at n = 20,000).
#ifndef RFX_RG_NT
#define RFX_RG_NT 1
#endif
__device__ __forceinline__ double rg_load(const double* p) {
  return RFX_RG_NT ? __builtin_nontemporal_load(p) : *p;
}
__device__ __forceinline__ void rg_store(double* p, double x) {
  if (RFX_RG_NT)
    __builtin_nontemporal_store(x, p);
  else
    *p = x;
}

__global__ __launch_bounds__(256) void rg_symmetric_kernel(double* __restrict__ w, const uint32_t* __restrict__ oob,
                                                           int64_t n, int64_t nt) {
  // p_ij is rounded before the sum, as in the reference: a fused
  // multiply-add of w_ij * (1 / S_i) + p_ji would differ between the two
  // mirrored entries of the diagonal tile, which are computed separately
#pragma clang fp contract(off)
  int64_t I, J;  // workgroup -> tile pair I <= J of the upper triangle of nt x nt tiles
  upper_tile(blockIdx.x, nt, I, J);
  __shared__ double L[kRgTile][kRgTile + 1];  // L[b][a] = element (J0 + b, I0 + a)
  const int64_t i0 = I * kRgTile, j0 = J * kRgTile;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  constexpr int R = kRgTile / 4;  // tile rows per thread; each pass loads all of its rows first
  const double inv_b = (j0 + tx < n && oob[j0 + tx] > 0) ? 1.0 / static_cast<double>(oob[j0 + tx]) : 1.0;
  const double inv_a = (i0 + tx < n && oob[i0 + tx] > 0) ? 1.0 / static_cast<double>(oob[i0 + tx]) : 1.0;
  const bool row_b = j0 + tx < n, row_a = i0 + tx < n;
  double v[R];
  // (J, I) tile: row J0 + b (b = tx, contiguous), column I0 + a; scaled by its row's 1 / S
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int64_t c = i0 + ty + 4 * k;
    v[k] = (row_b && c < n) ? rg_load(&w[(j0 + tx) + c * n]) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) L[tx][ty + 4 * k] = v[k] * inv_b;
  __syncthreads();
  // (I, J) tile: row I0 + a (a = tx, contiguous), column J0 + b
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int64_t c = j0 + ty + 4 * k;
    v[k] = (row_a && c < n) ? rg_load(&w[(i0 + tx) + c * n]) : 0.0;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int b = ty + 4 * k;
    const int64_t r = i0 + tx, c = j0 + b;
    if (!row_a || c >= n) continue;
    double x;
    if (r == c)
      x = oob[r] > 0 ? v[k] * inv_a : 1.0;
    else
      x = (v[k] * inv_a + L[b][tx]) / 2.0;
    rg_store(&w[r + c * n], x);
    L[b][tx] = x;
  }
  __syncthreads();
  if (I == J) return;  // the diagonal tile is its own mirror: written above
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int64_t c = i0 + ty + 4 * k;
    if (row_b && c < n) rg_store(&w[(j0 + tx) + c * n], L[tx][ty + 4 * k]);
  }
}

// The leaf sums of W: bucket build, leaf draws, then the column gather.
template <bool kRecip>
int weighted_leaf_sums(hipStream_t st, const int32_t* keys, const int32_t* nin, int64_t n, int32_t ntree,
                       int64_t total_nodes, int32_t accumulate, double* out, void* ws) {
  const BucketWs w = bucket_ws_carve(ws, n, ntree, total_nodes);
  uint32_t* draws = reinterpret_cast<uint32_t*>(static_cast<char*>(ws) + bucket_ws_bytes(n, ntree, total_nodes));
  int rc = bucket_build(st, keys, n, ntree, total_nodes, w);
  if (!rc) rc = fill_u32(st, draws, total_nodes, 0u);
  if (rc) return rc;
  const int64_t slots = n * static_cast<int64_t>(ntree);
  hipLaunchKernelGGL(cw_leaf_draws_kernel, dim3(rfx_ceil_div(slots, 256)), dim3(256), 0, st, w.mbucket, w.V[0], nin,
                     slots, static_cast<uint32_t>(total_nodes), draws);
  const int64_t nchunks = (n + kCwMaxRows - 1) / kCwMaxRows;
  const int64_t chunk_len = (n + nchunks - 1) / nchunks;
  const dim3 grid(static_cast<unsigned>(xcd_rows_padded(n)), static_cast<unsigned>(nchunks));
  const size_t lds = 8 * static_cast<size_t>(chunk_len);
  if (w.members16)
    hipLaunchKernelGGL((cw_line_kernel<kRecip, uint16_t>), grid, dim3(kCwThreads), lds, st, w.pe, w.bstart, w.mbucket,
                       w.members16, draws, nin, n, ntree, chunk_len, accumulate, out);
  else
    hipLaunchKernelGGL((cw_line_kernel<kRecip, uint32_t>), grid, dim3(kCwThreads), lds, st, w.pe, w.bstart, w.mbucket,
                       w.members, draws, nin, n, ntree, chunk_len, accumulate, out);
  return rfx_launch_status();
}

bool weighted_args_ok(const int32_t* keys, const int32_t* nin, const double* out, const void* ws, int64_t n,
                      int32_t ntree, int64_t total_nodes) {
  return keys && nin && out && ws && n > 0 && n <= RFX_MAX_SAMPLES && ntree > 0 && total_nodes > 0 &&
         total_nodes <= RFX_MAX_TOTAL_NODES;
}

}  // namespace

extern "C" size_t rfx_prox_casewise_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes) {
  if (n <= 0 || ntree <= 0 || total_nodes <= 0) return 0;
  return bucket_ws_bytes(n, ntree, total_nodes) + align256(4 * static_cast<size_t>(total_nodes));
}

extern "C" int rfx_prox_casewise(hipStream_t st, const int32_t* keys, const int32_t* nin, int64_t n, int32_t ntree,
                                 int64_t total_nodes, int32_t accumulate, double* out, void* ws, size_t ws_bytes) {
  if (!weighted_args_ok(keys, nin, out, ws, n, ntree, total_nodes)) return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  if (ws_bytes < rfx_prox_casewise_workspace_bytes(n, ntree, total_nodes)) return RFX_ERR_WORKSPACE;
  return weighted_leaf_sums<false>(st, keys, nin, n, ntree, total_nodes, accumulate, out, ws);
}

extern "C" size_t rfx_prox_rfgap_workspace_bytes(int64_t n, int32_t ntree, int64_t total_nodes) {
  if (n <= 0 || ntree <= 0 || total_nodes <= 0) return 0;
  return rfx_prox_casewise_workspace_bytes(n, ntree, total_nodes) + align256(4 * static_cast<size_t>(n)) +
         align256(4 * static_cast<size_t>(n) * static_cast<size_t>(ntree));
}

extern "C" int rfx_prox_rfgap(hipStream_t st, const int32_t* nin, const int32_t* nodextr, const int64_t* node_off,
                              int64_t n, int32_t ntree, int64_t total_nodes, double* out, void* ws, size_t ws_bytes) {
  if (!nodextr || !node_off || !weighted_args_ok(nin, nin, out, ws, n, ntree, total_nodes)) return RFX_ERR_INVALID;
  if (n * static_cast<int64_t>(ntree) >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  if (ws_bytes < rfx_prox_rfgap_workspace_bytes(n, ntree, total_nodes)) return RFX_ERR_WORKSPACE;
  char* tail = static_cast<char*>(ws) + rfx_prox_casewise_workspace_bytes(n, ntree, total_nodes);
  uint32_t* oob = reinterpret_cast<uint32_t*>(tail);
  int32_t* keys = reinterpret_cast<int32_t*>(tail + align256(4 * static_cast<size_t>(n)));
  int rc = fill_u32(st, oob, n, 0u);
  if (rc) return rc;
  hipLaunchKernelGGL(rg_keys_kernel, dim3(rfx_ceil_div(n, 256), rfx_ceil_div(ntree, kRgTreeGroup)), dim3(256), 0, st,
                     nin, nodextr, node_off, n, ntree, keys, oob);
  rc = weighted_leaf_sums<true>(st, keys, nin, n, ntree, total_nodes, 0, out, ws);
  if (rc) return rc;
  const unsigned nt = static_cast<unsigned>(rfx_ceil_div(n, kRgTile));
  // one workgroup per tile pair I <= J
  const int64_t pairs = static_cast<int64_t>(nt) * (nt + 1) / 2;
  if (pairs >= (1LL << 31)) return RFX_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(rg_symmetric_kernel, dim3(static_cast<unsigned>(pairs)), dim3(256), 0, st, out, oob, n,
                     static_cast<int64_t>(nt));
  return rfx_launch_status();
}
