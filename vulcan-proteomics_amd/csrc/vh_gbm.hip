// Gradient-boosted tree scorer: sklearn HistGradientBoosting raw predictions (one tree per iteration) in f64,
// the device counterpart of the reference's GpuHGB (vulcan/rescore/fast.py:198-254).
//
// A workgroup owns up to gbm_spw(nf) samples (one per thread). Their features are staged once in LDS, column-
// major (feature f of the workgroup's row r at xs[f * spw + r], so the lanes of a wave read consecutive
// doubles), plus one column of -inf. The trees are staged in chunks of whole trees (at most
// VH_GBM_CHUNK_NODES nodes, 16 B each, children as chunk-local indexes). Every thread walks the trees of the
// chunk for its sample GBM_WALK at a time and adds the leaf values to its f64 sum in tree order -- the
// summation order of sklearn (0 + baseline, then += tree by tree), so the result is bit-identical.
//
// The walk has no branches per node: a leaf points to itself and its feature (0xFFFF) is clamped to the -inf
// column, so "x <= v" holds and the walker stays on the leaf; a wave stops when none of its walkers moved
// (masking the walkers that stopped, to spare their LDS reads, measured slower).
// Child and feature indexes are clamped to the staged arrays (any node array stays in bounds), walks are
// bounded by max_depth + 1 steps, and a walk that does not end on a leaf sets status bit 1.
#include "vh_common.h"

namespace {

constexpr int GBM_X_BYTES = 32768;   // LDS for the staged features (+ the -inf column)
constexpr int GBM_WALK = 4;          // trees walked at once per thread

// samples per workgroup: the largest of 256 / 128 / 64 whose nf + 1 columns fit GBM_X_BYTES
__host__ __device__ inline int gbm_spw(int nf) {
  const int col = (nf + 1) * 8;
  return col * 256 <= GBM_X_BYTES ? 256 : col * 128 <= GBM_X_BYTES ? 128 : 64;
}

__global__ __launch_bounds__(256) void gbm_kernel(const VhGbmNode* __restrict__ nodes,
                                                  const int32_t* __restrict__ tree_root,
                                                  const int32_t* __restrict__ chunk_node0,
                                                  const int32_t* __restrict__ chunk_tree0, int n_chunks,
                                                  int64_t n_nodes, int n_trees, double baseline, const double* __restrict__ X, int64_t n,
                                                  int nf, int max_steps, double* __restrict__ out,
                                                  int32_t* status) {
  __shared__ __align__(16) VhGbmNode sn[VH_GBM_CHUNK_NODES];
  __shared__ __align__(16) double xs[GBM_X_BYTES / 8];
  const int tid = threadIdx.x;
  const int spw = blockDim.x;
  const int64_t s0 = (int64_t)blockIdx.x * spw;
  const int ns = (int)(n - s0 < spw ? n - s0 : spw);
  {
    const double* src = X + s0 * nf;
    const int tot = ns * nf;
    for (int i = tid; i < tot; i += spw) {
      const int r = i / nf;
      xs[(i - r * nf) * spw + r] = src[i];
    }
    xs[nf * spw + tid] = -__builtin_inf();
  }
  const double* xcol = xs + tid;
  double acc = 0.0;
  acc += baseline;
  bool bad = false;
  bool clamped = false;
  for (int c = 0; c < n_chunks; ++c) {
    // a malformed chunk table (a chunk beyond the node array or larger than the LDS chunk, trees outside
    // [0, n_trees)) is clamped into bounds and reported through status bit 2
    int64_t c0l = chunk_node0[c], c1l = chunk_node0[c + 1];
    int64_t t0l = chunk_tree0[c], t1l = chunk_tree0[c + 1];
    const int64_t c0c = c0l < 0 ? 0 : (c0l > n_nodes ? n_nodes : c0l);
    int64_t cnl = c1l - c0c;
    cnl = cnl < 0 ? 0 : cnl;
    cnl = cnl > VH_GBM_CHUNK_NODES ? VH_GBM_CHUNK_NODES : cnl;
    cnl = c0c + cnl > n_nodes ? n_nodes - c0c : cnl;
    const int64_t t0c = t0l < 0 ? 0 : (t0l > n_trees ? n_trees : t0l);
    const int64_t t1c = t1l < t0c ? t0c : (t1l > n_trees ? n_trees : t1l);
    clamped |= c0c != c0l || c0c + cnl != c1l || t0c != t0l || t1c != t1l;
    const int c0 = (int)c0c, cn = (int)cnl;
    const int t0 = (int)t0c, t1 = (int)t1c;
    const unsigned last = (unsigned)(cn > 0 ? cn - 1 : 0);
    __syncthreads();   // the previous chunk's walks are done (and, on the first chunk, the rows are staged)
    {
      const uint4* g = (const uint4*)(nodes + c0);
      uint4* l = (uint4*)sn;
      for (int i = tid; i < cn; i += spw) l[i] = g[i];
    }
    __syncthreads();
    if (tid < ns && cn > 0) {
      for (int t = t0; t < t1; t += GBM_WALK) {
        unsigned li[GBM_WALK];
#pragma unroll
        for (int u = 0; u < GBM_WALK; ++u) {
          const unsigned r = (unsigned)(tree_root[t + u < t1 ? t + u : t1 - 1] - c0);
          li[u] = r < last ? r : last;
        }
        for (int step = 0; step < max_steps; ++step) {
          bool moved = false;
#pragma unroll
          for (int u = 0; u < GBM_WALK; ++u) {
            const VhGbmNode nd = sn[li[u]];
            const unsigned f = (uint16_t)nd.feature;
            const double x = xcol[(f < (unsigned)nf ? f : (unsigned)nf) * spw];
            const bool go_left = (x <= nd.v) || (x != x && nd.missing_left);
            const unsigned nx = (unsigned)nd.left + (go_left ? 0u : 1u);
            const unsigned nc = nx < last ? nx : last;
            moved |= nc != li[u];
            li[u] = nc;
          }
          if (!__any(moved)) break;
        }
#pragma unroll
        for (int u = 0; u < GBM_WALK; ++u) {
          const VhGbmNode nd = sn[li[u]];
          if (t + u < t1) {
            bad |= !nd.is_leaf;
            acc += nd.v;   // tree order
          }
        }
      }
    }
  }
  if (tid < ns) {
    out[s0 + tid] = acc;
    if (bad) atomicOr(status, 1);
  }
  if (clamped && tid == 0) atomicOr(status, 2);
}

}  // namespace

extern "C" int vh_gbm_decision(const VhGbmNode* nodes, int64_t n_nodes, const int32_t* tree_root, int32_t n_trees,
                               const int32_t* chunk_node0, const int32_t* chunk_tree0, int32_t n_chunks,
                               int32_t max_depth, double baseline, const double* X, int64_t n, int32_t n_features,
                               double* out, int32_t* status, void* stream) {
  if (n < 0 || n_trees < 0 || n_chunks < 0 || n_nodes < 0 || max_depth < 0 || n_features < 1 || !status ||
      (n && (!X || !out)) || (n_trees && (!nodes || !tree_root || !chunk_node0 || !chunk_tree0 || !n_chunks)))
    return VH_E_ARG;
  if (n_features > VH_GBM_MAX_FEATURES || max_depth >= 1 << 16 || n > ((int64_t)1 << 40)) return VH_E_LIMIT;
  if (n == 0) return 0;
  const int spw = gbm_spw(n_features);
  const int64_t blocks = (n + spw - 1) / spw;
  hipLaunchKernelGGL(gbm_kernel, dim3((unsigned)blocks), dim3(spw), 0, (hipStream_t)stream, nodes, tree_root,
                     chunk_node0, chunk_tree0, n_trees ? n_chunks : 0, n_nodes, n_trees, baseline, X, n, n_features, max_depth + 1,
                     out, status);
  VH_CHECK_LAUNCH();
  return 0;
}
