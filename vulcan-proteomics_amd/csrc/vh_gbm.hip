// Gradient-boosted tree scorer: sklearn HistGradientBoosting raw predictions (one tree per iteration) in f64,
// the device counterpart of the reference's GpuHGB (vulcan/rescore/fast.py:198-254).
//
// A workgroup owns up to GBM_SPW(nf) samples (one per thread). Their feature rows are staged once in LDS
// (row stride nf | 1 doubles, so lanes reading different rows spread over the banks); the trees are staged in
// chunks of whole trees (at most VH_GBM_CHUNK_NODES nodes, 16 B each). Every thread walks the trees of the
// chunk for its sample GBM_WALK at a time (independent chains of LDS reads) and adds the leaf values to its f64
// sum in tree order -- the summation order of sklearn (0 + baseline, then += tree by tree), so the result is
// bit-identical. Walks are bounded by max_depth + 1
// node visits and every node / feature index is checked, so a malformed model sets status bit 1 instead of
// reading outside the staged arrays or looping.
#include "vh_common.h"

namespace {

constexpr int GBM_X_BYTES = 32768;   // LDS for the staged feature rows
constexpr int GBM_WALK = 4;          // trees walked at once per thread

__host__ __device__ inline int gbm_nfp(int nf) { return nf | 1; }
// samples per workgroup: the largest of 256 / 128 / 64 whose rows fit GBM_X_BYTES
__host__ __device__ inline int gbm_spw(int nf) {
  const int row = gbm_nfp(nf) * 8;
  return row * 256 <= GBM_X_BYTES ? 256 : row * 128 <= GBM_X_BYTES ? 128 : 64;
}

__global__ __launch_bounds__(256) void gbm_kernel(const VhGbmNode* __restrict__ nodes,
                                                  const int32_t* __restrict__ tree_root,
                                                  const int32_t* __restrict__ chunk_node0,
                                                  const int32_t* __restrict__ chunk_tree0, int n_chunks,
                                                  double baseline, const double* __restrict__ X, int64_t n,
                                                  int nf, int max_steps, double* __restrict__ out,
                                                  int32_t* status) {
  __shared__ __align__(16) VhGbmNode sn[VH_GBM_CHUNK_NODES];
  __shared__ __align__(16) double xs[GBM_X_BYTES / 8];
  const int tid = threadIdx.x;
  const int spw = blockDim.x;
  const int nfp = gbm_nfp(nf);
  const int64_t s0 = (int64_t)blockIdx.x * spw;
  const int ns = (int)(n - s0 < spw ? n - s0 : spw);
  {
    const double* src = X + s0 * nf;
    const int tot = ns * nf;
    for (int i = tid; i < tot; i += spw) {
      const int r = i / nf;
      xs[r * nfp + (i - r * nf)] = src[i];
    }
  }
  const double* xr = xs + tid * nfp;
  double acc = 0.0;
  acc += baseline;
  bool bad = false;
  for (int c = 0; c < n_chunks; ++c) {
    const int c0 = chunk_node0[c], cn = chunk_node0[c + 1] - c0;
    const int t0 = chunk_tree0[c], t1 = chunk_tree0[c + 1];
    __syncthreads();   // the previous chunk's walks are done (and, on the first chunk, the rows are staged)
    {
      const uint4* g = (const uint4*)(nodes + c0);
      uint4* l = (uint4*)sn;
      for (int i = tid; i < cn; i += spw) l[i] = g[i];
    }
    __syncthreads();
    if (tid < ns) {
      // GBM_WALK trees at once: independent chains of dependent LDS reads. A walker that has reached its leaf
      // re-reads it (no branch around the loads); walkers past the chunk's last tree walk that tree and are
      // not added. Child and feature indexes are checked before use; a walk that leaves the chunk or ends
      // without a leaf marks the sample bad.
      for (int t = t0; t < t1; t += GBM_WALK) {
        int li[GBM_WALK];
#pragma unroll
        for (int u = 0; u < GBM_WALK; ++u) li[u] = tree_root[t + u < t1 ? t + u : t1 - 1] - c0;
        bool at_leaf[GBM_WALK];
        double v[GBM_WALK];
#pragma unroll
        for (int u = 0; u < GBM_WALK; ++u) {
          if ((unsigned)li[u] >= (unsigned)cn) {
            li[u] = 0;
            bad = true;
          }
          at_leaf[u] = false;
          v[u] = 0.0;
        }
        for (int step = 0; step < max_steps; ++step) {
          VhGbmNode nd[GBM_WALK];
#pragma unroll
          for (int u = 0; u < GBM_WALK; ++u) nd[u] = sn[li[u]];
          double x[GBM_WALK];
#pragma unroll
          for (int u = 0; u < GBM_WALK; ++u) {
            const bool fok = (unsigned)nd[u].feature < (unsigned)nf;
            bad |= !fok && !nd[u].is_leaf;
            x[u] = xr[fok ? nd[u].feature : 0];
          }
          bool all = true;
#pragma unroll
          for (int u = 0; u < GBM_WALK; ++u) {
            if (nd[u].is_leaf) {
              at_leaf[u] = true;
              v[u] = nd[u].v;
            } else {
              const bool go_left = (x[u] != x[u]) ? (nd[u].missing_left != 0) : (x[u] <= nd[u].v);
              const int nx = nd[u].left - c0 + (go_left ? 0 : 1);
              if ((unsigned)nx < (unsigned)cn) {
                li[u] = nx;
              } else {
                bad = true;
              }
              all = false;
            }
          }
          if (all) break;
        }
#pragma unroll
        for (int u = 0; u < GBM_WALK; ++u) {
          if (t + u < t1) {
            bad |= !at_leaf[u];
            acc += v[u];   // tree order
          }
        }
      }
    }
  }
  if (tid < ns) {
    out[s0 + tid] = acc;
    if (bad) atomicOr(status, 1);
  }
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
                     chunk_node0, chunk_tree0, n_trees ? n_chunks : 0, baseline, X, n, n_features, max_depth + 1,
                     out, status);
  VH_CHECK_LAUNCH();
  return 0;
}
