// Parameter plumbing, device-wide prefix scans and the calibration copy kernel.
#include "vh_common.h"

DevP vh_make_devp(const VhParams* p) {
  DevP d;
  for (int i = 0; i < 26; ++i) {
    // same single f32 rounding as resmass() in kernels.cu:47-51 (mono + static / mono + var)
    d.rm[0][i] = p->mono[i] + p->stat[i];
    d.rm[1][i] = p->mono[i] + p->var[i];
  }
  d.nterm = p->nterm;
  d.tol_da = p->tol_da;
  d.tol_lo = p->tol_lo;
  d.tol_hi = p->tol_hi;
  d.fmin = p->fmin;
  d.fmax = p->fmax;
  d.min_ion_index = p->min_ion_index;
  d.prec_lo = p->prec_lo;
  d.prec_hi = p->prec_hi;
  d.iso_lo = p->iso_lo;
  d.iso_hi = p->iso_hi;
  d.z_lo = p->z_lo;
  d.z_hi = p->z_hi;
  d.max_fragment_charge = p->max_fragment_charge;
  d.max_peaks = p->max_peaks;
  d.min_peaks = p->min_peaks;
  d.min_matched_peaks = p->min_matched_peaks;
  d.topk = p->topk;
  d.k1_threshold = p->k1_threshold;
  d.pair_closure = p->pair_closure;
  d.tile = p->tile;
  d.xcd_group = p->xcd_group;
  d.n_offsets = p->n_offsets;
  for (int i = 0; i < VH_MAX_OFFSETS; ++i) d.offsets[i] = p->offsets[i];
  d.open_search = p->open_search;
  d.open_lo = p->open_lo;
  d.open_hi = p->open_hi;
  d.k1_threshold_open = p->k1_threshold_open;
  d.peak_window_k = p->peak_window_k;
  d.peak_window_th = p->peak_window_th;
  return d;
}

DevIdx vh_make_devidx(const VhIndexView* ix) {
  DevIdx d;
  d.n_rows = ix->n_rows;
  d.seq_off = ix->seq_off;
  d.length = ix->length;
  d.mask = ix->mask;
  d.mass32 = ix->mass32;
  d.residues = ix->residues;
  d.partner = ix->partner;
  d.bucket_log2 = ix->bucket_log2;
  d.bin_shift = ix->bin_shift;
  d.bin_base = ix->bin_base;
  d.n_bins = ix->n_bins;
  d.n_buckets = ix->n_buckets;
  d.offsets = ix->offsets;
  d.postings = ix->postings;
  return d;
}

// ------------------------------------------------------------------------------------------------ scans
// Three-phase reduce / scan-partials / scan-apply over tiles of 2048 items (256 threads x 8).
#define SCAN_THREADS 256
#define SCAN_ITEMS 8
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

template <typename T>
__device__ T block_exclusive(T v, T* lds_w, T* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) lds_w[w] = x;
  __syncthreads();
  T before = 0, tot = 0;
  for (int i = 0; i < nw; ++i) {
    T t = lds_w[i];
    if (i < w) before += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void scan_reduce_kernel(const T* a, int64_t n, T* part) {
  __shared__ T lw[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  T s = 0;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
    int64_t j = base + i;
    if (j < n) s += a[j];
  }
  T tot;
  block_exclusive<T>(s, lw, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// single block: exclusive scan of part[0..nb) in place
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void scan_partials_kernel(T* part, int64_t nb) {
  __shared__ T lw[SCAN_THREADS / 64];
  T carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += SCAN_THREADS) {
    int64_t j = b0 + threadIdx.x;
    T v = j < nb ? part[j] : 0;
    T tot;
    T ex = block_exclusive<T>(v, lw, &tot);
    if (j < nb) part[j] = carry + ex;
    carry += tot;
  }
}

template <typename T, bool INCLUSIVE>
__global__ __launch_bounds__(SCAN_THREADS) void scan_apply_kernel(T* a, int64_t n, const T* part) {
  __shared__ T tile[SCAN_TILE];
  __shared__ T lw[SCAN_THREADS / 64];
  int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
    int64_t j = base + i;
    tile[i] = j < n ? a[j] : (T)0;
  }
  __syncthreads();
  T v[SCAN_ITEMS];
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = tile[threadIdx.x * SCAN_ITEMS + k];
    s += v[k];
  }
  T tot;
  T run = block_exclusive<T>(s, lw, &tot) + part[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    T x = v[k];
    if (INCLUSIVE) { run += x; tile[threadIdx.x * SCAN_ITEMS + k] = run; }
    else { tile[threadIdx.x * SCAN_ITEMS + k] = run; run += x; }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_THREADS) {
    int64_t j = base + i;
    if (j < n) a[j] = tile[i];
  }
}

static size_t scan_ws(int64_t n, size_t elem) {
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  return (size_t)(nb < 1 ? 1 : nb) * elem + 256;
}

size_t vh_scan_ws_u32(int64_t n) { return scan_ws(n, sizeof(uint32_t)); }
size_t vh_scan_ws_i64(int64_t n) { return scan_ws(n, sizeof(int64_t)); }

template <typename T, bool INCLUSIVE>
static int scan_any(T* a, int64_t n, void* ws, size_t ws_bytes, hipStream_t st) {
  if (n <= 0) return 0;
  if (!a || !ws) return VH_E_ARG;
  if (ws_bytes < scan_ws(n, sizeof(T))) return VH_E_WORKSPACE;
  int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
  T* part = (T*)ws;
  scan_reduce_kernel<T><<<(unsigned)nb, SCAN_THREADS, 0, st>>>(a, n, part);
  scan_partials_kernel<T><<<1, SCAN_THREADS, 0, st>>>(part, nb);
  scan_apply_kernel<T, INCLUSIVE><<<(unsigned)nb, SCAN_THREADS, 0, st>>>(a, n, part);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_scan_inclusive_i64(int64_t* a, int64_t n, void* ws, size_t ws_bytes, hipStream_t st) {
  return scan_any<int64_t, true>(a, n, ws, ws_bytes, st);
}

int vh_scan_exclusive_u32(uint32_t* a, int64_t n, void* ws, size_t ws_bytes, hipStream_t st) {
  return scan_any<uint32_t, false>(a, n, ws, ws_bytes, st);
}

// ------------------------------------------------------------------------------------------------ copy
__global__ __launch_bounds__(256) void copy_u32_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                       int64_t n4) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

__global__ void copy_tail_kernel(const uint32_t* src, uint32_t* dst, int64_t a, int64_t n) {
  int64_t i = a + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void add_i32_kernel(int32_t* a, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] += v;
}

extern "C" {

int vh_version(void) { return VH_ABI_VERSION; }

size_t vh_abi_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(VhParams);
    case 1: return sizeof(VhIndexView);
    case 2: return sizeof(VhBlockView);
    case 3: return sizeof(VhK2Cols);
    case 4: return sizeof(VhPsmCols);
    case 5: return sizeof(VhGbmNode);
    case 6: return sizeof(VhFamilyCols);
    default: return 0;
  }
}

size_t vh_scan_workspace_bytes(int64_t n) { return scan_ws(n, sizeof(int64_t)); }

int vh_index_rebase_rows(int32_t* row0, int64_t n, int32_t base, void* stream) {
  if (n < 0 || (n && !row0)) return VH_E_ARG;
  if (n == 0 || base == 0) return 0;
  const unsigned g = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
  add_i32_kernel<<<g, 256, 0, (hipStream_t)stream>>>(row0, n, base);
  VH_CHECK_LAUNCH();
  return 0;
}

int vh_copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, void* stream) {
  if (n < 0 || (n && (!src || !dst))) return VH_E_ARG;
  if (n == 0) return 0;
  if (((uintptr_t)src | (uintptr_t)dst) & 15) return VH_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  int64_t n4 = n / 4;
  if (n4) copy_u32_kernel<<<2048, 256, 0, st>>>((const uint4*)src, (uint4*)dst, n4);
  if (n4 * 4 < n) copy_tail_kernel<<<1, 64, 0, st>>>(src, dst, n4 * 4, n);
  VH_CHECK_LAUNCH();
  return 0;
}

}  // extern "C"
